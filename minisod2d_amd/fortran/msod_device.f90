! iso_c_binding interface to the C ABI of the GPU kernels (include/msod2d.h).
!
! The data-region calls replace the reference driver's OpenACC directives
! (app/sources/main.f90:197-425,539-542):
!   !$acc enter data copyin(x)   ->  call msod_copyin(x, c_sizeof(x))
!   !$acc enter data create(x)   ->  call msod_create(x, c_sizeof(x))
!   !$acc update host(x)         ->  call msod_update_host(x, c_sizeof(x))
!   !$acc update device(x)       ->  call msod_update_device(x, c_sizeof(x))
!   !$acc exit data delete(x)    ->  call msod_delete(x)
! Every wrapper stops the program with the library's message on failure, as
! the reference's run-time would on a missing `present` array.
module msod_device

   use iso_c_binding
   implicit none
   private

   public :: msod_start, msod_stop, msod_copyin, msod_create, msod_update_host, msod_update_device
   public :: msod_delete, msod_sync, msod_check, msod_set_async_mode, msod_set_assembly_mode
   public :: msod_full_convec_ijk, msod_full_diffusion_ijk, msod_full_convec_ijk_mixed
   public :: msod_full_diffusion_ijk_mixed, msod_set_diffusion_coefficients
   public :: msod_full_diffusion_ijk_closure, msod_jacobian_device
   public :: msod_tet_jacobian_device, msod_fem_convec, msod_fem_diffu
   public :: msod_full_convec_ijk_r4, msod_full_diffusion_ijk_r4

   interface
      integer(c_int) function msod_init(device) bind(C, name='msod_init')
         import :: c_int
         integer(c_int), value :: device
      end function msod_init
      integer(c_int) function msod_finalize() bind(C, name='msod_finalize')
         import :: c_int
      end function msod_finalize
      type(c_ptr) function msod_last_error() bind(C, name='msod_last_error')
         import :: c_ptr
      end function msod_last_error
      integer(c_int) function msod_data_copyin(host, bytes) bind(C, name='msod_data_copyin')
         import :: c_int, c_size_t
         type(*), dimension(*), intent(in) :: host
         integer(c_size_t), value :: bytes
      end function msod_data_copyin
      integer(c_int) function msod_data_create(host, bytes) bind(C, name='msod_data_create')
         import :: c_int, c_size_t
         type(*), dimension(*), intent(in) :: host
         integer(c_size_t), value :: bytes
      end function msod_data_create
      integer(c_int) function msod_data_update_host(host, bytes) bind(C, name='msod_data_update_host')
         import :: c_int, c_size_t
         type(*), dimension(*) :: host
         integer(c_size_t), value :: bytes
      end function msod_data_update_host
      integer(c_int) function msod_data_update_device(host, bytes) bind(C, name='msod_data_update_device')
         import :: c_int, c_size_t
         type(*), dimension(*), intent(in) :: host
         integer(c_size_t), value :: bytes
      end function msod_data_update_device
      integer(c_int) function msod_data_delete(host) bind(C, name='msod_data_delete')
         import :: c_int
         type(*), dimension(*), intent(in) :: host
      end function msod_data_delete
      integer(c_int) function msod_synchronize() bind(C, name='msod_synchronize')
         import :: c_int
      end function msod_synchronize
      integer(c_int) function msod_set_async(flag) bind(C, name='msod_set_async')
         import :: c_int
         integer(c_int), value :: flag
      end function msod_set_async
      integer(c_int) function msod_set_assembly(mode) bind(C, name='msod_set_assembly')
         import :: c_int
         integer(c_int), value :: mode
      end function msod_set_assembly
      integer(c_int) function msod_set_diffusion_coefficients(c_rho, c_ener) &
            bind(C, name='msod_set_diffusion_coefficients')
         import :: c_int, c_double
         real(c_double), value :: c_rho, c_ener
      end function msod_set_diffusion_coefficients

      integer(c_int) function msod_full_convec_ijk(porder, nelem, npoin, connec, Ngp, dNgp, He, gpvol, &
            dlxigp_ip, xgp, invAtoIJK, AtoI, AtoJ, AtoK, u, q, rho, pr, E, Rmass, Rmom, Rener) &
            bind(C, name='msod_full_convec_ijk')
         import :: c_int
         integer(c_int), value :: porder, nelem, npoin
         type(*), dimension(*), intent(in) :: connec, Ngp, dNgp, He, gpvol, dlxigp_ip, xgp
         type(*), dimension(*), intent(in) :: invAtoIJK, AtoI, AtoJ, AtoK, u, q, rho, pr, E
         type(*), dimension(*) :: Rmass, Rmom, Rener
      end function msod_full_convec_ijk

      integer(c_int) function msod_full_diffusion_ijk(porder, nelem, npoin, connec, Ngp, dNgp, He, gpvol, &
            dlxigp_ip, xgp, invAtoIJK, AtoI, AtoJ, AtoK, Cp, Pr, rho, u, Tem, mu_fluid, mu_e, mu_sgs, &
            Rmass, Rmom, Rener) bind(C, name='msod_full_diffusion_ijk')
         import :: c_int, c_double
         integer(c_int), value :: porder, nelem, npoin
         type(*), dimension(*), intent(in) :: connec, Ngp, dNgp, He, gpvol, dlxigp_ip, xgp
         type(*), dimension(*), intent(in) :: invAtoIJK, AtoI, AtoJ, AtoK
         real(c_double), value :: Cp, Pr
         type(*), dimension(*), intent(in) :: rho, u, Tem, mu_fluid, mu_e, mu_sgs
         type(*), dimension(*) :: Rmass, Rmom, Rener
      end function msod_full_diffusion_ijk

      ! the drop-in at the reference's default real kind (rp = 4)
      integer(c_int) function msod_full_convec_ijk_r4(porder, nelem, npoin, connec, Ngp, dNgp, He, gpvol, &
            dlxigp_ip, xgp, invAtoIJK, AtoI, AtoJ, AtoK, u, q, rho, pr, E, Rmass, Rmom, Rener) &
            bind(C, name='msod_full_convec_ijk_r4')
         import :: c_int
         integer(c_int), value :: porder, nelem, npoin
         type(*), dimension(*), intent(in) :: connec, Ngp, dNgp, He, gpvol, dlxigp_ip, xgp
         type(*), dimension(*), intent(in) :: invAtoIJK, AtoI, AtoJ, AtoK, u, q, rho, pr, E
         type(*), dimension(*) :: Rmass, Rmom, Rener
      end function msod_full_convec_ijk_r4

      integer(c_int) function msod_full_diffusion_ijk_r4(porder, nelem, npoin, connec, Ngp, dNgp, He, gpvol, &
            dlxigp_ip, xgp, invAtoIJK, AtoI, AtoJ, AtoK, Cp, Pr, rho, u, Tem, mu_fluid, mu_e, mu_sgs, &
            Rmass, Rmom, Rener) bind(C, name='msod_full_diffusion_ijk_r4')
         import :: c_int, c_float
         integer(c_int), value :: porder, nelem, npoin
         type(*), dimension(*), intent(in) :: connec, Ngp, dNgp, He, gpvol, dlxigp_ip, xgp
         type(*), dimension(*), intent(in) :: invAtoIJK, AtoI, AtoJ, AtoK
         real(c_float), value :: Cp, Pr
         type(*), dimension(*), intent(in) :: rho, u, Tem, mu_fluid, mu_e, mu_sgs
         type(*), dimension(*) :: Rmass, Rmom, Rener
      end function msod_full_diffusion_ijk_r4

      integer(c_int) function msod_full_convec_ijk_mixed(porder, nelem, npoin, connec, He, gpvol, &
            dlxigp_ip, invAtoIJK, AtoI, AtoJ, AtoK, u, q, rho, E, gamma_gas, Rmass, Rmom, Rener) &
            bind(C, name='msod_full_convec_ijk_mixed')
         import :: c_int, c_double
         integer(c_int), value :: porder, nelem, npoin
         type(*), dimension(*), intent(in) :: connec, He, gpvol, dlxigp_ip, invAtoIJK, AtoI, AtoJ, AtoK
         type(*), dimension(*), intent(in) :: u, q, rho, E
         real(c_double), value :: gamma_gas
         type(*), dimension(*) :: Rmass, Rmom, Rener
      end function msod_full_convec_ijk_mixed

      integer(c_int) function msod_full_diffusion_ijk_mixed(porder, nelem, npoin, connec, He, gpvol, &
            dlxigp_ip, invAtoIJK, AtoI, AtoJ, AtoK, Cp, Pr, rho, u, Tem, mu_fluid, mu_e, mu_sgs, &
            Rmass, Rmom, Rener) bind(C, name='msod_full_diffusion_ijk_mixed')
         import :: c_int, c_double
         integer(c_int), value :: porder, nelem, npoin
         type(*), dimension(*), intent(in) :: connec, He, gpvol, dlxigp_ip, invAtoIJK, AtoI, AtoJ, AtoK
         real(c_double), value :: Cp, Pr
         type(*), dimension(*), intent(in) :: rho, u, Tem, mu_fluid, mu_e, mu_sgs
         type(*), dimension(*) :: Rmass, Rmom, Rener
      end function msod_full_diffusion_ijk_mixed

      integer(c_int) function msod_full_diffusion_ijk_closure(porder, nelem, npoin, connec, Ngp, dNgp, He, &
            gpvol, dlxigp_ip, xgp, invAtoIJK, AtoI, AtoJ, AtoK, Cp, Pr, rho, u, Tem, mu_fluid, mu_e, &
            mlp_params, used_model, Rmass, Rmom, Rener) bind(C, name='msod_full_diffusion_ijk_closure')
         import :: c_int, c_double
         integer(c_int), value :: porder, nelem, npoin
         type(*), dimension(*), intent(in) :: connec, Ngp, dNgp, He, gpvol, dlxigp_ip, xgp
         type(*), dimension(*), intent(in) :: invAtoIJK, AtoI, AtoJ, AtoK
         real(c_double), value :: Cp, Pr
         type(*), dimension(*), intent(in) :: rho, u, Tem, mu_fluid, mu_e, mlp_params
         integer(c_int), intent(out) :: used_model
         type(*), dimension(*) :: Rmass, Rmom, Rener
      end function msod_full_diffusion_ijk_closure

      integer(c_int) function msod_jacobian_device(porder, nelem, npoin, connec, coord, dNgp, wgp, &
            gpvol, He, nbad) bind(C, name='msod_elem_jacobian')
         import :: c_int, c_int64_t
         integer(c_int), value :: porder, nelem, npoin
         type(*), dimension(*), intent(in) :: connec, coord, dNgp, wgp
         type(*), dimension(*) :: gpvol, He
         integer(c_int64_t), intent(out) :: nbad
      end function msod_jacobian_device

      integer(c_int) function msod_tet_jacobian_device(nelem, npoin, connec, coord, dNgp, wgp, gpvol, He, &
            nbad) bind(C, name='msod_tet_jacobian')
         import :: c_int, c_int64_t
         integer(c_int), value :: nelem, npoin
         type(*), dimension(*), intent(in) :: connec, coord, dNgp, wgp
         type(*), dimension(*) :: gpvol, He
         integer(c_int64_t), intent(out) :: nbad
      end function msod_tet_jacobian_device

      integer(c_int) function msod_fem_convec(nelem, npoin, connec, Ngp, dNgp, He, gpvol, u, q, rho, pr, E, &
            Rmass, Rmom, Rener) bind(C, name='msod_fem_convec')
         import :: c_int
         integer(c_int), value :: nelem, npoin
         type(*), dimension(*), intent(in) :: connec, Ngp, dNgp, He, gpvol, u, q, rho, pr, E
         type(*), dimension(*) :: Rmass, Rmom, Rener
      end function msod_fem_convec

      integer(c_int) function msod_fem_diffu(nelem, npoin, connec, Ngp, dNgp, He, gpvol, Cp, Pr, rho, u, Tem, &
            mu_fluid, mu_e, mu_sgs, Rmass, Rmom, Rener) bind(C, name='msod_fem_diffu')
         import :: c_int, c_double
         integer(c_int), value :: nelem, npoin
         type(*), dimension(*), intent(in) :: connec, Ngp, dNgp, He, gpvol
         real(c_double), value :: Cp, Pr
         type(*), dimension(*), intent(in) :: rho, u, Tem, mu_fluid, mu_e, mu_sgs
         type(*), dimension(*) :: Rmass, Rmom, Rener
      end function msod_fem_diffu
   end interface

contains

   ! Stops with the library's last message when status /= 0.
   subroutine msod_check(status, what)
      integer(c_int), intent(in) :: status
      character(*), intent(in) :: what
      character(kind=c_char), pointer :: msg(:)
      character(len=512) :: text
      integer :: i
      if (status == 0) return
      call c_f_pointer(msod_last_error(), msg, [512])
      text = ''
      do i = 1, 512
         if (msg(i) == c_null_char) exit
         text(i:i) = msg(i)
      end do
      write(*,'(a,a,a,i0,a,a)') 'msod: ', what, ' failed (status ', status, '): ', trim(text)
      error stop 2
   end subroutine msod_check

   subroutine msod_start(device)
      integer, intent(in) :: device
      call msod_check(msod_init(int(device, c_int)), 'msod_init')
   end subroutine msod_start

   subroutine msod_stop()
      call msod_check(msod_finalize(), 'msod_finalize')
   end subroutine msod_stop

   subroutine msod_copyin(a, bytes)
      type(*), dimension(*), intent(in) :: a
      integer(c_size_t), intent(in) :: bytes
      call msod_check(msod_data_copyin(a, bytes), 'enter data copyin')
   end subroutine msod_copyin

   subroutine msod_create(a, bytes)
      type(*), dimension(*), intent(in) :: a
      integer(c_size_t), intent(in) :: bytes
      call msod_check(msod_data_create(a, bytes), 'enter data create')
   end subroutine msod_create

   subroutine msod_update_host(a, bytes)
      type(*), dimension(*) :: a
      integer(c_size_t), intent(in) :: bytes
      call msod_check(msod_data_update_host(a, bytes), 'update host')
   end subroutine msod_update_host

   subroutine msod_update_device(a, bytes)
      type(*), dimension(*), intent(in) :: a
      integer(c_size_t), intent(in) :: bytes
      call msod_check(msod_data_update_device(a, bytes), 'update device')
   end subroutine msod_update_device

   subroutine msod_delete(a)
      type(*), dimension(*), intent(in) :: a
      call msod_check(msod_data_delete(a), 'exit data delete')
   end subroutine msod_delete

   subroutine msod_sync()
      call msod_check(msod_synchronize(), 'synchronize')
   end subroutine msod_sync

   subroutine msod_set_async_mode(flag)
      logical, intent(in) :: flag
      call msod_check(msod_set_async(merge(1_c_int, 0_c_int, flag)), 'set_async')
   end subroutine msod_set_async_mode

   subroutine msod_set_assembly_mode(mode)
      integer, intent(in) :: mode
      call msod_check(msod_set_assembly(int(mode, c_int)), 'set_assembly')
   end subroutine msod_set_assembly_mode

end module msod_device
