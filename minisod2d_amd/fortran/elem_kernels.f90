! Drop-in replacements of the reference's kernel modules:
!   module elem_convec    (kernels/sources/elem_convec.f90:198-211, fem_convec :839-854)
!   module elem_diffu     (kernels/sources/elem_diffu.f90:10-23, fem_diffu :199-212)
!   module jacobian_oper  (fem/sources/jacobian_oper.f90:7-19, tet_jacobian :138-151)
! Same module and procedure names and the same argument lists, so a caller
! written for the reference compiles unchanged; the bodies run the gfx950
! kernels through the C ABI.  Differences a caller can see:
!   * porder is read from the shape of invAtoIJK (run time) instead of
!     mod_constants (compile time); the dummies are assumed-shape, and every
!     array must be contiguous and `present` on the device (msod_copyin /
!     msod_create of module msod_device, the `enter data` counterparts).
!   * full_convec_ijk and full_diffusion_ijk are generic: a caller's real(8)
!     arrays reach the FP64 kernels, real(4) arrays (the reference built at its
!     default rp = 4, mod_constants.f90:5) the float drop-in, which computes in
!     FP64 and returns real(4) residuals.  The other procedures take real(8).
!   * c_rho and c_ener, module globals of mod_numerical_params in the
!     reference (mod_numerical_params.f90:26-27), are module variables of
!     elem_diffu here, with the same defaults.
!   * Rener of full_convec_ijk is the well-defined energy residual of the
!     reference's split form (elem_convec.f90:694-697); the fused reference
!     kernel leaves two of its flux arrays uninitialised (:218,:248-250).
!   * shapes are checked; a mismatch stops with a message.
! Shape check shared by the TET04 procedures below.
module msod_tet_shapes

   implicit none
   private
   public :: check_tet_shapes

contains

   ! connec(nelem,4), Ngp(4,4), dNgp(3,4,4), He(3,3,4,nelem), gpvol(1,4,nelem),
   ! nodal arrays of at least npoin entries
   subroutine check_tet_shapes(who, nelem, npoin, connec, Ngp, dNgp, He, gpvol, nu, nr1, nr2, nr3)
      character(*), intent(in) :: who
      integer, intent(in) :: nelem, npoin, nu, nr1, nr2, nr3
      integer(4), intent(in) :: connec(:,:)
      real(8), intent(in) :: Ngp(:,:), dNgp(:,:,:), He(:,:,:,:), gpvol(:,:,:)
      if (size(connec,1) /= nelem .or. size(connec,2) /= 4 .or. size(Ngp,1) /= 4 .or. size(Ngp,2) /= 4 .or. &
          size(dNgp,1) /= 3 .or. size(dNgp,2) /= 4 .or. size(dNgp,3) /= 4 .or. size(He,3) /= 4 .or. &
          size(He,4) < nelem .or. size(gpvol,2) /= 4 .or. size(gpvol,3) < nelem .or. &
          nu < npoin .or. nr1 < npoin .or. nr2 < npoin .or. nr3 < npoin) then
         write(*,'(a,a,a,i0,a,i0)') 'msod: ', who, ': array shapes do not match nelem=', nelem, ' npoin=', npoin
         error stop 2
      end if
   end subroutine check_tet_shapes

end module msod_tet_shapes

module elem_convec

   use iso_c_binding
   use msod_device, only: msod_full_convec_ijk, msod_full_convec_ijk_r4, msod_fem_convec, msod_check
   use msod_tet_shapes, only: check_tet_shapes
   implicit none
   private
   public :: full_convec_ijk, fem_convec

   ! elem_convec.f90:198-211 at either real kind of the reference
   interface full_convec_ijk
      module procedure full_convec_ijk_r8, full_convec_ijk_r4
   end interface full_convec_ijk

contains

   subroutine full_convec_ijk_r8(nelem, npoin, connec, Ngp, dNgp, He, gpvol, dlxigp_ip, xgp, invAtoIJK, &
                                 gmshAtoI, gmshAtoJ, gmshAtoK, u, q, rho, pr, E, Rmass, Rmom, Rener)
      integer(4), intent(in) :: nelem, npoin
      integer(4), intent(in), contiguous :: connec(:,:)
      real(8), intent(in), contiguous :: Ngp(:,:), dNgp(:,:,:), He(:,:,:,:), xgp(:,:), dlxigp_ip(:,:,:)
      real(8), intent(in), contiguous :: gpvol(:,:,:)
      integer(4), intent(in), contiguous :: invAtoIJK(:,:,:), gmshAtoI(:), gmshAtoJ(:), gmshAtoK(:)
      real(8), intent(in), contiguous :: q(:,:), u(:,:), rho(:), pr(:), E(:)
      real(8), intent(out), contiguous :: Rmass(:), Rmom(:,:), Rener(:)
      integer :: porder
      porder = size(invAtoIJK, 1) - 1
      call check_shapes('full_convec_ijk', nelem, npoin, porder, connec, He, gpvol, size(u,1), &
                        size(Rmass), size(Rmom,1), size(Rener))
      call msod_check(msod_full_convec_ijk(int(porder, c_int), nelem, npoin, connec, Ngp, dNgp, He, &
                      gpvol, dlxigp_ip, xgp, invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK, u, q, rho, pr, E, &
                      Rmass, Rmom, Rener), 'full_convec_ijk')
   end subroutine full_convec_ijk_r8

   ! The same at rp = 4: real(4) arrays and residuals.
   subroutine full_convec_ijk_r4(nelem, npoin, connec, Ngp, dNgp, He, gpvol, dlxigp_ip, xgp, invAtoIJK, &
                                 gmshAtoI, gmshAtoJ, gmshAtoK, u, q, rho, pr, E, Rmass, Rmom, Rener)
      integer(4), intent(in) :: nelem, npoin
      integer(4), intent(in), contiguous :: connec(:,:)
      real(4), intent(in), contiguous :: Ngp(:,:), dNgp(:,:,:), He(:,:,:,:), xgp(:,:), dlxigp_ip(:,:,:)
      real(4), intent(in), contiguous :: gpvol(:,:,:)
      integer(4), intent(in), contiguous :: invAtoIJK(:,:,:), gmshAtoI(:), gmshAtoJ(:), gmshAtoK(:)
      real(4), intent(in), contiguous :: q(:,:), u(:,:), rho(:), pr(:), E(:)
      real(4), intent(out), contiguous :: Rmass(:), Rmom(:,:), Rener(:)
      integer :: porder, nn
      porder = size(invAtoIJK, 1) - 1
      nn = (porder+1)**3
      if (size(connec,1) /= nelem .or. size(connec,2) /= nn .or. size(He,3) /= nn .or. &
          size(He,4) < nelem .or. size(gpvol,2) /= nn .or. size(gpvol,3) < nelem .or. &
          size(u,1) < npoin .or. size(Rmass) < npoin .or. size(Rmom,1) < npoin .or. size(Rener) < npoin) then
         write(*,'(a,i0,a,i0,a,i0)') 'msod: full_convec_ijk (real(4)): array shapes do not match nelem=', &
              nelem, ' npoin=', npoin, ' porder=', porder
         error stop 2
      end if
      call msod_check(msod_full_convec_ijk_r4(int(porder, c_int), nelem, npoin, connec, Ngp, dNgp, He, &
                      gpvol, dlxigp_ip, xgp, invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK, u, q, rho, pr, E, &
                      Rmass, Rmom, Rener), 'full_convec_ijk (real(4))')
   end subroutine full_convec_ijk_r4

   subroutine check_shapes(who, nelem, npoin, porder, connec, He, gpvol, nu, nr1, nr2, nr3)
      character(*), intent(in) :: who
      integer, intent(in) :: nelem, npoin, porder, nu, nr1, nr2, nr3
      integer(4), intent(in) :: connec(:,:)
      real(8), intent(in) :: He(:,:,:,:), gpvol(:,:,:)
      integer :: nn
      nn = (porder+1)**3
      if (size(connec,1) /= nelem .or. size(connec,2) /= nn .or. size(He,3) /= nn .or. &
          size(He,4) < nelem .or. size(gpvol,2) /= nn .or. size(gpvol,3) < nelem .or. &
          nu < npoin .or. nr1 < npoin .or. nr2 < npoin .or. nr3 < npoin) then
         write(*,'(a,a,a,i0,a,i0,a,i0)') 'msod: ', who, ': array shapes do not match nelem=', nelem, &
              ' npoin=', npoin, ' porder=', porder
         error stop 2
      end if
   end subroutine check_shapes

   ! Linear tetrahedra (elem_convec.f90:839-948).  On a mesh where no node is
   ! shared between element nodes (the reference's disjoint TET04 mesh) the
   ! contributions are stored; otherwise they meet at the nodes through FP64
   ! atomics, like the reference's atomic updates.
   subroutine fem_convec(nelem, npoin, connec, Ngp, dNgp, He, gpvol, u, q, rho, pr, E, Rmass, Rmom, Rener)
      integer(4), intent(in) :: nelem, npoin
      integer(4), intent(in), contiguous :: connec(:,:)
      real(8), intent(in), contiguous :: Ngp(:,:), dNgp(:,:,:), He(:,:,:,:), gpvol(:,:,:)
      real(8), intent(in), contiguous :: q(:,:), u(:,:), rho(:), pr(:), E(:)
      real(8), intent(out), contiguous :: Rmass(:), Rmom(:,:), Rener(:)
      call check_tet_shapes('fem_convec', nelem, npoin, connec, Ngp, dNgp, He, gpvol, size(u,1), &
                            size(Rmass), size(Rmom,1), size(Rener))
      call msod_check(msod_fem_convec(nelem, npoin, connec, Ngp, dNgp, He, gpvol, u, q, rho, pr, E, &
                      Rmass, Rmom, Rener), 'fem_convec')
   end subroutine fem_convec

end module elem_convec



module elem_diffu

   use iso_c_binding
   use msod_device, only: msod_full_diffusion_ijk, msod_set_diffusion_coefficients, msod_check, &
                          msod_full_diffusion_ijk_closure, msod_fem_diffu, msod_full_diffusion_ijk_r4
   use msod_tet_shapes, only: check_tet_shapes
   implicit none
   private
   public :: full_diffusion_ijk, full_diffusion_ijk_closure, fem_diffu, c_rho, c_ener

   real(8) :: c_rho = 1.0d0
   real(8) :: c_ener = 1.0d0

   ! elem_diffu.f90:10-23 at either real kind of the reference
   interface full_diffusion_ijk
      module procedure full_diffusion_ijk_r8, full_diffusion_ijk_r4
   end interface full_diffusion_ijk

contains

   subroutine full_diffusion_ijk_r8(nelem, npoin, connec, Ngp, dNgp, He, gpvol, dlxigp_ip, xgp, invAtoIJK, &
                                    gmshAtoI, gmshAtoJ, gmshAtoK, Cp, Pr, rho, u, Tem, mu_fluid, mu_e, mu_sgs, &
                                    Rmass, Rmom, Rener)
      integer(4), intent(in) :: nelem, npoin
      integer(4), intent(in), contiguous :: connec(:,:)
      real(8), intent(in), contiguous :: Ngp(:,:), dNgp(:,:,:), He(:,:,:,:), xgp(:,:), dlxigp_ip(:,:,:)
      real(8), intent(in), contiguous :: gpvol(:,:,:)
      integer(4), intent(in), contiguous :: invAtoIJK(:,:,:), gmshAtoI(:), gmshAtoJ(:), gmshAtoK(:)
      real(8), intent(in) :: Cp, Pr
      real(8), intent(in), contiguous :: rho(:), u(:,:), Tem(:), mu_fluid(:), mu_e(:,:), mu_sgs(:,:)
      real(8), intent(out), contiguous :: Rmass(:), Rmom(:,:), Rener(:)
      integer :: porder, nn
      porder = size(invAtoIJK, 1) - 1
      nn = (porder+1)**3
      if (size(connec,1) /= nelem .or. size(connec,2) /= nn .or. size(He,4) < nelem .or. &
          size(mu_e,1) /= nelem .or. size(mu_e,2) /= nn .or. size(mu_sgs,1) /= nelem .or. &
          size(u,1) < npoin .or. size(Rmass) < npoin .or. size(Rmom,1) < npoin) then
         write(*,'(a,i0,a,i0,a,i0)') 'msod: full_diffusion_ijk: array shapes do not match nelem=', nelem, &
              ' npoin=', npoin, ' porder=', porder
         error stop 2
      end if
      call msod_check(msod_set_diffusion_coefficients(c_rho, c_ener), 'set_diffusion_coefficients')
      call msod_check(msod_full_diffusion_ijk(int(porder, c_int), nelem, npoin, connec, Ngp, dNgp, He, &
                      gpvol, dlxigp_ip, xgp, invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK, Cp, Pr, rho, u, Tem, &
                      mu_fluid, mu_e, mu_sgs, Rmass, Rmom, Rener), 'full_diffusion_ijk')
   end subroutine full_diffusion_ijk_r8

   ! The same at rp = 4: real(4) arrays, Cp, Pr and residuals.
   subroutine full_diffusion_ijk_r4(nelem, npoin, connec, Ngp, dNgp, He, gpvol, dlxigp_ip, xgp, invAtoIJK, &
                                    gmshAtoI, gmshAtoJ, gmshAtoK, Cp, Pr, rho, u, Tem, mu_fluid, mu_e, mu_sgs, &
                                    Rmass, Rmom, Rener)
      integer(4), intent(in) :: nelem, npoin
      integer(4), intent(in), contiguous :: connec(:,:)
      real(4), intent(in), contiguous :: Ngp(:,:), dNgp(:,:,:), He(:,:,:,:), xgp(:,:), dlxigp_ip(:,:,:)
      real(4), intent(in), contiguous :: gpvol(:,:,:)
      integer(4), intent(in), contiguous :: invAtoIJK(:,:,:), gmshAtoI(:), gmshAtoJ(:), gmshAtoK(:)
      real(4), intent(in) :: Cp, Pr
      real(4), intent(in), contiguous :: rho(:), u(:,:), Tem(:), mu_fluid(:), mu_e(:,:), mu_sgs(:,:)
      real(4), intent(out), contiguous :: Rmass(:), Rmom(:,:), Rener(:)
      integer :: porder, nn
      porder = size(invAtoIJK, 1) - 1
      nn = (porder+1)**3
      if (size(connec,1) /= nelem .or. size(connec,2) /= nn .or. size(He,4) < nelem .or. &
          size(mu_e,1) /= nelem .or. size(mu_e,2) /= nn .or. size(mu_sgs,1) /= nelem .or. &
          size(u,1) < npoin .or. size(Rmass) < npoin .or. size(Rmom,1) < npoin) then
         write(*,'(a,i0,a,i0,a,i0)') 'msod: full_diffusion_ijk (real(4)): array shapes do not match nelem=', &
              nelem, ' npoin=', npoin, ' porder=', porder
         error stop 2
      end if
      call msod_check(msod_set_diffusion_coefficients(c_rho, c_ener), 'set_diffusion_coefficients')
      call msod_check(msod_full_diffusion_ijk_r4(int(porder, c_int), nelem, npoin, connec, Ngp, dNgp, He, &
                      gpvol, dlxigp_ip, xgp, invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK, Cp, Pr, rho, u, Tem, &
                      mu_fluid, mu_e, mu_sgs, Rmass, Rmom, Rener), 'full_diffusion_ijk (real(4))')
   end subroutine full_diffusion_ijk_r4

   ! full_diffusion_ijk with mu_sgs predicted by the learned closure inside the
   ! kernel (no counterpart in the reference).  mlp_params: the 1530 real(4)
   ! weights, present on the device or host memory.  used_model = .false. when
   ! the weights are unusable (non-finite, negative scale): mu_sgs = 0 then.
   subroutine full_diffusion_ijk_closure(nelem, npoin, connec, Ngp, dNgp, He, gpvol, dlxigp_ip, xgp, &
                                         invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK, Cp, Pr, rho, u, Tem, &
                                         mu_fluid, mu_e, mlp_params, used_model, Rmass, Rmom, Rener)
      integer(4), intent(in) :: nelem, npoin
      integer(4), intent(in), contiguous :: connec(:,:)
      real(8), intent(in), contiguous :: Ngp(:,:), dNgp(:,:,:), He(:,:,:,:), xgp(:,:), dlxigp_ip(:,:,:)
      real(8), intent(in), contiguous :: gpvol(:,:,:)
      integer(4), intent(in), contiguous :: invAtoIJK(:,:,:), gmshAtoI(:), gmshAtoJ(:), gmshAtoK(:)
      real(8), intent(in) :: Cp, Pr
      real(8), intent(in), contiguous :: rho(:), u(:,:), Tem(:), mu_fluid(:), mu_e(:,:)
      real(4), intent(in), contiguous :: mlp_params(:)
      logical, intent(out) :: used_model
      real(8), intent(out), contiguous :: Rmass(:), Rmom(:,:), Rener(:)
      integer :: porder, nn
      integer(c_int) :: used
      porder = size(invAtoIJK, 1) - 1
      nn = (porder+1)**3
      if (size(connec,1) /= nelem .or. size(connec,2) /= nn .or. size(He,4) < nelem .or. &
          size(mu_e,1) /= nelem .or. size(mu_e,2) /= nn .or. size(mlp_params) < 1530 .or. &
          size(u,1) < npoin .or. size(Rmass) < npoin .or. size(Rmom,1) < npoin) then
         write(*,'(a,i0,a,i0,a,i0)') 'msod: full_diffusion_ijk_closure: array shapes do not match nelem=', &
              nelem, ' npoin=', npoin, ' porder=', porder
         error stop 2
      end if
      call msod_check(msod_set_diffusion_coefficients(c_rho, c_ener), 'set_diffusion_coefficients')
      call msod_check(msod_full_diffusion_ijk_closure(int(porder, c_int), nelem, npoin, connec, Ngp, dNgp, &
                      He, gpvol, dlxigp_ip, xgp, invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK, Cp, Pr, rho, u, &
                      Tem, mu_fluid, mu_e, mlp_params, used, Rmass, Rmom, Rener), 'full_diffusion_ijk_closure')
      used_model = used /= 0
   end subroutine full_diffusion_ijk_closure

   ! Linear tetrahedra (elem_diffu.f90:199-317): stores on a mesh without
   ! shared nodes (the reference adds without atomics there too,
   ! elem_diffu.f90:306-312), FP64 atomics otherwise.
   subroutine fem_diffu(nelem, npoin, connec, Ngp, dNgp, He, gpvol, Cp, Pr, rho, u, Tem, mu_fluid, mu_e, &
                        mu_sgs, Rmass, Rmom, Rener)
      integer(4), intent(in) :: nelem, npoin
      integer(4), intent(in), contiguous :: connec(:,:)
      real(8), intent(in), contiguous :: Ngp(:,:), dNgp(:,:,:), He(:,:,:,:), gpvol(:,:,:)
      real(8), intent(in) :: Cp, Pr
      real(8), intent(in), contiguous :: rho(:), u(:,:), Tem(:), mu_fluid(:), mu_e(:,:), mu_sgs(:,:)
      real(8), intent(out), contiguous :: Rmass(:), Rmom(:,:), Rener(:)
      call check_tet_shapes('fem_diffu', nelem, npoin, connec, Ngp, dNgp, He, gpvol, size(u,1), &
                            size(Rmass), size(Rmom,1), size(Rener))
      if (size(mu_e,1) /= nelem .or. size(mu_e,2) /= 4 .or. size(mu_sgs,1) /= nelem .or. size(mu_sgs,2) /= 4) then
         write(*,'(a,i0)') 'msod: fem_diffu: mu_e, mu_sgs must be (nelem,4), nelem=', nelem
         error stop 2
      end if
      call msod_check(msod_set_diffusion_coefficients(c_rho, c_ener), 'set_diffusion_coefficients')
      call msod_check(msod_fem_diffu(nelem, npoin, connec, Ngp, dNgp, He, gpvol, Cp, Pr, rho, u, Tem, &
                      mu_fluid, mu_e, mu_sgs, Rmass, Rmom, Rener), 'fem_diffu')
   end subroutine fem_diffu

end module elem_diffu

module jacobian_oper

   use iso_c_binding
   use msod_device, only: msod_jacobian_device, msod_tet_jacobian_device, msod_check
   implicit none
   private
   public :: elem_jacobian, tet_jacobian

contains

   ! He and gpvol of every element node on the GPU.  Like the reference,
   ! which runs it on its `enter data` copies (main.f90:280-295), every
   ! array must be present on the device (msod_copyin of connec, coord,
   ! dNgp, wgp; msod_copyin or msod_create of gpvol and He); the results
   ! stay there until msod_update_host.  porder is read from the shape of
   ! dNgp(3,nnode,ngaus).  A non-positive det(Je), which the reference does
   ! not check, stops with a message.
   subroutine elem_jacobian(nelem, npoin, connec, coord, dNgp, wgp, gpvol, He)
      integer(4), intent(in) :: nelem, npoin
      integer(4), intent(in), contiguous :: connec(:,:)
      real(8), intent(in), contiguous :: coord(:,:), dNgp(:,:,:), wgp(:)
      real(8), intent(out), contiguous :: gpvol(:,:,:), He(:,:,:,:)
      integer :: nn, porder
      integer(c_int64_t) :: nbad
      nn = size(dNgp, 2)
      porder = nint(real(nn, 8)**(1d0/3d0)) - 1
      if ((porder+1)**3 /= nn .or. size(connec,1) /= nelem .or. size(connec,2) /= nn .or. &
          size(coord,1) < npoin .or. size(He,3) /= nn .or. size(He,4) < nelem .or. &
          size(gpvol,2) /= nn .or. size(gpvol,3) < nelem .or. size(wgp) < nn) then
         write(*,'(a,i0,a,i0)') 'msod: elem_jacobian: array shapes do not match nelem=', nelem, &
              ' npoin=', npoin
         error stop 2
      end if
      call msod_check(msod_jacobian_device(int(porder, c_int), nelem, npoin, connec, coord, dNgp, wgp, &
                      gpvol, He, nbad), 'elem_jacobian')
      if (nbad > 0) then
         write(*,'(a,i0,a)') 'msod: elem_jacobian: ', nbad, ' element nodes with det(Je) <= 0'
         error stop 3
      end if
   end subroutine elem_jacobian

   ! He(3,3,4,nelem) and gpvol(1,4,nelem) of linear tetrahedra on the GPU
   ! (jacobian_oper.f90:138-267); arrays present on the device, as above.
   subroutine tet_jacobian(nelem, npoin, connec, coord, dNgp, wgp, gpvol, He)
      integer(4), intent(in) :: nelem, npoin
      integer(4), intent(in), contiguous :: connec(:,:)
      real(8), intent(in), contiguous :: coord(:,:), dNgp(:,:,:), wgp(:)
      real(8), intent(out), contiguous :: gpvol(:,:,:), He(:,:,:,:)
      integer(c_int64_t) :: nbad
      if (size(connec,1) /= nelem .or. size(connec,2) /= 4 .or. size(coord,1) < npoin .or. &
          size(dNgp,2) /= 4 .or. size(wgp) < 4 .or. size(He,3) /= 4 .or. size(He,4) < nelem .or. &
          size(gpvol,2) /= 4 .or. size(gpvol,3) < nelem) then
         write(*,'(a,i0,a,i0)') 'msod: tet_jacobian: array shapes do not match nelem=', nelem, ' npoin=', npoin
         error stop 2
      end if
      call msod_check(msod_tet_jacobian_device(nelem, npoin, connec, coord, dNgp, wgp, gpvol, He, nbad), &
                      'tet_jacobian')
      if (nbad > 0) then
         write(*,'(a,i0,a)') 'msod: tet_jacobian: ', nbad, ' points with det(Je) <= 0'
         error stop 3
      end if
   end subroutine tet_jacobian

end module jacobian_oper
