! miniSOD timer driver on the gfx950 kernels: counterpart of the reference's
! program main (app/sources/main.f90).  Builds the element tables, a mesh
! with its geometry and fields, mirrors them on the GPU, times nruns calls of
! full_convec_ijk, fem_convec, full_diffusion_ijk and fem_diffu (each call
! synchronous, the reference's timed region, main.f90:444-496), writes
! timers.dat in the reference's format (main.f90:495,499) and prints the
! reference's summary (main.f90:504-581) plus node-update rates.
!
! The TET04 kernels run, as in the reference (main.f90:364-407), on npoin/4
! disjoint tetrahedra over the same nodal fields, with mu_e = mu_sgs = 1 and
! tet_jacobian on the device copies.  The reference stops when 4 does not
! divide npoin; here the TET part is skipped then (its timers.dat columns are
! 0), and also with --mixed or --no-tet.
!
! Usage: miniSOD [--n N] [--nelem NE] [--porder P] [--nruns R] [--jitter J]
!                [--rho-jump F] [--mixed] [--assembly colored|atomic] [--quiet]
!                [--gpu-geometry] [--no-tet]
!   --n N      conforming N x N x N box of the unit cube (default 8)
!   --nelem NE the reference's own case instead: NE disjoint elements with
!              its initial conditions (main.f90:257-413)
!   --mixed    time the mixed-precision kernels (FP32 state and geometry)
!   --gpu-geometry  compute He and gpvol on the device (elem_jacobian of the
!              drop-in module jacobian_oper) on the device copies, as the
!              reference does (main.f90:280-295); FP64 kernels only
!   --assembly atomic (default, the library's default) or colored
program miniSOD

   use iso_c_binding
   use msod_fem, only: msod_hex_tables, msod_tet_tables
   use msod_mesh, only: msod_box_mesh, msod_box_sizes, msod_disjoint_mesh
   use msod_geom, only: msod_elem_jacobian
   use msod_fields, only: msod_flow_fields, msod_viscosity_fields
   use msod_device
   use elem_convec, only: full_convec_ijk, fem_convec
   use elem_diffu, only: full_diffusion_ijk, fem_diffu
   use jacobian_oper, only: elem_jacobian, tet_jacobian
   implicit none

   integer :: n = 8, nelem_ref = 0, porder = 4, nruns = 10, nn, q, i, nbad, assembly = 1
   real(8) :: jitter = 0d0, rho_jump = 0d0
   logical :: mixed = .false., quiet = .false., gpu_geometry = .false., tet = .true.
   integer(8) :: nelem8, npoin8
   integer :: nelem, npoin
   integer(c_int), allocatable :: connec(:,:), invAtoIJK(:,:,:), AtoIJK(:), AtoI(:), AtoJ(:), AtoK(:)
   real(c_double), allocatable :: xgp(:,:), wgp(:), Ngp(:,:), dNgp(:,:,:), dlxigp_ip(:,:,:)
   real(c_double), allocatable :: coord(:,:), He(:,:,:,:), gpvol(:,:,:)
   real(c_double), allocatable :: u(:,:), qm(:,:), rho(:), pr(:), E(:), Tem(:), mu_fluid(:), mu_e(:,:), mu_sgs(:,:)
   real(c_double), allocatable :: Rmass(:), Rmom(:,:), Rener(:), Dmass(:), Dmom(:,:), Dener(:)
   real(c_double), allocatable :: Rmass_s(:), Rmom_s(:,:), Rener_s(:)
   real(c_float), allocatable :: He4(:,:,:,:), gpvol4(:,:,:), u4(:,:), q4(:,:), rho4(:), E4(:), Tem4(:)
   real(c_float), allocatable :: muf4(:), mue4(:,:), mus4(:,:)
   real(8), parameter :: gamma = 1.4d0, Rgas = 287d0
   real(8) :: Cp, Pra, tc, td, tavg_c, tavg_d, tmax_c, tmax_d, tmin_c, tmin_d
   real(8) :: tcs, tavg_cs, tmax_cs, tmin_cs
   ! TET04 data (main.f90:364-407)
   integer :: nelem_t
   integer(c_int), allocatable :: connec_t(:,:)
   real(c_double), allocatable :: coord_t(:,:), xgp_t(:,:), wgp_t(:), Ngp_t(:,:), dNgp_t(:,:,:)
   real(c_double), allocatable :: He_t(:,:,:,:), gpvol_t(:,:,:), mu_e_t(:,:), mu_sgs_t(:,:)
   real(c_double), allocatable :: Rmass_t(:), Rmom_t(:,:), Rener_t(:), Dmass_t(:), Dmom_t(:,:), Dener_t(:)
   real(8) :: tct, tdt, tavg_ct, tavg_dt, tmax_ct, tmax_dt, tmin_ct, tmin_dt
   real(c_double), parameter :: xyzTET(4,3) = reshape([0d0, 1d0, 0d0, 0d0, 0d0, 0d0, 1d0, 0d0, &
                                                       0d0, 0d0, 0d0, 1d0], [4, 3])
   integer(8) :: c0, c1, c2, rate

   call parse_args()
   q = porder + 1
   nn = q**3

   ! ---- tables ---------------------------------------------------------------
   allocate(invAtoIJK(q,q,q), AtoIJK(nn), AtoI(nn), AtoJ(nn), AtoK(nn))
   allocate(xgp(nn,3), wgp(nn), Ngp(nn,nn), dNgp(3,nn,nn), dlxigp_ip(nn,3,q))
   call msod_hex_tables(porder, invAtoIJK, AtoIJK, AtoI, AtoJ, AtoK, xgp, wgp, Ngp, dNgp, dlxigp_ip)

   ! ---- mesh and geometry ----------------------------------------------------
   if (nelem_ref > 0) then
      nelem8 = nelem_ref
      npoin8 = int(nelem_ref, 8)*nn
   else
      call msod_box_sizes(n, n, n, porder, nelem8, npoin8)
   end if
   if (npoin8 >= 2_8**31 .or. nelem8*nn >= 2_8**31) stop 'mesh too large for 32-bit indices'
   nelem = int(nelem8)
   npoin = int(npoin8)
   write(*,*) 'Number of elements          = ', nelem
   write(*,*) 'Element order               = ', porder
   write(*,*) 'Number of nodes per element = ', nn
   write(*,*) 'Nodes on mesh               = ', npoin
   write(*,*) 'Number of runs              = ', nruns
   write(*,*)
   allocate(connec(nelem,nn), coord(npoin,3), He(3,3,nn,nelem), gpvol(1,nn,nelem))
   if (nelem_ref > 0) then
      call msod_disjoint_mesh(nelem, npoin, nn, xgp, connec, coord)
   else
      call msod_box_mesh(n, n, n, porder, jitter, 1004_c_int64_t, connec, coord)
   end if
   call msod_elem_jacobian(nelem, npoin, nn, connec, coord, dNgp, wgp, gpvol, He, nbad)
   if (nbad > 0) stop 'non-positive Jacobian determinant'

   ! ---- fields -----------------------------------------------------------------
   allocate(u(npoin,3), qm(npoin,3), rho(npoin), pr(npoin), E(npoin), Tem(npoin), mu_fluid(npoin))
   allocate(mu_e(nelem,nn), mu_sgs(nelem,nn))
   if (nelem_ref > 0) then
      call reference_fields()
      Cp = 1d0
      Pra = 1d0
   else
      call msod_flow_fields(npoin8, coord, gamma, Rgas, 1001_c_int64_t, rho_jump, u, qm, rho, pr, E, Tem, mu_fluid)
      call msod_viscosity_fields(nelem8, nn, 1002_c_int64_t, 1003_c_int64_t, 1d-3, 1d-4, mu_e, mu_sgs)
      Cp = 1004.5d0
      Pra = 0.71d0
   end if
   allocate(Rmass(npoin), Rmom(npoin,3), Rener(npoin), Dmass(npoin), Dmom(npoin,3), Dener(npoin))
   allocate(Rmass_s(npoin), Rmom_s(npoin,3), Rener_s(npoin))

   ! ---- device data (the reference's `enter data` block) ------------------------
   call msod_start(0)
   call msod_set_assembly_mode(assembly)
   call msod_copyin(connec, c_sizeof(connec))
   call msod_create(Rmass, c_sizeof(Rmass))
   call msod_create(Rmom, c_sizeof(Rmom))
   call msod_create(Rener, c_sizeof(Rener))
   call msod_create(Dmass, c_sizeof(Dmass))
   call msod_create(Dmom, c_sizeof(Dmom))
   call msod_create(Dener, c_sizeof(Dener))
   call msod_create(Rmass_s, c_sizeof(Rmass_s))
   call msod_create(Rmom_s, c_sizeof(Rmom_s))
   call msod_create(Rener_s, c_sizeof(Rener_s))
   if (mixed) then
      allocate(He4, source=real(He, c_float))
      allocate(gpvol4, source=real(gpvol, c_float))
      allocate(u4, source=real(u, c_float))
      allocate(q4, source=real(qm, c_float))
      allocate(rho4, source=real(rho, c_float))
      allocate(E4, source=real(E, c_float))
      allocate(Tem4, source=real(Tem, c_float))
      allocate(muf4, source=real(mu_fluid, c_float))
      allocate(mue4, source=real(mu_e, c_float))
      allocate(mus4, source=real(mu_sgs, c_float))
      call msod_copyin(He4, c_sizeof(He4))
      call msod_copyin(gpvol4, c_sizeof(gpvol4))
      call msod_copyin(u4, c_sizeof(u4))
      call msod_copyin(q4, c_sizeof(q4))
      call msod_copyin(rho4, c_sizeof(rho4))
      call msod_copyin(E4, c_sizeof(E4))
      call msod_copyin(Tem4, c_sizeof(Tem4))
      call msod_copyin(muf4, c_sizeof(muf4))
      call msod_copyin(mue4, c_sizeof(mue4))
      call msod_copyin(mus4, c_sizeof(mus4))
   else
      call msod_copyin(He, c_sizeof(He))
      call msod_copyin(gpvol, c_sizeof(gpvol))
      if (gpu_geometry) then
         ! the device copies of He and gpvol are recomputed from the coordinates
         call msod_copyin(coord, c_sizeof(coord))
         call msod_copyin(dNgp, c_sizeof(dNgp))
         call msod_copyin(wgp, c_sizeof(wgp))
         call elem_jacobian(nelem, npoin, connec, coord, dNgp, wgp, gpvol, He)
      end if
      call msod_copyin(u, c_sizeof(u))
      call msod_copyin(qm, c_sizeof(qm))
      call msod_copyin(rho, c_sizeof(rho))
      call msod_copyin(pr, c_sizeof(pr))
      call msod_copyin(E, c_sizeof(E))
      call msod_copyin(Tem, c_sizeof(Tem))
      call msod_copyin(mu_fluid, c_sizeof(mu_fluid))
      call msod_copyin(mu_e, c_sizeof(mu_e))
      call msod_copyin(mu_sgs, c_sizeof(mu_sgs))
   end if

   ! ---- TET04 data (main.f90:364-407) ----------------------------------------------
   if (mod(npoin, 4) /= 0 .or. mixed) tet = .false.
   if (tet) then
      nelem_t = npoin/4
      allocate(xgp_t(4,3), wgp_t(4), Ngp_t(4,4), dNgp_t(3,4,4))
      call msod_tet_tables(xgp_t, wgp_t, Ngp_t, dNgp_t)
      allocate(connec_t(nelem_t,4), coord_t(npoin,3))
      call msod_disjoint_mesh(nelem_t, npoin, 4, xyzTET, connec_t, coord_t)
      allocate(He_t(3,3,4,nelem_t), gpvol_t(1,4,nelem_t), mu_e_t(nelem_t,4), mu_sgs_t(nelem_t,4))
      mu_e_t = 1d0
      mu_sgs_t = 1d0
      allocate(Rmass_t(npoin), Rmom_t(npoin,3), Rener_t(npoin), Dmass_t(npoin), Dmom_t(npoin,3), Dener_t(npoin))
      call msod_copyin(connec_t, c_sizeof(connec_t))
      call msod_copyin(coord_t, c_sizeof(coord_t))
      call msod_copyin(dNgp_t, c_sizeof(dNgp_t))
      call msod_copyin(wgp_t, c_sizeof(wgp_t))
      call msod_create(He_t, c_sizeof(He_t))
      call msod_create(gpvol_t, c_sizeof(gpvol_t))
      call tet_jacobian(nelem_t, npoin, connec_t, coord_t, dNgp_t, wgp_t, gpvol_t, He_t)
      call msod_copyin(mu_e_t, c_sizeof(mu_e_t))
      call msod_copyin(mu_sgs_t, c_sizeof(mu_sgs_t))
      call msod_create(Rmass_t, c_sizeof(Rmass_t))
      call msod_create(Rmom_t, c_sizeof(Rmom_t))
      call msod_create(Rener_t, c_sizeof(Rener_t))
      call msod_create(Dmass_t, c_sizeof(Dmass_t))
      call msod_create(Dmom_t, c_sizeof(Dmom_t))
      call msod_create(Dener_t, c_sizeof(Dener_t))
   else
      write(*,*) 'TET04 kernels skipped (npoin not divisible by 4, or --mixed / --no-tet)'
   end if

   ! ---- timing loop (main.f90:427-499) ------------------------------------------
   open(unit=1, file='timers.dat', status='replace')
   tavg_c = 0d0; tavg_d = 0d0; tmax_c = 0d0; tmax_d = 0d0; tmin_c = huge(1d0); tmin_d = huge(1d0)
   tavg_cs = 0d0; tmax_cs = 0d0; tmin_cs = huge(1d0)
   tavg_ct = 0d0; tavg_dt = 0d0; tmax_ct = 0d0; tmax_dt = 0d0; tmin_ct = huge(1d0); tmin_dt = huge(1d0)
   tct = 0d0; tdt = 0d0
   call system_clock(count_rate=rate)
   do i = 1, nruns
      call system_clock(c0)
      call convec(Rmass, Rmom, Rener)
      call system_clock(c1)
      tc = dble(c1 - c0)/dble(rate)
      ! the reference's second convective call, full_convec_ijk_split
      ! (main.f90:455-463), into its own outputs: here the same kernel, which
      ! computes the split form's residuals (its Rener is the split kernel's,
      ! elem_convec.f90:694-697)
      call system_clock(c0)
      call convec(Rmass_s, Rmom_s, Rener_s)
      call system_clock(c1)
      tcs = dble(c1 - c0)/dble(rate)
      if (tet) then
         call system_clock(c0)
         call fem_convec(nelem_t, npoin, connec_t, Ngp_t, dNgp_t, He_t, gpvol_t, u, qm, rho, pr, E, &
                         Rmass_t, Rmom_t, Rener_t)
         call system_clock(c1)
         tct = dble(c1 - c0)/dble(rate)
      end if
      call system_clock(c1)
      call diffu()
      call system_clock(c2)
      td = dble(c2 - c1)/dble(rate)
      if (tet) then
         call system_clock(c1)
         call fem_diffu(nelem_t, npoin, connec_t, Ngp_t, dNgp_t, He_t, gpvol_t, Cp, Pra, rho, u, Tem, &
                        mu_fluid, mu_e_t, mu_sgs_t, Dmass_t, Dmom_t, Dener_t)
         call system_clock(c2)
         tdt = dble(c2 - c1)/dble(rate)
      end if
      tavg_c = tavg_c + tc; tmax_c = max(tmax_c, tc); tmin_c = min(tmin_c, tc)
      tavg_cs = tavg_cs + tcs; tmax_cs = max(tmax_cs, tcs); tmin_cs = min(tmin_cs, tcs)
      tavg_d = tavg_d + td; tmax_d = max(tmax_d, td); tmin_d = min(tmin_d, td)
      tavg_ct = tavg_ct + tct; tmax_ct = max(tmax_ct, tct); tmin_ct = min(tmin_ct, tct)
      tavg_dt = tavg_dt + tdt; tmax_dt = max(tmax_dt, tdt); tmin_dt = min(tmin_dt, tdt)
      write(1,10) i, tc, tct, td, tdt
   end do
   close(1)
10 format(i3,1x,4(f16.8,1x))
   tavg_c = tavg_c/dble(nruns)
   tavg_cs = tavg_cs/dble(nruns)
   tavg_d = tavg_d/dble(nruns)
   tavg_ct = tavg_ct/dble(nruns)
   tavg_dt = tavg_dt/dble(nruns)
   write(*,*)
   write(*,*) 'Timings:'
   write(*,*) '----------------------------------------'
   write(*,*) 'Avg. convective time     = ', tavg_c
   write(*,*) 'Avg. split convective time = ', tavg_cs
   if (tet) write(*,*) 'Avg. TET convective time = ', tavg_ct
   write(*,*) 'Avg. diffusive time      = ', tavg_d
   if (tet) write(*,*) 'Avg. TET diffusive time  = ', tavg_dt
   write(*,*) 'Max. convective time     = ', tmax_c
   write(*,*) 'Max. split convective time = ', tmax_cs
   if (tet) write(*,*) 'Max. TET convective time = ', tmax_ct
   write(*,*) 'Max. diffusive time      = ', tmax_d
   if (tet) write(*,*) 'Max. TET diffusive time  = ', tmax_dt
   write(*,*) 'Min. convective time     = ', tmin_c
   write(*,*) 'Min. split convective time = ', tmin_cs
   if (tet) write(*,*) 'Min. TET convective time = ', tmin_ct
   write(*,*) 'Min. diffusive time      = ', tmin_d
   if (tet) write(*,*) 'Min. TET diffusive time  = ', tmin_dt
   write(*,*) '----------------------------------------'
   write(*,*) 'Variation convec.        = ', (tmax_c - tmin_c)/tavg_c
   write(*,*) 'Variation split convec.  = ', (tmax_cs - tmin_cs)/tavg_cs
   if (tet) write(*,*) 'Variation TET convec.    = ', (tmax_ct - tmin_ct)/tavg_ct
   write(*,*) 'Variation diffu.         = ', (tmax_d - tmin_d)/tavg_d
   if (tet) write(*,*) 'Variation TET diffu.     = ', (tmax_dt - tmin_dt)/tavg_dt
   write(*,*) 'GLL-node updates/s (min) = ', dble(npoin)/(tmin_c + tmin_d)
   write(*,*) '----------------------------------------'

   ! ---- results (main.f90:539-566) -------------------------------------------------
   call msod_update_host(Rmass, c_sizeof(Rmass))
   call msod_update_host(Rmom, c_sizeof(Rmom))
   call msod_update_host(Rener, c_sizeof(Rener))
   call msod_update_host(Dmass, c_sizeof(Dmass))
   call msod_update_host(Dmom, c_sizeof(Dmom))
   call msod_update_host(Dener, c_sizeof(Dener))
   if (.not. quiet) then
      write(*,*)
      write(*,*) 'Basic results:'
      write(*,*) '----------------------------------------'
      write(*,*) '--| Convec'
      write(*,*) 'Max Rmass     = ', maxval(Rmass), 'Min Rmass     = ', minval(Rmass)
      write(*,*) 'Max Rmom(:,1) = ', maxval(Rmom(:,1)), 'Min Rmom(:,1) = ', minval(Rmom(:,1))
      write(*,*) 'Max Rmom(:,2) = ', maxval(Rmom(:,2)), 'Min Rmom(:,2) = ', minval(Rmom(:,2))
      write(*,*) 'Max Rmom(:,3) = ', maxval(Rmom(:,3)), 'Min Rmom(:,3) = ', minval(Rmom(:,3))
      write(*,*) 'Max Rener     = ', maxval(Rener), 'Min Rener     = ', minval(Rener)
      write(*,*) '----------------------------------------'
      write(*,*) '--| Diffu'
      write(*,*) 'Max Dmass     = ', maxval(Dmass), 'Min Dmass     = ', minval(Dmass)
      write(*,*) 'Max Dmom(:,1) = ', maxval(Dmom(:,1)), 'Min Dmom(:,1) = ', minval(Dmom(:,1))
      write(*,*) 'Max Dmom(:,2) = ', maxval(Dmom(:,2)), 'Min Dmom(:,2) = ', minval(Dmom(:,2))
      write(*,*) 'Max Dmom(:,3) = ', maxval(Dmom(:,3)), 'Min Dmom(:,3) = ', minval(Dmom(:,3))
      write(*,*) 'Max Dener     = ', maxval(Dener), 'Min Dener     = ', minval(Dener)
   end if
   if (tet) then
      call msod_update_host(Rmass_t, c_sizeof(Rmass_t))
      call msod_update_host(Rmom_t, c_sizeof(Rmom_t))
      call msod_update_host(Rener_t, c_sizeof(Rener_t))
      call msod_update_host(Dmass_t, c_sizeof(Dmass_t))
      call msod_update_host(Dmom_t, c_sizeof(Dmom_t))
      call msod_update_host(Dener_t, c_sizeof(Dener_t))
      if (.not. quiet) then
         write(*,*) 'Basic TET results:'
         write(*,*) '----------------------------------------'
         write(*,*) '--| Convec'
         write(*,*) 'Max Rmass     = ', maxval(Rmass_t), 'Min Rmass     = ', minval(Rmass_t)
         write(*,*) 'Max Rmom(:,1) = ', maxval(Rmom_t(:,1)), 'Min Rmom(:,1) = ', minval(Rmom_t(:,1))
         write(*,*) 'Max Rmom(:,2) = ', maxval(Rmom_t(:,2)), 'Min Rmom(:,2) = ', minval(Rmom_t(:,2))
         write(*,*) 'Max Rmom(:,3) = ', maxval(Rmom_t(:,3)), 'Min Rmom(:,3) = ', minval(Rmom_t(:,3))
         write(*,*) 'Max Rener     = ', maxval(Rener_t), 'Min Rener     = ', minval(Rener_t)
         write(*,*) '----------------------------------------'
         write(*,*) '--| Diffu'
         write(*,*) 'Max Dmass     = ', maxval(Dmass_t), 'Min Dmass     = ', minval(Dmass_t)
         write(*,*) 'Max Dmom(:,1) = ', maxval(Dmom_t(:,1)), 'Min Dmom(:,1) = ', minval(Dmom_t(:,1))
         write(*,*) 'Max Dmom(:,2) = ', maxval(Dmom_t(:,2)), 'Min Dmom(:,2) = ', minval(Dmom_t(:,2))
         write(*,*) 'Max Dmom(:,3) = ', maxval(Dmom_t(:,3)), 'Min Dmom(:,3) = ', minval(Dmom_t(:,3))
         write(*,*) 'Max Dener     = ', maxval(Dener_t), 'Min Dener     = ', minval(Dener_t)
      end if
   end if
   call write_results()
   call msod_stop()

contains

   subroutine convec(Rmass, Rmom, Rener)
      ! explicit shape: the actual arrays are passed by address, the key of
      ! their device mirrors
      real(c_double), intent(inout) :: Rmass(npoin), Rmom(npoin,3), Rener(npoin)
      if (mixed) then
         call msod_check(msod_full_convec_ijk_mixed(int(porder, c_int), nelem, npoin, connec, He4, gpvol4, &
                         dlxigp_ip, invAtoIJK, AtoI, AtoJ, AtoK, u4, q4, rho4, E4, gamma, Rmass, Rmom, Rener), &
                         'full_convec_ijk_mixed')
      else
         call full_convec_ijk(nelem, npoin, connec, Ngp, dNgp, He, gpvol, dlxigp_ip, xgp, invAtoIJK, &
                              AtoI, AtoJ, AtoK, u, qm, rho, pr, E, Rmass, Rmom, Rener)
      end if
   end subroutine convec

   subroutine diffu()
      if (mixed) then
         call msod_check(msod_full_diffusion_ijk_mixed(int(porder, c_int), nelem, npoin, connec, He4, gpvol4, &
                         dlxigp_ip, invAtoIJK, AtoI, AtoJ, AtoK, Cp, Pra, rho4, u4, Tem4, muf4, mue4, mus4, &
                         Dmass, Dmom, Dener), 'full_diffusion_ijk_mixed')
      else
         call full_diffusion_ijk(nelem, npoin, connec, Ngp, dNgp, He, gpvol, dlxigp_ip, xgp, invAtoIJK, &
                                 AtoI, AtoJ, AtoK, Cp, Pra, rho, u, Tem, mu_fluid, mu_e, mu_sgs, &
                                 Dmass, Dmom, Dener)
      end if
   end subroutine diffu

   ! The reference driver's initial conditions (main.f90:326-361).
   subroutine reference_fields()
      integer :: iel, ip, d
      real(8) :: fmag, eint
      u = 1d0; rho = 1d0; pr = 1d0; E = 1d0; Tem = 1d0; mu_fluid = 1d0; mu_e = 1d0; mu_sgs = 1d0
      do iel = 1, nelem
         u(connec(iel,1),:) = 10d0
         u(connec(iel,4),:) = 20d0
         rho(connec(iel,1)) = 10d0
         rho(connec(iel,4)) = 20d0
         pr(connec(iel,1)) = 100d0
         pr(connec(iel,4)) = 200d0
      end do
      do ip = 1, npoin
         fmag = 0d0
         do d = 1, 3
            qm(ip,d) = rho(ip)*u(ip,d)
            fmag = fmag + qm(ip,d)*u(ip,d)
         end do
         eint = pr(ip)/((gamma - 1d0)*rho(ip))
         E(ip) = rho(ip)*eint + 0.5d0*fmag
         Tem(ip) = pr(ip)/(rho(ip)*Rgas)
      end do
   end subroutine reference_fields

   ! results_convec.dat / results_diffu.dat (main.f90:586-599), small meshes
   ! only; one record per node (explicit format: list-directed output may wrap)
   subroutine write_results()
      integer :: ip
      character(*), parameter :: fmt = '(i0,5(1x,es24.16e3))'
      if (npoin > 200000) return
      open(unit=10, file='results_convec.dat', status='replace')
      do ip = 1, npoin
         write(10,fmt) ip, Rmass(ip), Rmom(ip,1), Rmom(ip,2), Rmom(ip,3), Rener(ip)
      end do
      close(10)
      open(unit=11, file='results_diffu.dat', status='replace')
      do ip = 1, npoin
         write(11,fmt) ip, Dmass(ip), Dmom(ip,1), Dmom(ip,2), Dmom(ip,3), Dener(ip)
      end do
      close(11)
   end subroutine write_results

   subroutine parse_args()
      integer :: k, nargs
      character(len=64) :: a, v
      nargs = command_argument_count()
      k = 1
      do while (k <= nargs)
         call get_command_argument(k, a)
         select case (trim(a))
         case ('--mixed')
            mixed = .true.
         case ('--quiet')
            quiet = .true.
         case ('--gpu-geometry')
            gpu_geometry = .true.
         case ('--no-tet')
            tet = .false.
         case ('--n', '--nelem', '--porder', '--nruns', '--jitter', '--rho-jump', '--assembly')
            if (k == nargs) stop 'missing value'
            call get_command_argument(k+1, v)
            select case (trim(a))
            case ('--n'); read(v,*) n
            case ('--nelem'); read(v,*) nelem_ref
            case ('--porder'); read(v,*) porder
            case ('--nruns'); read(v,*) nruns
            case ('--jitter'); read(v,*) jitter
            case ('--rho-jump'); read(v,*) rho_jump
            case ('--assembly')
               if (trim(v) == 'atomic') then
                  assembly = 1
               else
                  assembly = 0
               end if
            end select
            k = k + 1
         case default
            write(*,*) 'unknown argument ', trim(a)
            stop 1
         end select
         k = k + 1
      end do
      if (porder < 1 .or. porder > 8) stop 'porder must be 1..8'
      if (gpu_geometry .and. mixed) stop '--gpu-geometry computes FP64 geometry: not with --mixed'
   end subroutine parse_args

end program miniSOD
