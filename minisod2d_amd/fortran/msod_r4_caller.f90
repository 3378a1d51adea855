! A caller built at the reference's default real kind (rp = 4,
! src/utils/sources/mod_constants.f90:5): every real array is real(4), and the
! calls of full_convec_ijk and full_diffusion_ijk are written exactly as in
! the reference driver (main.f90:447,477); the generic drop-in modules
! elem_convec and elem_diffu route them to the float entry points.
!
! Usage: msod_r4_caller <input file> <output file>
! Input (stream, little endian): int32 porder, nelem, npoin; real(4) Cp, Pra;
!   then, in Fortran order, connec(nelem,nnode) int32, invAtoIJK(q,q,q),
!   AtoI, AtoJ, AtoK(nnode) int32, dlxigp_ip(ngaus,3,q), xgp(ngaus,3),
!   Ngp(ngaus,nnode), dNgp(3,nnode,ngaus), He(3,3,ngaus,nelem),
!   gpvol(1,ngaus,nelem), u(npoin,3), q(npoin,3), rho, pr, E, Tem,
!   mu_fluid(npoin), mu_e, mu_sgs(nelem,ngaus), all real(4).
! Output (stream): Rmass, Rmom, Rener of the convection, then of the
!   diffusion, real(4).
program msod_r4_caller

   use iso_c_binding
   use msod_device, only: msod_start, msod_stop, msod_copyin, msod_create, msod_update_host, msod_delete
   use elem_convec, only: full_convec_ijk
   use elem_diffu, only: full_diffusion_ijk

   implicit none

   integer(4) :: porder, nelem, npoin, q, nnode, ngaus, lin, lout
   real(4) :: Cp, Pra
   integer(4), allocatable :: connec(:,:), invAtoIJK(:,:,:), AtoI(:), AtoJ(:), AtoK(:)
   real(4), allocatable :: dlxigp_ip(:,:,:), xgp(:,:), Ngp(:,:), dNgp(:,:,:), He(:,:,:,:), gpvol(:,:,:)
   real(4), allocatable :: u(:,:), qq(:,:), rho(:), pr(:), E(:), Tem(:), mu_fluid(:), mu_e(:,:), mu_sgs(:,:)
   real(4), allocatable :: Rmass(:), Rmom(:,:), Rener(:), Dmass(:), Dmom(:,:), Dener(:)
   character(len=4096) :: fin, fout

   if (command_argument_count() /= 2) then
      write(*,'(a)') 'usage: msod_r4_caller <input file> <output file>'
      error stop 1
   end if
   call get_command_argument(1, fin)
   call get_command_argument(2, fout)

   open(newunit=lin, file=trim(fin), access='stream', form='unformatted', status='old', action='read')
   read(lin) porder, nelem, npoin, Cp, Pra
   q = porder + 1
   nnode = q**3
   ngaus = nnode
   allocate(connec(nelem,nnode), invAtoIJK(q,q,q), AtoI(nnode), AtoJ(nnode), AtoK(nnode))
   allocate(dlxigp_ip(ngaus,3,q), xgp(ngaus,3), Ngp(ngaus,nnode), dNgp(3,nnode,ngaus))
   allocate(He(3,3,ngaus,nelem), gpvol(1,ngaus,nelem))
   allocate(u(npoin,3), qq(npoin,3), rho(npoin), pr(npoin), E(npoin), Tem(npoin), mu_fluid(npoin))
   allocate(mu_e(nelem,ngaus), mu_sgs(nelem,ngaus))
   allocate(Rmass(npoin), Rmom(npoin,3), Rener(npoin), Dmass(npoin), Dmom(npoin,3), Dener(npoin))
   read(lin) connec, invAtoIJK, AtoI, AtoJ, AtoK, dlxigp_ip, xgp, Ngp, dNgp, He, gpvol
   read(lin) u, qq, rho, pr, E, Tem, mu_fluid, mu_e, mu_sgs
   close(lin)

   ! the reference driver's data region (main.f90:197-425)
   call msod_start(0)
   call msod_copyin(connec, c_sizeof(connec))
   call msod_copyin(He, c_sizeof(He))
   call msod_copyin(gpvol, c_sizeof(gpvol))
   call msod_copyin(u, c_sizeof(u))
   call msod_copyin(qq, c_sizeof(qq))
   call msod_copyin(rho, c_sizeof(rho))
   call msod_copyin(pr, c_sizeof(pr))
   call msod_copyin(E, c_sizeof(E))
   call msod_copyin(Tem, c_sizeof(Tem))
   call msod_copyin(mu_fluid, c_sizeof(mu_fluid))
   call msod_copyin(mu_e, c_sizeof(mu_e))
   call msod_copyin(mu_sgs, c_sizeof(mu_sgs))
   call msod_create(Rmass, c_sizeof(Rmass))
   call msod_create(Rmom, c_sizeof(Rmom))
   call msod_create(Rener, c_sizeof(Rener))
   call msod_create(Dmass, c_sizeof(Dmass))
   call msod_create(Dmom, c_sizeof(Dmom))
   call msod_create(Dener, c_sizeof(Dener))

   call full_convec_ijk(nelem, npoin, connec, Ngp, dNgp, He, gpvol, dlxigp_ip, xgp, invAtoIJK, &
                        AtoI, AtoJ, AtoK, u, qq, rho, pr, E, Rmass, Rmom, Rener)
   call full_diffusion_ijk(nelem, npoin, connec, Ngp, dNgp, He, gpvol, dlxigp_ip, xgp, invAtoIJK, &
                           AtoI, AtoJ, AtoK, Cp, Pra, rho, u, Tem, mu_fluid, mu_e, mu_sgs, &
                           Dmass, Dmom, Dener)

   call msod_update_host(Rmass, c_sizeof(Rmass))
   call msod_update_host(Rmom, c_sizeof(Rmom))
   call msod_update_host(Rener, c_sizeof(Rener))
   call msod_update_host(Dmass, c_sizeof(Dmass))
   call msod_update_host(Dmom, c_sizeof(Dmom))
   call msod_update_host(Dener, c_sizeof(Dener))
   call msod_stop()

   open(newunit=lout, file=trim(fout), access='stream', form='unformatted', status='replace', action='write')
   write(lout) Rmass, Rmom, Rener, Dmass, Dmom, Dener
   close(lout)
   write(*,'(a,i0,a,i0,a,i0)') 'msod_r4_caller: porder=', porder, ' nelem=', nelem, ' npoin=', npoin

end program msod_r4_caller
