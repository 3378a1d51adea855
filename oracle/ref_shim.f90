! bind(C) shim over the miniSOD2D reference modules.
!
! TEST INFRASTRUCTURE ONLY.  This file is compiled together with the
! reference's own Fortran sources (taken where they lie, under
! /root/reference/src, by oracle/Makefile) into oracle/_ref/libsodref_pP.so.
! It adds no arithmetic of its own except the gmsh->SOD index remap of the
! reference driver (main.f90:131-212), which lives in main.f90 -- a file that
! cannot be compiled here because it needs MPI -- and is restated below.
!
! The real arrays of the kernel and geometry wrappers have the reference's
! own kind real(rp): double in the rp=8 libraries (oracle/_ref/libsodref_pP.so),
! float in the rp=4 ones (oracle/_ref/r4/libsodref_r4_pP.so, the reference's
! default kind, mod_constants.f90:5).  The tables are returned in double
! (exact: they are computed in real(rp)).
!
! Every entry point is a thin C-callable wrapper:
!   sodref_porder / sodref_rp        compile-time porder and real kind
!   sodref_tables                    main.f90:110-255 (tables + index maps)
!   sodref_create_mesh               mod_mesh.f90:11-65 (disjoint mesh)
!   sodref_elem_jacobian             jacobian_oper.f90:7-136
!   sodref_full_convec_ijk           elem_convec.f90:198-431
!   sodref_full_convec_ijk_split     elem_convec.f90:433-836
!   sodref_full_diffusion_ijk        elem_diffu.f90:10-197
!   sodref_tet_tables                main.f90:374-383 (tet_4points, tet_04)
!   sodref_gen_tet_mesh              mod_mesh.f90:71-125
!   sodref_tet_jacobian              jacobian_oper.f90:138-267
!   sodref_fem_convec                elem_convec.f90:839-948
!   sodref_fem_diffu                 elem_diffu.f90:199-317
module sodref_shim

   use iso_c_binding
   use mod_constants
   use mod_numerical_params, only: c_rho, c_ener
   use elem_qua, only: init_basic_qua, quad_order_edges
   use elem_hex, only: init_basic_hex, hex_highorder, hex_order_edges
   use mod_gmsh_indices, only: genHighOrderHex
   use elem_tet, only: tet_04
   use quadrature_rules, only: GaussLobattoLegendre_hex, tet_4points
   use mod_mesh, only: create_mesh, gen_tet_mesh
   use jacobian_oper, only: elem_jacobian, tet_jacobian
   use elem_convec, only: full_convec_ijk, full_convec_ijk_split, fem_convec
   use elem_diffu, only: full_diffusion_ijk, fem_diffu

   implicit none

contains

   function sodref_porder() bind(C, name='sodref_porder') result(p)
      integer(c_int) :: p
      p = porder
   end function sodref_porder

   function sodref_rp() bind(C, name='sodref_rp') result(k)
      integer(c_int) :: k
      k = rp
   end function sodref_rp

   ! Tables of the reference driver: main.f90:110-255.
   ! Outputs (Fortran layout):
   !   gmshIJK(nnode,3) [SOD 1-D indices after the remap], invAtoIJK(q,q,q),
   !   AtoIJK(nnode), AtoI/J/K(nnode), xgp(ngaus,3), wgp(ngaus),
   !   Ngp(ngaus,nnode), dNgp(3,nnode,ngaus), dlxigp_ip(ngaus,3,q)
   subroutine sodref_tables(gmshIJK, invAtoIJK, AtoIJK, AtoI, AtoJ, AtoK, xgp, wgp, Ngp, dNgp, dlxigp_ip) &
         bind(C, name='sodref_tables')
      integer(c_int), intent(out) :: gmshIJK(nnode,3), invAtoIJK(porder+1,porder+1,porder+1)
      integer(c_int), intent(out) :: AtoIJK(nnode), AtoI(nnode), AtoJ(nnode), AtoK(nnode)
      real(c_double), intent(out) :: xgp(ngaus,ndime), wgp(ngaus), Ngp(ngaus,nnode), dNgp(ndime,nnode,ngaus)
      real(c_double), intent(out) :: dlxigp_ip(ngaus,ndime,porder+1)
      integer(4) :: gmsh2sod(porder+1,2), inode, i, j, k, iorder, jnode, igaus
      integer(4) :: tab(nnode,3)
      real(rp), allocatable :: xgp_r(:,:), wgp_r(:), Ngp_r(:,:), dNgp_r(:,:,:)
      real(rp), allocatable :: Ngp_l(:,:), dNgp_l(:,:,:), dlx_r(:,:,:)

      if (.not. allocated(quad_order_edges)) call init_basic_qua()
      if (.not. allocated(hex_order_edges)) call init_basic_hex()

      call genHighOrderHex(porder, tab)

      ! gmsh 1-D position -> SOD 1-D index: 0 -> 1, p -> 2, m -> m+1 (restates main.f90:133-137)
      gmsh2sod(1,:) = [0,1]
      gmsh2sod(2,:) = [porder,2]
      do inode = 3,porder+1
         gmsh2sod(inode,:) = [inode-2,inode]
      end do
      ! remap (restates main.f90:148-157)
      do inode = 1,nnode
         do i = 1,3
            do iorder = 1,porder+1
               if (tab(inode,i) == gmsh2sod(iorder,1)) then
                  tab(inode,i) = gmsh2sod(iorder,2)
                  exit
               end if
            end do
         end do
      end do
      ! invAtoIJK / AtoIJK with the driver's (k,i,j) loop order (restates main.f90:168-185)
      jnode = 1
      do k = 1,porder+1
         do i = 1,porder+1
            do j = 1,porder+1
               do inode = 1,nnode
                  if (tab(inode,1) == i .and. tab(inode,2) == j .and. tab(inode,3) == k) then
                     invAtoIJK(i,j,k) = inode
                     exit
                  end if
               end do
               AtoIJK(jnode) = invAtoIJK(i,j,k)
               jnode = jnode + 1
            end do
         end do
      end do
      do inode = 1,nnode
         AtoI(inode) = tab(inode,1)
         AtoJ(inode) = tab(inode,2)
         AtoK(inode) = tab(inode,3)
      end do
      gmshIJK = tab

      allocate(xgp_r(ngaus,ndime), wgp_r(ngaus), Ngp_r(ngaus,nnode), dNgp_r(ndime,nnode,ngaus))
      allocate(Ngp_l(ngaus,nnode), dNgp_l(ndime,nnode,ngaus), dlx_r(ngaus,ndime,porder+1))
      call GaussLobattoLegendre_hex(porder, ngaus, AtoIJK, xgp_r, wgp_r)
      do igaus = 1,ngaus
         call hex_highorder(porder, nnode, xgp_r(igaus,1), xgp_r(igaus,2), xgp_r(igaus,3), &
                            AtoIJK, Ngp_r(igaus,:), dNgp_r(:,:,igaus), &
                            Ngp_l(igaus,:), dNgp_l(:,:,igaus), dlx_r(igaus,:,:))
      end do
      xgp = real(xgp_r, c_double)
      wgp = real(wgp_r, c_double)
      Ngp = real(Ngp_r, c_double)
      dNgp = real(dNgp_r, c_double)
      dlxigp_ip = real(dlx_r, c_double)
   end subroutine sodref_tables

   ! The reference's own disjoint mesh: mod_mesh.f90:11-65.
   subroutine sodref_create_mesh(nelem, npoin, xyzBase, connec, xyz) bind(C, name='sodref_create_mesh')
      integer(c_int), value :: nelem, npoin
      real(rp), intent(in) :: xyzBase(nnode,ndime)
      integer(c_int), intent(out) :: connec(nelem,nnode)
      real(rp), intent(out) :: xyz(npoin,ndime)
      call create_mesh(nelem, npoin, xyzBase, connec, xyz)
   end subroutine sodref_create_mesh

   subroutine sodref_elem_jacobian(nelem, npoin, connec, coord, dNgp, wgp, gpvol, He) &
         bind(C, name='sodref_elem_jacobian')
      integer(c_int), value :: nelem, npoin
      integer(c_int), intent(in) :: connec(nelem,nnode)
      real(rp), intent(in) :: coord(npoin,ndime), dNgp(ndime,nnode,ngaus), wgp(ngaus)
      real(rp), intent(out) :: gpvol(1,ngaus,nelem), He(ndime,ndime,ngaus,nelem)
      call elem_jacobian(nelem, npoin, connec, coord, dNgp, wgp, gpvol, He)
   end subroutine sodref_elem_jacobian

   subroutine sodref_full_convec_ijk(nelem, npoin, connec, Ngp, dNgp, He, gpvol, dlxigp_ip, xgp, &
         invAtoIJK, AtoI, AtoJ, AtoK, u, q, rho, pr, E, Rmass, Rmom, Rener) &
         bind(C, name='sodref_full_convec_ijk')
      integer(c_int), value :: nelem, npoin
      integer(c_int), intent(in) :: connec(nelem,nnode)
      real(rp), intent(in) :: Ngp(ngaus,nnode), dNgp(ndime,nnode,ngaus)
      real(rp), intent(in) :: He(ndime,ndime,ngaus,nelem), xgp(ngaus,ndime), dlxigp_ip(ngaus,ndime,porder+1)
      real(rp), intent(in) :: gpvol(1,ngaus,nelem)
      integer(c_int), intent(in) :: invAtoIJK(porder+1,porder+1,porder+1), AtoI(nnode), AtoJ(nnode), AtoK(nnode)
      real(rp), intent(in) :: q(npoin,ndime), u(npoin,ndime), rho(npoin), pr(npoin), E(npoin)
      real(rp), intent(out) :: Rmass(npoin), Rmom(npoin,ndime), Rener(npoin)
      call full_convec_ijk(nelem, npoin, connec, Ngp, dNgp, He, gpvol, dlxigp_ip, xgp, invAtoIJK, &
                           AtoI, AtoJ, AtoK, u, q, rho, pr, E, Rmass, Rmom, Rener)
   end subroutine sodref_full_convec_ijk

   subroutine sodref_full_convec_ijk_split(nelem, npoin, connec, Ngp, dNgp, He, gpvol, dlxigp_ip, xgp, &
         invAtoIJK, AtoI, AtoJ, AtoK, u, q, rho, pr, E, Rmass, Rmom, Rener) &
         bind(C, name='sodref_full_convec_ijk_split')
      integer(c_int), value :: nelem, npoin
      integer(c_int), intent(in) :: connec(nelem,nnode)
      real(rp), intent(in) :: Ngp(ngaus,nnode), dNgp(ndime,nnode,ngaus)
      real(rp), intent(in) :: He(ndime,ndime,ngaus,nelem), xgp(ngaus,ndime), dlxigp_ip(ngaus,ndime,porder+1)
      real(rp), intent(in) :: gpvol(1,ngaus,nelem)
      integer(c_int), intent(in) :: invAtoIJK(porder+1,porder+1,porder+1), AtoI(nnode), AtoJ(nnode), AtoK(nnode)
      real(rp), intent(in) :: q(npoin,ndime), u(npoin,ndime), rho(npoin), pr(npoin), E(npoin)
      real(rp), intent(out) :: Rmass(npoin), Rmom(npoin,ndime), Rener(npoin)
      call full_convec_ijk_split(nelem, npoin, connec, Ngp, dNgp, He, gpvol, dlxigp_ip, xgp, invAtoIJK, &
                                 AtoI, AtoJ, AtoK, u, q, rho, pr, E, Rmass, Rmom, Rener)
   end subroutine sodref_full_convec_ijk_split

   ! c_rho and c_ener are module globals in the reference (mod_numerical_params.f90:26-27);
   ! they are set from the arguments before the call.
   subroutine sodref_full_diffusion_ijk(nelem, npoin, connec, Ngp, dNgp, He, gpvol, dlxigp_ip, xgp, &
         invAtoIJK, AtoI, AtoJ, AtoK, Cp, Pr, c_rho_in, c_ener_in, rho, u, Tem, mu_fluid, mu_e, mu_sgs, &
         Rmass, Rmom, Rener) bind(C, name='sodref_full_diffusion_ijk')
      integer(c_int), value :: nelem, npoin
      integer(c_int), intent(in) :: connec(nelem,nnode)
      real(rp), intent(in) :: Ngp(ngaus,nnode), dNgp(ndime,nnode,ngaus)
      real(rp), intent(in) :: He(ndime,ndime,ngaus,nelem), xgp(ngaus,ndime), dlxigp_ip(ngaus,ndime,porder+1)
      real(rp), intent(in) :: gpvol(1,ngaus,nelem)
      integer(c_int), intent(in) :: invAtoIJK(porder+1,porder+1,porder+1), AtoI(nnode), AtoJ(nnode), AtoK(nnode)
      real(c_double), value :: Cp, Pr, c_rho_in, c_ener_in
      real(rp), intent(in) :: rho(npoin), u(npoin,ndime), Tem(npoin), mu_fluid(npoin)
      real(rp), intent(in) :: mu_e(nelem,ngaus), mu_sgs(nelem,ngaus)
      real(rp), intent(out) :: Rmass(npoin), Rmom(npoin,ndime), Rener(npoin)
      real(rp) :: Cp_r, Pr_r
      c_rho = c_rho_in
      c_ener = c_ener_in
      Cp_r = Cp
      Pr_r = Pr
      call full_diffusion_ijk(nelem, npoin, connec, Ngp, dNgp, He, gpvol, dlxigp_ip, xgp, invAtoIJK, &
                              AtoI, AtoJ, AtoK, Cp_r, Pr_r, rho, u, Tem, mu_fluid, mu_e, mu_sgs, &
                              Rmass, Rmom, Rener)
   end subroutine sodref_full_diffusion_ijk

   ! TET04 tables as the driver builds them (main.f90:374-383).
   subroutine sodref_tet_tables(xgp, wgp, Ngp, dNgp) bind(C, name='sodref_tet_tables')
      real(c_double), intent(out) :: xgp(4,ndime), wgp(4), Ngp(4,4), dNgp(ndime,4,4)
      real(rp) :: x(4,ndime), w(4), N(4,4), dN(ndime,4,4)
      integer(4) :: igaus
      call tet_4points(x, w)
      do igaus = 1,4
         call tet_04(x(igaus,1), x(igaus,2), x(igaus,3), N(igaus,:), dN(:,:,igaus))
      end do
      xgp = real(x, c_double)
      wgp = real(w, c_double)
      Ngp = real(N, c_double)
      dNgp = real(dN, c_double)
   end subroutine sodref_tet_tables

   subroutine sodref_gen_tet_mesh(nelem, npoin, xyzBase, connec, xyz) bind(C, name='sodref_gen_tet_mesh')
      integer(c_int), value :: nelem, npoin
      real(rp), intent(in) :: xyzBase(4,ndime)
      integer(c_int), intent(out) :: connec(nelem,4)
      real(rp), intent(out) :: xyz(npoin,ndime)
      call gen_tet_mesh(nelem, npoin, xyzBase, connec, xyz)
   end subroutine sodref_gen_tet_mesh

   subroutine sodref_tet_jacobian(nelem, npoin, connec, coord, dNgp, wgp, gpvol, He) &
         bind(C, name='sodref_tet_jacobian')
      integer(c_int), value :: nelem, npoin
      integer(c_int), intent(in) :: connec(nelem,4)
      real(rp), intent(in) :: coord(npoin,ndime), dNgp(ndime,4,4), wgp(4)
      real(rp), intent(out) :: gpvol(1,4,nelem), He(ndime,ndime,4,nelem)
      call tet_jacobian(nelem, npoin, connec, coord, dNgp, wgp, gpvol, He)
   end subroutine sodref_tet_jacobian

   subroutine sodref_fem_convec(nelem, npoin, connec, Ngp, dNgp, He, gpvol, u, q, rho, pr, E, &
         Rmass, Rmom, Rener) bind(C, name='sodref_fem_convec')
      integer(c_int), value :: nelem, npoin
      integer(c_int), intent(in) :: connec(nelem,4)
      real(rp), intent(in) :: Ngp(4,4), dNgp(ndime,4,4), He(ndime,ndime,4,nelem), gpvol(1,4,nelem)
      real(rp), intent(in) :: q(npoin,ndime), u(npoin,ndime), rho(npoin), pr(npoin), E(npoin)
      real(rp), intent(out) :: Rmass(npoin), Rmom(npoin,ndime), Rener(npoin)
      call fem_convec(nelem, npoin, connec, Ngp, dNgp, He, gpvol, u, q, rho, pr, E, Rmass, Rmom, Rener)
   end subroutine sodref_fem_convec

   subroutine sodref_fem_diffu(nelem, npoin, connec, Ngp, dNgp, He, gpvol, Cp, Pr, c_rho_in, c_ener_in, &
         rho, u, Tem, mu_fluid, mu_e, mu_sgs, Rmass, Rmom, Rener) bind(C, name='sodref_fem_diffu')
      integer(c_int), value :: nelem, npoin
      integer(c_int), intent(in) :: connec(nelem,4)
      real(rp), intent(in) :: Ngp(4,4), dNgp(ndime,4,4), He(ndime,ndime,4,nelem), gpvol(1,4,nelem)
      real(c_double), value :: Cp, Pr, c_rho_in, c_ener_in
      real(rp), intent(in) :: rho(npoin), u(npoin,ndime), Tem(npoin), mu_fluid(npoin)
      real(rp), intent(in) :: mu_e(nelem,4), mu_sgs(nelem,4)
      real(rp), intent(out) :: Rmass(npoin), Rmom(npoin,ndime), Rener(npoin)
      real(rp) :: Cp_r, Pr_r
      c_rho = c_rho_in
      c_ener = c_ener_in
      Cp_r = Cp
      Pr_r = Pr
      call fem_diffu(nelem, npoin, connec, Ngp, dNgp, He, gpvol, Cp_r, Pr_r, rho, u, Tem, mu_fluid, &
                     mu_e, mu_sgs, Rmass, Rmom, Rener)
   end subroutine sodref_fem_diffu

end module sodref_shim
