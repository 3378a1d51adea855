! C-callable entry points of the host-side Fortran library (libmsod_host.so):
! tables, synthetic meshes, geometry and fields, for hosts that are not
! Fortran programs (the Python mirror in minisod2d_amd/ uses them through
! ctypes).  All arrays use the reference's Fortran layouts.
module msod_capi

   use iso_c_binding
   use omp_lib, only: omp_set_num_threads, omp_get_max_threads
   use msod_fem, only: msod_gll, msod_hex_tables, msod_gmsh_hex, msod_tet_tables
   use msod_mesh, only: msod_box_mesh, msod_disjoint_mesh, msod_box_sizes
   use msod_geom, only: msod_elem_jacobian
   use msod_fields, only: msod_flow_fields, msod_viscosity_fields
   use msod_interp, only: msod_interp_1d, msod_prolong_fields, msod_restrict_residual
   implicit none

contains

   subroutine msodh_set_threads(n) bind(C, name='msodh_set_threads')
      integer(c_int), value :: n
      if (n > 0) call omp_set_num_threads(n)
   end subroutine msodh_set_threads

   integer(c_int) function msodh_get_threads() bind(C, name='msodh_get_threads')
      msodh_get_threads = omp_get_max_threads()
   end function msodh_get_threads

   subroutine msodh_gll(p, xi, w) bind(C, name='msodh_gll')
      integer(c_int), value :: p
      real(c_double), intent(out) :: xi(p+1), w(p+1)
      call msod_gll(p, xi, w)
   end subroutine msodh_gll

   ! Gmsh (i,j,k) positions 0..p of the local nodes, t(3,nnode).
   subroutine msodh_gmsh_hex(p, t) bind(C, name='msodh_gmsh_hex')
      integer(c_int), value :: p
      integer(c_int), intent(out) :: t(3,(p+1)**3)
      integer, allocatable :: tt(:,:)
      allocate(tt(3,(p+1)**3))
      call msod_gmsh_hex(p, tt)
      t = tt
   end subroutine msodh_gmsh_hex

   subroutine msodh_hex_tables(p, invAtoIJK, AtoIJK, AtoI, AtoJ, AtoK, xgp, wgp, Ngp, dNgp, dlxigp_ip) &
         bind(C, name='msodh_hex_tables')
      integer(c_int), value :: p
      integer(c_int), intent(out) :: invAtoIJK(p+1,p+1,p+1), AtoIJK((p+1)**3)
      integer(c_int), intent(out) :: AtoI((p+1)**3), AtoJ((p+1)**3), AtoK((p+1)**3)
      real(c_double), intent(out) :: xgp((p+1)**3,3), wgp((p+1)**3)
      real(c_double), intent(out) :: Ngp((p+1)**3,(p+1)**3), dNgp(3,(p+1)**3,(p+1)**3)
      real(c_double), intent(out) :: dlxigp_ip((p+1)**3,3,p+1)
      call msod_hex_tables(p, invAtoIJK, AtoIJK, AtoI, AtoJ, AtoK, xgp, wgp, Ngp, dNgp, dlxigp_ip)
   end subroutine msodh_hex_tables

   ! TET04 tables: xgp(4,3), wgp(4), Ngp(4,4), dNgp(3,4,4).
   subroutine msodh_tet_tables(xgp, wgp, Ngp, dNgp) bind(C, name='msodh_tet_tables')
      real(c_double), intent(out) :: xgp(4,3), wgp(4), Ngp(4,4), dNgp(3,4,4)
      call msod_tet_tables(xgp, wgp, Ngp, dNgp)
   end subroutine msodh_tet_tables

   subroutine msodh_box_sizes(nx, ny, nz, p, nelem, npoin) bind(C, name='msodh_box_sizes')
      integer(c_int), value :: nx, ny, nz, p
      integer(c_int64_t), intent(out) :: nelem, npoin
      call msod_box_sizes(nx, ny, nz, p, nelem, npoin)
   end subroutine msodh_box_sizes

   subroutine msodh_box_mesh(nx, ny, nz, p, jitter, seed, connec, coord) bind(C, name='msodh_box_mesh')
      integer(c_int), value :: nx, ny, nz, p
      real(c_double), value :: jitter
      integer(c_int64_t), value :: seed
      integer(c_int), intent(out) :: connec(*)
      real(c_double), intent(out) :: coord(*)
      call msod_box_mesh(nx, ny, nz, p, jitter, seed, connec, coord)
   end subroutine msodh_box_mesh

   subroutine msodh_disjoint_mesh(nelem, npoin, nnode, xyzBase, connec, xyz) bind(C, name='msodh_disjoint_mesh')
      integer(c_int), value :: nelem, npoin, nnode
      real(c_double), intent(in) :: xyzBase(nnode,3)
      integer(c_int), intent(out) :: connec(nelem,nnode)
      real(c_double), intent(out) :: xyz(npoin,3)
      call msod_disjoint_mesh(nelem, npoin, nnode, xyzBase, connec, xyz)
   end subroutine msodh_disjoint_mesh

   ! Returns the number of points with a non-positive Jacobian determinant.
   integer(c_int) function msodh_elem_jacobian(nelem, npoin, nnode, connec, coord, dNgp, wgp, gpvol, He) &
         bind(C, name='msodh_elem_jacobian')
      integer(c_int), value :: nelem, npoin, nnode
      integer(c_int), intent(in) :: connec(nelem,nnode)
      real(c_double), intent(in) :: coord(npoin,3), dNgp(3,nnode,nnode), wgp(nnode)
      real(c_double), intent(out) :: gpvol(1,nnode,nelem), He(3,3,nnode,nelem)
      integer :: nbad
      call msod_elem_jacobian(nelem, npoin, nnode, connec, coord, dNgp, wgp, gpvol, He, nbad)
      msodh_elem_jacobian = nbad
   end function msodh_elem_jacobian

   subroutine msodh_flow_fields(npoin, coord, gamma, Rgas, seed_rho, rho_jump, u, q, rho, pr, E, Tem, mu_fluid) &
         bind(C, name='msodh_flow_fields')
      integer(c_int64_t), value :: npoin, seed_rho
      real(c_double), value :: gamma, Rgas, rho_jump
      real(c_double), intent(in) :: coord(npoin,3)
      real(c_double), intent(out) :: u(npoin,3), q(npoin,3), rho(npoin), pr(npoin), E(npoin)
      real(c_double), intent(out) :: Tem(npoin), mu_fluid(npoin)
      call msod_flow_fields(npoin, coord, gamma, Rgas, seed_rho, rho_jump, u, q, rho, pr, E, Tem, mu_fluid)
   end subroutine msodh_flow_fields

   subroutine msodh_viscosity_fields(nelem, ngaus, seed_mue, seed_musgs, mue_max, musgs_max, mu_e, mu_sgs) &
         bind(C, name='msodh_viscosity_fields')
      integer(c_int64_t), value :: nelem, seed_mue, seed_musgs
      integer(c_int), value :: ngaus
      real(c_double), value :: mue_max, musgs_max
      real(c_double), intent(out) :: mu_e(nelem,ngaus), mu_sgs(nelem,ngaus)
      call msod_viscosity_fields(nelem, ngaus, seed_mue, seed_musgs, mue_max, musgs_max, mu_e, mu_sgs)
   end subroutine msodh_viscosity_fields

   subroutine msodh_interp_1d(pf, pt, M) bind(C, name='msodh_interp_1d')
      integer(c_int), value :: pf, pt
      real(c_double), intent(out) :: M(pt+1, pf+1)
      real(8) :: Mt(pt+1, pf+1)
      call msod_interp_1d(pf, pt, Mt)
      M = Mt
   end subroutine msodh_interp_1d

   subroutine msodh_prolong_fields(pf, pt, nelem, npf, npt, ncomp, connec_f, connec_t, ff, ft) &
         bind(C, name='msodh_prolong_fields')
      integer(c_int), value :: pf, pt, nelem, npf, npt, ncomp
      integer(c_int), intent(in) :: connec_f(nelem,(pf+1)**3), connec_t(nelem,(pt+1)**3)
      real(c_double), intent(in) :: ff(npf, ncomp)
      real(c_double), intent(out) :: ft(npt, ncomp)
      call msod_prolong_fields(pf, pt, nelem, npf, npt, ncomp, connec_f, connec_t, ff, ft)
   end subroutine msodh_prolong_fields

   subroutine msodh_restrict_residual(pf, pt, nelem, npf, npt, ncomp, connec_f, connec_t, rf, rt) &
         bind(C, name='msodh_restrict_residual')
      integer(c_int), value :: pf, pt, nelem, npf, npt, ncomp
      integer(c_int), intent(in) :: connec_f(nelem,(pf+1)**3), connec_t(nelem,(pt+1)**3)
      real(c_double), intent(in) :: rf(npf, ncomp)
      real(c_double), intent(out) :: rt(npt, ncomp)
      call msod_restrict_residual(pf, pt, nelem, npf, npt, ncomp, connec_f, connec_t, rf, rt)
   end subroutine msodh_restrict_residual

end module msod_capi
