! Synthetic meshes for the residual kernels.
!
! msod_box_mesh: conforming structured mesh of the unit cube [0,1]^3 with
!   nx*ny*nz hexahedra of order p whose nodes sit on the (binary32-rounded)
!   GLL points, npoin = (nx*p+1)*(ny*p+1)*(nz*p+1).  Neighbouring elements
!   share their face/edge/corner nodes, so the nodal assembly of the kernels
!   really sums contributions.  Optionally, element-interior nodes (unshared,
!   (p-1)^3 per element) are moved by a seeded jitter of at most `jitter`
!   times the local 1-D GLL spacing on each axis, which makes He and gpvol
!   non-affine while keeping the mesh conforming.
! msod_disjoint_mesh: the reference's own mesh (geometry/sources/
!   mod_mesh.f90:11-65): every element is a private copy of [-1,1]^3 shifted
!   by 3 along the diagonal; npoin = nelem*nnode.
!
! Numbering: elements lexicographic (x fastest); global nodes lexicographic
! on the (nx*p+1)x(ny*p+1)x(nz*p+1) lattice.  connec holds 1-based node ids
! in the reference layout connec(nelem,nnode).
module msod_mesh

   use iso_c_binding, only: c_int, c_double, c_int64_t
   use msod_fem, only: msod_gll, msod_sod_order, msod_hex_tables
   implicit none
   private

   public :: msod_box_mesh, msod_disjoint_mesh, msod_box_sizes

   interface
      real(c_double) function msod_rng_uniform(seed, i) bind(C, name='msod_rng_uniform')
         import :: c_double, c_int64_t
         integer(c_int64_t), value :: seed, i
      end function msod_rng_uniform
   end interface

contains

   pure subroutine msod_box_sizes(nx, ny, nz, p, nelem, npoin)
      integer, intent(in) :: nx, ny, nz, p
      integer(8), intent(out) :: nelem, npoin
      nelem = int(nx,8)*int(ny,8)*int(nz,8)
      npoin = (int(nx,8)*p+1)*(int(ny,8)*p+1)*(int(nz,8)*p+1)
   end subroutine msod_box_sizes

   subroutine msod_box_mesh(nx, ny, nz, p, jitter, seed, connec, coord)
      integer, intent(in) :: nx, ny, nz, p
      real(8), intent(in) :: jitter
      integer(c_int64_t), intent(in) :: seed
      integer(c_int), intent(out) :: connec(int(nx,8)*ny*nz, (p+1)**3)
      real(c_double), intent(out) :: coord((int(nx,8)*p+1)*(int(ny,8)*p+1)*(int(nz,8)*p+1), 3)
      integer :: q, nn, a, ex, ey, ez, pi, pj, pk
      integer(8) :: nelem, npoin, e, n1x, n1y, n1z, gi, gj, gk, ip, gidx(3)
      integer, allocatable :: AtoI(:), AtoJ(:), AtoK(:)
      integer(c_int), allocatable :: inv(:,:,:), aijk(:)
      real(c_double), allocatable :: xgp(:,:), wgp(:), Ngp(:,:), dNgp(:,:,:), dl(:,:,:)
      real(8) :: xi(0:p), w(0:p), h(3), hloc(0:p)
      integer :: d, pos(3), nvec(3)

      q = p + 1
      nn = q**3
      call msod_box_sizes(nx, ny, nz, p, nelem, npoin)
      n1x = int(nx,8)*p + 1; n1y = int(ny,8)*p + 1; n1z = int(nz,8)*p + 1
      nvec = [nx, ny, nz]
      h = 1d0/dble(nvec)
      allocate(AtoI(nn), AtoJ(nn), AtoK(nn), inv(q,q,q), aijk(nn))
      allocate(xgp(nn,3), wgp(nn), Ngp(nn,nn), dNgp(3,nn,nn), dl(nn,3,q))
      call msod_hex_tables(p, inv, aijk, AtoI, AtoJ, AtoK, xgp, wgp, Ngp, dNgp, dl)
      call msod_gll(p, xi, w)
      ! smallest neighbour gap of each interior 1-D position
      hloc = 0d0
      do a = 1, p-1
         hloc(a) = min(xi(a+1) - xi(a), xi(a) - xi(a-1))
      end do

      !$omp parallel do schedule(static) private(ex, ey, ez, a, pi, pj, pk)
      do e = 1, nelem
         ex = int(mod(e-1, int(nx,8)))
         ey = int(mod((e-1)/nx, int(ny,8)))
         ez = int((e-1)/(int(nx,8)*ny))
         do a = 1, nn
            pi = msod_sod_order(p, AtoI(a))
            pj = msod_sod_order(p, AtoJ(a))
            pk = msod_sod_order(p, AtoK(a))
            connec(e, a) = int(1 + (int(ex,8)*p + pi) + n1x*((int(ey,8)*p + pj) + n1y*(int(ez,8)*p + pk)), c_int)
         end do
      end do
      !$omp end parallel do

      !$omp parallel do schedule(static) private(gi, gj, gk, gidx, d, pos)
      do ip = 1, npoin
         gi = mod(ip-1, n1x)
         gj = mod((ip-1)/n1x, n1y)
         gk = (ip-1)/(n1x*n1y)
         gidx = [gi, gj, gk]
         do d = 1, 3
            pos(d) = int(mod(gidx(d), int(p,8)))
            coord(ip, d) = h(d)*(dble(gidx(d)/p) + 0.5d0*(xi(pos(d)) + 1d0))
         end do
         if (jitter > 0d0 .and. all(pos /= 0)) then
            do d = 1, 3
               coord(ip, d) = coord(ip, d) + jitter*0.5d0*h(d)*hloc(pos(d)) * &
                    (2d0*msod_rng_uniform(seed, 3_c_int64_t*(ip-1) + (d-1)) - 1d0)
            end do
         end if
      end do
      !$omp end parallel do
   end subroutine msod_box_mesh

   ! The reference's disjoint mesh (mod_mesh.f90:26-65): element e owns nodes
   ! (e-1)*nnode+1 .. e*nnode, placed at xyzBase shifted by 3*(e-1).
   subroutine msod_disjoint_mesh(nelem, npoin, nnode, xyzBase, connec, xyz)
      integer, intent(in) :: nelem, npoin, nnode
      real(c_double), intent(in) :: xyzBase(nnode,3)
      integer(c_int), intent(out) :: connec(nelem,nnode)
      real(c_double), intent(out) :: xyz(npoin,3)
      integer :: e, a
      real(8) :: shift
      shift = 0d0
      do e = 1, nelem
         do a = 1, nnode
            connec(e, a) = (e-1)*nnode + a
            xyz(connec(e,a), :) = xyzBase(a, :) + shift
         end do
         shift = shift + 3d0
      end do
   end subroutine msod_disjoint_mesh

end module msod_mesh
