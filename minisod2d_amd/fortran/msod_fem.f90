! High-order hexahedron tables for the spectral-element kernels.
!
! Replaces the table set the reference driver builds before its timing loop
! (app/sources/main.f90:110-255 with utils/sources/mod_maths.f90:18-297,
! fem/sources/quadrature_rules.f90:50-82, fem/sources/elem_hex.f90:124-146,
! fem/sources/mod_gmsh_indices.f90:11-108).  porder is a run-time argument
! here instead of a compile-time parameter.
!
! Conventions kept from the reference because the kernels' input format
! depends on them:
!  * 1-D "SOD" node order: index 1 <-> xi=-1, 2 <-> xi=+1, 3..p+1 <-> the
!    interior GLL roots in ascending order.
!  * Element-local node numbering a = 1..(p+1)^3 is Gmsh's recursive
!    high-order hexahedron order (corners, edges, faces, interior).
!  * GLL roots and weights are computed in double precision and then rounded
!    through binary32, exactly as the reference does with a kind-less real()
!    (mod_maths.f90:25,39-40); every derived table inherits those values.
module msod_fem

   use iso_c_binding, only: c_int, c_double
   implicit none
   private

   public :: msod_gll, msod_sod_order, msod_lagrange, msod_lagrange_deriv
   public :: msod_gmsh_hex, msod_hex_tables, msod_deriv_matrix
   public :: msod_tet_tables

   ! Corner pairs of the 12 hex edges and corner quadruples of the 6 faces,
   ! in Gmsh numbering (the same tables as elem_hex.f90:119-120).
   integer, parameter :: hex_edge(2,12) = reshape([1,2, 1,4, 1,5, 2,3, 2,6, 3,4, &
                                                   3,7, 4,8, 5,6, 5,8, 6,7, 7,8], [2,12])
   integer, parameter :: hex_face(4,6) = reshape([1,4,3,2, 1,2,6,5, 1,5,8,4, &
                                                  2,3,7,6, 3,4,8,7, 5,6,7,8], [4,6])
   integer, parameter :: quad_edge(2,4) = reshape([1,2, 2,3, 3,4, 4,1], [2,4])

contains

   ! Legendre P_n and its first two derivatives at x (three-term recurrence).
   pure subroutine legendre(n, x, y, dy, d2y)
      integer, intent(in) :: n
      real(8), intent(in) :: x
      real(8), intent(out) :: y, dy, d2y
      real(8) :: y0, dy0, d2y0, y1, dy1, d2y1, a, b
      integer :: k
      y = 1d0; dy = 0d0; d2y = 0d0
      if (n == 0) return
      y = x; dy = 1d0; d2y = 0d0
      if (n == 1) return
      y0 = 1d0; dy0 = 0d0; d2y0 = 0d0
      do k = 2, n
         a = 2d0*dble(k) - 1d0
         b = dble(k) - 1d0
         y1 = y; dy1 = dy; d2y1 = d2y
         y = (a*x*y1 - b*y0)/dble(k)
         dy = (a*x*dy1 - b*dy0 + a*y1)/dble(k)
         d2y = (a*x*d2y1 - b*d2y0 + 2d0*a*dy1)/dble(k)
         y0 = y1; dy0 = dy1; d2y0 = d2y1
      end do
   end subroutine legendre

   ! GLL roots (ascending) and weights of order p, rounded through binary32.
   ! Interior roots: zeros of P'_p by 8 Newton steps from the Chebyshev-Gauss-
   ! Lobatto guesses cos(pi*i/p), the same fixed iteration count as the
   ! reference, so the rounded values coincide.  Weights c/P_p(x)^2 with
   ! c = 2/(p(p+1)).
   subroutine msod_gll(p, xi, w)
      integer, intent(in) :: p
      real(8), intent(out) :: xi(0:p), w(0:p)
      real(8) :: x, y, dy, d2y, pn(0:p), pi8, c
      integer :: i, it, half
      pi8 = 2d0*asin(1d0)
      xi(0) = -1d0; xi(p) = 1d0
      call legendre(p, -1d0, pn(0), dy, d2y)
      call legendre(p, 1d0, pn(p), dy, d2y)
      half = (p-1)/2
      if (p >= 2) then
         xi(half+1) = 0d0
         call legendre(p, 0d0, pn(half+1), dy, d2y)
      end if
      do i = 1, half
         x = cos(pi8/dble(p)*dble(i))
         do it = 1, 8
            call legendre(p, x, y, dy, d2y)
            x = x - dy/d2y
         end do
         xi(i) = -x; xi(p-i) = x
         pn(p-i) = y
         pn(i) = y*merge(1d0, -1d0, mod(p,2) == 0)
      end do
      c = 2d0/(dble(p)*(dble(p) + 1d0))
      do i = 0, p
         w(i) = c/(pn(i)*pn(i))
      end do
      do i = 0, p
         xi(i) = dble(real(xi(i), 4))
         w(i) = dble(real(w(i), 4))
      end do
   end subroutine msod_gll

   ! SOD 1-D index (1-based) -> 0-based position in the ascending grid.
   pure integer function msod_sod_order(p, isod) result(pos)
      integer, intent(in) :: p, isod
      if (isod == 1) then
         pos = 0
      else if (isod == 2) then
         pos = p
      else
         pos = isod - 2
      end if
   end function msod_sod_order

   ! Lagrange basis l_n(x), n in SOD order, on the ascending grid xi(0:p).
   pure subroutine msod_lagrange(p, xi, x, l)
      integer, intent(in) :: p
      real(8), intent(in) :: xi(0:p), x
      real(8), intent(out) :: l(p+1)
      integer :: n, j, c
      do n = 1, p+1
         c = msod_sod_order(p, n)
         l(n) = 1d0
         do j = 0, p
            if (j /= c) l(n) = l(n)*((x - xi(j))/(xi(c) - xi(j)))
         end do
      end do
   end subroutine msod_lagrange

   ! Derivative l_n'(x) = sum_j 1/(x_c-x_j) prod_{m/=j,c} (x-x_m)/(x_c-x_m).
   pure subroutine msod_lagrange_deriv(p, xi, x, dl)
      integer, intent(in) :: p
      real(8), intent(in) :: xi(0:p), x
      real(8), intent(out) :: dl(p+1)
      integer :: n, j, m, c
      real(8) :: t
      do n = 1, p+1
         c = msod_sod_order(p, n)
         dl(n) = 0d0
         do j = 0, p
            if (j == c) cycle
            t = 1d0
            do m = 0, p
               if (m /= j .and. m /= c) t = t*(x - xi(m))/(xi(c) - xi(m))
            end do
            dl(n) = dl(n) + (1d0/(xi(c) - xi(j)))*t
         end do
      end do
   end subroutine msod_lagrange_deriv

   ! 1-D derivative matrix in SOD order: D(i,n) = l_n'(x_i).
   subroutine msod_deriv_matrix(p, D)
      integer, intent(in) :: p
      real(8), intent(out) :: D(p+1,p+1)
      real(8) :: xi(0:p), w(0:p)
      integer :: i
      call msod_gll(p, xi, w)
      do i = 1, p+1
         call msod_lagrange_deriv(p, xi, xi(msod_sod_order(p, i)), D(i,:))
      end do
   end subroutine msod_deriv_matrix

   ! Gmsh high-order quadrilateral (i,j) positions, 0..p, nodes in Gmsh order.
   recursive subroutine gmsh_quad(p, t)
      integer, intent(in) :: p
      integer, intent(out) :: t(2,(p+1)**2)
      integer :: nxt, e, s, a0, a1, dir(2)
      integer, allocatable :: sub(:,:)
      t(:,1) = 0
      if (p < 1) return
      t(:,2) = [p,0]; t(:,3) = [p,p]; t(:,4) = [0,p]
      if (p < 2) return
      nxt = 5
      do e = 1, 4
         a0 = quad_edge(1,e); a1 = quad_edge(2,e)
         dir = (t(:,a1) - t(:,a0))/p
         do s = 1, p-1
            t(:,nxt) = t(:,a0) + s*dir
            nxt = nxt + 1
         end do
      end do
      allocate(sub(2,(p-1)**2))
      call gmsh_quad(p-2, sub)
      t(:,nxt:) = sub + 1
   end subroutine gmsh_quad

   ! Gmsh high-order hexahedron (i,j,k) positions, 0..p, in Gmsh order.
   recursive subroutine msod_gmsh_hex(p, t)
      integer, intent(in) :: p
      integer, intent(out) :: t(3,(p+1)**3)
      integer :: nxt, e, f, s, a0, a1, a3, du(3), dv(3)
      integer, allocatable :: fq(:,:), sub(:,:)
      t(:,1) = 0
      if (p < 1) return
      t(:,2) = [p,0,0]; t(:,3) = [p,p,0]; t(:,4) = [0,p,0]
      t(:,5) = [0,0,p]; t(:,6) = [p,0,p]; t(:,7) = [p,p,p]; t(:,8) = [0,p,p]
      if (p < 2) return
      nxt = 9
      do e = 1, 12
         a0 = hex_edge(1,e); a1 = hex_edge(2,e)
         du = (t(:,a1) - t(:,a0))/p
         do s = 1, p-1
            t(:,nxt) = t(:,a0) + s*du
            nxt = nxt + 1
         end do
      end do
      allocate(fq(2,(p-1)**2))
      call gmsh_quad(p-2, fq)
      fq = fq + 1
      do f = 1, 6
         a0 = hex_face(1,f); a1 = hex_face(2,f); a3 = hex_face(4,f)
         du = (t(:,a1) - t(:,a0))/p
         dv = (t(:,a3) - t(:,a0))/p
         do s = 1, (p-1)**2
            t(:,nxt) = t(:,a0) + du*fq(1,s) + dv*fq(2,s)
            nxt = nxt + 1
         end do
      end do
      allocate(sub(3,(p-1)**3))
      call msod_gmsh_hex(p-2, sub)
      t(:,nxt:) = sub + 1
   end subroutine msod_gmsh_hex

   ! Full table set of the drop-in kernel interface, for run-time porder.
   !   invAtoIJK(q,q,q): local node at SOD indices (i,j,k)
   !   AtoIJK(nnode): invAtoIJK listed in (k,i,j) loop order
   !   AtoI/J/K(nnode): SOD indices of each local node
   !   xgp(ngaus,3), wgp(ngaus): GLL points and tensor weights
   !   Ngp(ngaus,nnode), dNgp(3,nnode,ngaus): shape functions at the points
   !   dlxigp_ip(ngaus,3,q): 1-D derivative l_n'(xi) of each point's coordinates
   subroutine msod_hex_tables(p, invAtoIJK, AtoIJK, AtoI, AtoJ, AtoK, xgp, wgp, Ngp, dNgp, dlxigp_ip)
      integer, intent(in) :: p
      integer(c_int), intent(out) :: invAtoIJK(p+1,p+1,p+1), AtoIJK((p+1)**3)
      integer(c_int), intent(out) :: AtoI((p+1)**3), AtoJ((p+1)**3), AtoK((p+1)**3)
      real(c_double), intent(out) :: xgp((p+1)**3,3), wgp((p+1)**3)
      real(c_double), intent(out) :: Ngp((p+1)**3,(p+1)**3), dNgp(3,(p+1)**3,(p+1)**3)
      real(c_double), intent(out) :: dlxigp_ip((p+1)**3,3,p+1)
      integer :: q, nn, a, i, j, k, c, d, g
      integer, allocatable :: t(:,:)
      real(8) :: xi(0:p), w(0:p), l(p+1,3), dl(p+1,3)
      q = p + 1
      nn = q**3
      allocate(t(3,nn))
      call msod_gmsh_hex(p, t)
      ! Gmsh position 0 -> SOD 1, p -> SOD 2, m -> SOD m+2
      do a = 1, nn
         do d = 1, 3
            if (t(d,a) == 0) then
               t(d,a) = 1
            else if (t(d,a) == p) then
               t(d,a) = 2
            else
               t(d,a) = t(d,a) + 2
            end if
         end do
         AtoI(a) = t(1,a); AtoJ(a) = t(2,a); AtoK(a) = t(3,a)
         invAtoIJK(t(1,a), t(2,a), t(3,a)) = a
      end do
      c = 0
      do k = 1, q
         do i = 1, q
            do j = 1, q
               c = c + 1
               AtoIJK(c) = invAtoIJK(i,j,k)
            end do
         end do
      end do
      call msod_gll(p, xi, w)
      do a = 1, nn
         xgp(a,1) = xi(msod_sod_order(p, AtoI(a)))
         xgp(a,2) = xi(msod_sod_order(p, AtoJ(a)))
         xgp(a,3) = xi(msod_sod_order(p, AtoK(a)))
         wgp(a) = w(msod_sod_order(p, AtoI(a)))*w(msod_sod_order(p, AtoJ(a)))*w(msod_sod_order(p, AtoK(a)))
      end do
      do g = 1, nn
         do d = 1, 3
            call msod_lagrange(p, xi, xgp(g,d), l(:,d))
            call msod_lagrange_deriv(p, xi, xgp(g,d), dl(:,d))
            dlxigp_ip(g,d,:) = dl(:,d)
         end do
         do a = 1, nn
            i = AtoI(a); j = AtoJ(a); k = AtoK(a)
            Ngp(g,a) = l(i,1)*l(j,2)*l(k,3)
            dNgp(1,a,g) = dl(i,1)*l(j,2)*l(k,3)
            dNgp(2,a,g) = l(i,1)*dl(j,2)*l(k,3)
            dNgp(3,a,g) = l(i,1)*l(j,2)*dl(k,3)
         end do
      end do
   end subroutine msod_hex_tables

   ! Linear tetrahedron (TET04) with the 4-point rule of the reference driver
   ! (main.f90:374-383): tet_4points (quadrature_rules.f90:84-99) gives the
   ! points (a,b,b), (b,a,b), (b,b,a), (b,b,b) and weights 1/4; the P1 basis
   ! (elem_tet.f90:9-36) N = (1-xi-eta-zeta, xi, eta, zeta) is evaluated there.
   ! The point coordinates are the reference's 16-digit literals (not the
   ! exact (5 +- 3 sqrt 5)/20 roots): they are an input format.
   !   xgp(4,3), wgp(4), Ngp(4,4) (point, node), dNgp(3,4,4)
   subroutine msod_tet_tables(xgp, wgp, Ngp, dNgp)
      real(c_double), intent(out) :: xgp(4,3), wgp(4), Ngp(4,4), dNgp(3,4,4)
      real(8), parameter :: pa = 0.5854101966249680d0, pb = 0.1381966011250110d0
      integer :: g
      xgp(1,:) = [pa, pb, pb]
      xgp(2,:) = [pb, pa, pb]
      xgp(3,:) = [pb, pb, pa]
      xgp(4,:) = [pb, pb, pb]
      wgp = 0.25d0
      do g = 1, 4
         Ngp(g,1) = 1d0 - xgp(g,1) - xgp(g,2) - xgp(g,3)
         Ngp(g,2) = xgp(g,1)
         Ngp(g,3) = xgp(g,2)
         Ngp(g,4) = xgp(g,3)
         dNgp(:,1,g) = [-1d0, -1d0, -1d0]
         dNgp(:,2,g) = [1d0, 0d0, 0d0]
         dNgp(:,3,g) = [0d0, 1d0, 0d0]
         dNgp(:,4,g) = [0d0, 0d0, 1d0]
      end do
   end subroutine msod_tet_tables

end module msod_fem
