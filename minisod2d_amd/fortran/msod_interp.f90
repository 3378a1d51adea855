! Transfer between element orders on the same hexahedra, used to build the
! training pairs of the learned closure from a fine-order run:
!
!  msod_interp_1d(pf, pt, M):  M(i,n) = l^pf_n(xi^pt_i), the 1-D Lagrange
!     basis of order pf (SOD node order, binary32-rounded GLL nodes as
!     everywhere else) at the order-pt nodes -- the 1-D factor of the
!     reference's shape functions (utils/sources/mod_maths.f90:177-213) and
!     of var_interpolate (:337-364).
!  msod_prolong_fields: nodal fields of order pf evaluated at the order-pt
!     nodes of every element (tensor product of M); shared nodes receive the
!     same value from each of their elements up to rounding.
!  msod_restrict_residual: weak-form residuals of order pf moved to order pt
!     test functions: R_t(j) = sum_i phi^pt_j(x_i) R_f(i), the exact
!     restriction of integrals against phi^pf_i to the order-pt space
!     (phi^pt_j = sum_i phi^pt_j(x_i) phi^pf_i when pt <= pf).  Each fine
!     node contributes once, through any one element holding it.
! connec_f / connec_t: the two meshes' connectivities over the same elements
! (reference layout, 1-based); node-local numbering in Gmsh order via
! AtoI/J/K of each order.
module msod_interp

   use iso_c_binding, only: c_int, c_double
   use msod_fem, only: msod_gll, msod_lagrange, msod_hex_tables
   implicit none
   private

   public :: msod_interp_1d, msod_prolong_fields, msod_restrict_residual

contains

   subroutine msod_interp_1d(pf, pt, M)
      integer, intent(in) :: pf, pt
      real(8), intent(out) :: M(pt+1, pf+1)
      real(8) :: xf(0:pf), wf(0:pf), xt(0:pt), wt(0:pt)
      integer :: i, pos
      call msod_gll(pf, xf, wf)
      call msod_gll(pt, xt, wt)
      do i = 1, pt+1
         if (i == 1) then
            pos = 0
         else if (i == 2) then
            pos = pt
         else
            pos = i - 2
         end if
         call msod_lagrange(pf, xf, xt(pos), M(i,:))
      end do
   end subroutine msod_interp_1d

   ! local node -> SOD (i,j,k) of one order
   subroutine sod_ijk(p, ijk)
      integer, intent(in) :: p
      integer, intent(out) :: ijk(3, (p+1)**3)
      integer(c_int) :: inv(p+1,p+1,p+1), aijk((p+1)**3), ai((p+1)**3), aj((p+1)**3), ak((p+1)**3)
      real(c_double), allocatable :: xgp(:,:), wgp(:), Ngp(:,:), dNgp(:,:,:), dl(:,:,:)
      integer :: nn
      nn = (p+1)**3
      allocate(xgp(nn,3), wgp(nn), Ngp(nn,nn), dNgp(3,nn,nn), dl(nn,3,p+1))
      call msod_hex_tables(p, inv, aijk, ai, aj, ak, xgp, wgp, Ngp, dNgp, dl)
      ijk(1,:) = ai; ijk(2,:) = aj; ijk(3,:) = ak
   end subroutine sod_ijk

   ! T(it,jt,kt) = sum M(it,if) M(jt,jf) M(kt,kf) F(if,jf,kf), by three 1-D contractions
   pure subroutine tensor_apply(qf, qt, M, F, T)
      integer, intent(in) :: qf, qt
      real(8), intent(in) :: M(qt,qf), F(qf,qf,qf)
      real(8), intent(out) :: T(qt,qt,qt)
      real(8) :: A(qt,qf,qf), B(qt,qt,qf)
      integer :: i, j, k, n
      A = 0d0
      do k = 1, qf
         do j = 1, qf
            do n = 1, qf
               do i = 1, qt
                  A(i,j,k) = A(i,j,k) + M(i,n)*F(n,j,k)
               end do
            end do
         end do
      end do
      B = 0d0
      do k = 1, qf
         do n = 1, qf
            do j = 1, qt
               do i = 1, qt
                  B(i,j,k) = B(i,j,k) + M(j,n)*A(i,n,k)
               end do
            end do
         end do
      end do
      T = 0d0
      do n = 1, qf
         do k = 1, qt
            do j = 1, qt
               do i = 1, qt
                  T(i,j,k) = T(i,j,k) + M(k,n)*B(i,j,n)
               end do
            end do
         end do
      end do
   end subroutine tensor_apply

   subroutine msod_prolong_fields(pf, pt, nelem, npf, npt, ncomp, connec_f, connec_t, ff, ft)
      integer, intent(in) :: pf, pt, nelem, npf, npt, ncomp
      integer(c_int), intent(in) :: connec_f(nelem,(pf+1)**3), connec_t(nelem,(pt+1)**3)
      real(c_double), intent(in) :: ff(npf, ncomp)
      real(c_double), intent(out) :: ft(npt, ncomp)
      real(8) :: M(pt+1, pf+1), F(pf+1,pf+1,pf+1), T(pt+1,pt+1,pt+1)
      integer :: ijkf(3,(pf+1)**3), ijkt(3,(pt+1)**3), e, a, b, c
      call msod_interp_1d(pf, pt, M)
      call sod_ijk(pf, ijkf)
      call sod_ijk(pt, ijkt)
      !$omp parallel do schedule(static) private(a, b, c, F, T)
      do e = 1, nelem
         do c = 1, ncomp
            do a = 1, (pf+1)**3
               F(ijkf(1,a), ijkf(2,a), ijkf(3,a)) = ff(connec_f(e,a), c)
            end do
            call tensor_apply(pf+1, pt+1, M, F, T)
            do b = 1, (pt+1)**3
               ft(connec_t(e,b), c) = T(ijkt(1,b), ijkt(2,b), ijkt(3,b))
            end do
         end do
      end do
      !$omp end parallel do
   end subroutine msod_prolong_fields

   subroutine msod_restrict_residual(pf, pt, nelem, npf, npt, ncomp, connec_f, connec_t, rf, rt)
      integer, intent(in) :: pf, pt, nelem, npf, npt, ncomp
      integer(c_int), intent(in) :: connec_f(nelem,(pf+1)**3), connec_t(nelem,(pt+1)**3)
      real(c_double), intent(in) :: rf(npf, ncomp)
      real(c_double), intent(out) :: rt(npt, ncomp)
      real(8) :: Rm(pf+1, pt+1), RmT(pt+1, pf+1), F(pf+1,pf+1,pf+1), T(pt+1,pt+1,pt+1)
      integer :: ijkf(3,(pf+1)**3), ijkt(3,(pt+1)**3), e, a, b, c, nnf, nnt
      integer, allocatable :: owner(:)
      real(8), allocatable :: loc(:,:,:)
      nnf = (pf+1)**3
      nnt = (pt+1)**3
      ! Rm(i,n) = l^pt_n(xi^pf_i): the order-pt basis at the order-pf nodes
      call msod_interp_1d(pt, pf, Rm)
      RmT = transpose(Rm)
      call sod_ijk(pf, ijkf)
      call sod_ijk(pt, ijkt)
      ! every fine node is counted through the first element that holds it
      allocate(owner(npf))
      owner = 0
      do e = nelem, 1, -1
         do a = 1, nnf
            owner(connec_f(e,a)) = e
         end do
      end do
      allocate(loc(nnt, ncomp, nelem))
      !$omp parallel do schedule(static) private(a, b, c, F, T)
      do e = 1, nelem
         do c = 1, ncomp
            do a = 1, nnf
               if (owner(connec_f(e,a)) == e) then
                  F(ijkf(1,a), ijkf(2,a), ijkf(3,a)) = rf(connec_f(e,a), c)
               else
                  F(ijkf(1,a), ijkf(2,a), ijkf(3,a)) = 0d0
               end if
            end do
            ! T(b) = sum_a phi^pt_b(x_a) F(a)
            call tensor_apply(pf+1, pt+1, RmT, F, T)
            do b = 1, nnt
               loc(b, c, e) = T(ijkt(1,b), ijkt(2,b), ijkt(3,b))
            end do
         end do
      end do
      !$omp end parallel do
      rt = 0d0
      do e = 1, nelem
         do c = 1, ncomp
            do b = 1, nnt
               rt(connec_t(e,b), c) = rt(connec_t(e,b), c) + loc(b, c, e)
            end do
         end do
      end do
   end subroutine msod_restrict_residual

end module msod_interp
