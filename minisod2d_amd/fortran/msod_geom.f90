! Element geometry: inverse Jacobian He and quadrature volume gpvol.
!
! Same contract as the reference's elem_jacobian
! (fem/sources/jacobian_oper.f90:7-136):
!   Je(i,j) = sum_a dNgp(i,a,g) * x_j(connec(e,a))      (Je = (dx/dxi)^T)
!   He = Je^{-1} by cofactors, so He(i,k) = d xi_k / d x_i
!   gpvol(1,g,e) = wgp(g) * det(Je)
! At GLL collocation points dNgp(i,:,g) is non-zero only on the q nodes of
! the 1-D line through g in direction i, so the sum visits just those
! (pre-listed once, in increasing node order: skipping exact zeros leaves a
! sequential sum unchanged).  Elements run in parallel with OpenMP, which is
! what makes the 2.1M-element meshes practical to generate on the host.
module msod_geom

   use iso_c_binding, only: c_int, c_double
   implicit none
   private

   public :: msod_elem_jacobian

contains

   subroutine msod_elem_jacobian(nelem, npoin, nnode, connec, coord, dNgp, wgp, gpvol, He, nbad)
      integer, intent(in) :: nelem, npoin, nnode
      integer(c_int), intent(in) :: connec(nelem,nnode)
      real(c_double), intent(in) :: coord(npoin,3), dNgp(3,nnode,nnode), wgp(nnode)
      real(c_double), intent(out) :: gpvol(1,nnode,nelem), He(3,3,nnode,nelem)
      integer, intent(out) :: nbad          ! number of points with det(Je) <= 0
      integer, allocatable :: nz(:,:,:), nnz(:,:)
      integer :: e, g, i, j, m, a, maxnz
      real(8) :: Je(3,3), det, c(3,3), s
      ! sparsity of dNgp: nz(1:nnz(i,g), i, g) = nodes with dNgp(i,a,g) /= 0
      maxnz = 0
      allocate(nnz(3,nnode))
      do g = 1, nnode
         do i = 1, 3
            nnz(i,g) = count(dNgp(i,:,g) /= 0d0)
            maxnz = max(maxnz, nnz(i,g))
         end do
      end do
      allocate(nz(maxnz,3,nnode))
      do g = 1, nnode
         do i = 1, 3
            m = 0
            do a = 1, nnode
               if (dNgp(i,a,g) /= 0d0) then
                  m = m + 1
                  nz(m,i,g) = a
               end if
            end do
         end do
      end do
      nbad = 0
      !$omp parallel do schedule(static) private(g, i, j, m, a, Je, det, c, s) reduction(+:nbad)
      do e = 1, nelem
         do g = 1, nnode
            do j = 1, 3
               do i = 1, 3
                  s = 0d0
                  do m = 1, nnz(i,g)
                     a = nz(m,i,g)
                     s = s + dNgp(i,a,g)*coord(connec(e,a), j)
                  end do
                  Je(i,j) = s
               end do
            end do
            det = Je(1,1)*Je(2,2)*Je(3,3) + Je(1,2)*Je(2,3)*Je(3,1) + Je(1,3)*Je(2,1)*Je(3,2) &
                - Je(3,1)*Je(2,2)*Je(1,3) - Je(3,2)*Je(2,3)*Je(1,1) - Je(3,3)*Je(2,1)*Je(1,2)
            ! cofactor matrix c(i,j) = (-1)^(i+j) * minor(i,j); inverse = c^T / det
            c(1,1) =   Je(2,2)*Je(3,3) - Je(3,2)*Je(2,3)
            c(1,2) = -(Je(2,1)*Je(3,3) - Je(3,1)*Je(2,3))
            c(1,3) =   Je(2,1)*Je(3,2) - Je(3,1)*Je(2,2)
            c(2,1) = -(Je(1,2)*Je(3,3) - Je(3,2)*Je(1,3))
            c(2,2) =   Je(1,1)*Je(3,3) - Je(3,1)*Je(1,3)
            c(2,3) = -(Je(1,1)*Je(3,2) - Je(3,1)*Je(1,2))
            c(3,1) =   Je(1,2)*Je(2,3) - Je(2,2)*Je(1,3)
            c(3,2) = -(Je(1,1)*Je(2,3) - Je(2,1)*Je(1,3))
            c(3,3) =   Je(1,1)*Je(2,2) - Je(2,1)*Je(1,2)
            do i = 1, 3
               do j = 1, 3
                  He(i,j,g,e) = (1d0/det)*c(j,i)
               end do
            end do
            gpvol(1,g,e) = wgp(g)*det
            if (.not. (det > 0d0)) nbad = nbad + 1
         end do
      end do
      !$omp end parallel do
   end subroutine msod_elem_jacobian

end module msod_geom
