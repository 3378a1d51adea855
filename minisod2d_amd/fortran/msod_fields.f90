! Synthetic flow fields for the residual kernels.
!
! Per global node at (x,y,z):
!   u   = ( sin2pix cos2piy cos2piz, -cos2pix sin2piy cos2piz, 0 )
!   rho = 1 + 0.05 N(0,1)               (seeded; optionally x1.1 for x > 0.5)
!   pr  = 1/gamma + 0.1 cos4pix
!   q   = rho u,  E = rho e + 1/2 q.u with e = pr/((gamma-1) rho)
!         (the reference's own derived-field relations, app/sources/
!         main.f90:350-361, which make E = pr/(gamma-1) + 1/2 rho|u|^2)
!   Tem = pr/(rho Rgas),  mu_fluid = 1e-3
! Per element node: mu_e ~ U[0,1e-3], mu_sgs ~ U[0,1e-4] (seeded), stored in
! the reference layout mu(nelem,ngaus).
module msod_fields

   use iso_c_binding, only: c_int, c_double, c_int64_t
   implicit none
   private

   public :: msod_flow_fields, msod_viscosity_fields

   interface
      real(c_double) function msod_rng_uniform(seed, i) bind(C, name='msod_rng_uniform')
         import :: c_double, c_int64_t
         integer(c_int64_t), value :: seed, i
      end function msod_rng_uniform
      real(c_double) function msod_rng_normal(seed, i) bind(C, name='msod_rng_normal')
         import :: c_double, c_int64_t
         integer(c_int64_t), value :: seed, i
      end function msod_rng_normal
   end interface

contains

   subroutine msod_flow_fields(npoin, coord, gamma, Rgas, seed_rho, rho_jump, u, q, rho, pr, E, Tem, mu_fluid)
      integer(8), intent(in) :: npoin
      real(c_double), intent(in) :: coord(npoin,3), gamma, Rgas, rho_jump
      integer(c_int64_t), intent(in) :: seed_rho
      real(c_double), intent(out) :: u(npoin,3), q(npoin,3), rho(npoin), pr(npoin), E(npoin)
      real(c_double), intent(out) :: Tem(npoin), mu_fluid(npoin)
      real(8), parameter :: twopi = 6.283185307179586476925286766559d0
      real(8) :: x, y, z, fmag, eint
      integer(8) :: ip
      integer :: d
      !$omp parallel do schedule(static) private(x, y, z, fmag, eint, d)
      do ip = 1, npoin
         x = coord(ip,1); y = coord(ip,2); z = coord(ip,3)
         u(ip,1) = sin(twopi*x)*cos(twopi*y)*cos(twopi*z)
         u(ip,2) = -cos(twopi*x)*sin(twopi*y)*cos(twopi*z)
         u(ip,3) = 0d0
         rho(ip) = 1d0 + 0.05d0*msod_rng_normal(seed_rho, ip-1)
         if (rho_jump /= 0d0 .and. x > 0.5d0) rho(ip) = rho(ip)*(1d0 + rho_jump)
         pr(ip) = 1d0/gamma + 0.1d0*cos(2d0*twopi*x)
         fmag = 0d0
         do d = 1, 3
            q(ip,d) = rho(ip)*u(ip,d)
            fmag = fmag + q(ip,d)*u(ip,d)
         end do
         eint = pr(ip)/((gamma - 1d0)*rho(ip))
         E(ip) = rho(ip)*eint + 0.5d0*fmag
         Tem(ip) = pr(ip)/(rho(ip)*Rgas)
         mu_fluid(ip) = 1d-3
      end do
      !$omp end parallel do
   end subroutine msod_flow_fields

   subroutine msod_viscosity_fields(nelem, ngaus, seed_mue, seed_musgs, mue_max, musgs_max, mu_e, mu_sgs)
      integer(8), intent(in) :: nelem
      integer, intent(in) :: ngaus
      integer(c_int64_t), intent(in) :: seed_mue, seed_musgs
      real(c_double), intent(in) :: mue_max, musgs_max
      real(c_double), intent(out) :: mu_e(nelem,ngaus), mu_sgs(nelem,ngaus)
      integer(8) :: e
      integer :: g
      !$omp parallel do schedule(static) private(e)
      do g = 1, ngaus
         do e = 1, nelem
            mu_e(e,g) = mue_max*msod_rng_uniform(seed_mue, (e-1)*ngaus + (g-1))
            mu_sgs(e,g) = musgs_max*msod_rng_uniform(seed_musgs, (e-1)*ngaus + (g-1))
         end do
      end do
      !$omp end parallel do
   end subroutine msod_viscosity_fields

end module msod_fields
