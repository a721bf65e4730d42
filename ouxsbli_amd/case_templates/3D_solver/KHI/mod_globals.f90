! Kelvin-Helmholtz instability: a periodic slab of fast light gas (stream 2,
! 0.25 Ly < y < 0.75 Ly) inside slower heavy gas (stream 1), equal pressure,
! sinusoidal transverse perturbation of amplitude amp; read by ouxsbli_amd.
module mod_globals
  implicit none
  integer, parameter  :: dimension = 3
  real(8), parameter  :: threshold = 0.9d0
  real(8), parameter  :: Lx = 5.d-3
  real(8), parameter  :: Ly = 5.d-3
  real(8), parameter  :: Lz = 5.d-3
  integer, parameter  :: nx = 130
  integer, parameter  :: ny = 130
  integer, parameter  :: nz = 130
  real(8), parameter  :: gamma = 1.4d0
  real(8), parameter  :: Pr    = 0.71d0
  real(8), parameter  :: Prt   = 0.9d0
  real(8), parameter  :: R     = 287.15d0
  real(8), parameter  :: p    = 2.72d3
  real(8), parameter  :: T1   = 162.7908d0
  real(8), parameter  :: rho1 = p / (R * T1)
  real(8), parameter  :: u1   = 2.d0 * sqrt(gamma * R * T1)
  real(8), parameter  :: T2   = 104.6303d0
  real(8), parameter  :: rho2 = p / (R * T2)
  real(8), parameter  :: u2   = 3.d0 * sqrt(gamma * R * T2)
  real(8), parameter  :: amp  = 0.05d0
  real(8), parameter  :: CFL  = 0.05d0
  real(8), parameter  :: dt   = CFL * Lx / (dble(nx-1) * abs(u1))
  real(8), parameter  :: endT = 10.d0 * Lx / u2
  integer, parameter  :: np   = 100
  integer, parameter  :: nt   = int(endT / (dble(np) * dt))
end module mod_globals
