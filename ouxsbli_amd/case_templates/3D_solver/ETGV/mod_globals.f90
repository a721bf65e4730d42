! Inviscid (Euler) Taylor-Green vortex in non-dimensional units: rho0 = 1,
! p0 = 1/gamma, peak velocity M0; read by ouxsbli_amd.
module mod_globals
  implicit none
  integer, parameter  :: dimension = 3
  real(8), parameter  :: threshold = 0.4d0
  real(8), parameter  :: pi = acos(-1.d0)
  real(8), parameter  :: Lx = 2.d0 * pi
  real(8), parameter  :: Ly = 2.d0 * pi
  real(8), parameter  :: Lz = 2.d0 * pi
  integer, parameter  :: nx = 66
  integer, parameter  :: ny = 66
  integer, parameter  :: nz = 66
  real(8), parameter  :: gamma = 1.4d0
  real(8), parameter  :: Pr    = 0.71d0
  real(8), parameter  :: Prt   = 0.9d0
  real(8), parameter  :: R     = 287.03d0
  real(8), parameter  :: M0    = 0.4d0
  real(8), parameter  :: RHO0  = 1.d0
  real(8), parameter  :: CFL   = 0.1d0
  real(8), parameter  :: dt    = CFL * (Lx / dble(nx-1)) / (M0 + 1.d0)
  integer, parameter  :: np    = 10
  integer, parameter  :: nt    = int(10.d0 / (dble(np) * dt))
end module mod_globals
