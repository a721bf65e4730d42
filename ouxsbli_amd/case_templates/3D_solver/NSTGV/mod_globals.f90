! Supersonic Navier-Stokes Taylor-Green vortex: grid, gas and run parameters
! read by ouxsbli_amd (Case.build parses every `parameter ::` line below).
module mod_globals
  implicit none
  integer, parameter  :: dimension = 3
  real(8), parameter  :: threshold = 0.4d0      ! hybrid-TVD switch on the Ducros sensor
  ! mesh: nx, ny, nz count ORDER/2 ghost layers on each side
  real(8), parameter  :: L0 = 1.524d-3
  real(8), parameter  :: pi = acos(-1.d0)
  real(8), parameter  :: Lx = 2.d0 * pi * L0
  real(8), parameter  :: Ly = 2.d0 * pi * L0
  real(8), parameter  :: Lz = 2.d0 * pi * L0
  integer, parameter  :: nx = 513
  integer, parameter  :: ny = 513
  integer, parameter  :: nz = 513
  ! gas
  real(8), parameter  :: gamma = 1.4d0
  real(8), parameter  :: Pr    = 0.71d0
  real(8), parameter  :: Prt   = 0.9d0
  real(8), parameter  :: R     = 287.03d0
  ! initial state
  real(8), parameter  :: Re   = 1600.d0
  real(8), parameter  :: M0   = 1.25d0
  real(8), parameter  :: T    = 530.d0 * 5.d0 / 9.d0
  real(8), parameter  :: S    = 111.d0
  real(8), parameter  :: mu0  = 1.716d-5 * (273.2d0 + S) / (T + S) * (T / 273.2d0)**1.5d0
  real(8), parameter  :: V0   = M0 * sqrt(gamma * R * T)
  real(8), parameter  :: RHO0 = mu0 * Re / (V0 * L0)
  real(8), parameter  :: p0   = RHO0 * R * T
  ! time: np outputs of nt steps each
  real(8), parameter  :: CFL  = 0.03d0
  real(8), parameter  :: dt   = CFL * (Lx / dble(nx-1)) / V0
  real(8), parameter  :: dtn  = V0 * dt / L0
  integer, parameter  :: np   = 100
  integer, parameter  :: nt   = int(20.d0 / (dble(np) * dtn))
end module mod_globals
