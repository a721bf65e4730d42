! Oblique-shock reflection, two-dimensional: a Mach M0 stream (isentropic from
! p_tot, T_tot) enters on the left; from x = 0.1 Lx on, the top boundary holds
! the state behind a shock of angle beta, which reflects off the slip wall at
! the bottom; the right boundary extrapolates.  Read by ouxsbli_amd.
module mod_globals
  implicit none
  integer, parameter :: dimension = 2
  real(8), parameter :: threshold = 0.4d0
  real(8), parameter :: blt = 1.d-3
  real(8), parameter :: Lx = 5.d0 * blt
  real(8), parameter :: Ly = 2.d0 * blt
  integer, parameter :: nx = 257
  integer, parameter :: ny = 129
  integer, parameter :: step_offset = 0
  real(8), parameter :: endT = 0.1d-3
  integer, parameter :: np = 10
  real(8), parameter :: dt = 3.d-9
  integer, parameter :: nt = int(endT / (dble(np) * dt))
  real(8), parameter :: R     = 287.03d0
  real(8), parameter :: gamma = 1.4d0
  real(8), parameter :: Pr    = 0.72d0
  real(8), parameter :: Prt   = 0.9d0
  real(8), parameter :: M0    = 2.d0
  real(8), parameter :: p_tot = 100.d3
  real(8), parameter :: T_tot = 295.d0
  real(8), parameter :: beta  = 37.2d0 * dacos(-1.d0) / 180.d0
end module mod_globals
