! Sod shock tube along z between slip walls (x and y) with zero-gradient z
! ends; read by ouxsbli_amd.  One GPU holds the whole tube: nz is the z extent
! of all the reference's z slabs together (its ranks' slabs meet at the
! diaphragm, Lz / 2).
module mod_globals
  implicit none
  integer, parameter  :: dimension = 3
  real(8), parameter  :: threshold = 0.4d0
  real(8), parameter  :: Lx = 0.1d0
  real(8), parameter  :: Ly = 0.1d0
  real(8), parameter  :: Lz = 1.0d0
  integer, parameter  :: nx = 6
  integer, parameter  :: ny = 6
  integer, parameter  :: nz = 258
  real(8), parameter  :: gamma = 1.4d0
  real(8), parameter  :: Pr    = 0.71d0
  real(8), parameter  :: Prt   = 0.9d0
  real(8), parameter  :: R     = 287.03d0
  real(8), parameter  :: rho_L = 1.0d0
  real(8), parameter  :: p_L   = 1.0d0
  real(8), parameter  :: rho_R = 0.125d0
  real(8), parameter  :: p_R   = 0.1d0
  real(8), parameter  :: CFL  = 0.01d0
  real(8), parameter  :: dt   = CFL * Lz / (dble(Nz-1) * sqrt(p_L / rho_L))
  real(8), parameter  :: endT = 0.1d0
  integer, parameter  :: np   = 50
  integer, parameter  :: nt   = int(endT / (dble(np) * dt))
end module mod_globals
