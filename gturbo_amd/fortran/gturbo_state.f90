! gturbo_state.f90 -- run state shared by the drop-in subroutines, in place of
! the reference's modules (src/modules.f90:1-65).  The grid size is a run-time
! value here (the reference fixes NX at compile time, src/modules.f90:4); the
! physics tables of commk/commforc live inside the library context.
module gturbo_state
  use iso_c_binding, only: c_int32_t
  implicit none
  integer :: nx = 8192          ! NX = NY
  integer :: ns = 0             ! (NX+2)*NY doubles per field
  integer :: nbin = 0           ! NBIN = NX/2
  ! commphy
  real(8) :: xlx, xly, nu, alpnu, mu, alpmu, alpv
  real(8) :: dt, dt2, dt3, dt6, sdt, t = 0.d0, dkx, dky, inpz, inpe
  ! commforc
  real(8) :: famp, k1f, k2f
  integer :: nf = 0
  ! commsim
  integer :: npas, iout, imix, split, rko
  ! commk
  real(8) :: rtrunc
  integer :: truncate
  ! plans (src/modules.f90:22-27)
  integer(c_int32_t) :: plan_dir = 0, plan_inv = 0
end module gturbo_state
