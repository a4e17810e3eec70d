! gturbo_sub.f90 -- the reference's subroutines, same names and arguments,
! as external procedures with implicit interfaces (callers pass arrays by
! address, as in the reference).  Each forwards to the C ABI; the arrays must
! be present on the device (gt_enter_data), as OpenACC requires in the
! reference.

! src/inieuler.f90:1-97: params.dat, split/RK rules, dt constants, defk,
! iniforcing (inside gt_create), restart field or zero field.
subroutine inieuler(z, ifr)
  use iso_c_binding
  use gturbo_hip
  use gturbo_state
  implicit none
  real(8), intent(out) :: z(nx + 2, nx)
  integer :: ifr
  real(c_double) :: p(17), dinfo(11)
  integer(c_int32_t) :: iinfo(7)

  open(unit=9, file='./params.dat', status='old')
  read(9, *) xlx, xly
  read(9, *) nu, alpnu, mu, alpmu, alpv
  read(9, *) dt, rko
  read(9, *) npas, iout, imix
  read(9, *) rtrunc, truncate
  read(9, *) famp, k1f, k2f
  close(9)

  p = [xlx, xly, nu, alpnu, mu, alpmu, alpv, dt, real(rko, 8), real(npas, 8), &
       real(iout, 8), real(imix, 8), rtrunc, real(truncate, 8), famp, k1f, k2f]
  call gtcheck(gt_create(int(nx, c_int32_t), p, gtctx), 'gt_create')
  call gtcheck(gt_info(gtctx, iinfo, dinfo), 'gt_info')
  nf = iinfo(1)
  split = iinfo(2)
  rko = iinfo(3)
  inpz = dinfo(1)
  inpe = dinfo(2)
  dt2 = dinfo(3)
  dt3 = dinfo(4)
  dt6 = dinfo(5)
  sdt = dinfo(6)
  dkx = dinfo(7)
  dky = dinfo(8)
  ns = (nx + 2) * nx
  nbin = nx / 2

  write(6, *) ' 2D alpha-turbulence simulation (MI355X)'
  write(6, *) ' Resolution       = ', nx, nx
  write(6, *) ' Viscosity        = ', nu
  write(6, *) ' Viscosity order  = ', alpnu
  write(6, *) ' Friction         = ', mu
  write(6, *) ' Friction order   = ', alpmu
  write(6, *) ' Alpha            = ', alpv
  if (truncate == 1) then
    write(6, *) ' Truncation at   = ', rtrunc
  else
    write(6, *) ' No truncation'
  end if
  if (split == 1) then
    write(6, *) ' Splitted linear integrator (only for Ekman-NSE)'
  else
    write(6, *) ' Unsplitted linear integrator'
  end if
  write(6, *) ' Runge-Kutta of order ', rko
  write(6, *) ' Time step           = ', dt
  write(6, *) ' Starting Frame      = ', ifr
  write(6, *) ' Total step number   = ', npas
  write(6, *) ' Outputs Frames/Diag = ', iout, imix
  call defk
  call iniforcing
  call flush(6)

  if (ifr > 0) then
    call readf(z, ifr)
  else
    z = 0.d0
  end if
end subroutine inieuler

! src/defk.f90: the tables are built inside the library context by gt_create.
subroutine defk
  use gturbo_state
  implicit none
  write(6, *) ' Wavenumber increments dkx, dky = ', dkx, dky
end subroutine defk

! src/iniforcing.f90:70-75 (the forced-mode table itself is built by
! gt_create, in the reference's order).
subroutine iniforcing
  use gturbo_state
  implicit none
  write(6, *) ' # of forced wavenumbers = ', nf
  write(6, *) ' Mean forcing amplitude  = ', famp
  write(6, *) ' Forcing range           : ', k1f, ' < k =<', k2f
  write(6, *) ' Mean energy input       = ', inpe
  write(6, *) ' Mean enstrophy input    = ', inpz
  write(6, *) ' '
end subroutine iniforcing

! src/fft_dir.f90
subroutine fft_dir(z, plan)
  use iso_c_binding
  use gturbo_hip
  implicit none
  real(8) :: z(*)
  integer :: plan
  call gtcheck(gt_fft_dir(gtctx, z, int(plan, c_int32_t)), 'fft_dir')
end subroutine fft_dir

! src/fft_inv.f90
subroutine fft_inv(z, plan)
  use iso_c_binding
  use gturbo_hip
  implicit none
  real(8) :: z(*)
  integer :: plan
  call gtcheck(gt_fft_inv(gtctx, z, int(plan, c_int32_t)), 'fft_inv')
end subroutine fft_inv

! src/produ.f90
subroutine produ(z, u, v)
  use gturbo_hip
  implicit none
  real(8) :: z(*), u(*), v(*)
  call gtcheck(gt_produ(gtctx, z, u, v), 'produ')
end subroutine produ

! src/trunc.f90:1-28
subroutine trunc(z)
  use gturbo_hip
  implicit none
  real(8) :: z(*)
  call gtcheck(gt_trunc(gtctx, z), 'trunc')
end subroutine trunc

! src/nlt.f90
subroutine nlt(z, u, v)
  use gturbo_hip
  implicit none
  real(8) :: z(*), u(*), v(*)
  call gtcheck(gt_nlt(gtctx, z, u, v), 'nlt')
end subroutine nlt

! src/rkstep.f90:2-76
subroutine step(z, znl, u, v)
  use gturbo_hip
  implicit none
  real(8) :: z(*), znl(*), u(*), v(*)
  call gtcheck(gt_step(gtctx, z, znl, u, v), 'step')
end subroutine step

! src/rkstep.f90:82-131
subroutine step4(z, znl, znew, u, v)
  use gturbo_hip
  implicit none
  real(8) :: z(*), znl(*), znew(*), u(*), v(*)
  call gtcheck(gt_step4(gtctx, z, znl, znew, u, v), 'step4')
end subroutine step4

! src/forcing.f90
subroutine forcing(z, seed)
  use iso_c_binding
  use gturbo_hip
  implicit none
  real(8) :: z(*)
  integer :: seed
  integer(c_int32_t) :: s
  s = int(seed, c_int32_t)
  call gtcheck(gt_forcing(gtctx, z, s), 'forcing')
  seed = int(s)
end subroutine forcing

! src/diagnostics.f90:5-56, written with a fixed, parseable format
subroutine spectrum(z, ifr)
  use gturbo_hip
  use gturbo_state
  implicit none
  real(8) :: z(*)
  integer :: ifr, ib
  real(8), allocatable :: sp(:), sz(:)
  allocate(sp(nbin), sz(nbin))
  call gtcheck(gt_spectrum(gtctx, z, sp, sz), 'spectrum')
  do ib = 1, nbin
    write(20, '(3es24.15e3)') dkx * (ib - 1), sp(ib), sz(ib)
  end do
  deallocate(sp, sz)
end subroutine spectrum

! src/diagnostics.f90:62-119
subroutine fluxes(z, znl, ifr)
  use gturbo_hip
  use gturbo_state
  implicit none
  real(8) :: z(*), znl(*)
  integer :: ifr, ib
  real(8), allocatable :: fle(:), flz(:)
  allocate(fle(nbin), flz(nbin))
  call gtcheck(gt_fluxes(gtctx, z, znl, fle, flz), 'fluxes')
  do ib = 1, nbin
    write(30, '(3es24.15e3)') dkx * ib, fle(ib), flz(ib)
  end do
  deallocate(fle, flz)
end subroutine fluxes

! src/diagnostics.f90:125-179
subroutine glob(z, ifr)
  use gturbo_hip
  use gturbo_state
  implicit none
  real(8) :: z(*)
  integer :: ifr
  real(8) :: g(6)
  call gtcheck(gt_glob(gtctx, z, g), 'glob')
  ! t, Z, E, Z-diss nu, Z-diss mu, Z-input, E-diss nu, E-diss mu, E-input
  write(18, '(9es24.15e3)') ifr * dt, g(2), g(1), g(3), g(4), inpz, g(5), g(6), inpe
  call flush(18)
  write(6, '(a,i10,a,es14.6,a,es14.6)') ' Nt = ', ifr, ', Zb = ', (g(3) + g(4)) / inpz, &
    ', Eb = ', (g(5) + g(6)) / inpe
  call flush(6)
end subroutine glob

! src/trunc.f90:32-49 (show_mem), through hipMemGetInfo
subroutine show_mem(fre, tot, step)
  use iso_c_binding
  use gturbo_hip
  implicit none
  integer(c_size_t) :: fre, tot
  integer :: step
  real :: r
  call gtcheck(gt_mem_info(fre, tot), 'mem_info')
  r = real(tot - fre) / 2.0**30
  write(6, "(A25,I4,A3,F12.6,A4,F9.3)") "Alloc mem (Gb) in step", step, " = ", r, " of ", &
    real(tot) / 2.0**30
  call flush(6)
end subroutine show_mem

! The adjoint (no reference counterpart): the final field wT of nsteps steps
! (step/step4, forcing) from the physical field w0, with a checkpoint every
! kchk steps kept for adjoint_backward.  Arrays must be present.
subroutine adjoint_forward(w0, wt, nsteps, kchk, seed)
  use iso_c_binding
  use gturbo_hip
  implicit none
  real(8) :: w0(*), wt(*)
  integer :: nsteps, kchk, seed
  integer(c_int32_t) :: s
  s = int(seed, c_int32_t)
  call gtcheck(gt_adj_forward(gtctx, w0, wt, int(nsteps, c_int32_t), int(kchk, c_int32_t), &
                              s), 'adjoint_forward')
  seed = int(s)
end subroutine adjoint_forward

! g0 = dL/dw0 from gt = dL/dwT for the run of the last adjoint_forward.
subroutine adjoint_backward(gt, g0)
  use gturbo_hip
  implicit none
  real(8) :: gt(*), g0(*)
  call gtcheck(gt_adj_backward(gtctx, gt, g0), 'adjoint_backward')
end subroutine adjoint_backward
