! invert.f90 -- initial-condition inversion with the library's adjoint.
!
!   gturbo_invert [N] [iterations]     (defaults: 8192, 10)
!
! In a run directory laid out as the reference's (params.dat, curframe.dat,
! files/seed.NNN, fields/w.NNN): the frame NNN of curframe.dat is the first
! guess of the initial field and fields/obs.NNN (same float32 record) an
! observation of the field npas steps later.  Minimises
!   L(w0) = 1/2 sum over every 4th grid point in x and y of (wT(w0) - obs)^2
! by steepest descent with a backtracking step, the gradient from
! adjoint_forward / adjoint_backward (checkpoints every 20 steps, or npas if
! fewer).  Prints one line per iteration and writes the estimate to
! fields/winv.NNN.
program gturbo_invert
  use iso_c_binding
  use gturbo_hip
  use gturbo_state
  implicit none
  real(8), allocatable, target :: w0(:), wt(:), gt(:), g0(:), obs(:), trial(:)
  integer :: ifr, seed, seed0, it, niter, k, i, j, kchk
  integer(c_int64_t) :: nbytes
  character(len=64) :: nome, arg
  real(8) :: loss, lnew, step, gnorm, wnorm
  logical :: accepted

  niter = 10
  if (command_argument_count() >= 1) then
    call get_command_argument(1, arg)
    read(arg, *) nx
  end if
  if (command_argument_count() >= 2) then
    call get_command_argument(2, arg)
    read(arg, *) niter
  end if
  open(unit=1, file='./curframe.dat', status='old')
  read(1, *) ifr, t
  close(1)
  write(nome, "('./files/seed.',i3.3)") ifr
  open(unit=2, file=nome, status='old')
  read(2, *) seed0
  close(2)

  ns = (nx + 2) * nx
  allocate(w0(ns), wt(ns), gt(ns), g0(ns), obs(ns), trial(ns))
  call inieuler(w0, ifr)
  call reado(obs, ifr)
  kchk = min(20, npas)
  nbytes = int(ns, c_int64_t) * 8
  call gtcheck(gt_enter_data(gtctx, w0, nbytes, 1), 'enter w0')
  call gtcheck(gt_enter_data(gtctx, trial, nbytes, 0), 'enter trial')
  call gtcheck(gt_enter_data(gtctx, wt, nbytes, 0), 'enter wT')
  call gtcheck(gt_enter_data(gtctx, gt, nbytes, 0), 'enter gT')
  call gtcheck(gt_enter_data(gtctx, g0, nbytes, 0), 'enter g0')

  loss = misfit(w0)
  step = -1.d0
  write(6, '(a,i4,a,es22.14)') ' iteration ', 0, '  loss ', loss
  do it = 1, niter
    ! gradient at w0: forward (checkpoints), misfit gradient, backward
    seed = seed0
    call adjoint_forward(w0, wt, npas, kchk, seed)
    call gtcheck(gt_update_host(gtctx, wt), 'update wT')
    gt = 0.d0
    do j = 1, nx, 4
      do i = 1, nx, 4
        k = i + (j - 1) * (nx + 2)
        gt(k) = wt(k) - obs(k)
      end do
    end do
    call gtcheck(gt_update_device(gtctx, gt), 'update gT')
    call adjoint_backward(gt, g0)
    call gtcheck(gt_update_host(gtctx, g0), 'update g0')
    gnorm = sqrt(sum(g0 * g0))
    if (step < 0.d0) then
      wnorm = sqrt(sum(w0 * w0))
      step = 0.05d0 * wnorm / max(gnorm, tiny(1.d0))
    end if
    accepted = .false.
    do k = 1, 20                                  ! backtracking
      trial = w0 - step * g0
      lnew = misfit(trial)
      if (lnew < loss) then
        accepted = .true.
        exit
      end if
      step = 0.5d0 * step
    end do
    if (.not. accepted) exit
    w0 = trial
    call gtcheck(gt_update_device(gtctx, w0), 'update w0')
    write(6, '(a,i4,a,es22.14,a,es12.4,a,es12.4)') ' iteration ', it, '  loss ', lnew, &
      '  |grad| ', gnorm, '  step ', step
    loss = lnew
    step = 1.5d0 * step
    call flush(6)
  end do
  write(nome, "('./fields/winv.',i3.3)") ifr
  call write_w(w0, nome)
  call gtcheck(gt_destroy(gtctx), 'destroy')

contains

  real(8) function misfit(w)
    real(8) :: w(ns)
    integer :: ii, jj, kk, sd
    call gtcheck(gt_update_device(gtctx, w), 'update field')
    sd = seed0
    call adjoint_forward(w, wt, npas, kchk, sd)
    call gtcheck(gt_update_host(gtctx, wt), 'update wT')
    misfit = 0.d0
    do jj = 1, nx, 4
      do ii = 1, nx, 4
        kk = ii + (jj - 1) * (nx + 2)
        misfit = misfit + 0.5d0 * (wt(kk) - obs(kk))**2
      end do
    end do
  end function misfit

end program gturbo_invert

! fields/obs.NNN, the record format of src/read.f90.
subroutine reado(z, ifr)
  use gturbo_state
  implicit none
  real(8) :: z(nx + 2, nx)
  integer :: ifr, i, j
  real(4), allocatable :: w(:, :)
  character(len=40) :: nome
  allocate(w(nx, nx))
  write(nome, "('./fields/obs.',i3.3)") ifr
  open(unit=91, file=nome, form='unformatted', action='read', status='old')
  read(91) w
  close(91)
  do j = 1, nx
    do i = 1, nx
      z(i, j) = dble(w(i, j))
    end do
    z(nx + 1, j) = 0.d0
    z(nx + 2, j) = 0.d0
  end do
  deallocate(w)
end subroutine reado
