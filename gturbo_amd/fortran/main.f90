! main.f90 -- the driver of src/main.f90:1-195 on the MI355X library.
!
!   gturbo_run [N] [exact]   (N = NX = NY, default 8192 as src/modules.f90:4)
!
! Reads curframe.dat, files/seed.NNN, params.dat and (when the frame is > 0)
! fields/w.NNN; runs npas steps with diagnostics every imix steps and
! snapshots every iout steps; writes fields/w.(NNN+1), files/seed.(NNN+1)
! and curframe.dat.  The OpenACC data region of the reference is the
! library's present table (gt_enter_data ... gt_exit_data).
! It also writes the FP64 spectral state as fields/z.(NNN+1); with "exact" a
! run starting at frame NNN resumes from fields/z.NNN when it exists (bit for
! bit the continuation of the run that wrote it) instead of the float32
! frame.
program gturbo_main
  use iso_c_binding
  use gturbo_hip
  use gturbo_state
  implicit none
  real(8), allocatable, target :: z(:), znl(:), znew(:), u(:), v(:)
  integer :: ipas, ifr, jfr, seed
  integer(8) :: start_time, stop_time, trate
  integer(c_size_t) :: fre, tota
  integer(c_int64_t) :: nbytes
  character(len=64) :: nome, arg
  real :: s
  logical :: exact, spectral
  logical, external :: havez

  call system_clock(start_time, trate)
  if (command_argument_count() >= 1) then
    call get_command_argument(1, arg)
    read(arg, *) nx
  end if
  exact = .false.
  if (command_argument_count() >= 2) then
    call get_command_argument(2, arg)
    exact = trim(arg) == 'exact'
  end if
  call show_mem(fre, tota, 0)

  ! src/main.f90:37-45
  open(unit=1, file='./curframe.dat', status='old')
  read(1, *) ifr, t
  close(1)
  write(nome, "('./files/seed.',i3.3)") ifr
  open(unit=2, file=nome, status='old')
  read(2, *) seed
  close(2)

  ns = (nx + 2) * nx
  allocate(z(ns), znl(ns), znew(ns), u(ns), v(ns))
  call inieuler(z, ifr)
  spectral = exact .and. ifr > 0
  if (spectral) spectral = havez(ifr)
  if (spectral) then
    call readz(z, ifr)
    write(6, *) ' Resuming from the FP64 state of frame ', ifr
  end if
  call gtcheck(gt_plan_create(gtctx, GT_PLAN_Z2D, plan_inv), 'plan Z2D')
  call gtcheck(gt_plan_create(gtctx, GT_PLAN_D2Z, plan_dir), 'plan D2Z')

  write(nome, "('./files/global.',i3.3)") ifr
  open(unit=18, file=nome, form='formatted', status='unknown', access='append')
  write(nome, "('./files/spectra.',i3.3)") ifr
  open(unit=20, file=nome, form='formatted', status='unknown', access='append')
  write(nome, "('./files/fluxes.',i3.3)") ifr
  open(unit=30, file=nome, form='formatted', status='unknown', access='append')
  call show_mem(fre, tota, 1)

  ! src/main.f90:67-83
  nbytes = int(ns, c_int64_t) * 8
  call gtcheck(gt_enter_data(gtctx, znl, nbytes, 0), 'enter znl')
  call gtcheck(gt_enter_data(gtctx, u, nbytes, 0), 'enter u')
  call gtcheck(gt_enter_data(gtctx, v, nbytes, 0), 'enter v')
  if (rko == 4) call gtcheck(gt_enter_data(gtctx, znew, nbytes, 0), 'enter znew')
  call gtcheck(gt_enter_data(gtctx, z, nbytes, 1), 'enter z')
  call show_mem(fre, tota, 2)

  ! src/main.f90:89-92 (an FP64 spectral state is already transformed)
  if (.not. spectral) then
    call fft_dir(z, plan_dir)
    if (truncate == 1) call trunc(z)
  end if

  jfr = 1
  do ipas = 1, npas                                      ! src/main.f90:97-137
    t = t + dt
    if (rko == 4) then
      call step4(z, znl, znew, u, v)
    else
      call step(z, znl, u, v)
    end if
    if (mod(ipas, imix) == 0) then
      call gtcheck(gt_copy(gtctx, znl, z), 'znl=z')
      call nlt(znl, u, v)
      call spectrum(z, ipas)
      call fluxes(z, znl, ipas)
      call glob(z, ipas)
    end if
    if (mod(ipas, iout) == 0 .and. ipas /= npas) then
      call gtcheck(gt_copy(gtctx, znl, z), 'znl=z')
      call fft_inv(znl, plan_inv)
      call gtcheck(gt_update_host(gtctx, znl), 'update host znl')
      call writef2(znl, ifr, jfr)
      jfr = jfr + 1
      write(6, *) 'Saved w field, seed and curframe at timestep = ', ipas
      call flush(6)
    end if
    call forcing(z, seed)
  end do

  ! src/main.f90:140-154
  ifr = ifr + 1
  call gtcheck(gt_update_host(gtctx, z), 'update host z')
  call writez(z, ifr)
  call fft_inv(z, plan_inv)
  write(nome, "('./files/seed.',i3.3)") ifr
  open(unit=2, file=nome, status='replace')
  write(2, *) seed
  close(2)
  open(unit=1, file='./curframe.dat', status='replace')
  write(1, *) ifr, t
  close(1)
  close(18)
  close(20)
  close(30)

  ! src/main.f90:164-182
  if (rko == 4) call gtcheck(gt_exit_data(gtctx, znew, 0), 'exit znew')
  call gtcheck(gt_exit_data(gtctx, u, 0), 'exit u')
  call gtcheck(gt_exit_data(gtctx, v, 0), 'exit v')
  call gtcheck(gt_exit_data(gtctx, znl, 0), 'exit znl')
  call gtcheck(gt_exit_data(gtctx, z, 1), 'exit z')
  call writef(z, ifr)
  call gtcheck(gt_destroy(gtctx), 'destroy')

  ! src/main.f90:190-193
  call system_clock(stop_time)
  s = real(stop_time - start_time) / (real(npas) * real(trate))
  write(6, *) 'Mean time per iteration      = ', s
  call flush(6)
end program gturbo_main
