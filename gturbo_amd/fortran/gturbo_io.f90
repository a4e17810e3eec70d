! gturbo_io.f90 -- field files as the reference writes them: one Fortran
! unformatted sequential record holding float32 w(NX,NY) (src/read.f90:3-29,
! src/write.f90:3-51).

subroutine readf(z, ifr)
  use gturbo_state
  implicit none
  real(8) :: z(nx + 2, nx)
  integer :: ifr, i, j
  real(4), allocatable :: w(:, :)
  character(len=40) :: nome
  allocate(w(nx, nx))
  write(nome, "('./fields/w.',i3.3)") ifr
  open(unit=90, file=nome, form='unformatted', action='read', status='old')
  read(90) w
  close(90)
  do j = 1, nx
    do i = 1, nx
      z(i, j) = dble(w(i, j))
    end do
    z(nx + 1, j) = 0.d0
    z(nx + 2, j) = 0.d0
  end do
  deallocate(w)
end subroutine readf

subroutine writef(u, ifr)
  use gturbo_state
  implicit none
  real(8) :: u(nx + 2, nx)
  integer :: ifr
  character(len=40) :: nome
  write(nome, "('./fields/w.',i3.3)") ifr
  call write_w(u, nome)
end subroutine writef

subroutine writef2(u, ifr, jfr)
  use gturbo_state
  implicit none
  real(8) :: u(nx + 2, nx)
  integer :: ifr, jfr
  character(len=40) :: nome
  write(nome, "('./fields/w.',i3.3,'.',i3.3)") ifr, jfr
  call write_w(u, nome)
end subroutine writef2

subroutine write_w(u, nome)
  use gturbo_state
  implicit none
  real(8) :: u(nx + 2, nx)
  character(len=*) :: nome
  real(4), allocatable :: w(:, :)
  integer :: i, j
  allocate(w(nx, nx))
  do j = 1, nx
    do i = 1, nx
      w(i, j) = real(u(i, j))
    end do
  end do
  open(unit=90, file=nome, form='unformatted', status='replace')
  write(90) w
  close(90)
  deallocate(w)
end subroutine write_w

! FP64 checkpoint (no reference counterpart): the spectral state z(NX+2,NY)
! in double precision, fields/z.NNN, one unformatted record.  A run that
! resumes from it continues bit for bit; the float32 frame w.NNN of
! src/write.f90 rounds the state.
subroutine writez(z, ifr)
  use gturbo_state
  implicit none
  real(8) :: z(nx + 2, nx)
  integer :: ifr
  character(len=40) :: nome
  write(nome, "('./fields/z.',i3.3)") ifr
  open(unit=92, file=nome, form='unformatted', status='replace')
  write(92) z
  close(92)
end subroutine writez

logical function havez(ifr)
  implicit none
  integer :: ifr
  character(len=40) :: nome
  write(nome, "('./fields/z.',i3.3)") ifr
  inquire(file=nome, exist=havez)
end function havez

subroutine readz(z, ifr)
  use gturbo_state
  implicit none
  real(8) :: z(nx + 2, nx)
  integer :: ifr
  character(len=40) :: nome
  write(nome, "('./fields/z.',i3.3)") ifr
  open(unit=92, file=nome, form='unformatted', action='read', status='old')
  read(92) z
  close(92)
end subroutine readz
