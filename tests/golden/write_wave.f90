! Writes one small orbital in the wavefunction file layout of BigDFT's
! writeonewave (bigdft/src/modules/io.f90:2639-2732): the binary file (Fortran
! unformatted sequential records) and the plain-text file, with points in
! key order.  Used by tests/test_wavefunction_io.py as an independent
! producer of the layout (compiled with the image's Fortran compiler).
! Values are exact in binary floating point (small integers over powers of
! two), so they need no mathematical library.
program write_wave
  implicit none
  integer, parameter :: n1 = 5, n2 = 3, n3 = 4, nat = 2, iorb = 7
  character(len=512) :: fbin, ftxt
  integer :: i1, i2, i3, k, iat, nc, nf, j
  real(8) :: rxyz(3, nat), hx, hy, hz, ev, tt, t(7)

  call get_command_argument(1, fbin)
  call get_command_argument(2, ftxt)
  hx = 0.45d0
  hy = 0.4d0
  hz = 0.5d0
  ev = -0.3125d0
  rxyz = reshape([1.d0, 2.d0, 3.d0, -0.5d0, 0.25d0, 1.125d0], [3, nat])
  nc = 0
  nf = 0
  do i3 = 0, n3
     do i2 = 0, n2
        do i1 = 0, n1
           if (coarse(i1, i2, i3)) nc = nc + 1
           if (fine(i1, i2, i3)) nf = nf + 1
        end do
     end do
  end do

  open(10, file=trim(fbin), form='unformatted', access='sequential', status='replace')
  write(10) iorb, ev
  write(10) hx, hy, hz
  write(10) n1, n2, n3
  write(10) nat
  do iat = 1, nat
     write(10) (rxyz(j, iat), j = 1, 3)
  end do
  write(10) nc, nf
  k = 0
  do i3 = 0, n3
     do i2 = 0, n2
        do i1 = 0, n1
           if (coarse(i1, i2, i3)) then
              k = k + 1
              tt = vc(k)
              write(10) i1, i2, i3, tt
           end if
        end do
     end do
  end do
  k = 0
  do i3 = 0, n3
     do i2 = 0, n2
        do i1 = 0, n1
           if (fine(i1, i2, i3)) then
              k = k + 1
              do j = 1, 7
                 t(j) = vf(k, j)
              end do
              write(10) i1, i2, i3, t(1), t(2), t(3), t(4), t(5), t(6), t(7)
           end if
        end do
     end do
  end do
  close(10)

  open(11, file=trim(ftxt), form='formatted', status='replace')
  write(11, *) iorb, ev
  write(11, *) hx, hy, hz
  write(11, *) n1, n2, n3
  write(11, *) nat
  do iat = 1, nat
     write(11, '(3(1x,e24.17))') (rxyz(j, iat), j = 1, 3)
  end do
  write(11, *) nc, nf
  k = 0
  do i3 = 0, n3
     do i2 = 0, n2
        do i1 = 0, n1
           if (coarse(i1, i2, i3)) then
              k = k + 1
              write(11, '(3(i4),1x,e19.12)') i1, i2, i3, vc(k)
           end if
        end do
     end do
  end do
  k = 0
  do i3 = 0, n3
     do i2 = 0, n2
        do i1 = 0, n1
           if (fine(i1, i2, i3)) then
              k = k + 1
              write(11, '(3(i4),7(1x,e17.10))') i1, i2, i3, (vf(k, j), j = 1, 7)
           end if
        end do
     end do
  end do
  close(11)

contains

  logical function coarse(a, b, c)
    integer, intent(in) :: a, b, c
    coarse = mod(a + 2 * b + 3 * c, 4) /= 0
  end function coarse

  logical function fine(a, b, c)
    integer, intent(in) :: a, b, c
    fine = coarse(a, b, c) .and. mod(a * b + c, 3) == 0
  end function fine

  real(8) function vc(kk)
    integer, intent(in) :: kk
    vc = dble(mod(kk * 37, 101) - 50) / 64.0d0 * 2.0d0**(mod(kk, 5) - 2)
  end function vc

  real(8) function vf(kk, jj)
    integer, intent(in) :: kk, jj
    vf = dble(mod((7 * kk + jj) * 53, 97) - 48) / 32.0d0 * 2.0d0**(mod(kk + jj, 7) - 3)
  end function vf

end program write_wave
