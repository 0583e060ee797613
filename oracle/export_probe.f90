! TEST INFRASTRUCTURE ONLY.  Lets the Fortran runtime itself show what the
! export statements of DataExport2Sub produce, for the tests to compare with
! the oracle's export restatement (oracle/export.py):
!  - the file names built from I5 edits with adjustl/adjustr into
!    character(50) variables (_8_0_2_DataExport2.f90:32-41, 129-141;
!    declarations _KineticMainParams.f90:35-47, 346-358), printed one a line
!    for the (nn, s, f, Qind) given on the command line, then a spin-up name
!    (:706-712);
!  - the bytes of open(status='replace', form='unformatted', access='stream')
!    + write + close of a real(dp) scalar, an integer(dp) scalar, a real(dp)
!    (3, 4) array and a real(dp) (2, 3, 4) array written whole as a(:, :, :)
!    (N2PerpphRTp, d3vCTp), to probe_scalar.bin, probe_int.bin,
!    probe_array.bin, probe_array3.bin in the working directory.
program export_probe
  implicit none
  integer, parameter :: dp = kind(1.d0)
  character(50) :: nnstring, sstring, fstring, Qindstring, expstring, fname
  character(32) :: arg
  integer(kind=dp) :: nn, s, f, Qind, i, j, l
  real(kind=dp) :: x, a(3, 4), b(2, 3, 4)
  integer(kind=dp) :: k

  call get_command_argument(1, arg); read (arg, *) nn
  call get_command_argument(2, arg); read (arg, *) s
  call get_command_argument(3, arg); read (arg, *) f
  call get_command_argument(4, arg); read (arg, *) Qind

  write (nnstring, '(I5)') nn
  write (sstring, '(I5)') s
  write (fstring, '(I5)') f
  write (Qindstring, '(I5)') Qind

  expstring = adjustl(adjustr(nnstring) // '_' // adjustl(adjustr(sstring) // '_' // &
              adjustl(adjustr(fstring) // '_')))
  fname = adjustl(adjustr(expstring) // adjustl(adjustr('TimeTfort.bin')))
  print '(A)', trim(fname)
  fname = adjustl(adjustr(expstring) // adjustl(adjustr('M1Perp2FiltAvrgRTfort.bin')))
  print '(A)', trim(fname)

  expstring = adjustl(adjustr(nnstring) // '_' // adjustl(adjustr(sstring) // '_' // &
              adjustl(adjustr(fstring) // '_' // adjustl(adjustr(Qindstring) // '_'))))
  fname = adjustl(adjustr(expstring) // adjustl(adjustr('NphRTfort.bin')))
  print '(A)', trim(fname)

  ! spin-up files (_8_0_2_DataExport2.f90:706-712): species and flux tube only
  expstring = adjustl(adjustr(sstring) // '_' // adjustl(adjustr(fstring) // '_'))
  fname = adjustl(adjustr(expstring) // adjustl(adjustr('TemperatureOutputRTfort.bin')))
  print '(A)', trim(fname)

  x = 0.1_dp * real(nn, dp)
  k = 7_dp * nn
  do j = 1, 4
    do i = 1, 3
      a(i, j) = real(10 * i + j, dp) + 0.25_dp
    end do
  end do
  open (unit=10, file='probe_scalar.bin', status='replace', form='unformatted', access='stream')
  write (10) x
  close (10)
  open (unit=10, file='probe_int.bin', status='replace', form='unformatted', access='stream')
  write (10) k
  close (10)
  open (unit=10, file='probe_array.bin', status='replace', form='unformatted', access='stream')
  write (10) a(:, :)
  close (10)
  do l = 1, 4
    do j = 1, 3
      do i = 1, 2
        b(i, j, l) = real(100 * i + 10 * j + l, dp) + 0.5_dp
      end do
    end do
  end do
  open (unit=10, file='probe_array3.bin', status='replace', form='unformatted', access='stream')
  write (10) b(:, :, :)
  close (10)
end program export_probe
