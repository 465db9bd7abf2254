! Writes records of doubles with the edit descriptors of the reference's
! text output, so that the library's formatter can be checked against a
! Fortran run-time library.  Input (stream): int32 n, then n records of six
! doubles.  Output: for each record one line in the snapshot format of
! output.f90 (x, y, p as F10.3; vx, vy as ES10.3; con + 0.5 as F10.5) and
! one line in the probe-series format (F10.3, 1X, F10.3) of its first two
! values.
program fortran_records
    use iso_fortran_env, only: real64, int32
    implicit none
    integer(int32) :: n, k
    real(real64) :: v(6)
    character(len=256) :: fin
    integer :: u
    call get_command_argument(1, fin)
    open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
    read(u) n
    do k = 1, n
        read(u) v
        write(*, '(F10.3,1X,F10.3,1X,F10.3,1X,ES10.3,1X,ES10.3,1X,F10.5)') v
        write(*, '(F10.3,1X,F10.3)') v(1), v(2)
    end do
    close(u)
end program fortran_records
