! Writes REAL(8) values in the E15.6 edit descriptor (the FORMAT 50 fields of the Tecplot
! field file), each after its bit pattern as a decimal INTEGER(8), for the check of the
! library's Ew.d writer. Compiled and run by make_e15_fixture.sh; output e15_flang.txt.
program e15_fixture
    implicit none
    real(8) :: v(24), x
    integer(8) :: bits, state
    integer :: i
    v = (/ 0.0d0, 1.0d0, -1.0d0, 0.5d0, 1234.5678d0, -9.99999949d-3, 9.9999995d5, 9.9999994d5, &
           1.0d-300, 1.0d300, -2.5d-100, 123456789.0d0, 3.14159265358979d0, 4.9406564584124654d-324, &
           -0.0d0, 1.0d-99, 9.9999996d-100, 0.1d0, 0.9999995d0, 0.99999949d0, -1.0d100, 2.0d-5, 7.0d0, -3.0d-3 /)
    open(10, file='e15_flang.txt')
    do i = 1, size(v)
        bits = transfer(v(i), bits)
        write(10, '(I21, 1X, E15.6)') bits, v(i)
    end do
    state = 12345_8
    do i = 1, 200
        state = mod(state * 1103515245_8 + 12345_8, 2147483648_8)
        x = (dble(state) / 2147483648.0d0 - 0.5d0) * 10.0d0 ** (mod(state, 41_8) - 20)
        bits = transfer(x, bits)
        write(10, '(I21, 1X, E15.6)') bits, x
    end do
    close(10)
end program e15_fixture
