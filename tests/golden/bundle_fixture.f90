! Writes a three-particle particle-bundle file with the record structure of the MCC_jw
! bundle dump: WRITE(unit) NPar, NParNormal / WRITE(unit) Charge, Mass, Weight / one
! WRITE(unit) of each particle record (X, Y, Z, Vx, Vy, Vz, Ax, Ay, Az, Wq as REAL(8) and
! Location as the default INTEGER). Compiled and run by make_bundle_fixture.sh; its output,
! bundle_flang.dat, is the fixture that pins the Fortran unformatted record layout. On a
! 24 x 24 mesh of unit cells at n_ref = 1e21 m^-3, T_ref = 1 eV (L_ref = 2.3509e-7 m) the
! particles lie in elements 97, 193 and 289, the Locations written.
program bundle_fixture
    implicit none
    type :: rec_t
        real(8) :: X, Y, Z, Vx, Vy, Vz, Ax, Ay, Az, Wq
        integer :: Location
    end type rec_t
    type(rec_t) :: po(3)
    integer(4) :: npar, npar_normal, i
    real(8) :: charge, mass, weight
    npar = 3
    npar_normal = 1000
    charge = -1.60217663d-19
    mass = 9.1093837d-31
    weight = 2.5d6
    do i = 1, 3
        po(i)%X = 1.0d-6 * i
        po(i)%Y = 0.25d0 * i
        po(i)%Z = 0.5d0 * i
        po(i)%Vx = 1.0d5 * i
        po(i)%Vy = -2.0d5 * i
        po(i)%Vz = 3.0d5 * i
        po(i)%Ax = 0.0d0
        po(i)%Ay = 0.0d0
        po(i)%Az = 0.0d0
        po(i)%Wq = 0.0d0
        po(i)%Location = 96 * i + 1
    end do
    open(10, form='UNFORMATTED', file='bundle_flang.dat')
    write(10) npar, npar_normal
    write(10) charge, mass, weight
    do i = 1, npar
        write(10) po(i)
    end do
    close(10)
end program bundle_fixture
