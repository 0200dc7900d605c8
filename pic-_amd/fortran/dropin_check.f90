! Check of the binding technique the drop-in shims of INTEGRATION.md rely on, run by
! tests/test_gpu_fortran.py: an array of a non-interoperable derived type with the storage and
! the type-bound procedures of the reference's TYPE ParticleOne (MCC_jw/code/base/
! ParticleOne.f90:4-6: ten REAL(8), one INTEGER) goes through pic_load_bundle_loc /
! pic_store_bundle_loc by C_LOC of its first element, an array with the storage of TYPE SPARSE
! (IFE_MAIN_PARAM.f90:8-11: REAL(8) K, INTEGER JCOL, SROW) through the TYPE(*) dummy of
! pic_MY_JPCG_Solver, and the species index is read through a SpecyOne pointer component as
! PB%SO%SpecyIndex (ParticleBundle.f90:18, SpecyOne.f90:50). The local types below mirror only that storage;
! the reference's own modules do not compile with this image's flang (ControlFlow.f90:36).
module dropin_check_types
    implicit none
    type :: specy_like
        integer(4) :: SpecyIndex = -1, GasIndex = -1
    end type specy_like
    type :: particle_like
        real(8) :: X, Y, Z, Vx, Vy, Vz, Ax, Ay, Az, Wq
        integer :: Location
    contains
        procedure :: speed2 => particle_speed2
    end type particle_like
    type :: bundle_like
        integer(4) :: NPar = 0, NParNormal = 1000
        type(specy_like), pointer :: SO => null()
        type(particle_like), allocatable :: PO(:)
    end type bundle_like
    type :: sparse_like
        real(8) :: K
        integer :: JCOL, SROW
    end type sparse_like
contains
    real(8) function particle_speed2(self)
        class(particle_like), intent(in) :: self
        particle_speed2 = self%Vx**2 + self%Vy**2 + self%Vz**2
    end function particle_speed2
end module dropin_check_types

program dropin_check
    use iso_c_binding
    use pic_hip_mod
    use dropin_check_types
    implicit none
    type(pic_domain_config) :: cfg
    type(pic_species_params) :: sp
    type(bundle_like), target :: PB
    type(specy_like), target :: electrons
    type(sparse_like), allocatable :: matrix(:)
    real(c_double) :: x(8), y(8), vx(8), vy(8), vz(8), R(6), X6(6), M6(6), r0(6)
    integer(c_int32_t) :: loc(8)
    integer(c_int64_t) :: n, npar
    integer(c_int) :: iter
    integer :: i, k, nz
    real(8), parameter :: A(3) = [-1.d0, 4.d0, -1.d0]

    call pic_check(pic_init(0_c_int), 'pic_init')
    call pic_check(pic_set_precision(64_c_int, 0_c_int), 'precision')
    cfg%nnx = 25; cfg%nny = 25; cfg%delta = 0; cfg%n_objects = 0
    cfg%xmin = 0.d0; cfg%xmax = 24.d0; cfg%ymin = 0.d0; cfg%ymax = 24.d0
    cfg%wall_dirichlet = 1; cfg%wall_phi = 0.d0
    cfg%n_ref = 1.d21; cfg%T_ref_eV = 1.d0; cfg%dt = 0.05d0
    call pic_check(pic_setup_domain(cfg), 'setup')
    sp = pic_species_params(-1.6022d-19, 9.1095d-31, 1.d4, 2.d0 * 11605.d0)
    call pic_check(pic_set_species(0_c_int, sp, 100_c_int64_t), 'species')

    ! a bundle of 3 electrons in the reference's layout; Location = the element holding (X, Y)
    electrons%SpecyIndex = 0
    PB%SO => electrons
    PB%NPar = 3
    allocate(PB%PO(PB%NPar))
    do i = 1, 3
        PB%PO(i) = particle_like(4.25d0 * i, 0.5d0 + i, 0.d0, 0.1d0 * i, -0.2d0 * i, 0.3d0 * i, 0.d0, 0.d0, &
                                 0.d0, 0.d0, int(0.5d0 + i) + 1 + int(4.25d0 * i) * 24)
    end do
    call pic_check(pic_load_bundle_loc(int(PB%SO%SpecyIndex, c_int), c_loc(PB%PO(1)), int(PB%NPar, c_int64_t)), &
                   'load_bundle')
    call pic_check(pic_get_particles(0_c_int, x, y, vx, vy, vz, loc, 8_c_int64_t, n), 'get_particles')
    if (n /= 3) stop 'dropin_check: particle count'
    do i = 1, 3
        if (x(i) /= PB%PO(i)%X .or. y(i) /= PB%PO(i)%Y .or. vx(i) /= PB%PO(i)%Vx .or. vy(i) /= PB%PO(i)%Vy &
            .or. vz(i) /= PB%PO(i)%Vz .or. loc(i) /= PB%PO(i)%Location) stop 'dropin_check: bundle values'
        if (abs(PB%PO(i)%speed2() - (vx(i)**2 + vy(i)**2 + vz(i)**2)) > 0.d0) stop 'dropin_check: speed'
    end do
    ! and back into the reference layout
    PB%PO(:)%X = 0.d0
    call pic_check(pic_store_bundle_loc(0_c_int, c_loc(PB%PO(1)), int(size(PB%PO), c_int64_t), npar), 'store_bundle')
    if (npar /= 3 .or. PB%PO(2)%X /= 8.5d0 .or. PB%PO(3)%Location /= loc(3)) stop 'dropin_check: store'

    ! MY_JPCG_Solver on a 6 x 6 tridiagonal system in the reference's SPARSE layout (row
    ! pointers in SROW of entries 1..N+1, Sparse_Structure.f90:163-175)
    nz = 3 * 6 - 2
    allocate(matrix(nz))
    matrix(:)%SROW = 0
    k = 0
    do i = 1, 6
        matrix(i)%SROW = k + 1
        if (i > 1) then
            k = k + 1; matrix(k)%K = A(1); matrix(k)%JCOL = i - 1
        end if
        k = k + 1; matrix(k)%K = A(2); matrix(k)%JCOL = i
        if (i < 6) then
            k = k + 1; matrix(k)%K = A(3); matrix(k)%JCOL = i + 1
        end if
    end do
    matrix(7)%SROW = k + 1
    R = [(dble(i), i = 1, 6)]
    r0 = R
    X6 = 0.d0
    M6 = 4.d0
    call pic_check(pic_set_pcg_params(1.d-12, 100_c_int32_t), 'pcg params')
    call pic_check(pic_MY_JPCG_Solver(matrix, R, X6, M6, int(nz, c_int), 6_c_int, iter), 'MY_JPCG_Solver')
    do i = 1, 6
        if (abs(4.d0 * X6(i) - merge(X6(max(i - 1, 1)), 0.d0, i > 1) - merge(X6(min(i + 1, 6)), 0.d0, i < 6) &
                - r0(i)) > 1.d-10) stop 'dropin_check: solve'
    end do
    write(*, '(a,i4,a)') 'dropin_check: ParticleOne-shaped bundle by address, SPARSE array by TYPE(*); PCG', iter, &
        ' iterations'
    call pic_check(pic_finalize(), 'pic_finalize')
end program dropin_check
