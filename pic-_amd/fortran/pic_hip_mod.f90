! ISO_C_BINDING interface to libpichip.so (include/pic_hip.h).
!
! Fortran callers of the PIC-IFE_GEC time loop bind the GPU drop-ins through this
! module. The derived types are layout-identical to the reference's TYPE SPARSE
! (IFE_MAIN_PARAM.f90:8-11) and TYPE ParticleOne (MCC_jw/code/base/ParticleOne.f90:4-6).
! The reference's own arrays go through without a copy: TYPE SPARSE (no type-bound
! procedures) through the TYPE(*) dummies of MY_JPCG_Solver, and PB%PO(1:NPar) -- a
! non-interoperable ParticleOne with type-bound procedures, which a TYPE(*) dummy may not
! take, but the same storage (10 REAL(8), one INTEGER, 88-byte stride) -- by its address,
! pic_load_bundle_loc(isp, c_loc(PB%PO(1)), npar) and pic_store_bundle_loc (PB a TARGET).
module pic_hip_mod
    use iso_c_binding
    implicit none

    integer(c_int), parameter :: PIC_OK = 0

    type, bind(c) :: pic_sparse
        real(c_double) :: K
        integer(c_int32_t) :: JCOL, SROW
    end type pic_sparse

    type, bind(c) :: pic_particle_one
        real(c_double) :: X, Y, Z, Vx, Vy, Vz, Ax, Ay, Az, Wq
        integer(c_int32_t) :: Location, pad_
    end type pic_particle_one

    type, bind(c) :: pic_object
        integer(c_int32_t) :: shape, conductor
        real(c_double) :: dimensions(2)
        real(c_double) :: locations(2, 4)      ! locations(d,k) holds the reference's Locations(k,d)
        real(c_double) :: phi, eps
    end type pic_object

    type, bind(c) :: pic_domain_config
        integer(c_int32_t) :: nnx, nny, delta, n_objects
        real(c_double) :: xmin, xmax, ymin, ymax
        type(pic_object) :: objects(8)
        integer(c_int32_t) :: wall_dirichlet(4)
        real(c_double) :: wall_phi(4)
        real(c_double) :: n_ref, T_ref_eV, dt
    end type pic_domain_config

    type, bind(c) :: pic_species_params
        real(c_double) :: charge_C, mass_kg, weight, init_temperature_K
    end type pic_species_params

    type, bind(c) :: pic_gas_params
        real(c_double) :: mass_kg, temperature_K, density_m3
    end type pic_gas_params

    ! pic_injection (pic.inp injection block, Input_2D.f90:333-344)
    type, bind(c) :: pic_injection
        real(c_double) :: density_m3, temp_eV
        real(c_double) :: drift_ms(3)
        real(c_double) :: r_start, y_lo, y_hi
    end type pic_injection

    interface
        integer(c_int) function pic_init(device) bind(c, name="pic_init")
            import :: c_int
            integer(c_int), value :: device
        end function
        integer(c_int) function pic_finalize() bind(c, name="pic_finalize")
            import :: c_int
        end function
        type(c_ptr) function pic_last_error() bind(c, name="pic_last_error")
            import :: c_ptr
        end function
        integer(c_int) function pic_set_precision(bits, solver_mode) bind(c, name="pic_set_precision")
            import :: c_int
            integer(c_int), value :: bits, solver_mode
        end function
        integer(c_int) function pic_setup_domain(cfg) bind(c, name="pic_setup_domain")
            import :: pic_domain_config, c_int
            type(pic_domain_config), intent(in) :: cfg
        end function
        integer(c_int) function pic_set_species(isp, p, capacity) bind(c, name="pic_set_species")
            import :: pic_species_params, c_int, c_int64_t
            integer(c_int), value :: isp
            type(pic_species_params), intent(in) :: p
            integer(c_int64_t), value :: capacity
        end function
        integer(c_int) function pic_load_maxwellian(isp, ppc, temp_eV, seed) bind(c, name="pic_load_maxwellian")
            import :: c_int, c_int32_t, c_double, c_int64_t
            integer(c_int), value :: isp
            integer(c_int32_t), value :: ppc
            real(c_double), value :: temp_eV
            integer(c_int64_t), value :: seed
        end function
        integer(c_int) function pic_load_bundle(isp, po, npar) bind(c, name="pic_load_bundle")
            import :: c_int, c_int64_t
            integer(c_int), value :: isp
            type(*), intent(in) :: po(*)          ! pic_particle_one, or any type of that storage without procedures
            integer(c_int64_t), value :: npar
        end function
        integer(c_int) function pic_store_bundle(isp, po, capacity, npar) bind(c, name="pic_store_bundle")
            import :: c_int, c_int64_t
            integer(c_int), value :: isp
            type(*) :: po(*)
            integer(c_int64_t), value :: capacity
            integer(c_int64_t) :: npar
        end function
        integer(c_int) function pic_get_particles(isp, x, y, vx, vy, vz, loc, capacity, n) &
                bind(c, name="pic_get_particles")
            import :: c_int, c_double, c_int32_t, c_int64_t
            integer(c_int), value :: isp
            real(c_double) :: x(*), y(*), vx(*), vy(*), vz(*)
            integer(c_int32_t) :: loc(*)
            integer(c_int64_t), value :: capacity
            integer(c_int64_t) :: n
        end function
        integer(c_int) function pic_get_counters(isp, counters, n) bind(c, name="pic_get_counters")
            import :: c_int, c_int64_t
            integer(c_int), value :: isp
            integer(c_int64_t) :: counters(*)
            integer(c_int), value :: n
        end function
        ! files in the reference's formats; names are NUL-terminated (trim(name)//c_null_char)
        integer(c_int) function pic_DumpParticleBundle(isp, path) bind(c, name="pic_DumpParticleBundle")
            import :: c_int, c_char
            integer(c_int), value :: isp
            character(kind=c_char) :: path(*)
        end function
        integer(c_int) function pic_LoadParticleBundle(isp, path, npar) bind(c, name="pic_LoadParticleBundle")
            import :: c_int, c_char, c_int64_t
            integer(c_int), value :: isp
            character(kind=c_char) :: path(*)
            integer(c_int64_t) :: npar
        end function
        integer(c_int) function pic_Dump_2D(lap, dir) bind(c, name="pic_Dump_2D")
            import :: c_int, c_char
            integer(c_int), value :: lap
            character(kind=c_char) :: dir(*)
        end function
        integer(c_int) function pic_Restart_2D(lap, dir) bind(c, name="pic_Restart_2D")
            import :: c_int, c_char
            integer(c_int), value :: lap
            character(kind=c_char) :: dir(*)
        end function
        integer(c_int) function pic_Output_To_Tecplot_IJ(it, dir) bind(c, name="pic_Output_To_Tecplot_IJ")
            import :: c_int, c_char
            integer(c_int), value :: it
            character(kind=c_char) :: dir(*)
        end function
        integer(c_int) function pic_set_azimuth(isp, n, th) bind(c, name="pic_set_azimuth")
            import :: c_int, c_int64_t, c_double
            integer(c_int), value :: isp
            integer(c_int64_t), value :: n
            real(c_double) :: th(*)
        end function
        integer(c_int) function pic_get_azimuth(isp, th, capacity, n) bind(c, name="pic_get_azimuth")
            import :: c_int, c_int64_t, c_double
            integer(c_int), value :: isp
            real(c_double) :: th(*)
            integer(c_int64_t), value :: capacity
            integer(c_int64_t) :: n
        end function
        integer(c_int) function pic_get_counts(isp, nlive, nslots, counters) bind(c, name="pic_get_counts")
            import :: c_int, c_int64_t
            integer(c_int), value :: isp
            integer(c_int64_t) :: nlive, nslots, counters(8)
        end function
        integer(c_int) function pic_MoveAll(isp, delta, PushFlag) bind(c, name="pic_MoveAll")
            import :: c_int
            integer(c_int), value :: isp, delta, PushFlag
        end function
        integer(c_int) function pic_GetParChg_2D(delta) bind(c, name="pic_GetParChg_2D")
            import :: c_int
            integer(c_int), value :: delta
        end function
        integer(c_int) function pic_IFE_Solve_2D(Rho, Phi, nnx, nny, delta, time) bind(c, name="pic_IFE_Solve_2D")
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: Rho, Phi
            integer(c_int), value :: nnx, nny, delta
            real(c_double), value :: time
        end function
        integer(c_int) function pic_last_solve_info(iterations, relres, outer) bind(c, name="pic_last_solve_info")
            import :: c_int32_t, c_double, c_int
            integer(c_int32_t) :: iterations, outer
            real(c_double) :: relres
        end function
        integer(c_int) function pic_set_pcg_params(tol, maxit) bind(c, name="pic_set_pcg_params")
            import :: c_double, c_int32_t, c_int
            real(c_double), value :: tol
            integer(c_int32_t), value :: maxit
        end function
        integer(c_int) function pic_GetEField_SIDG_PPR() bind(c, name="pic_GetEField_SIDG_PPR")
            import :: c_int
        end function
        integer(c_int) function pic_MCC() bind(c, name="pic_MCC")
            import :: c_int
        end function
        ! InjectBeams_2D(it, delta, PB), IMPIC/InjectBeams_2D.f90:1
        integer(c_int) function pic_set_injection(isp, p) bind(c, name="pic_set_injection")
            import :: pic_injection, c_int
            integer(c_int), value :: isp
            type(pic_injection), intent(in) :: p
        end function
        integer(c_int) function pic_InjectBeams_2D(it, delta, added) bind(c, name="pic_InjectBeams_2D")
            import :: c_int, c_int64_t
            integer(c_int), value :: it, delta
            integer(c_int64_t), intent(out) :: added(4)
        end function
        integer(c_int) function pic_MCCBundleInit(gas, electron_species, ion_species) bind(c, name="pic_MCCBundleInit")
            import :: pic_gas_params, c_int
            type(pic_gas_params), intent(in) :: gas
            integer(c_int), value :: electron_species, ion_species
        end function
        integer(c_int) function pic_set_run_params(sort_every, seed, mcc_enabled) bind(c, name="pic_set_run_params")
            import :: c_int32_t, c_int64_t, c_int
            integer(c_int32_t), value :: sort_every, mcc_enabled
            integer(c_int64_t), value :: seed
        end function
        integer(c_int) function pic_sort_particles() bind(c, name="pic_sort_particles")
            import :: c_int
        end function
        integer(c_int) function pic_set_step(step) bind(c, name="pic_set_step")
            import :: c_int64_t, c_int
            integer(c_int64_t), value :: step
        end function
        integer(c_int) function pic_get_field(which, dst, n) bind(c, name="pic_get_field")
            import :: c_int, c_double, c_int64_t
            integer(c_int), value :: which
            real(c_double) :: dst(*)
            integer(c_int64_t), value :: n
        end function
        integer(c_int) function pic_MY_JPCG_Solver(matrix, R, X, M, NZ, N, ITER) bind(c, name="pic_MY_JPCG_Solver")
            import :: c_double, c_int
            type(*), intent(in) :: matrix(*)      ! pic_sparse or the reference's TYPE SPARSE (16 B each)
            real(c_double) :: R(*), X(*)
            real(c_double), intent(in) :: M(*)
            integer(c_int), value :: NZ, N
            integer(c_int) :: ITER
        end function
        integer(c_int) function pic_MY_JPCG_Solver_noalloc(matrix, R, X, M, NZ, N, ITER) bind(c, name="pic_MY_JPCG_Solver_noalloc")
            import :: c_double, c_int
            type(*), intent(in) :: matrix(*)      ! pic_sparse or the reference's TYPE SPARSE (16 B each)
            real(c_double) :: R(*), X(*)
            real(c_double), intent(in) :: M(*)
            integer(c_int), value :: NZ, N
            integer(c_int) :: ITER
        end function
        integer(c_int) function pic_Get_ParticleEField_PPR(n, x, y, EleIndex, Ex, Ey) &
                bind(c, name="pic_Get_ParticleEField_PPR")
            import :: c_int, c_int64_t, c_int32_t, c_double
            integer(c_int64_t), value :: n
            real(c_double), intent(in) :: x(*), y(*)
            integer(c_int32_t), intent(in) :: EleIndex(*)
            real(c_double) :: Ex(*), Ey(*)
        end function
        integer(c_int) function pic_InjectPositioning(n, x, y, n_element) bind(c, name="pic_InjectPositioning")
            import :: c_int, c_int64_t, c_int32_t, c_double
            integer(c_int64_t), value :: n
            real(c_double), intent(in) :: x(*), y(*)
            integer(c_int32_t) :: n_element(*)
        end function
        integer(c_int) function pic_ParticlePositioning(n, x, y, n_element, traverse_flag) &
                bind(c, name="pic_ParticlePositioning")
            import :: c_int, c_int64_t, c_int32_t, c_double
            integer(c_int64_t), value :: n
            real(c_double), intent(in) :: x(*), y(*)
            integer(c_int32_t) :: n_element(*), traverse_flag(*)
        end function
        integer(c_int) function pic_generate_Elementmap(Element_Map, repeat_refinement) &
                bind(c, name="pic_generate_Elementmap")
            import :: c_int, c_int32_t
            integer(c_int32_t) :: Element_Map(*)
            integer(c_int), value :: repeat_refinement
        end function
    end interface

contains

    ! STOP with the library's message on a failed call (the reference STOPs on errors).
    ! pic_load_bundle / pic_store_bundle on the address of the first particle record of an
    ! array with pic_particle_one's storage (the reference's PB%PO(1) with C_LOC)
    integer(c_int) function pic_load_bundle_loc(isp, po_first, npar)
        integer(c_int), intent(in) :: isp
        type(c_ptr), intent(in) :: po_first
        integer(c_int64_t), intent(in) :: npar
        type(pic_particle_one), pointer :: po(:)
        call c_f_pointer(po_first, po, [max(npar, 1_c_int64_t)])
        pic_load_bundle_loc = pic_load_bundle(isp, po, npar)
    end function pic_load_bundle_loc

    integer(c_int) function pic_store_bundle_loc(isp, po_first, capacity, npar)
        integer(c_int), intent(in) :: isp
        type(c_ptr), intent(in) :: po_first
        integer(c_int64_t), intent(in) :: capacity
        integer(c_int64_t), intent(out) :: npar
        type(pic_particle_one), pointer :: po(:)
        call c_f_pointer(po_first, po, [max(capacity, 1_c_int64_t)])
        pic_store_bundle_loc = pic_store_bundle(isp, po, capacity, npar)
    end function pic_store_bundle_loc

    subroutine pic_check(rc, where)
        integer(c_int), intent(in) :: rc
        character(*), intent(in) :: where
        character(kind=c_char), pointer :: msg(:)
        integer :: k
        if (rc == PIC_OK) return
        call c_f_pointer(pic_last_error(), msg, [512])
        write(*, '(a,a,a,i0,a)', advance='no') 'pic_hip error in ', where, ' (status ', rc, '): '
        do k = 1, 512
            if (msg(k) == c_null_char) exit
            write(*, '(a)', advance='no') msg(k)
        end do
        write(*, *)
        stop 1
    end subroutine pic_check

end module pic_hip_mod
