! Fortran driver replaying the explicit branch of the reference time loop
! (Main_IFE_Test_2.f90:221-233 before the loop, :329-341 inside it) with MCC placed
! after the field solve as in MCC_jw/code/OneStep.F90:49-67, every call going through
! the ISO_C_BINDING drop-ins of pic_hip_mod.
!
! usage: pic_driver <ncells> <ppc> <nsteps> <seed> [fields.bin]
! Prints one line per step: step, live electrons, live ions, PCG iterations, relres;
! with a fifth argument, writes phi then rho (real(8), node order) to that file.
program pic_driver
    use iso_c_binding
    use pic_hip_mod
    implicit none
    type(pic_domain_config) :: cfg
    type(pic_species_params) :: sp_e, sp_i
    type(pic_gas_params) :: gas
    integer :: ncells, ppc, nsteps, it, isp
    integer(c_int64_t) :: seed, nlive(0:1), nslots, counters(8), cap
    integer(c_int32_t) :: iters, outer
    real(c_double) :: relres, L_ref, weight
    real(c_double), allocatable :: phi(:), rho(:)
    character(len=32) :: arg
    character(len=512) :: outpath
    integer :: u
    real(c_double), parameter :: e_charge = 1.6022d-19, e_mass = 9.1095d-31, eV_K = 11605.d0
    real(c_double), parameter :: eps0 = 8.8542d-12, kB = 1.3807d-23

    call get_command_argument(1, arg); read(arg, *) ncells
    call get_command_argument(2, arg); read(arg, *) ppc
    call get_command_argument(3, arg); read(arg, *) nsteps
    call get_command_argument(4, arg); read(arg, *) seed

    ! --- Input_2D / Setup_IFE_Mesh_2D / IFE_Start_2D equivalents
    cfg%nnx = ncells + 1; cfg%nny = ncells + 1; cfg%delta = 0; cfg%n_objects = 1
    cfg%xmin = 0.d0; cfg%xmax = dble(ncells); cfg%ymin = 0.d0; cfg%ymax = dble(ncells)
    cfg%objects(1)%shape = 1; cfg%objects(1)%conductor = 1
    cfg%objects(1)%dimensions = [0.15d0 * ncells, 0.d0]
    cfg%objects(1)%locations = 0.d0
    cfg%objects(1)%locations(1, 1) = 0.5d0 * ncells
    cfg%objects(1)%locations(2, 1) = 0.5d0 * ncells
    cfg%objects(1)%phi = -1.d0; cfg%objects(1)%eps = 1.d0
    cfg%wall_dirichlet = 1; cfg%wall_phi = 0.d0
    cfg%n_ref = 1.d21; cfg%T_ref_eV = 1.d0; cfg%dt = 0.05d0

    call pic_check(pic_init(0_c_int), 'pic_init')
    call pic_check(pic_set_precision(64_c_int, 0_c_int), 'pic_set_precision')
    call pic_check(pic_setup_domain(cfg), 'pic_setup_domain')
    call pic_check(pic_set_pcg_params(1.d-10, 50000_c_int32_t), 'pic_set_pcg_params')

    ! --- AllInitialization: electrons (Te = 2 eV) and Ar+ (0.026 eV), weight for n = n_ref
    L_ref = sqrt(eps0 * kB * eV_K / cfg%n_ref / e_charge / e_charge)
    weight = cfg%n_ref * L_ref * L_ref / dble(ppc)
    cap = int(ncells, c_int64_t) * ncells * ppc * 3 / 2 + 1024
    sp_e = pic_species_params(-e_charge, e_mass, weight, 2.d0 * eV_K)
    sp_i = pic_species_params(e_charge, 73440.d0 * e_mass, weight, 0.026d0 * eV_K)
    call pic_check(pic_set_species(0_c_int, sp_e, cap), 'pic_set_species')
    call pic_check(pic_set_species(1_c_int, sp_i, cap), 'pic_set_species')
    call pic_check(pic_load_maxwellian(0_c_int, int(ppc, c_int32_t), 2.d0, seed), 'pic_load_maxwellian')
    call pic_check(pic_load_maxwellian(1_c_int, int(ppc, c_int32_t), 0.026d0, seed + 1), 'pic_load_maxwellian')
    gas = pic_gas_params(73440.d0 * e_mass, 300.d0, 3.2d20)
    call pic_check(pic_MCCBundleInit(gas, 0_c_int, 1_c_int), 'MCCBundleInit')
    call pic_check(pic_set_run_params(20_c_int32_t, seed, 1_c_int32_t), 'pic_set_run_params')
    call pic_check(pic_set_step(0_c_int64_t), 'pic_set_step')

    ! --- Main_IFE_Test_2.f90:221-233
    call pic_check(pic_GetParChg_2D(0_c_int), 'GetParChg_2D')
    call pic_check(pic_IFE_Solve_2D(c_null_ptr, c_null_ptr, cfg%nnx, cfg%nny, 0_c_int, 0.d0), 'IFE_Solve_2D')
    call pic_check(pic_GetEField_SIDG_PPR(), 'GetEField_SIDG_PPR')

    ! --- Main_IFE_Test_2.f90:278-341 (+ MCC, + periodic cell sort)
    do it = 1, nsteps
        do isp = 0, 1
            call pic_check(pic_MoveAll(int(isp, c_int), 0_c_int, 0_c_int), 'MoveAll')
        end do
        call pic_check(pic_GetParChg_2D(0_c_int), 'GetParChg_2D')
        call pic_check(pic_IFE_Solve_2D(c_null_ptr, c_null_ptr, cfg%nnx, cfg%nny, 0_c_int, dble(it)), 'IFE_Solve_2D')
        call pic_check(pic_GetEField_SIDG_PPR(), 'GetEField_SIDG_PPR')
        call pic_check(pic_MCC(), 'MCC')     ! ends the step: the random-stream counter advances
        if (mod(it, 20) == 0) call pic_check(pic_sort_particles(), 'sort')
        do isp = 0, 1
            call pic_check(pic_get_counts(int(isp, c_int), nlive(isp), nslots, counters), 'counts')
        end do
        call pic_check(pic_last_solve_info(iters, relres, outer), 'solve info')
        write(*, '(a,i6,2i12,i8,es12.4)') 'step', it, nlive(0), nlive(1), iters, relres
    end do

    allocate(phi(cfg%nnx * cfg%nny), rho(cfg%nnx * cfg%nny))
    call pic_check(pic_get_field(2_c_int, phi, int(size(phi), c_int64_t)), 'get phi')
    call pic_check(pic_get_field(1_c_int, rho, int(size(rho), c_int64_t)), 'get rho')
    write(*, '(a,es26.17,a,es26.17)') 'sum_phi', sum(phi), ' sum_rho', sum(rho)
    if (command_argument_count() >= 5) then
        call get_command_argument(5, outpath)
        open(newunit=u, file=trim(outpath), access='stream', form='unformatted', status='replace')
        write(u) phi, rho
        close(u)
    end if
    call pic_check(pic_finalize(), 'pic_finalize')
end program pic_driver
