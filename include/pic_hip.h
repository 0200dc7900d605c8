/*
 * pic_hip.h -- C-ABI of the MI355X (gfx950) per-timestep IFE-PIC hot path.
 *
 * Drop-in boundary for the serial Fortran subroutines of PIC-IFE_GEC that form one
 * explicit time step (Main_IFE_Test_2.f90:329-341 plus MCC_jw's MCC). Every entry
 * point below names the reference interface it replaces (path:line, relative to the
 * reference checkout). The Fortran side binds these through ISO_C_BINDING (module
 * pic_hip_mod in pic-_amd/fortran/pic_hip_mod.f90).
 *
 * Conventions
 *   - Plain C types only: pointers, sizes, doubles, 32/64-bit integers.
 *   - Element, node and location ids crossing this boundary are 1-based like the
 *     Fortran arrays they mirror; array-valued arguments are Fortran column-major.
 *   - Every function returns an int status (PIC_OK = 0) and never aborts;
 *     pic_last_error() gives the message of the last failure.
 *   - Single host thread, one HIP stream, no re-entrancy (the reference is serial).
 *   - The state the reference keeps in module globals (IFE_Data, Domain_2D, Field_2D,
 *     Particle_2D, ModuleMCCInterface::ParticleGlobal) lives on the device, owned by
 *     this library. Host copies are made only on explicit pic_get_* / pic_store_* calls.
 */
#ifndef PIC_HIP_H
#define PIC_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status codes */
enum {
    PIC_OK = 0,
    PIC_ERR_ARG = 1,           /* bad argument */
    PIC_ERR_STATE = 2,         /* call order (e.g. no mesh yet) */
    PIC_ERR_HIP = 3,           /* HIP runtime error */
    PIC_ERR_CAPACITY = 4,      /* particle buffer capacity exceeded (nothing written) */
    PIC_ERR_PCG_MAXITER = 5,   /* MY_PCG IERR=1 */
    PIC_ERR_PCG_BREAKDOWN = 6, /* MY_PCG IERR=2/3 (rho0 == 0 or p.Ap == 0) */
    PIC_ERR_NONFINITE = 7,     /* non-finite particle state met in a kernel */
    PIC_ERR_POSITION = 8,      /* located element out of range */
    PIC_ERR_UNSUPPORTED = 9,   /* input the GPU path does not implement */
    PIC_ERR_BOUNDARY = 10,     /* boundary loop did not settle (reference: PAUSE/STOP) */
    PIC_ERR_NO_DEVICE = 11,    /* no HIP device / library built without GPU code */
    PIC_ERR_OVERFLOW = 12      /* fixed-point deposition accumulator out of range */
};

/* TYPE SPARSE, IFE_MAIN_PARAM.f90:8-11: value K, column JCOL, and the CSR row pointer
 * stored in SROW of entries 1..N+1 of the same array (Sparse_Structure.f90:163-175). */
typedef struct { double K; int32_t JCOL; int32_t SROW; } pic_sparse;

/* TYPE ParticleOne, MCC_jw/code/base/ParticleOne.f90:4-6 (88 bytes with padding). */
typedef struct {
    double X, Y, Z, Vx, Vy, Vz, Ax, Ay, Az, Wq;
    int32_t Location;
    int32_t pad_;
} pic_particle_one;

/* Subset of TYPE ObjectType, Object_Data_2D.f90:30-80 used on the hot path:
 * shape 1 = circle (Locations(1,:) centre, Dimensions(1) radius),
 * shape 3 = box (Locations(1,:) lower-left, Locations(2,:) upper-right). */
typedef struct {
    int32_t shape;
    int32_t conductor;        /* 1: fixed-potential conductor (interior nodes Dirichlet at phi) */
    double dimensions[2];
    double locations[4][2];   /* C order: locations[k][d] == Fortran Locations(k+1,d+1) */
    double phi;               /* normalised potential of a conductor */
    double eps;               /* relative permittivity (Global_Beta of the object region); an
                                 object with eps != 1 must be a circle on square elements: its
                                 cut elements take the bilinear IFE element
                                 (Generate_stiffness_matrix_local_IFE.f90, IFE_RHS_2D.f90), a
                                 conductor keeping its inside nodes Dirichlet */
} pic_object;

/* Synthetic domain description (what INPUT/mesh.inp + object.inp + pic.inp carry,
 * Input_2D.f90:62-347, for the paths this library implements). */
typedef struct {
    int32_t nnx, nny;          /* nodes in x and y (mesh.inp line 3) */
    int32_t delta;             /* 0: Cartesian, 1: axisymmetric r-z (ife.inp line 12) */
    int32_t n_objects;         /* <= 8 */
    double xmin, xmax, ymin, ymax;
    pic_object objects[8];
    int32_t wall_dirichlet[4]; /* left, right, bottom, top: 1 = Dirichlet, 0 = natural */
    double wall_phi[4];        /* normalised Dirichlet value per wall */
    double n_ref;              /* reference density, m^-3 (pic.inp line 5) */
    double T_ref_eV;           /* reference temperature, eV (pic.inp line 6) */
    double dt;                 /* normalised time step (pic.inp line 12) */
} pic_domain_config;

/* One species (SpecyOne + the Particle_2D normalisation of Input_2D.f90:479-483). */
typedef struct {
    double charge_C;           /* qs before normalisation, C */
    double mass_kg;            /* xm before normalisation, kg */
    double weight;             /* real particles per macro-particle (affp_bjw) */
    double init_temperature_K; /* SpecyOne%InitTemperature (kappa re-emission) */
} pic_species_params;

/* Background gas and the analytic cross sections. */
typedef struct {
    double mass_kg;
    double temperature_K;
    double density_m3;
} pic_gas_params;

/* ------------------------------------------------------------ library lifetime */
int  pic_init(int device);           /* select device, create stream; idempotent */
int  pic_finalize(void);             /* free everything */
const char* pic_last_error(void);
int  pic_device_count(int* n);

/* Precision of the particle state (32 or 64) and of the field solve
 * (0 = FP64 Jacobi-PCG, the reference semantics; 1 = FP32 inner PCG inside FP64
 * iterative refinement; 2 = as 1 with the learned flexible CG as the inner solver once
 * pic_learn_setup has run, falling back to the plain inner CG on stagnation; 3 = FP64 PCG
 * with a multicolour IC(0) preconditioner, a comparison solver with no counterpart in the
 * reference, to the same relative residual). Must precede pic_setup_domain / pic_set_species. */
int  pic_set_precision(int particle_bits, int solver_mode);

/* ------------------------------------------------------------- setup (host FP64)
 * Builds what Setup_IFE_Mesh_2D (Setup_IFE_Mesh_2D.f90:1-660), Partition_Cube_2D,
 * Classify_Nodes_2D (Classify_Nodes_2D.f90:33-80), SetupGrids_2D_QLL cell volumes
 * (SetupGrids_2D_QLL.f90:196-395), IFE_Start_2D stiffness + Jacobi diagonal
 * (IFE_Start_2D.f90:313-483), the per-step RHS operators of Global_RHS_PDE_2D /
 * Global_RHS_EBC_2D_bjw (as a precomputed matrix and constant vector) and the PPR
 * point patches of Get_PointPatch (Get_PointPatch.f90:83-590) produce, and uploads
 * them to the device. Uniform rectangular Q1 meshes without refinement (the shipped
 * IDG_inf.inp: repeat_refinement = 0); elements cut by a dielectric circle carry IFE
 * element matrices. PIC_ERR_UNSUPPORTED: a dielectric box, a dielectric on non-square
 * elements, an element touching two objects or cut through two diagonal vertices.
 * A new domain starts a new run: species, MCC tables, the learned network and the step
 * counter are reset (pic_set_step sets the counter for a restart). */
int  pic_setup_domain(const pic_domain_config* cfg);

/* Sizes of the set-up problem. */
int  pic_get_sizes(int32_t* nnodes, int32_t* nelem, int32_t* nunknowns, int64_t* nnz_A, int64_t* nnz_B);

/* Export set-up arrays for tests and for Fortran callers that keep module copies.
 * which: see PIC_SETUP_* below. n = element count of dst. */
enum {
    PIC_SETUP_HP = 1,            /* double[2*nnodes]  HP(2,:) */
    PIC_SETUP_HT = 2,            /* int32 [5*nelem]   HT(5,:), 1-based node ids, HT(5)=0 */
    PIC_SETUP_ELEMENT_INDEX = 3, /* int32 [nelem]     >0 interface id, -1 vacuum, -2 object */
    PIC_SETUP_NODE_INDEX = 4,    /* int32 [nnodes]    -1 vacuum, -2 object */
    PIC_SETUP_NODE_TYPE = 5,     /* int32 [2*nnodes]  node_type(2,:): row 2 unknown id >0 or <0 */
    PIC_SETUP_CELL_VOLUME = 6,   /* double[2*nnodes]  Cell_Volume_zwz(2,:) */
    PIC_SETUP_A_ROWPTR = 7,      /* int32 [N+1] 0-based CSR of the stiffness on unknowns */
    PIC_SETUP_A_COL = 8,         /* int32 [nnz_A] 0-based */
    PIC_SETUP_A_VAL = 9,         /* double[nnz_A] */
    PIC_SETUP_A_DIAG = 10,       /* double[N] */
    PIC_SETUP_B_ROWPTR = 11,     /* int32 [N+1] rhs = B*rho - c, columns = all nodes */
    PIC_SETUP_B_COL = 12,
    PIC_SETUP_B_VAL = 13,
    PIC_SETUP_RHS_CONST = 14,    /* double[N] c = rhs_FIX + rhs_EBC */
    PIC_SETUP_DIRICHLET_PHI = 15,/* double[nnodes] phi on Dirichlet nodes, 0 elsewhere */
    PIC_SETUP_POINT_PATCH = 16,  /* int32 [nnodes*50] point_patch(nnodes,50), Fortran order */
    PIC_SETUP_UNKNOWN_NODE = 17  /* int32 [N] 1-based node id of each unknown */
};
int  pic_get_setup_array(int which, void* dst, int64_t n);

/* --------------------------------------------------------------------- species
 * ParticleBundle initialisation (AllInitializationParticleBundleIFE,
 * MCCInterface.f90:48-214): capacity is the SoA slot count kept on the device. */
int  pic_set_species(int isp, const pic_species_params* p, int64_t capacity);

/* Uniform positions per cell and Maxwellian (Box-Muller) velocities at temp_eV from
 * the Philox stream (seed, species): ppc particles in every element, generated in
 * element order. Replaces the RANDOM_NUMBER loaders of MCCInterface.f90:146-188. On an
 * axisymmetric domain the radius is uniform in r^2 over the cell, r = sqrt(r0^2 + (r1 +
 * r0)(r1 - r0) R), and the azimuth uniform on [0, 2 pi) (ParticleOne.f90:97-101). */
int  pic_load_maxwellian(int isp, int32_t ppc, double temp_eV, uint64_t seed);
/* The same load with the velocity law the reference selects by MaxKappa
 * (MCCInterface.f90:168-176): kind 1 the Maxwellian above; kind 0 VelKappaInit
 * (ParticleOne.f90:154-204: kappa = 2, Vx by rejection, Vy = Vz = 0); kind 2 VelPolyInit
 * (ParticleOne.f90:206-245, gamma = 3: Vx uniform on +-sqrt(3) s, Vy = Vz = 0). For kinds 0
 * and 2, Vx is stored as the reference draws it, in units of sqrt(2 kB 11605 K / Me) (its
 * rescale to v_ref is commented out), from the stream (seed; slot, 0, load-velocity purpose). */
int  pic_load_distribution(int isp, int32_t ppc, double temp_eV, uint64_t seed, int kind);

/* Load / store a ParticleBundle in the reference's AoS record layout
 * (ParticleBundle%PO(1:NPar), ParticleBundle.f90:7-43). Load converts to device SoA;
 * Location (1-based element) must be the element GetParChg_2D stores for the position
 * (PIC_ERR_POSITION otherwise, nothing loaded). */
int  pic_load_bundle(int isp, const pic_particle_one* po, int64_t npar);
int  pic_store_bundle(int isp, pic_particle_one* po, int64_t capacity, int64_t* npar);

/* SoA load/store (x, y, vx, vy, vz in normalised units; loc 1-based, may be NULL
 * on load: then every particle is located as GetParChg_2D would; a given loc must equal
 * that element, PIC_ERR_POSITION otherwise). Store writes the live particles in slot
 * order. */
int  pic_load_particles(int isp, int64_t n, const double* x, const double* y,
                        const double* vx, const double* vy, const double* vz,
                        const int32_t* loc);
int  pic_get_particles(int isp, double* x, double* y, double* vx, double* vy, double* vz,
                       int32_t* loc, int64_t capacity, int64_t* n);

/* The azimuth PO%Z (ParticleOne.f90:4-6; radians) of an axisymmetric domain (delta = 1),
 * where MoveOne moves in Cartesian coordinates and returns to (r, PO%Z)
 * (MoveOne.f90:246-263). Set: slots 0..n-1 in load order (after pic_load_particles);
 * get: the live particles in the order of pic_get_particles. PIC_ERR_STATE on a Cartesian
 * domain, which keeps no azimuth. pic_load_bundle / pic_store_bundle carry Z through these. */
int  pic_set_azimuth(int isp, int64_t n, const double* th);
int  pic_get_azimuth(int isp, double* th, int64_t capacity, int64_t* n);

/* --------------------------------------------------------------------- files
 * The reference's on-disk formats (little-endian Fortran unformatted records framed by
 * 4-byte lengths where the reference writes FORM='UNFORMATTED').
 * DumpParticleBundle / LoadParticleBundle (MCC_jw ParticleBundle.f90:286-375): records
 * [NPar, NParNormal], [Charge, Mass, Weight], then [X, Y, Z, Vx, Vy, Vz, Ax, Ay, Az, Wq,
 * Location] per particle, X in metres and velocities in m/s (the bundle's PosRes/VelRes:
 * only X is rescaled, Y and Z stay normalised); the load converts back and needs the file's
 * charge, mass and weight to match the species. Dump_2D / Restart_2D (code/PIC/Dump_2D.f90,
 * Restart_2D.f90): <dir>/var<lap:7>dump (list-directed counters) and <dir>/phi<lap:7>dump
 * (Phi). Output_To_Tecplot_IJK_2D option 1: <dir>/field_IJ_<it:6>.dat. */
int  pic_DumpParticleBundle(int isp, const char* path);
int  pic_LoadParticleBundle(int isp, const char* path, int64_t* npar);
int  pic_Dump_2D(int lap, const char* dir);
int  pic_Restart_2D(int lap, const char* dir);
int  pic_Output_To_Tecplot_IJ(int it, const char* dir);
/* The Fortran Ew.d edit descriptor these writers use ([-]0.ddd E+ee, no E past exponent 99),
 * into out (at least w + 1 bytes). Host only. */
int  pic_format_fortran_e(double v, int w, int d, char* out);

/* Live particles, occupied slots, and event counters of a species:
 * counters[0] nLoss(1) left-wall re-emissions, [1] nLoss(2) right-wall absorptions,
 * [2] object absorptions, [3] wall reflections, [4] MCC selected (null incl.),
 * [5] MCC real collisions, [6] particles created by MCC, [7] particles deleted. */
int  pic_get_counts(int isp, int64_t* nlive, int64_t* nslots, int64_t counters[8]);
/* The first n counters of a species: [0..7] as pic_get_counts, [8] particles added by
 * InjectBeams_2D (the wall source, InjectBeams_2D.f90:56-143) since pic_set_species. The MCC
 * creation counter [6] of the electron species counts both products of an ionisation (one
 * electron, one ion), so per step nlive(e) grows by injected + [6]/2 - deleted, nlive(ion)
 * by injected + [6](e)/2 - deleted. */
int  pic_get_counters(int isp, int64_t* counters, int n);
/* MCC events by reaction index 1..nreact of the species' table (exact counts). */
int  pic_get_mcc_events(int isp, int64_t* events, int nmax);

/* ------------------------------------------------------------ per-step (device)
 * MoveAll, MCCInterface.f90:574-640 (+ MoveOne, IMPIC/MoveOne.f90:1-266;
 * Get_ParticleEField_PPR, PPR/Get_ParticleEField_PPR.f90:1-50;
 * AdjustParticleBoundary, MCCInterface.f90:217-534). Species isp (0-based, like
 * ParticleGlobal(isp)); PushFlag must be 0 (explicit mover; the implicit pre/post
 * pushers are not on this path). Objects come from pic_setup_domain. The same
 * kernel also computes each particle's new element and its deposition weights, kept
 * pending until GetParChg_2D commits them (the element seen by the next gather is
 * the one GetParChg_2D stores, exactly as in the reference). delta = 1 is the
 * axisymmetric mover (x axial, y radial, velocities in a fixed Cartesian frame, the
 * azimuth PO%Z updated; MoveOne.f90:94-104, :221-236, :246-263). */
int  pic_MoveAll(int isp, int delta, int PushFlag);

/* GetParChg_2D, code/PIC/GetParChg_2D.f90:1-952: per-particle element lookup
 * (InjectPositioning / ParticlePositioning), charge and kinetic-energy deposition to
 * element corners, corner -> node sums, normalisation by Cell_Volume_zwz. Results stay
 * on the device (pic_get_field). */
int  pic_GetParChg_2D(int delta);

/* IFE_Solve_2D, IFE-Rectangular/IFE_Solve_2D.f90:1-525: rhs = B*rho - c, then the
 * Jacobi-PCG solve warm-started from the current potential. Rho / Phi may be NULL
 * (device-resident rho / phi). A non-NULL Rho (nnodes) replaces the device rho; a
 * non-NULL Phi (nnodes) is the warm start on entry and the result on exit. */
int  pic_IFE_Solve_2D(const double* Rho, double* Phi, int nnx, int nny, int delta, double time);
/* Iterations and final relative residual ||b-Ax||_2/||b||_2 (recomputed in FP64) of the
 * last solve; outer refinement steps for solver_mode 1. */
int  pic_last_solve_info(int32_t* iterations, double* relres, int32_t* outer);
/* Device time (HIP events on the library's stream) of the solver in the last
 * pic_IFE_Solve_2D, RHS assembly excluded. No reference counterpart. */
int  pic_last_solve_time(double* ms);
/* PCG_Tol and PCG_MaxIT (ife.inp lines 6-7, IFE_MAIN_PARAM). */
int  pic_set_pcg_params(double tol, int32_t maxit);

/* GetEField_SIDG_PPR, PPR/GetEfield_SIDG_PPR.f90:1-304: per-node quadratic recovery
 * coef_ppr(1:6,:) and efx = -coef(2), efy = -coef(3). */
int  pic_GetEField_SIDG_PPR(void);

/* MCC, MCC_jw/code/mc/MCCPublic.F90:35-111, against the registered gas for all
 * species with a table; new particles appended as in MCCAddNewParticle (:173-189).
 * MCC ends a time step: on success the step counter (the key of every random stream)
 * advances by one, so a caller of the drop-ins in the reference's order draws new numbers
 * every step, the same numbers pic_step draws. */
int  pic_MCC(void);

/* MCCBundleInit (MCCInit.F90:14-152) for the analytic argon cross sections: builds the
 * cumulative probability tables on the reference's log grids (electrons: 10001 points
 * on log10 E in [-3,3]; ions: 10001 points on [-1,3]) and CollisionRatio. */
int  pic_MCCBundleInit(const pic_gas_params* gas, int electron_species, int ion_species);
/* Register an externally built table (e.g. the reference's own MCCBundle): prob is
 * Probility(nreact, nsigma) column-major, reaction_type the ReactionType codes,
 * threshold in eV. */
int  pic_mcc_set_table(int isp, int nreact, int nsigma, double energy_min_log10,
                       double energy_interval_log10, const double* prob,
                       const int32_t* reaction_type, const double* threshold,
                       double sigma_max, double collision_ratio, int product_species);
int  pic_mcc_get_table(int isp, double* prob, int64_t n, double* sigma_max, double* collision_ratio);

/* Wall-injection source of one species: the injection block of pic.inp as Input_2D.f90:333-344
 * reads it and :489-494 normalises it (density dens0 in m^-3, temperature tmpj in eV, drift
 * vd in m/s), the slab start R_start and the wall's y extent [f_left_wall(2), f_right_wall(2)]
 * in mesh units. density_m3 = 0 removes the source. On an axisymmetric domain (delta = 1)
 * the wall is the disk annulus of radii [y_lo, y_hi] at x = r_start: InjectBeams_2D.f90:59-61
 * counts with PI (y_hi - y_lo)^2 L_ref^2 (the squared radial extent, as written), and PosInit
 * draws r uniform in r^2 and the azimuth PO%Z uniform (ParticleOne.f90:95-98). */
typedef struct {
    double density_m3;
    double temp_eV;
    double drift_ms[3];
    double r_start;
    double y_lo, y_hi;
} pic_injection;
int  pic_set_injection(int isp, const pic_injection* p);
/* InjectBeams_2D(it, delta, PB) (IMPIC/InjectBeams_2D.f90:1-160): for every registered
 * source, add_N = INT(n |b_amb| v_ref dt t_ref (y_hi-y_lo) L_ref / Weight + R) new particles
 * with the inward velocity of a drifting Maxwellian flux (Bird's sampler, :89-106), the
 * transverse ones from Loadv (+ the ion drift), a uniform position in the slab
 * [R_start, R_start + b_amb dt / 2] x [y_lo, y_hi], located by InjectPositioning and appended
 * after the occupied slots. `it` is the random stream's step word. added (optional, 4
 * entries, one per species slot) receives the counts. */
int  pic_InjectBeams_2D(int it, int delta, int64_t* added);

/* One full explicit step (Main_IFE_Test_2.f90:329-341 followed by MCC, the order of
 * MCC_jw/code/OneStep.F90:49-67): MoveAll for every species, GetParChg_2D,
 * IFE_Solve_2D, GetEField_SIDG_PPR, InjectBeams_2D (when a source is registered, at the
 * place of the reference's call, Main_IFE_Test_2.f90:344), MCC (if tables are registered),
 * and the periodic cell sort. nsteps steps back to back. */
int  pic_step(int nsteps);
/* Initial field: GetParChg_2D + IFE_Solve_2D + GetEField_SIDG_PPR (Main:221-233). */
int  pic_initial_field(void);
/* Cell sort / compaction period in steps (0 = never) and the RNG seed of the run. */
int  pic_set_run_params(int32_t sort_every, uint64_t seed, int32_t mcc_enabled);
int  pic_sort_particles(void);
/* Slots per group of the cell sort's layout for the current mesh and precision (256, or 512
 * for FP32 particles on a Cartesian mesh with power-of-two spacing and exact FP32 node lines):
 * inside a full group of 64 K slots the sorted position K l + k sits at slot 64 k + l. No
 * reference counterpart (the reference keeps particles unsorted); for callers that read the
 * slot order. */
int  pic_sort_group(int32_t* slots);
/* Current step counter (the RNG counter word) and its setter (restart). The counter
 * advances after MCC (pic_MCC, or inside pic_step), and also when a species is moved a
 * second time at the same counter value (a caller that runs no MCC). */
int  pic_get_step(int64_t* step);
int  pic_set_step(int64_t step);

/* --------------------------------------------------------------- field access */
enum {
    PIC_F_RHO = 1,       /* rho(nnodes) */
    PIC_F_PHI = 2,       /* Phi(nnodes) */
    PIC_F_EFX = 3,       /* efx(nnodes) */
    PIC_F_EFY = 4,       /* efy(nnodes) */
    PIC_F_COEF_PPR = 5,  /* coef_ppr(6, nnodes) column-major */
    PIC_F_RHO_S = 6,     /* rho_s(nnodes, nspecies) */
    PIC_F_RHO_P = 7,     /* rho_p(nnodes, nspecies) */
    PIC_F_EK_TOT = 8,    /* Ek_tot(nnodes, nspecies) */
    PIC_F_RHS = 9,       /* rhs(nunknowns) of the last solve */
    PIC_F_EK_S = 10      /* Ek_s(nnodes, nspecies): Ek_tot / rho before normalisation (eV per
                          * particle; GetParChg_2D.f90:885-886), 0 where either sum is 0 */
};
int  pic_get_field(int which, double* dst, int64_t n);
int  pic_set_field(int which, const double* src, int64_t n);

/* Per-phase device time of the last pic_step call, ms: [0] push+locate+deposit,
 * [1] deposit finalise, [2] rhs, [3] PCG, [4] PPR, [5] MCC, [6] sort, [7] total; with
 * n >= 12 also kernel statistics: [8] ms in the CG iteration kernel, [9] its launches,
 * [10] ms in the fused push kernel alone (events on both sides of its launch; the wall
 * kernel after it is not included), [11] its launches (HIP events on the library's
 * stream). */
int  pic_get_phase_times(double* ms, int n);

/* --------------------------------------------------- stand-alone drop-ins (host I/O)
 * MY_JPCG_Solver, IFE-Solver/MY_JPCG_Solver.f90:1-51 (-> MY_PCG.f90:1-142): matrix is
 * the reference's SPARSE array of NZ entries (row pointers in SROW(1..N+1), 1-based),
 * R the RHS (on return the final residual), X the initial guess (on return the
 * solution), M the Jacobi diagonal. Solved on the GPU in FP64 with PCG_Tol / PCG_MaxIT.
 * Returns IERR as status (PIC_ERR_PCG_*) and ITER as MY_PCG reports it. */
int  pic_MY_JPCG_Solver(const pic_sparse* matrix, double* R, double* X, const double* M,
                        int NZ, int N, int* ITER);

/* No counterpart in the reference: pic_MY_JPCG_Solver for callers that cannot allocate
 * device memory (work set up around HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc). Same arguments, results and status codes as pic_MY_JPCG_Solver;
 * the matrix, the vectors and the partial sums are staged in a fixed device scratch array
 * of the library (54.6 MB per GPU, held until the process ends) instead of new
 * allocations, and no allocating routine is called. Limits: N <= 262144 unknowns and
 * NZ <= 2359296 stored entries (PIC_ERR_CAPACITY beyond, nothing run). Every call on one
 * GPU shares that GPU's scratch, which keeps the last call's contents: calls that overlap
 * on one GPU (several host threads) race. */
int  pic_MY_JPCG_Solver_noalloc(const pic_sparse* matrix, double* R, double* X, const double* M,
                                int NZ, int N, int* ITER);

/* Get_ParticleEField_PPR (PPR/Get_ParticleEField_PPR.f90:1-50) for n points against the
 * current device coef_ppr; EleIndex 1-based. */
int  pic_Get_ParticleEField_PPR(int64_t n, const double* x, const double* y,
                                const int32_t* EleIndex, double* Ex, double* Ey);

/* InjectPositioning (InjectPositioning.f90:1-24) for n points; n_element 1-based. */
int  pic_InjectPositioning(int64_t n, const double* x, const double* y, int32_t* n_element);

/* ParticlePositioning (ParticlePositioning.f90:1-366) for n points on the unrefined
 * mesh: n_element (in: the element of the rxp/ryp estimate, out: final element) and
 * traverse_flag (1 when the point lies on an inner edge). */
int  pic_ParticlePositioning(int64_t n, const double* x, const double* y,
                             int32_t* n_element, int32_t* traverse_flag);

/* The element lookup of GetParChg_2D.f90:137-163 (estimate, Mod test, then Inject- or
 * ParticlePositioning) for n points. */
int  pic_locate(int64_t n, const double* x, const double* y, int32_t* n_element);

/* generate_Elementmap (generate_Elementmap.f90:1-60): Element_Map of size
 * (2^r (nx-1)) x (2^r (ny-1)), column-major, 1-based element ids. */
int  pic_generate_Elementmap(int32_t* Element_Map, int repeat_refinement);

/* ------------------------------------------- learned preconditioner (no reference counterpart)
 * A multigrid V-cycle on the node grid: a symmetric V(1,1)-cycle over Galerkin levels
 * (bilinear interpolation, full-weighting restriction, exact coarsest solve) whose smoothers
 * on the finest levels are learned convolutions K = D^-1/2 In^T (Q Q^T) In D^-1/2 (In a 3x3
 * convolution from one channel to C; applied as the collapsed 81 tap-pair weights in FP32;
 * damped Jacobi on the coarser levels). It is the preconditioner of a flexible CG on the Jacobi-scaled system
 * (solver_mode 2), trained on the GPU by the library's own forward and backward kernels on
 * residual/error pairs (r = A' e) that the plain inner CG harvests: the solve's RHS with its
 * solution correction, and the iterates every `every` chunks of 32 iterations. */
/* Build the network for the current domain: levels down to ~3x3 nodes (at most max_levels,
 * 2..12); the nchannels finest levels (0..4) get learned smoothers with channels[l] in
 * 1..64; initialised to damped Jacobi (the untrained network is the classical
 * V-cycle). */
int  pic_learn_setup(int32_t max_levels, const int32_t* channels, int32_t nchannels, uint64_t seed);
/* Harvest pairs from the plain inner solves into a device ring of `capacity` pairs. */
int  pic_learn_harvest(int32_t enable, int32_t capacity, int32_t every);
/* Adam on the energy-norm loss, minibatches drawn from the ring; loss of the first and
 * the last minibatch (mean per pair). */
int  pic_learn_train(int32_t steps, int32_t batch, double lr, uint64_t seed, double* loss_first, double* loss_last);
/* Share (0..1, default 0.25) of harvested pairs in every training minibatch; the rest are
 * synthetic pairs e = uniform noise after 0..8 damped-Jacobi sweeps, r = A' e. */
int  pic_learn_data_mix(double harvested_share);
/* Mean loss over pairs [0, npairs) of the ring, no update. */
int  pic_learn_eval(int32_t npairs, double* mean_loss);
/* Pairs in the ring, FCG iterations of the last learned solve, fallbacks to the plain CG so
 * far, last training loss. */
int  pic_learn_info(int32_t* dataset, int32_t* last_fcg_iterations, int64_t* fallbacks, double* last_loss);
/* Online fine-tuning inside pic_step (solver_mode 2): every `every` steps (0 = off) that step's
 * solve runs the plain inner CG, harvesting (r, e) pairs into the ring, and after the step's
 * MCC the network takes train_steps Adam steps (minibatch `batch`, learning rate lr) on the
 * ring; the plain CG stays the fallback on stagnation. Info: fine-tunes done, milliseconds spent
 * training (host clock), loss of the last minibatch. */
int  pic_learn_finetune(int32_t every, int32_t train_steps, int32_t batch, double lr, uint64_t seed);
int  pic_learn_finetune_info(int64_t* count, double* train_ms, double* last_loss);
/* z = M r on host arrays of nnodes floats (node numbering). */
int  pic_learn_apply(const float* r, float* z, int64_t n);
int  pic_learn_get_params(float* dst, int64_t n, int64_t* nparams);
int  pic_learn_set_params(const float* src, int64_t n);
/* levels[0] = levels, levels[1] = learned smoother levels Ls; H, W per level (12 entries);
 * C per learned level; parameter offsets (2*Ls+1 entries: In_0..In_{Ls-1} [9][C],
 * Q_0..Q_{Ls-1} [C][C], total). */
int  pic_learn_layout(int32_t* levels, int32_t* H, int32_t* W, int32_t* C, int64_t* offsets);
/* Replace the ring by npairs host pairs (r, e), each nnodes floats. */
int  pic_learn_set_pairs(int32_t npairs, const float* r, const float* e);
int  pic_learn_get_pair(int32_t k, float* r, float* e, double* ea);
/* Summed loss and gradient (no update) over the ring pairs idx[0..npairs). */
int  pic_learn_gradient(int32_t npairs, const int32_t* idx, double* loss, float* grad, int64_t n);

/* ------------------------------------------- learned warm start (no reference counterpart)
 * A convolutional encoder-decoder on the node grid (levels l = 0..L-1 of nodes (I, J) <->
 * (2I, 2J), C channels, 3x3 convolutions, skips between the encoder and decoder levels, no
 * biases and no nonlinearity, so it is linear in its input) maps the Jacobi-scaled residual
 * of the solve's starting potential to a correction of it: for a cold start that residual is
 * the right-hand side of the deposited charge, inside the time loop the change of the charge
 * since the last solve. With solver_mode 2 and the warm start enabled, every IFE_Solve_2D
 * adds the A-norm-optimal multiple of that correction to its starting potential before the
 * learned flexible CG runs (never worse than no warm start in the A-norm). The C -> C
 * convolutions and their weight gradients run on BF16 matrix cores with FP32 accumulation.
 * It trains on the harvest ring's pairs that hold a solve's initial residual (the first pass
 * of a plain solve, its first snapshot), so pic_learn_setup and pic_learn_harvest come first. */
/* Build the network for the current domain: at most `levels` levels (2..8, down to about
 * 5 x 5 nodes), channels 16 or 32, weights from the seed. Disabled until pic_warm_enable. */
int  pic_warm_setup(int32_t levels, int32_t channels, uint64_t seed);
/* Adam on the energy-norm error (y - e)^T A' (y - e) / (e^T A' e) over minibatches of
 * initial-residual pairs; loss of the first and last minibatch (mean per pair). */
int  pic_warm_train(int32_t steps, int32_t batch, double lr, uint64_t seed, double* loss_first, double* loss_last);
/* 0 off; 1 on, the solve replays the application as one captured HIP graph (captured again
 * when its buffers move); 2 on, the same kernels launched one by one. */
int  pic_warm_enable(int32_t enable);
/* Mean loss over the first npairs initial-residual pairs, of the raw output and of its
 * A-norm-optimal multiple (the correction the solve applies). */
int  pic_warm_eval(int32_t npairs, double* mean_loss, double* mean_loss_scaled);
/* y = N(r) on host arrays of nnodes floats (node numbering). */
int  pic_warm_apply(const float* r, float* y, int64_t n);
int  pic_warm_get_params(float* dst, int64_t n, int64_t* nparams);
int  pic_warm_set_params(const float* src, int64_t n);
/* levels, H and W per level (8 entries), channels. Parameters (FP32, flat): lift [9][C],
 * out [9][C], encoder E_1..E_{L-1} [9][C][C], decoder D_0..D_{L-2} [9][C][C] ([tap][in][out],
 * tap t <-> node offset (t/3 - 1, t%3 - 1)). */
int  pic_warm_layout(int32_t* levels, int32_t* H, int32_t* W, int32_t* channels);
/* Summed loss and gradient (no update) over the ring pairs idx[0..npairs). */
int  pic_warm_gradient(int32_t npairs, const int32_t* idx, double* loss, float* grad, int64_t n);
/* Initial-residual pairs in the ring, whether the solve applies the warm start, alpha of the
 * last applied correction, last training loss, corrections applied so far. */
int  pic_warm_info(int32_t* initial_pairs, int32_t* enabled, double* last_alpha, double* last_loss, int64_t* applied);

#ifdef __cplusplus
}
#endif
#endif /* PIC_HIP_H */
