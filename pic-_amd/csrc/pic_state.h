// Internal state of the MI355X PIC hot-path library (not part of the C-ABI).
//
// Layout in HBM (one GPU):
//   mesh      : node x / y coordinate lines (nnx, nny doubles + float copies),
//               element_index (nelem int32), node_index (nnodes int32)
//   fields    : phi (nnodes f64), efx/efy (nnodes f64 + f32 copies),
//               coef (5 x nnodes f64 + f32: coef_ppr(2:6,:) node-major for the
//               interface-element gather), rho / rho_s / rho_p / Ek_tot (f64)
//   deposit   : per species two int64 fixed-point node accumulators (sum W, sum W*E)
//   particles : per species SoA slots x, y, vx, vy, vz (f32 or f64), cell and
//               cell_next (int32, 0-based element, -1 = dead slot)
//   solver    : CSR of the stiffness on unknowns (f64 + f32 values), Jacobi diagonal,
//               CSR of the RHS operator B (unknowns x nodes), constant c, PCG vectors
//   PPR       : per node a recovery class id; per class the patch offsets and the
//               6 x npatch least-squares rows (f64)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>
#include "../../include/pic_hip.h"

#define PIC_MAX_OBJECTS 8
#define PIC_MAX_SPECIES 4
#define PIC_PATCH_MAX 50
#define PIC_FIX_SHIFT 40                    // fixed-point scale 2^40 for deposition
#define PIC_FIX_SCALE 1099511627776.0       // 2^40
#define PIC_FIX_INV (1.0 / 1099511627776.0)

// Physical constants of the reference (MCC_jw/code/base/Constants.f90:3-16).
#define PIC_KB 1.3807e-23
#define PIC_EXP_CONST 2.71828
#define PIC_JTOEV 1.6022e-19
#define PIC_EVTOK 11605.0
#define PIC_PI 3.141592653589793238
#define PIC_MINREAL 1.0e-15
#define PIC_ELECTRON_MASS 9.1095e-31
#define PIC_ELECTRON_CHARGE 1.6022e-19
#define PIC_EPSILON0 8.8542e-12
// Pi_LY of SetupGrids_2D_QLL / MoveOne (3.14159265358979D0).
#define PIC_PI_LY 3.14159265358979

struct DevObject {
    int shape;      // 1 circle, 3 box
    double x0, y0;  // circle centre or box lower-left
    double x1, y1;  // box upper-right
    double r2;      // circle radius squared
};

// Everything a particle kernel needs about the mesh, passed by value.
struct MeshParams {
    int nnx, nny;            // nodes (nx, ny in Domain_2D)
    int nex, ney;            // elements per direction (nx-1, ny-1)
    int nnodes, nelem;
    int delta;
    double xmin, xmax, ymin, ymax;     // dxmin..dymax
    double hx, hy, hxi, hyi;            // hx(1:2), hxi
    double vox, voy;                    // Vert_o(1:2) = min - h
    double mapx, mapy;                  // hx / 2^r
    double imapx, imapy;                // 1 / mapx, 1 / mapy
    int pow2_map;                       // 1 if mapx, mapy are powers of two (exact division)
    int locate_f32;                     // 1 if pic_locate_inside_f32 equals pic_locate_ij (pow2 map
                                        // equal to the spacing, xmin / hx and ymin / hy integers)
    int nodes_f32;                      // 1 if (float)xs[i] == (float)xmin + (float)hx * (float)i for
                                        // every node line (and likewise y): FP32 lines recomputed
    int unif_r32, unif_r64;             // 1 if pow2_map and every element's spacing, from the node
                                        // lines the push recomputes in FP32 / FP64, equals hx (hy)
                                        // exactly: the lookup weights skip the far node line
    float ney_rcp;                      // 1 / ney rounded to FP32 (element -> column, pic_elem_ij)
    int ij_exact_div;                   // 1 when nelem >= 2^24: pic_elem_ij divides
    const double* xs; const double* ys; // node coordinate lines
    const float* xsf; const float* ysf;
    const int* element_index;           // nelem
    const int* node_index;              // nnodes
    const unsigned char* elem_clean;    // nelem: 1 if no object cuts or comes near the element
    const unsigned short* elem_nbr;     // nelem: bit 3 (di + 1) + (dj + 1) = elem_clean of element
                                        // (ci + di, cj + dj), di, dj in -1..1 (0 outside the mesh)
    int n_objects;
    DevObject obj[PIC_MAX_OBJECTS];
};

struct SpeciesHost {
    bool defined = false;
    pic_species_params p{};
    double qs = 0, xm = 0, qm = 0;      // normalised (Input_2D.f90:479-483)
    int64_t capacity = 0;
    int64_t nslots = 0;                 // occupied slots (live + dead)
    int64_t nlive = 0;
    void* x = nullptr; void* y = nullptr;
    void* vx = nullptr; void* vy = nullptr; void* vz = nullptr;
    int* cell = nullptr; int* cell_next = nullptr;
    void* th = nullptr;                 // azimuth PO%Z, axisymmetric domains only (delta = 1)
    void* spare[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};   // sort targets, swapped with x..vz, th
    long long* acc_w = nullptr;         // nnodes fixed-point sum W
    long long* acc_e = nullptr;         // nnodes fixed-point sum W*E
    long long* acc_h = nullptr;         // nnodes whole-eV parts of the W*E quanta >= 2^11 eV
    bool acc_dirty = false;
    bool pending = false;               // MoveAll has filled acc and cell_next
    unsigned long long* counters = nullptr;  // device [16]
    int64_t injected = 0;               // particles added by InjectBeams_2D since pic_set_species
    // MCC table
    bool has_mcc = false;
    int nreact = 0, nsigma = 0;
    double emin = 0, eint = 0, emax = 0;
    double sigma_max = 0, collision_ratio = 0;
    uint32_t select_threshold = 0;      // floor(collision_ratio * 2^32)
    int product_species = -1;           // ion species created by ionization
    std::vector<double> prob_host;      // nreact x nsigma
    std::vector<int> rtype_host;
    std::vector<double> thr_host;
    double* prob = nullptr;             // device, nreact x nsigma (column-major like Probility)
    int* rtype = nullptr;
    double* thr = nullptr;
    unsigned long long* mcc_events = nullptr; // device [32]
};

// Wall-injection source of one species (normalised as Input_2D.f90:489-494) with the
// constants of InjectBeams_2D.f90 precomputed on the host.
struct InjectHost {
    bool on = false;
    double dens = 0, tmpj = 0, vt = 0, vd[3] = {0, 0, 0};   // normalised
    double r_start = 0, y_lo = 0, y_hi = 0;
    double b_amb = 0;                   // mean inward velocity of the drifting flux (:42-50)
    double v_r = 0, fs1 = 0, fs2 = 0, qa = 0, sq2vt = 0;     // Bird's sampler (:91-95, :106)
    double vmp = 0;                     // Loadv's VMP_Sheath (Loadv.f90:14)
    double xl = 0, xu = 0;              // slab (:73-76)
};

struct PprClass {
    int n;                        // patch size
    int di[PIC_PATCH_MAX], dj[PIC_PATCH_MAX];
    double R[6][PIC_PATCH_MAX];   // (A^T A)^{-1} A^T of the scaled patch
    double h;                     // patch scale
};

struct PicState {
    bool inited = false;
    int device = 0;
    hipStream_t stream = nullptr;
    // side stream of pic_step: wall injection and MCC run there while the field solve runs on
    // `stream` (neither reads the fields); its own pinned read-back words
    hipStream_t stream2 = nullptr;
    int* h_side_pinned = nullptr;
    hipEvent_t ev_side = nullptr;            // end of the side stream's work of a step
    int (*overlap_hook)() = nullptr;         // run once at a solver's first host wait (overlap_point)
    int overlap_rc = 0;
    int particle_bits = 64;
    int solver_mode = 0;
    std::string err;

    // domain
    bool have_domain = false;
    pic_domain_config cfg{};
    MeshParams mesh{};
    int nunknowns = 0;
    int n_ife_elements = 0;          // elements cut by a dielectric, IFE element matrices
    double L_ref = 0, time_ref = 0, v_ref = 0, phi_ref = 0;

    // host copies of set-up arrays (Fortran conventions where exported)
    std::vector<double> h_xs, h_ys;
    std::vector<int> h_element_index, h_node_index, h_node_type2, h_unknown_node;
    std::vector<double> h_cell_volume, h_dirichlet_phi;
    std::vector<int> h_A_rowptr, h_A_col, h_B_rowptr, h_B_col;
    std::vector<double> h_A_val, h_A_diag, h_B_val, h_c;
    std::vector<int> h_point_patch;   // nnodes x 50, Fortran order
    std::vector<int> h_ppr_class;     // per node class id (-1 = no patch)
    std::vector<PprClass> h_ppr_classes;
    std::vector<unsigned char> h_coef_needed;

    // device mesh arrays
    double *d_xs = nullptr, *d_ys = nullptr;
    float *d_xsf = nullptr, *d_ysf = nullptr;
    int *d_element_index = nullptr, *d_node_index = nullptr;
    double* d_cell_volume = nullptr;   // nnodes (row 2 of Cell_Volume_zwz)
    DevObject* d_objects = nullptr;    // objects for the particle boundary test
    double* d_dirichlet_phi = nullptr;
    int* d_node_type2 = nullptr;
    int* d_unknown_node = nullptr;     // 0-based node of each unknown

    // fields
    double *d_phi = nullptr, *d_efx = nullptr, *d_efy = nullptr, *d_coef = nullptr; // coef: 6 x nnodes
    double* d_coef5 = nullptr;                                                         // c2..c6 node-major (5 x nnodes)
    float *d_efxf = nullptr, *d_efyf = nullptr, *d_coeff = nullptr;                  // coeff: 5 x nnodes (c2..c6)
    double *d_rho = nullptr, *d_rho_s = nullptr, *d_rho_p = nullptr, *d_ek = nullptr;
    double* d_eks = nullptr;                 // Ek_s per species (eV per particle at the node)
    int* d_live_bc = nullptr; int64_t live_bc_cap = 0;   // count_live_slots per-block counts

    // PPR
    int* d_ppr_class = nullptr;
    PprClass* d_ppr_classes = nullptr;
    int n_ppr_classes = 0;

    // solver
    int *d_A_rowptr = nullptr, *d_A_col = nullptr;
    double* d_A_val = nullptr; float* d_A_valf = nullptr;
    double* d_A_diag = nullptr; float* d_A_diagf = nullptr;
    int *d_B_rowptr = nullptr, *d_B_col = nullptr;
    double* d_B_val = nullptr; double* d_c = nullptr;
    double *d_b = nullptr, *d_x = nullptr, *d_r = nullptr, *d_z = nullptr, *d_p = nullptr, *d_p2 = nullptr, *d_q = nullptr;
    float *d_bf = nullptr, *d_xf = nullptr, *d_rf = nullptr, *d_zf = nullptr, *d_pf = nullptr, *d_p2f = nullptr, *d_qf = nullptr;
    // node-space PCG: nine stiffness diagonals (offsets dia_off) and padded vectors
    double* d_dia64[9] = {nullptr}; float* d_dia32[9] = {nullptr};
    int dia_off[9] = {0};
    void* d_pcgvec = nullptr;          // backing store of the padded node-space vectors
    double *d_bn = nullptr, *d_xn = nullptr, *d_rn = nullptr, *d_zn = nullptr, *d_pn = nullptr, *d_p2n = nullptr, *d_qn = nullptr;
    float *d_bfn = nullptr, *d_xfn = nullptr, *d_rfn = nullptr, *d_zfn = nullptr, *d_pfn = nullptr, *d_p2fn = nullptr, *d_qfn = nullptr;
    int pcg_pstride = 0;               // partial-sum region length
    // one-kernel CG on the Jacobi-scaled system: eight scaled off-diagonals, D^-1/2, vectors
    float* d_cg_off[8] = {nullptr};
    int cg_off[8] = {0};
    double* d_sinv = nullptr;
    void* d_cgvec = nullptr;
    float *d_cg_r[2] = {nullptr, nullptr}, *d_cg_w[2] = {nullptr, nullptr}, *d_cg_s[2] = {nullptr, nullptr};
    float *d_cg_p = nullptr, *d_cg_x = nullptr;
    int cg_rpt = 1;                    // rows per thread of the CG kernels
    // multicolour IC(0) factor of the node-space stiffness (solver_mode 3), built on the first
    // such solve of a domain: strictly lower (L) and upper (U = L^T read at the row) factor
    // entries per stencil direction, 1 / l_ii, and a padded FP64 work vector
    void* d_ic = nullptr;
    const double* ic_L[8] = {nullptr}; const double* ic_U[8] = {nullptr};
    const double* ic_dinv = nullptr; double* ic_y = nullptr;
    int64_t ic_epoch = -1, domain_epoch = 0;
    int ic_shifted = 0;                // pivots replaced by the diagonal (IC(0) breakdown guard)
    double last_solve_ms = 0.0;        // device time of the last IFE_Solve_2D solve
    int* d_wall_list = nullptr;        // particles the push hands to the wall kernel
    unsigned int* d_wall_count = nullptr;
    int64_t wall_cap = 0;
    int learn_maxit = 400;             // FCG iterations per inner solve before handing over
    // online fine-tuning inside pic_step: every ft_every steps the step's solve runs the plain
    // inner CG harvesting pairs, then ft_train Adam steps train the network on the ring
    int ft_every = 0, ft_train = 0, ft_batch = 4;
    double ft_lr = 1e-3;
    uint64_t ft_seed = 1;
    int64_t ft_count = 0;
    double ft_ms = 0.0, ft_last_loss = 0.0;
    // learned refinement: passes and FCG iterations per pass of the last learned solve, the
    // launch budget of the next one (all its passes go out before one host synchronisation)
    int lp_passes = 2;
    int lp_iters[6] = {8, 8, 8, 8, 8, 8};
    // one spare FCG iteration in the batch's last pass, so that a solve needing one iteration
    // more than the last one finishes without another batch; dropped after 8 batches in a row
    // that did not use it (in a steady time loop it only costs launches that exit at once),
    // restored when a pass runs out of iterations
    int lp_spare = 1, lp_calm = 0;
    int last_learned_iters = 0, last_fallback = 0;
    // kernel statistics of the last pic_step: [0] ms in CG iteration kernels (chunk events),
    // [1] their launches, [2] ms in push kernels (events around each launch), [3] launches,
    // [4] ms in learned-FCG iterations (chunk events), [5] FCG iterations launched
    double kstat[6] = {0, 0, 0, 0, 0, 0};
    hipEvent_t ev_push[PIC_MAX_SPECIES][3];   // before the push kernel, after it, after the wall kernel
    bool ev_push_pending[PIC_MAX_SPECIES] = {false};
    double* d_scal = nullptr;          // PCG scalars (device)
    double* d_partials = nullptr;      // block partial sums
    unsigned int* d_sync = nullptr;    // last-block counters
    double* h_scal_pinned = nullptr;
    double pcg_tol = 1e-9;
    int pcg_maxit = 50000;
    int last_iter = 0, last_outer = 0;
    double last_relres = 0;

    // species
    int nspecies = 0;
    SpeciesHost sp[PIC_MAX_SPECIES];

    // MCC / run
    bool mcc_ready = false;
    pic_gas_params gas{};
    int e_species = 0, i_species = 1;
    int sort_every = 20;
    uint64_t seed = 12345;
    int mcc_enabled = 1;
    int64_t step = 0;
    // species moved at the current step counter: a second MoveAll of the same species before
    // the counter advances starts a new step (the counter keys every random stream)
    bool moved[PIC_MAX_SPECIES] = {false, false, false, false};
    unsigned char* d_elem_clean = nullptr;   // MeshParams::elem_clean
    unsigned short* d_elem_nbr = nullptr;    // MeshParams::elem_nbr
    InjectHost inj[PIC_MAX_SPECIES];
    unsigned int* d_err = nullptr;     // device error flags
    void* d_errinfo = nullptr;         // PicErrInfo of the first failing particle
    // scratch for selection / compaction
    int* d_sel = nullptr; int64_t sel_cap = 0;
    int* d_blockcount = nullptr; int64_t blockcount_cap = 0;
    uint32_t* d_selmask = nullptr; int64_t selmask_cap = 0;   // MCC selection bits, both species
    int* d_pcount = nullptr; int64_t pcount_cap = 0;          // MCC product block counts
    int* d_mcc_info = nullptr; int64_t mcc_info_cap = 0;      // selected e / i, pairs created
    void* d_newbuf = nullptr; int64_t newbuf_cap = 0;
    void* d_sort_tmp = nullptr; size_t sort_tmp_bytes = 0;
    void* d_sort_buf = nullptr; size_t sort_buf_bytes = 0;

    hipEvent_t ev[16];
    double phase_ms[8] = {0};
};

extern PicState g;

// error helpers
int pic_fail(int code, const std::string& msg);
#define HIPCHK(call) do { hipError_t e_ = (call); if (e_ != hipSuccess) \
    return pic_fail(PIC_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); } while (0)

// host-side setup (setup.cpp)
int setup_build_domain(const pic_domain_config* cfg);
int setup_build_ppr_classes();
void setup_analytic_tables(int isp, bool electron, const pic_gas_params& gas, double mass_kg, double dt_SI);
void setup_maxwellian(int isp, int ppc, double temp_eV, uint64_t seed,
                      std::vector<double>& x, std::vector<double>& y,
                      std::vector<double>& vx, std::vector<double>& vy, std::vector<double>& vz,
                      std::vector<int>& cell, std::vector<double>* th, int kind = 1);

// device launchers
int launch_push_deposit(int isp, bool do_push);
// slots per thread of the moving push kernel for the current mesh and precision (4 or 8); the
// cell sort lays its groups out for it
int pic_push_ppt();
int launch_finalize_deposit();
int launch_rhs();
int launch_ppr();
int run_pcg_fp64(bool reference_semantics);
int run_pcg_mixed();
int run_pcg_ic();
// Called by the solvers right before their first host wait, with a chunk of solver work queued:
// runs the pending side-stream work of pic_step (if any) so that it overlaps the solve.
void overlap_point();
int run_mcc();
int run_sort_all();
int launch_zero_acc(int isp);
int count_live_slots(int64_t n, const int* cell, int64_t* out);
int launch_locate_points(int64_t n, const double* dx, const double* dy, int* dout, int mode);
int launch_ppr_gather_points(int64_t n, const double* dx, const double* dy, const int* ele, double* ex, double* ey);
// learned preconditioner (k_learn.hip)
// records of the learned refinement passes in the device scalar block d_scal (64 doubles):
// slot k at FCG_REC_BASE + FCG_RECW * k = iterations, done, error code, ||r||^2/||r0||^2, skipped
#define FCG_REC_BASE 32
#define FCG_RECW 5
#define FCG_MAX_PASS 6
int run_injection(int it, int64_t* added);
int learn_setup(int max_levels, const int32_t* channels, int nch, uint64_t seed);
int learn_set_harvest(int enable, int capacity, int every);
bool learn_harvesting();
int learn_snapshot(const float* x, int chunk);
int learn_harvest_finish(const float* xfinal);
int learn_train(int steps, int batch, double lr, uint64_t seed, double* loss_first, double* loss_last);
int learn_eval(int npairs, double* mean_loss);
int learn_fcg_pass(int n, const double* b64, double inner_tol, int K, double otol2, const double* outer_rrbb,
                   int first, int slot, double* xn);
bool learn_trace_on();
void learn_note_iterations(int it);
bool learn_ready();
int learn_set_mix(double harvested_share);
double learn_inner_tol();
int learn_apply_host(const float* r, float* z, int64_t n);
int learn_params(float* dst, const float* src, int64_t n, int64_t* nparams);
int learn_layout(int32_t* levels, int32_t* H, int32_t* W, int32_t* C, int64_t* offsets);
int learn_set_pairs(int npairs, const float* r, const float* e);
int learn_get_pair(int k, float* r, float* e, double* ea);
int learn_gradient(int np, const int32_t* idx, double* loss, float* grad, int64_t n);
int learn_info(int32_t* dataset, int32_t* last_iters, int64_t* fallbacks, double* last_loss);
void learn_count_fallback();
void learn_free();
// access for the warm-start network (k_warm.hip)
struct LearnBatch { const float* r; const float* e; float* y; float* gy; };
void learn_set_solve_pass(int pass);
void learn_ring_initial(std::vector<int>& idx);
struct Dia8;
Dia8 learn_scaled_op();
int learn_batch_pairs(int batch, const int* idx, LearnBatch* out);
int learn_batch_loss(int batch, double* loss);
int learn_batch_loss_scaled(const float* y, double* loss, double* scaled);
void learn_adam_step(int64_t n, float* p, const float* gr, float* m, float* v, float lr, int t, float gscale);
// warm-start network (k_warm.hip)
int warm_setup(int levels, int channels, uint64_t seed);
bool warm_active();
int warm_enable(int enable);
int warm_solve_start(int n, const double* rn, double* xn);
int warm_train(int steps, int batch, double lr, uint64_t seed, double* loss_first, double* loss_last);
int warm_eval(int npairs, double* mean_loss, double* mean_scaled);
int warm_apply_host(const float* r, float* y, int64_t n);
int warm_params(float* dst, const float* src, int64_t n, int64_t* nparams);
int warm_layout(int32_t* levels, int32_t* H, int32_t* W, int32_t* channels);
int warm_gradient(int np, const int32_t* idx, double* loss, float* grad, int64_t n);
int warm_info(int32_t* initial_pairs, int32_t* enabled, double* last_alpha, double* last_loss, int64_t* applied);
void warm_free();
int pcg_generic_fp64(int n, const int* rowptr, const int* col, const double* val, const double* diag,
                     const double* b, double* x, double tol, int maxit, int* iters, int* ierr,
                     double* r_out);
int pcg_sparse_aos_fp64(int n, const pic_sparse* dmat, const double* diag, const double* b, double* x, double* r,
                        double* z, double* p, double* p2, double* q, double* partials, int pstride, double* scal,
                        double tol, int maxit, int* iters, int* ierr);
