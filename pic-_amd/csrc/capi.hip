// extern "C" entry points of include/pic_hip.h: library state, set-up upload, species,
// the per-step drop-ins and the stand-alone drop-ins.
#include "pic_device.h"
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

PicState g;

// ---------------------------------------------------------------- allocation-free scratch
// Device scratch of pic_MY_JPCG_Solver_noalloc: the staged SPARSE array, the eight PCG
// vectors, the block partial sums and the scalar block, for systems of up to
// PIC_NOALLOC_MAX_N unknowns and PIC_NOALLOC_MAX_NZ stored entries (54.6 MB, held by every
// GPU the library runs on until the process ends). In host code `scratch` names a host-side
// stand-in, not the device array: the device address comes from hipGetSymbolAddress, once
// per device at pic_init, and only that address is handed to copies and kernels.
#define PIC_NOALLOC_MAX_N 262144
#define PIC_NOALLOC_MAX_NZ (9 * PIC_NOALLOC_MAX_N)
#define PIC_NOALLOC_PSTRIDE (PIC_NOALLOC_MAX_N / 256 + 2)
static constexpr size_t kNoallocMatBytes = (size_t)PIC_NOALLOC_MAX_NZ * sizeof(pic_sparse);
static constexpr size_t kNoallocVecBytes = (size_t)PIC_NOALLOC_MAX_N * sizeof(double);
static constexpr size_t kNoallocPartBytes = ((size_t)4 * PIC_NOALLOC_PSTRIDE * sizeof(double) + 255) / 256 * 256;
static constexpr size_t kNoallocScalBytes = 64 * sizeof(double);
static constexpr size_t kNoallocBytes = kNoallocMatBytes + 8 * kNoallocVecBytes + kNoallocPartBytes + kNoallocScalBytes;
static __device__ __align__(256) double scratch[kNoallocBytes / sizeof(double)];
#define PIC_MAX_DEVICES 64
static void* g_scratch_dev[PIC_MAX_DEVICES] = {nullptr};   // device address of `scratch` per device

static int scratch_lookup(int device) {
    if (device < 0 || device >= PIC_MAX_DEVICES) return pic_fail(PIC_ERR_ARG, "device index beyond the scratch table");
    if (g_scratch_dev[device]) return PIC_OK;
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return pic_fail(PIC_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_dev[device] = ptr;
    return PIC_OK;
}

int pic_fail(int code, const std::string& msg) {
    g.err = msg;
    return code;
}

#define REQUIRE_INIT() do { if (!g.inited) return pic_fail(PIC_ERR_STATE, "pic_init not called"); } while (0)
#define REQUIRE_DOMAIN() do { REQUIRE_INIT(); if (!g.have_domain) return pic_fail(PIC_ERR_STATE, "pic_setup_domain not called"); } while (0)

template <typename T>
static int dupload(T** d, const std::vector<T>& h) {
    if (*d) { hipFree(*d); *d = nullptr; }
    size_t n = std::max<size_t>(h.size(), 1);
    HIPCHK(hipMalloc((void**)d, n * sizeof(T)));
    if (!h.empty()) HIPCHK(hipMemcpy(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return PIC_OK;
}
template <typename T>
static int dalloc(T** d, size_t n) {
    if (*d) { hipFree(*d); *d = nullptr; }
    HIPCHK(hipMalloc((void**)d, std::max<size_t>(n, 1) * sizeof(T)));
    HIPCHK(hipMemset(*d, 0, std::max<size_t>(n, 1) * sizeof(T)));
    return PIC_OK;
}
#define TRY(x) do { int rc_ = (x); if (rc_) return rc_; } while (0)

static int check_err_flags() {
    unsigned int e = 0;
    HIPCHK(hipMemcpyAsync(&e, g.d_err, sizeof(unsigned int), hipMemcpyDeviceToHost, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    if (!e) return PIC_OK;
    PicErrInfo ei;
    HIPCHK(hipMemcpy(&ei, g.d_errinfo, sizeof(ei), hipMemcpyDeviceToHost));
    HIPCHK(hipMemset(g.d_err, 0, sizeof(unsigned int)));
    HIPCHK(hipMemset(g.d_errinfo, 0, sizeof(PicErrInfo)));
    std::string where;
    if (ei.taken) {
        char buf[320];
        snprintf(buf, sizeof buf, " [step %lld, species %d, slot %lld: x=%.17g y=%.17g vx=%.17g vy=%.17g iterations=%g "
                 "element %g -> %g]", (long long)g.step, (int)ei.v[7], ei.slot, ei.v[0], ei.v[1], ei.v[2], ei.v[3],
                 ei.v[4], ei.v[5], ei.v[6]);
        where = buf;
    }
    if (e & PIC_EBIT_NONFINITE) return pic_fail(PIC_ERR_NONFINITE, "non-finite particle state" + where);
    if (e & (PIC_EBIT_BOUNDARY | PIC_EBIT_WRONGCROSS)) return pic_fail(PIC_ERR_BOUNDARY, "particle boundary loop did not settle" + where);
    if (e & PIC_EBIT_POSITION) return pic_fail(PIC_ERR_POSITION, "particle located outside the mesh" + where);
    if (e & PIC_EBIT_EKRHO) return pic_fail(PIC_ERR_POSITION, "Ek_tot /= 0 with rho == 0 (GetParChg_2D.f90:887-890)");
    if (e & PIC_EBIT_MCCLOOP) return pic_fail(PIC_ERR_BOUNDARY, "MCC rejection loop did not end");
    if (e & PIC_EBIT_SAMPLER) return pic_fail(PIC_ERR_BOUNDARY, "injection rejection sampler did not end");
    if (e & PIC_EBIT_OVERFLOW)
        return pic_fail(PIC_ERR_OVERFLOW, "deposition out of the fixed-point range (a weight x energy term that is "
                                          "not finite or >= 2^62 eV, or a node sum past its int64 range)" + where);
    return pic_fail(PIC_ERR_HIP, "device error flag " + std::to_string(e));
}

template <typename R>
__global__ void k_load(int64_t n, const double* x, const double* y, const double* vx, const double* vy, const double* vz,
                       const int* loc, R* X, R* Y, R* VX, R* VY, R* VZ, int* cell, MeshParams m, unsigned int* err) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    R xr = (R)x[i], yr = (R)y[i];
    X[i] = xr; Y[i] = yr; VX[i] = (R)vx[i]; VY[i] = (R)vy[i]; VZ[i] = (R)vz[i];
    if (loc) cell[i] = loc[i] - 1;
    else {
        int bad = 0;
        int e = pic_locate_elem((double)xr, (double)yr, m, &bad);
        if (bad) atomicOr(err, PIC_EBIT_POSITION);
        cell[i] = e;
    }
}

static MeshParams dev_mesh() {
    MeshParams m = g.mesh;
    m.xs = g.d_xs; m.ys = g.d_ys; m.xsf = g.d_xsf; m.ysf = g.d_ysf;
    m.elem_clean = g.d_elem_clean;
    m.elem_nbr = g.d_elem_nbr;
    m.element_index = g.d_element_index; m.node_index = g.d_node_index;
    return m;
}

template <typename R>
static void to_host_double(const void* dsrc, int64_t n, std::vector<double>& out) {
    std::vector<R> tmp(n);
    hipMemcpy(tmp.data(), dsrc, n * sizeof(R), hipMemcpyDeviceToHost);
    out.assign(tmp.begin(), tmp.end());
}

// live-slot count: per-block counts, then a one-block sum (no same-address atomics)
__global__ __launch_bounds__(256) void k_count_live_blocks(int64_t n, const int* cell, int* blockcount) {
    __shared__ int sh[4];
    int64_t base = (int64_t)blockIdx.x * 1024 + threadIdx.x;
    int c = 0;
    for (int k = 0; k < 4; ++k) { int64_t i = base + k * 256; if (i < n && cell[i] >= 0) c++; }
    for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) blockcount[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}
__global__ __launch_bounds__(1024) void k_sum_counts(int nb, const int* blockcount, long long* out) {
    __shared__ long long sh[1024];
    long long s = 0;
    for (int i = threadIdx.x; i < nb; i += 1024) s += blockcount[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int off = 512; off > 0; off >>= 1) { if (threadIdx.x < (unsigned)off) sh[threadIdx.x] += sh[threadIdx.x + off]; __syncthreads(); }
    if (threadIdx.x == 0) *out = sh[0];
}

// number of slots with cell >= 0 among the first n
// Live (cell >= 0) slots among the first n: per-block counts into a scratch array kept
// between calls (regrown only when a larger population needs more blocks), one reduction,
// one pinned read-back.
int count_live_slots(int64_t n, const int* cell, int64_t* out) {
    if (n == 0) { *out = 0; return PIC_OK; }
    int nb = (int)((n + 1023) / 1024);
    const int64_t need = (int64_t)nb * (int64_t)sizeof(int) + 16;
    if (g.live_bc_cap < need) {
        if (g.d_live_bc) (void)hipFree(g.d_live_bc);
        g.d_live_bc = nullptr;
        const int64_t bytes = need + need / 4 + 4096;
        HIPCHK(hipMalloc(&g.d_live_bc, (size_t)bytes));
        g.live_bc_cap = bytes;
    }
    long long* d = (long long*)g.d_partials;
    hipLaunchKernelGGL(k_count_live_blocks, dim3(nb), dim3(256), 0, g.stream, n, cell, g.d_live_bc);
    hipLaunchKernelGGL(k_sum_counts, dim3(1), dim3(1024), 0, g.stream, nb, g.d_live_bc, d);
    HIPCHK(hipGetLastError());
    long long* h = (long long*)g.h_scal_pinned;
    HIPCHK(hipMemcpyAsync(h, d, 8, hipMemcpyDeviceToHost, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    *out = (int64_t)h[0];
    return PIC_OK;
}

__global__ void k_impose_dirichlet(int n, const int* nt2, const double* dphi, double* phi) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && nt2[i] <= 0) phi[i] = dphi[i];
}

__global__ void k_coef_split(int nn, const double* coef6, double* coef5, float* coeff) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nn) return;
    for (int r = 0; r < 5; ++r) { double v = coef6[(size_t)n * 6 + r + 1]; coef5[(size_t)n * 5 + r] = v; coeff[(size_t)n * 5 + r] = (float)v; }
}
__global__ void k_d2f_n(int64_t n, const double* a, float* b) { int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; if (i < n) b[i] = (float)a[i]; }

static void free_species(int s) {
    SpeciesHost& S = g.sp[s];
    void* ptrs[] = {S.x, S.y, S.vx, S.vy, S.vz, S.th, S.cell, S.cell_next, S.acc_w, S.acc_e, S.acc_h, S.counters, S.prob,
                    S.rtype, S.thr, S.mcc_events, S.spare[0], S.spare[1], S.spare[2], S.spare[3], S.spare[4],
                    S.spare[5]};
    for (void* p : ptrs) if (p) hipFree(p);
    S = SpeciesHost();
    g.inj[s] = InjectHost();
}

extern "C" {

const char* pic_last_error(void) { return g.err.c_str(); }

int pic_device_count(int* n) {
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess) { *n = 0; return pic_fail(PIC_ERR_NO_DEVICE, hipGetErrorString(e)); }
    *n = c;
    return PIC_OK;
}

int pic_init(int device) {
    if (g.inited) return PIC_OK;
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess || c == 0) return pic_fail(PIC_ERR_NO_DEVICE, "no HIP device visible");
    if (device < 0 || device >= c) return pic_fail(PIC_ERR_ARG, "device index out of range");
    HIPCHK(hipSetDevice(device));
    g.device = device;
    // A blocking stream: the synchronous hipMemcpy / hipMemset calls of the API (null stream)
    // are then ordered with the library's kernels on both sides.
    HIPCHK(hipStreamCreateWithFlags(&g.stream, hipStreamDefault));
    HIPCHK(hipStreamCreateWithFlags(&g.stream2, hipStreamDefault));
    HIPCHK(hipHostMalloc((void**)&g.h_scal_pinned, 64 * sizeof(double), hipHostMallocDefault));
    HIPCHK(hipHostMalloc((void**)&g.h_side_pinned, 16 * sizeof(int), hipHostMallocDefault));
    HIPCHK(hipEventCreateWithFlags(&g.ev_side, hipEventDisableTiming));
    HIPCHK(hipMalloc((void**)&g.d_scal, 64 * sizeof(double)));
    HIPCHK(hipMemset(g.d_scal, 0, 64 * sizeof(double)));
    HIPCHK(hipMalloc((void**)&g.d_sync, 64 * sizeof(unsigned int)));
    HIPCHK(hipMemset(g.d_sync, 0, 64 * sizeof(unsigned int)));
    HIPCHK(hipMalloc((void**)&g.d_err, sizeof(unsigned int)));
    HIPCHK(hipMemset(g.d_err, 0, sizeof(unsigned int)));
    HIPCHK(hipMalloc(&g.d_errinfo, sizeof(PicErrInfo)));
    HIPCHK(hipMemset(g.d_errinfo, 0, sizeof(PicErrInfo)));
    for (int i = 0; i < 16; ++i) HIPCHK(hipEventCreate(&g.ev[i]));
    for (int s = 0; s < PIC_MAX_SPECIES; ++s) for (int i = 0; i < 3; ++i) HIPCHK(hipEventCreate(&g.ev_push[s][i]));
    {
        int rc = scratch_lookup(device);   // with `device` current
        if (rc) return rc;
    }
    g.inited = true;
    return PIC_OK;
}

int pic_finalize(void) {
    if (!g.inited) return PIC_OK;
    hipDeviceSynchronize();
    learn_free();
    warm_free();
    hipDeviceReset();
    // the reset unloads the code object: the scratch array gets a new address next time
    if (g.device >= 0 && g.device < PIC_MAX_DEVICES) g_scratch_dev[g.device] = nullptr;
    g = PicState();
    return PIC_OK;
}

int pic_set_precision(int particle_bits, int solver_mode) {
    if (particle_bits != 32 && particle_bits != 64) return pic_fail(PIC_ERR_ARG, "particle_bits must be 32 or 64");
    if (solver_mode < 0 || solver_mode > 3) return pic_fail(PIC_ERR_ARG, "solver_mode must be 0, 1, 2 or 3");
    if (particle_bits != g.particle_bits) {
        // particle storage changes width: existing species are discarded
        if (g.inited) HIPCHK(hipStreamSynchronize(g.stream));
        for (int s = 0; s < PIC_MAX_SPECIES; ++s) free_species(s);
        g.nspecies = 0;
        g.mcc_ready = false;
    }
    g.particle_bits = particle_bits;
    g.solver_mode = solver_mode;
    return PIC_OK;
}


int pic_setup_domain(const pic_domain_config* cfg) {
    REQUIRE_INIT();
    if (!cfg) return pic_fail(PIC_ERR_ARG, "cfg is NULL");
    // species, MCC tables, the learned network and the step counter belong to a domain's run
    HIPCHK(hipStreamSynchronize(g.stream));
    learn_free();
    warm_free();
    for (int s = 0; s < PIC_MAX_SPECIES; ++s) free_species(s);
    g.nspecies = 0;
    g.step = 0;
    for (int s = 0; s < PIC_MAX_SPECIES; ++s) g.moved[s] = false;
    g.ft_every = 0; g.ft_count = 0; g.ft_ms = 0.0; g.ft_last_loss = 0.0;
    g.domain_epoch++;                                // a cached IC(0) factor belongs to the old domain
    if (g.d_ic) { (void)hipFree(g.d_ic); g.d_ic = nullptr; }
    g.mcc_ready = false;
    g.have_domain = false;
    TRY(setup_build_domain(cfg));
    TRY(setup_build_ppr_classes());
    const MeshParams& m = g.mesh;
    TRY(dupload(&g.d_xs, g.h_xs)); TRY(dupload(&g.d_ys, g.h_ys));
    std::vector<float> xf(g.h_xs.begin(), g.h_xs.end()), yf(g.h_ys.begin(), g.h_ys.end());
    TRY(dupload(&g.d_xsf, xf)); TRY(dupload(&g.d_ysf, yf));
    TRY(dupload(&g.d_element_index, g.h_element_index));
    TRY(dupload(&g.d_node_index, g.h_node_index));
    std::vector<double> vol(m.nnodes);
    for (int n = 0; n < m.nnodes; ++n) vol[n] = g.h_cell_volume[2 * (size_t)n + 1];
    TRY(dupload(&g.d_cell_volume, vol));
    TRY(dupload(&g.d_dirichlet_phi, g.h_dirichlet_phi));
    {
        std::vector<DevObject> ob(g.mesh.obj, g.mesh.obj + PIC_MAX_OBJECTS);
        TRY(dupload(&g.d_objects, ob));
    }
    {
        // elements where a particle can need the object branches of MoveOne /
        // AdjustParticleBoundary: cut elements and every element whose box, widened by
        // 1e-6 h, meets an object (the push kernel hands their particles to the full loop)
        std::vector<unsigned char> clean(m.nelem, 1);
        for (int e = 0; e < m.nelem; ++e) {
            const int ci = e / m.ney, cj = e - ci * m.ney;
            if (g.h_element_index[e] > 0) { clean[e] = 0; continue; }
            const double mx = 1e-6 * m.hx, my = 1e-6 * m.hy;
            const double bx0 = g.h_xs[ci] - mx, bx1 = g.h_xs[ci + 1] + mx;
            const double by0 = g.h_ys[cj] - my, by1 = g.h_ys[cj + 1] + my;
            for (int o = 0; o < m.n_objects; ++o) {
                const DevObject& d = m.obj[o];
                bool touch = false;
                if (d.shape == 3) {
                    touch = bx0 < d.x1 && bx1 > d.x0 && by0 < d.y1 && by1 > d.y0;
                } else if (d.shape == 1) {
                    const double dx = std::max(std::max(bx0 - d.x0, 0.0), d.x0 - bx1);
                    const double dy = std::max(std::max(by0 - d.y0, 0.0), d.y0 - by1);
                    touch = dx * dx + dy * dy < d.r2 * (1.0 + 1e-9);
                }
                if (touch) { clean[e] = 0; break; }
            }
        }
        TRY(dupload(&g.d_elem_clean, clean));
        std::vector<unsigned short> nbr(m.nelem, 0);
        for (int e = 0; e < m.nelem; ++e) {
            const int ci = e / m.ney, cj = e - ci * m.ney;
            unsigned short b = 0;
            for (int di = -1; di <= 1; ++di)
                for (int dj = -1; dj <= 1; ++dj) {
                    const int i = ci + di, j = cj + dj;
                    if (i < 0 || i >= m.nex || j < 0 || j >= m.ney) continue;
                    if (clean[(size_t)i * m.ney + j]) b |= (unsigned short)(1u << (3 * (di + 1) + (dj + 1)));
                }
            nbr[e] = b;
        }
        TRY(dupload(&g.d_elem_nbr, nbr));
    }
    TRY(dupload(&g.d_node_type2, g.h_node_type2));
    std::vector<int> un0(g.h_unknown_node.size());
    for (size_t i = 0; i < un0.size(); ++i) un0[i] = g.h_unknown_node[i] - 1;
    TRY(dupload(&g.d_unknown_node, un0));
    // fields: phi starts at the Dirichlet values (InitialPhi_2D / SetObjects_2D), 0 elsewhere
    TRY(dupload(&g.d_phi, g.h_dirichlet_phi));
    TRY(dalloc(&g.d_efx, m.nnodes)); TRY(dalloc(&g.d_efy, m.nnodes));
    TRY(dalloc(&g.d_efxf, m.nnodes)); TRY(dalloc(&g.d_efyf, m.nnodes));
    TRY(dalloc(&g.d_coef, (size_t)m.nnodes * 6)); TRY(dalloc(&g.d_coef5, (size_t)m.nnodes * 5));
    TRY(dalloc(&g.d_coeff, (size_t)m.nnodes * 5));
    TRY(dalloc(&g.d_rho, m.nnodes));
    TRY(dalloc(&g.d_rho_s, (size_t)m.nnodes * PIC_MAX_SPECIES));
    TRY(dalloc(&g.d_rho_p, (size_t)m.nnodes * PIC_MAX_SPECIES));
    TRY(dalloc(&g.d_ek, (size_t)m.nnodes * PIC_MAX_SPECIES));
    TRY(dalloc(&g.d_eks, (size_t)m.nnodes * PIC_MAX_SPECIES));
    // PPR
    TRY(dupload(&g.d_ppr_class, g.h_ppr_class));
    TRY(dupload(&g.d_ppr_classes, g.h_ppr_classes));
    g.n_ppr_classes = (int)g.h_ppr_classes.size();
    // solver
    int nu = g.nunknowns;
    TRY(dupload(&g.d_A_rowptr, g.h_A_rowptr)); TRY(dupload(&g.d_A_col, g.h_A_col));
    TRY(dupload(&g.d_A_val, g.h_A_val)); TRY(dupload(&g.d_A_diag, g.h_A_diag));
    std::vector<float> avf(g.h_A_val.begin(), g.h_A_val.end()), adf(g.h_A_diag.begin(), g.h_A_diag.end());
    TRY(dupload(&g.d_A_valf, avf)); TRY(dupload(&g.d_A_diagf, adf));
    TRY(dupload(&g.d_B_rowptr, g.h_B_rowptr)); TRY(dupload(&g.d_B_col, g.h_B_col));
    TRY(dupload(&g.d_B_val, g.h_B_val)); TRY(dupload(&g.d_c, g.h_c));
    TRY(dalloc(&g.d_b, nu)); TRY(dalloc(&g.d_x, nu)); TRY(dalloc(&g.d_r, nu)); TRY(dalloc(&g.d_z, nu));
    TRY(dalloc(&g.d_p, nu)); TRY(dalloc(&g.d_p2, nu)); TRY(dalloc(&g.d_q, nu));
    TRY(dalloc(&g.d_bf, nu)); TRY(dalloc(&g.d_xf, nu)); TRY(dalloc(&g.d_rf, nu)); TRY(dalloc(&g.d_zf, nu));
    TRY(dalloc(&g.d_pf, nu)); TRY(dalloc(&g.d_p2f, nu)); TRY(dalloc(&g.d_qf, nu));
    // node-space PCG: the nine diagonals of the stiffness (node offsets), padded vectors
    {
        const int off[9] = {-m.nny - 1, -m.nny, -m.nny + 1, -1, 0, 1, m.nny - 1, m.nny, m.nny + 1};
        std::vector<std::vector<double>> dia(9, std::vector<double>(m.nnodes, 0.0));
        for (int u = 0; u < nu; ++u) {
            int nu_node = g.h_unknown_node[u] - 1;
            for (int k = g.h_A_rowptr[u]; k < g.h_A_rowptr[u + 1]; ++k) {
                int nv = g.h_unknown_node[g.h_A_col[k]] - 1;
                int d = nv - nu_node, kk = -1;
                for (int q = 0; q < 9; ++q) if (off[q] == d) kk = q;
                if (kk < 0) return pic_fail(PIC_ERR_UNSUPPORTED, "stiffness entry off the nine-point stencil");
                dia[kk][nu_node] = g.h_A_val[k];
            }
        }
        for (int q = 0; q < 9; ++q) {
            g.dia_off[q] = off[q];
            TRY(dupload(&g.d_dia64[q], dia[q]));
            std::vector<float> f(dia[q].begin(), dia[q].end());
            TRY(dupload(&g.d_dia32[q], f));
        }
        size_t pad = ((size_t)m.nny + 2 + 63) / 64 * 64;
        size_t len = (size_t)m.nnodes + 2 * pad;
        if (g.d_pcgvec) { (void)hipFree(g.d_pcgvec); g.d_pcgvec = nullptr; }
        HIPCHK(hipMalloc(&g.d_pcgvec, len * 7 * (sizeof(double) + sizeof(float))));
        HIPCHK(hipMemset(g.d_pcgvec, 0, len * 7 * (sizeof(double) + sizeof(float))));
        double* d0 = (double*)g.d_pcgvec;
        double** dv[7] = {&g.d_bn, &g.d_xn, &g.d_rn, &g.d_zn, &g.d_pn, &g.d_p2n, &g.d_qn};
        for (int k = 0; k < 7; ++k) *dv[k] = d0 + k * len + pad;
        float* f0 = (float*)(d0 + 7 * len);
        float** fv[7] = {&g.d_bfn, &g.d_xfn, &g.d_rfn, &g.d_zfn, &g.d_pfn, &g.d_p2fn, &g.d_qfn};
        for (int k = 0; k < 7; ++k) *fv[k] = f0 + k * len + pad;
        g.pcg_pstride = (m.nnodes + 255) / 256 + 1;
        // Jacobi-scaled off-diagonals c_k[i] / sqrt(d_i d_j) for the FP32 CG (0 where d == 0)
        std::vector<double> sinv(m.nnodes, 0.0);
        for (int i = 0; i < m.nnodes; ++i) sinv[i] = dia[4][i] > 0.0 ? 1.0 / std::sqrt(dia[4][i]) : 0.0;
        for (int q = 0, kk = 0; q < 9; ++q) {
            if (q == 4) continue;
            std::vector<float> f(m.nnodes, 0.0f);
            for (int i = 0; i < m.nnodes; ++i) {
                int j = i + off[q];
                if (dia[q][i] != 0.0 && j >= 0 && j < m.nnodes) f[i] = (float)(dia[q][i] * sinv[i] * sinv[j]);
            }
            g.cg_off[kk] = off[q];
            TRY(dupload(&g.d_cg_off[kk], f));
            kk++;
        }
        // padded: eight FP32 vectors, then D^-1/2 (FP64; read at the neighbours of every row)
        if (g.d_cgvec) { (void)hipFree(g.d_cgvec); g.d_cgvec = nullptr; }
        size_t cg_bytes = len * 8 * sizeof(float) + len * sizeof(double);
        HIPCHK(hipMalloc(&g.d_cgvec, cg_bytes));
        HIPCHK(hipMemset(g.d_cgvec, 0, cg_bytes));
        float* c0 = (float*)g.d_cgvec;
        float** cv[8] = {&g.d_cg_r[0], &g.d_cg_r[1], &g.d_cg_w[0], &g.d_cg_w[1], &g.d_cg_s[0], &g.d_cg_s[1], &g.d_cg_p, &g.d_cg_x};
        for (int k = 0; k < 8; ++k) *cv[k] = c0 + k * len + pad;
        g.d_sinv = (double*)(c0 + 8 * len) + pad;
        HIPCHK(hipMemcpy(g.d_sinv, sinv.data(), sizeof(double) * m.nnodes, hipMemcpyHostToDevice));
        if (const char* e = std::getenv("PIC_CG_RPT")) { int v = std::atoi(e); if (v == 1 || v == 2 || v == 4) g.cg_rpt = v; }
    }
    size_t nblk = (size_t)std::max(m.nnodes, nu) / 256 + 2;
    TRY(dalloc(&g.d_partials, std::max<size_t>(4 * nblk + 64, (size_t)4 * g.pcg_pstride + 64)));
    g.have_domain = true;
    return PIC_OK;
}

int pic_get_sizes(int32_t* nnodes, int32_t* nelem, int32_t* nunknowns, int64_t* nnz_A, int64_t* nnz_B) {
    REQUIRE_DOMAIN();
    if (nnodes) *nnodes = g.mesh.nnodes;
    if (nelem) *nelem = g.mesh.nelem;
    if (nunknowns) *nunknowns = g.nunknowns;
    if (nnz_A) *nnz_A = (int64_t)g.h_A_col.size();
    if (nnz_B) *nnz_B = (int64_t)g.h_B_col.size();
    return PIC_OK;
}

int pic_get_setup_array(int which, void* dst, int64_t n) {
    REQUIRE_DOMAIN();
    const MeshParams& m = g.mesh;
    auto put = [&](const void* src, size_t cnt, size_t esz) -> int {
        if ((int64_t)cnt > n) return pic_fail(PIC_ERR_ARG, "destination too small");
        memcpy(dst, src, cnt * esz);
        return PIC_OK;
    };
    switch (which) {
    case PIC_SETUP_HP: {
        std::vector<double> hp(2 * (size_t)m.nnodes);
        for (int i = 0; i < m.nnx; ++i)
            for (int j = 0; j < m.nny; ++j) { size_t k = (size_t)i * m.nny + j; hp[2 * k] = g.h_xs[i]; hp[2 * k + 1] = g.h_ys[j]; }
        return put(hp.data(), hp.size(), 8);
    }
    case PIC_SETUP_HT: {
        std::vector<int> ht(5 * (size_t)m.nelem, 0);
        for (int e = 0; e < m.nelem; ++e) {
            int ci = e / m.ney, cj = e % m.ney;
            int n1 = ci * m.nny + cj;
            ht[5 * e] = n1 + 1; ht[5 * e + 1] = n1 + m.nny + 1; ht[5 * e + 2] = n1 + m.nny + 2; ht[5 * e + 3] = n1 + 2;
        }
        return put(ht.data(), ht.size(), 4);
    }
    case PIC_SETUP_ELEMENT_INDEX: return put(g.h_element_index.data(), g.h_element_index.size(), 4);
    case PIC_SETUP_NODE_INDEX: return put(g.h_node_index.data(), g.h_node_index.size(), 4);
    case PIC_SETUP_NODE_TYPE: {
        std::vector<int> nt(2 * (size_t)m.nnodes);
        for (int k = 0; k < m.nnodes; ++k) { nt[2 * k] = g.h_node_type2[k] > 0 ? 0 : 1; nt[2 * k + 1] = g.h_node_type2[k]; }
        return put(nt.data(), nt.size(), 4);
    }
    case PIC_SETUP_CELL_VOLUME: return put(g.h_cell_volume.data(), g.h_cell_volume.size(), 8);
    case PIC_SETUP_A_ROWPTR: return put(g.h_A_rowptr.data(), g.h_A_rowptr.size(), 4);
    case PIC_SETUP_A_COL: return put(g.h_A_col.data(), g.h_A_col.size(), 4);
    case PIC_SETUP_A_VAL: return put(g.h_A_val.data(), g.h_A_val.size(), 8);
    case PIC_SETUP_A_DIAG: return put(g.h_A_diag.data(), g.h_A_diag.size(), 8);
    case PIC_SETUP_B_ROWPTR: return put(g.h_B_rowptr.data(), g.h_B_rowptr.size(), 4);
    case PIC_SETUP_B_COL: return put(g.h_B_col.data(), g.h_B_col.size(), 4);
    case PIC_SETUP_B_VAL: return put(g.h_B_val.data(), g.h_B_val.size(), 8);
    case PIC_SETUP_RHS_CONST: return put(g.h_c.data(), g.h_c.size(), 8);
    case PIC_SETUP_DIRICHLET_PHI: return put(g.h_dirichlet_phi.data(), g.h_dirichlet_phi.size(), 8);
    case PIC_SETUP_POINT_PATCH: return put(g.h_point_patch.data(), g.h_point_patch.size(), 4);
    case PIC_SETUP_UNKNOWN_NODE: return put(g.h_unknown_node.data(), g.h_unknown_node.size(), 4);
    default: return pic_fail(PIC_ERR_ARG, "unknown set-up array");
    }
}

// ----------------------------------------------------------------------- species
int pic_set_species(int isp, const pic_species_params* p, int64_t capacity) {
    REQUIRE_DOMAIN();
    if (isp < 0 || isp >= PIC_MAX_SPECIES) return pic_fail(PIC_ERR_ARG, "species index out of range");
    if (!p || capacity <= 0 || capacity > (int64_t)INT32_MAX) return pic_fail(PIC_ERR_ARG, "bad species parameters");
    SpeciesHost& S = g.sp[isp];
    S.p = *p;
    S.qs = p->charge_C / PIC_ELECTRON_CHARGE;
    S.xm = p->mass_kg / PIC_ELECTRON_MASS;
    S.qm = S.qs / S.xm;
    S.capacity = capacity;
    S.nslots = 0; S.nlive = 0;
    size_t esz = g.particle_bits == 64 ? 8 : 4;
    void** arr[6] = {&S.x, &S.y, &S.vx, &S.vy, &S.vz, &S.th};
    for (int a = 0; a < 6; ++a) {
        if (*arr[a]) hipFree(*arr[a]);
        *arr[a] = nullptr;
        // the azimuth PO%Z is kept for axisymmetric domains only
        if (a < 5 || g.mesh.delta == 1) HIPCHK(hipMalloc(arr[a], (size_t)capacity * esz));
    }
    if (S.th) HIPCHK(hipMemset(S.th, 0, (size_t)capacity * esz));
    // the sort's spare arrays are sized for the capacity: drop them, the next sort reallocates
    for (int a = 0; a < 6; ++a) {
        if (S.spare[a]) (void)hipFree(S.spare[a]);
        S.spare[a] = nullptr;
    }
    TRY(dalloc(&S.cell, (size_t)capacity)); TRY(dalloc(&S.cell_next, (size_t)capacity));
    TRY(dalloc(&S.acc_w, (size_t)g.mesh.nnodes)); TRY(dalloc(&S.acc_e, (size_t)g.mesh.nnodes));
    TRY(dalloc(&S.acc_h, (size_t)g.mesh.nnodes));
    TRY(dalloc(&S.counters, 16)); TRY(dalloc(&S.mcc_events, 32));
    S.acc_dirty = false; S.pending = false; S.defined = true;
    g.nspecies = std::max(g.nspecies, isp + 1);
    return PIC_OK;
}


int pic_load_particles(int isp, int64_t n, const double* x, const double* y, const double* vx, const double* vy,
                       const double* vz, const int32_t* loc) {
    REQUIRE_DOMAIN();
    if (isp < 0 || isp >= PIC_MAX_SPECIES || !g.sp[isp].defined) return pic_fail(PIC_ERR_ARG, "species not defined");
    SpeciesHost& S = g.sp[isp];
    if (n < 0 || n > S.capacity) return pic_fail(PIC_ERR_CAPACITY, "more particles than capacity");
    if (n > 0 && (!x || !y || !vx || !vy || !vz)) return pic_fail(PIC_ERR_ARG, "NULL particle array");
    if (loc) {
        // A given Location must be the element GetParChg_2D would store for the position as
        // the device keeps it (rounded to the particle precision): every kernel indexes the
        // element arrays with it, and the gather of the next MoveOne uses it as its element.
        for (int64_t i = 0; i < n; ++i) {
            const double xr = g.particle_bits == 32 ? (double)(float)x[i] : x[i];
            const double yr = g.particle_bits == 32 ? (double)(float)y[i] : y[i];
            int bad = 0, ci = 0, cj = 0;
            const int e = pic_locate_ij(xr, yr, g.mesh, &bad, &ci, &cj);
            if (loc[i] < 1 || loc[i] > g.mesh.nelem || bad || loc[i] != e + 1) {
                char buf[256];
                snprintf(buf, sizeof buf, "particle %lld: Location %d is not the element containing (%.17g, %.17g)%s",
                         (long long)i, (int)loc[i], xr, yr, bad ? " (outside the mesh)" : "");
                S.nslots = 0; S.nlive = 0; S.pending = false;
                return pic_fail(PIC_ERR_POSITION, buf);
            }
        }
    }
    S.nslots = n; S.nlive = n; S.pending = false;
    if (n == 0) return PIC_OK;
    double* d = nullptr;
    int* dl = nullptr;
    HIPCHK(hipMalloc(&d, 5 * (size_t)n * sizeof(double)));
    const double* src[5] = {x, y, vx, vy, vz};
    for (int a = 0; a < 5; ++a) HIPCHK(hipMemcpy(d + (size_t)a * n, src[a], n * sizeof(double), hipMemcpyHostToDevice));
    if (loc) { HIPCHK(hipMalloc(&dl, n * sizeof(int))); HIPCHK(hipMemcpy(dl, loc, n * sizeof(int), hipMemcpyHostToDevice)); }
    int blocks = (int)((n + 255) / 256);
    MeshParams m = dev_mesh();
    if (g.particle_bits == 64)
        hipLaunchKernelGGL(k_load<double>, dim3(blocks), dim3(256), 0, g.stream, n, d, d + n, d + 2 * n, d + 3 * n, d + 4 * n, dl,
                           (double*)S.x, (double*)S.y, (double*)S.vx, (double*)S.vy, (double*)S.vz, S.cell, m, g.d_err);
    else
        hipLaunchKernelGGL(k_load<float>, dim3(blocks), dim3(256), 0, g.stream, n, d, d + n, d + 2 * n, d + 3 * n, d + 4 * n, dl,
                           (float*)S.x, (float*)S.y, (float*)S.vx, (float*)S.vy, (float*)S.vz, S.cell, m, g.d_err);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(g.stream));
    hipFree(d);
    if (dl) hipFree(dl);
    return check_err_flags();
}

int pic_load_maxwellian(int isp, int32_t ppc, double temp_eV, uint64_t seed) {
    return pic_load_distribution(isp, ppc, temp_eV, seed, 1);
}

int pic_load_distribution(int isp, int32_t ppc, double temp_eV, uint64_t seed, int kind) {
    REQUIRE_DOMAIN();
    if (kind < 0 || kind > 2) return pic_fail(PIC_ERR_ARG, "kind must be 0 (kappa), 1 (Maxwellian) or 2 (polytropic)");
    if (isp < 0 || isp >= PIC_MAX_SPECIES || !g.sp[isp].defined) return pic_fail(PIC_ERR_ARG, "species not defined");
    if (ppc <= 0 || !(temp_eV >= 0)) return pic_fail(PIC_ERR_ARG, "bad ppc / temperature");
    std::vector<double> x, y, vx, vy, vz;
    std::vector<int> cell;
    std::vector<double> th;
    const bool rz = g.mesh.delta == 1;
    setup_maxwellian(isp, ppc, temp_eV, seed, x, y, vx, vy, vz, cell, rz ? &th : nullptr, kind);
    if ((int64_t)x.size() > g.sp[isp].capacity) return pic_fail(PIC_ERR_CAPACITY, "load exceeds capacity");
    TRY(pic_load_particles(isp, (int64_t)x.size(), x.data(), y.data(), vx.data(), vy.data(), vz.data(), nullptr));
    if (rz) TRY(pic_set_azimuth(isp, (int64_t)th.size(), th.data()));
    return PIC_OK;
}


int pic_get_particles(int isp, double* x, double* y, double* vx, double* vy, double* vz, int32_t* loc,
                      int64_t capacity, int64_t* n) {
    REQUIRE_DOMAIN();
    if (isp < 0 || isp >= PIC_MAX_SPECIES || !g.sp[isp].defined) return pic_fail(PIC_ERR_ARG, "species not defined");
    SpeciesHost& S = g.sp[isp];
    HIPCHK(hipStreamSynchronize(g.stream));
    int64_t ns = S.nslots;
    std::vector<double> a[5];
    void* src[5] = {S.x, S.y, S.vx, S.vy, S.vz};
    for (int k = 0; k < 5; ++k) {
        if (g.particle_bits == 64) to_host_double<double>(src[k], ns, a[k]);
        else to_host_double<float>(src[k], ns, a[k]);
    }
    std::vector<int> c(ns);
    if (ns) HIPCHK(hipMemcpy(c.data(), S.pending ? S.cell : S.cell, ns * sizeof(int), hipMemcpyDeviceToHost));
    int64_t k = 0;
    for (int64_t i = 0; i < ns; ++i) {
        if (c[i] < 0) continue;
        if (k >= capacity) return pic_fail(PIC_ERR_CAPACITY, "destination too small");
        if (x) x[k] = a[0][i];
        if (y) y[k] = a[1][i];
        if (vx) vx[k] = a[2][i];
        if (vy) vy[k] = a[3][i];
        if (vz) vz[k] = a[4][i];
        if (loc) loc[k] = c[i] + 1;
        ++k;
    }
    *n = k;
    return PIC_OK;
}

int pic_set_azimuth(int isp, int64_t n, const double* th) {
    REQUIRE_DOMAIN();
    if (isp < 0 || isp >= PIC_MAX_SPECIES || !g.sp[isp].defined) return pic_fail(PIC_ERR_ARG, "species not defined");
    SpeciesHost& S = g.sp[isp];
    if (g.mesh.delta != 1 || !S.th) return pic_fail(PIC_ERR_STATE, "the azimuth is kept for axisymmetric domains only");
    if (n < 0 || n > S.nslots || (n > 0 && !th)) return pic_fail(PIC_ERR_ARG, "azimuth count exceeds the occupied slots");
    if (n == 0) return PIC_OK;
    HIPCHK(hipStreamSynchronize(g.stream));
    if (g.particle_bits == 64) {
        HIPCHK(hipMemcpy(S.th, th, (size_t)n * sizeof(double), hipMemcpyHostToDevice));
    } else {
        std::vector<float> f(th, th + n);
        HIPCHK(hipMemcpy(S.th, f.data(), (size_t)n * sizeof(float), hipMemcpyHostToDevice));
    }
    return PIC_OK;
}

int pic_get_azimuth(int isp, double* th, int64_t capacity, int64_t* n) {
    REQUIRE_DOMAIN();
    if (isp < 0 || isp >= PIC_MAX_SPECIES || !g.sp[isp].defined) return pic_fail(PIC_ERR_ARG, "species not defined");
    SpeciesHost& S = g.sp[isp];
    if (g.mesh.delta != 1 || !S.th) return pic_fail(PIC_ERR_STATE, "the azimuth is kept for axisymmetric domains only");
    HIPCHK(hipStreamSynchronize(g.stream));
    const int64_t ns = S.nslots;
    std::vector<double> a;
    if (g.particle_bits == 64) to_host_double<double>(S.th, ns, a);
    else to_host_double<float>(S.th, ns, a);
    std::vector<int> c(ns);
    if (ns) HIPCHK(hipMemcpy(c.data(), S.cell, ns * sizeof(int), hipMemcpyDeviceToHost));
    int64_t k = 0;
    for (int64_t i = 0; i < ns; ++i) {
        if (c[i] < 0) continue;
        if (k >= capacity) return pic_fail(PIC_ERR_CAPACITY, "destination too small");
        if (th) th[k] = a[i];
        ++k;
    }
    *n = k;
    return PIC_OK;
}

int pic_load_bundle(int isp, const pic_particle_one* po, int64_t npar) {
    if (npar < 0 || (npar > 0 && !po)) return pic_fail(PIC_ERR_ARG, "bad bundle");
    std::vector<double> x(npar), y(npar), z(npar), vx(npar), vy(npar), vz(npar);
    std::vector<int> loc(npar);
    for (int64_t i = 0; i < npar; ++i) {
        x[i] = po[i].X; y[i] = po[i].Y; z[i] = po[i].Z;
        vx[i] = po[i].Vx; vy[i] = po[i].Vy; vz[i] = po[i].Vz; loc[i] = po[i].Location;
    }
    TRY(pic_load_particles(isp, npar, x.data(), y.data(), vx.data(), vy.data(), vz.data(), loc.data()));
    if (g.mesh.delta == 1) TRY(pic_set_azimuth(isp, npar, z.data()));
    return PIC_OK;
}

int pic_store_bundle(int isp, pic_particle_one* po, int64_t capacity, int64_t* npar) {
    REQUIRE_DOMAIN();
    if (isp < 0 || isp >= PIC_MAX_SPECIES || !g.sp[isp].defined) return pic_fail(PIC_ERR_ARG, "species not defined");
    int64_t cap = g.sp[isp].nslots;
    std::vector<double> x(cap), y(cap), vx(cap), vy(cap), vz(cap);
    std::vector<int> loc(cap);
    int64_t n = 0;
    TRY(pic_get_particles(isp, x.data(), y.data(), vx.data(), vy.data(), vz.data(), loc.data(), cap, &n));
    if (n > capacity) return pic_fail(PIC_ERR_CAPACITY, "destination too small");
    std::vector<double> z(g.mesh.delta == 1 ? cap : 0);
    if (g.mesh.delta == 1) {
        int64_t nz = 0;
        TRY(pic_get_azimuth(isp, z.data(), cap, &nz));
    }
    for (int64_t i = 0; i < n; ++i) {
        pic_particle_one& o = po[i];
        memset(&o, 0, sizeof(o));
        o.X = x[i]; o.Y = y[i]; o.Vx = vx[i]; o.Vy = vy[i]; o.Vz = vz[i]; o.Location = loc[i];
        if (g.mesh.delta == 1) o.Z = z[i];
    }
    *npar = n;
    return PIC_OK;
}


int pic_get_counts(int isp, int64_t* nlive, int64_t* nslots, int64_t counters[8]) {
    REQUIRE_DOMAIN();
    if (isp < 0 || isp >= PIC_MAX_SPECIES || !g.sp[isp].defined) return pic_fail(PIC_ERR_ARG, "species not defined");
    SpeciesHost& S = g.sp[isp];
    unsigned long long h[16];
    HIPCHK(hipStreamSynchronize(g.stream));
    HIPCHK(hipMemcpy(h, S.counters, sizeof(h), hipMemcpyDeviceToHost));
    if (counters) for (int k = 0; k < 8; ++k) counters[k] = (int64_t)h[k];
    if (nslots) *nslots = S.nslots;
    if (nlive) {
        const int* c = S.pending ? S.cell_next : S.cell;
        TRY(count_live_slots(S.nslots, c, nlive));
    }
    return PIC_OK;
}

int pic_get_counters(int isp, int64_t* counters, int n) {
    REQUIRE_DOMAIN();
    if (isp < 0 || isp >= PIC_MAX_SPECIES || !g.sp[isp].defined) return pic_fail(PIC_ERR_ARG, "species not defined");
    if (!counters || n < 0) return pic_fail(PIC_ERR_ARG, "bad counter buffer");
    unsigned long long h[16];
    HIPCHK(hipStreamSynchronize(g.stream));
    HIPCHK(hipMemcpy(h, g.sp[isp].counters, sizeof(h), hipMemcpyDeviceToHost));
    for (int k = 0; k < n && k < 8; ++k) counters[k] = (int64_t)h[k];
    if (n > 8) counters[8] = g.sp[isp].injected;
    for (int k = 9; k < n; ++k) counters[k] = 0;
    return PIC_OK;
}

int pic_get_mcc_events(int isp, int64_t* events, int nmax) {
    REQUIRE_DOMAIN();
    if (isp < 0 || isp >= PIC_MAX_SPECIES || !g.sp[isp].defined) return pic_fail(PIC_ERR_ARG, "species not defined");
    unsigned long long h[32];
    HIPCHK(hipStreamSynchronize(g.stream));
    HIPCHK(hipMemcpy(h, g.sp[isp].mcc_events, sizeof(h), hipMemcpyDeviceToHost));
    for (int k = 0; k < nmax && k < 31; ++k) events[k] = (int64_t)h[k + 1];
    return PIC_OK;
}

// ----------------------------------------------------------------------- per step
// The step counter keys every random stream (kappa re-emission, MCC selection and
// collision). It advances where a time step ends: after MCC (the last random consumer of
// the reference's step, MCC_jw OneStep.F90:49-67), or, for a caller that runs no MCC, when
// a species moves a second time at the same counter value. pic_step and a caller of the
// drop-ins in the reference's order (MoveAll per species, GetParChg_2D, IFE_Solve_2D,
// GetEField_SIDG_PPR, InjectBeams_2D, MCC) therefore draw the same numbers.
static void advance_step() {
    g.step++;
    for (int s = 0; s < PIC_MAX_SPECIES; ++s) g.moved[s] = false;
}

int pic_MoveAll(int isp, int delta, int PushFlag) {
    REQUIRE_DOMAIN();
    if (isp < 0 || isp >= PIC_MAX_SPECIES || !g.sp[isp].defined) return pic_fail(PIC_ERR_ARG, "species not defined");
    if (PushFlag != 0) return pic_fail(PIC_ERR_UNSUPPORTED, "only the explicit mover (PushFlag = 0) is on this path");
    if (delta != g.mesh.delta) return pic_fail(PIC_ERR_ARG, "delta differs from the set-up domain");
    SpeciesHost& S = g.sp[isp];
    if (g.moved[isp]) advance_step();
    if (S.pending) S.acc_dirty = true;   // a repeated MoveAll discards the previous pending deposit
    TRY(launch_push_deposit(isp, true));
    S.pending = true;
    g.moved[isp] = true;
    return PIC_OK;
}

static int commit_locations() {
    for (int s = 0; s < g.nspecies; ++s) {
        SpeciesHost& S = g.sp[s];
        if (!S.defined) return pic_fail(PIC_ERR_STATE, "species indices must be contiguous from 0");
        if (!S.pending) TRY(launch_push_deposit(s, false));
        std::swap(S.cell, S.cell_next);
        S.pending = false;
    }
    return PIC_OK;
}

int pic_GetParChg_2D(int delta) {
    REQUIRE_DOMAIN();
    if (delta != g.mesh.delta) return pic_fail(PIC_ERR_ARG, "delta differs from the set-up domain");
    TRY(commit_locations());
    TRY(launch_finalize_deposit());
    return check_err_flags();
}


static int run_solver() {
    if (g.solver_mode == 0) return run_pcg_fp64(true);
    if (g.solver_mode == 3) return run_pcg_ic();
    return run_pcg_mixed();
}

static int solve_now() {
    TRY(launch_rhs());
    HIPCHK(hipEventRecord(g.ev[14], g.stream));
    const int rc = run_solver();
    HIPCHK(hipEventRecord(g.ev[15], g.stream));
    HIPCHK(hipEventSynchronize(g.ev[15]));
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, g.ev[14], g.ev[15]));
    g.last_solve_ms = ms;
    return rc;
}

int pic_IFE_Solve_2D(const double* Rho, double* Phi, int nnx, int nny, int delta, double time) {
    REQUIRE_DOMAIN();
    (void)time;
    if (nnx != g.mesh.nnx || nny != g.mesh.nny || delta != g.mesh.delta) return pic_fail(PIC_ERR_ARG, "mesh mismatch");
    int nn = g.mesh.nnodes;
    if (Rho) HIPCHK(hipMemcpy(g.d_rho, Rho, nn * sizeof(double), hipMemcpyHostToDevice));
    if (Phi) {
        HIPCHK(hipMemcpy(g.d_phi, Phi, nn * sizeof(double), hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_impose_dirichlet, dim3((nn + 255) / 256), dim3(256), 0, g.stream, nn, g.d_node_type2, g.d_dirichlet_phi, g.d_phi);
    }
    int rc = solve_now();
    if (Phi) {
        HIPCHK(hipStreamSynchronize(g.stream));
        HIPCHK(hipMemcpy(Phi, g.d_phi, nn * sizeof(double), hipMemcpyDeviceToHost));
    }
    return rc;
}

int pic_last_solve_info(int32_t* iterations, double* relres, int32_t* outer) {
    if (iterations) *iterations = g.last_iter;
    if (relres) *relres = g.last_relres;
    if (outer) *outer = g.last_outer;
    return PIC_OK;
}

int pic_last_solve_time(double* ms) {
    if (!ms) return pic_fail(PIC_ERR_ARG, "ms is NULL");
    *ms = g.last_solve_ms;
    return PIC_OK;
}

int pic_set_pcg_params(double tol, int32_t maxit) {
    if (!(tol > 0) || maxit <= 0) return pic_fail(PIC_ERR_ARG, "bad PCG parameters");
    g.pcg_tol = tol; g.pcg_maxit = maxit;
    return PIC_OK;
}

int pic_GetEField_SIDG_PPR(void) {
    REQUIRE_DOMAIN();
    TRY(launch_ppr());
    return PIC_OK;
}

int pic_MCCBundleInit(const pic_gas_params* gas, int electron_species, int ion_species) {
    REQUIRE_DOMAIN();
    if (!gas) return pic_fail(PIC_ERR_ARG, "gas is NULL");
    if (electron_species < 0 || electron_species >= PIC_MAX_SPECIES || !g.sp[electron_species].defined ||
        ion_species < 0 || ion_species >= PIC_MAX_SPECIES || !g.sp[ion_species].defined || ion_species == electron_species)
        return pic_fail(PIC_ERR_ARG, "species for MCC not defined");
    g.gas = *gas;
    g.e_species = electron_species; g.i_species = ion_species;
    double dt_SI = g.cfg.dt * g.time_ref;
    int sps[2] = {electron_species, ion_species};
    for (int k = 0; k < 2; ++k) {
        int s = sps[k];
        SpeciesHost& S = g.sp[s];
        setup_analytic_tables(s, k == 0, *gas, S.p.mass_kg, dt_SI);
        S.product_species = (k == 0) ? ion_species : -1;
        TRY(dupload(&S.prob, S.prob_host)); TRY(dupload(&S.rtype, S.rtype_host)); TRY(dupload(&S.thr, S.thr_host));
        double thr = std::floor(S.collision_ratio * 4294967296.0);
        S.select_threshold = thr >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)thr;
        S.has_mcc = true;
    }
    g.mcc_ready = true;
    return PIC_OK;
}

int pic_mcc_set_table(int isp, int nreact, int nsigma, double energy_min_log10, double energy_interval_log10,
                      const double* prob, const int32_t* reaction_type, const double* threshold, double sigma_max,
                      double collision_ratio, int product_species) {
    REQUIRE_DOMAIN();
    if (isp < 0 || isp >= PIC_MAX_SPECIES || !g.sp[isp].defined) return pic_fail(PIC_ERR_ARG, "species not defined");
    if (nreact < 1 || nreact > 4 || nsigma < 2 || !prob || !reaction_type || !threshold)
        return pic_fail(PIC_ERR_ARG, "bad table");
    SpeciesHost& S = g.sp[isp];
    S.nreact = nreact; S.nsigma = nsigma; S.emin = energy_min_log10; S.eint = energy_interval_log10;
    S.emax = std::pow(10.0, energy_min_log10 + energy_interval_log10 * (nsigma - 1));
    S.prob_host.assign(prob, prob + (size_t)nreact * nsigma);
    S.rtype_host.assign(reaction_type, reaction_type + nreact);
    S.thr_host.assign(threshold, threshold + nreact);
    S.sigma_max = sigma_max; S.collision_ratio = collision_ratio; S.product_species = product_species;
    TRY(dupload(&S.prob, S.prob_host)); TRY(dupload(&S.rtype, S.rtype_host)); TRY(dupload(&S.thr, S.thr_host));
    double thr = std::floor(collision_ratio * 4294967296.0);
    S.select_threshold = thr >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)thr;
    S.has_mcc = true;
    if (product_species >= 0) { g.e_species = isp; g.i_species = product_species; }
    g.mcc_ready = true;
    return PIC_OK;
}

int pic_mcc_get_table(int isp, double* prob, int64_t n, double* sigma_max, double* collision_ratio) {
    if (isp < 0 || isp >= PIC_MAX_SPECIES || !g.sp[isp].has_mcc) return pic_fail(PIC_ERR_ARG, "species has no table");
    SpeciesHost& S = g.sp[isp];
    if (prob) {
        if (n < (int64_t)S.prob_host.size()) return pic_fail(PIC_ERR_ARG, "destination too small");
        memcpy(prob, S.prob_host.data(), S.prob_host.size() * sizeof(double));
    }
    if (sigma_max) *sigma_max = S.sigma_max;
    if (collision_ratio) *collision_ratio = S.collision_ratio;
    return PIC_OK;
}

int pic_MCC(void) {
    REQUIRE_DOMAIN();
    if (!g.mcc_ready) return pic_fail(PIC_ERR_STATE, "no MCC tables registered");
    for (int s = 0; s < g.nspecies; ++s)
        if (g.sp[s].pending) return pic_fail(PIC_ERR_STATE, "MCC between MoveAll and GetParChg_2D");
    TRY(run_mcc());
    TRY(check_err_flags());
    advance_step();
    return PIC_OK;
}

int pic_set_run_params(int32_t sort_every, uint64_t seed, int32_t mcc_enabled) {
    if (sort_every < 0) return pic_fail(PIC_ERR_ARG, "sort_every < 0");
    g.sort_every = sort_every; g.seed = seed; g.mcc_enabled = mcc_enabled;
    return PIC_OK;
}

int pic_sort_particles(void) {
    REQUIRE_DOMAIN();
    for (int s = 0; s < g.nspecies; ++s)
        if (g.sp[s].pending) return pic_fail(PIC_ERR_STATE, "sort between MoveAll and GetParChg_2D");
    return run_sort_all();
}

int pic_sort_group(int32_t* slots) {
    REQUIRE_DOMAIN();
    if (!slots) return pic_fail(PIC_ERR_ARG, "null slots");
    *slots = 64 * pic_push_ppt();
    return PIC_OK;
}

int pic_get_step(int64_t* step) { *step = g.step; return PIC_OK; }
int pic_set_step(int64_t step) {
    if (step < 0) return pic_fail(PIC_ERR_ARG, "step < 0");
    g.step = step;
    for (int s = 0; s < PIC_MAX_SPECIES; ++s) g.moved[s] = false;
    return PIC_OK;
}

int pic_initial_field(void) {
    REQUIRE_DOMAIN();
    TRY(pic_GetParChg_2D(g.mesh.delta));
    TRY(solve_now());
    TRY(launch_ppr());
    return check_err_flags();
}

int pic_set_injection(int isp, const pic_injection* p) {
    REQUIRE_DOMAIN();
    if (isp < 0 || isp >= PIC_MAX_SPECIES || !g.sp[isp].defined) return pic_fail(PIC_ERR_ARG, "species not defined");
    if (!p) return pic_fail(PIC_ERR_ARG, "null injection parameters");
    InjectHost& q = g.inj[isp];
    q = InjectHost();
    if (p->density_m3 == 0.0) return PIC_OK;
    if (!(p->density_m3 > 0) || !(p->temp_eV > 0)) return pic_fail(PIC_ERR_ARG, "injection density and temperature must be > 0");
    if (!(p->y_hi > p->y_lo) || p->y_lo < g.mesh.ymin || p->y_hi > g.mesh.ymax || p->r_start < g.mesh.xmin ||
        p->r_start > g.mesh.xmax)
        return pic_fail(PIC_ERR_ARG, "injection slab outside the domain");
    const SpeciesHost& S = g.sp[isp];
    // Input_2D.f90:489-494
    q.dens = p->density_m3 / g.cfg.n_ref;
    q.tmpj = p->temp_eV * PIC_EVTOK / (g.cfg.T_ref_eV * PIC_EVTOK);
    q.vt = std::sqrt(q.tmpj / S.xm);
    for (int d = 0; d < 3; ++d) q.vd[d] = p->drift_ms[d] / g.v_ref;
    q.r_start = p->r_start; q.y_lo = p->y_lo; q.y_hi = p->y_hi;
    // InjectBeams_2D.f90:42-50 (SQRT(2.) is single precision there)
    const double s2 = (double)std::sqrt(2.0f);
    const double vd1 = q.vd[0], vt = q.vt;
    if (vd1 > 0.0)
        q.b_amb = vt / std::sqrt(2.0 * PIC_PI) * std::exp(-(vd1 * vd1) / (2.0 * vt * vt)) +
                  vd1 / 2.0 * (1.0 + std::erf(vd1 / (s2 * vt)));
    else
        q.b_amb = -vt / std::sqrt(2.0 * PIC_PI) * std::exp(-(vd1 * vd1) / (2.0 * vt * vt)) +
                  vd1 / 2.0 * (1.0 - std::erf(vd1 / (s2 * vt)));
    // :91-95, :106
    q.v_r = std::fabs(vd1) / (s2 * vt);
    q.fs1 = q.v_r + std::sqrt(q.v_r * q.v_r + 2.0);
    q.fs2 = 0.5 * (1.0 + q.v_r * (2.0 * q.v_r - q.fs1));
    q.qa = q.v_r < -3.0 ? std::fabs(q.v_r) + 1.0 : 3.0;
    q.sq2vt = s2 * vt;
    q.vmp = std::sqrt(2.0 * q.tmpj / S.xm);                  // Loadv.f90:14
    const double x2 = q.r_start + 0.5 * q.b_amb * g.cfg.dt;  // :73-76
    q.xl = std::min(q.r_start, x2); q.xu = std::max(q.r_start, x2);
    q.on = true;
    return PIC_OK;
}

int pic_InjectBeams_2D(int it, int delta, int64_t* added) {
    REQUIRE_DOMAIN();
    if (delta != g.mesh.delta) return pic_fail(PIC_ERR_ARG, "delta differs from the domain");
    for (int s = 0; s < g.nspecies; ++s)
        if (g.sp[s].pending) return pic_fail(PIC_ERR_STATE, "injection between MoveAll and GetParChg_2D");
    TRY(run_injection(it, added));
    return check_err_flags();
}

static bool injection_on() {
    for (int s = 0; s < PIC_MAX_SPECIES; ++s)
        if (g.inj[s].on) return true;
    return false;
}

// pic_step's wall injection and MCC on the side stream, after the deposit of the step (the
// event recorded after the finalisation) and in the reference's order (InjectBeams_2D, then
// MCC); neither reads the potential or the fields, so the results are those of the serial
// order. Run from overlap_point with solver work queued, so both overlap the solve.
static int side_inject_mcc() {
    int rc = PIC_OK;
    if (hipStreamWaitEvent(g.stream2, g.ev[2], 0) != hipSuccess) rc = pic_fail(PIC_ERR_HIP, "side stream wait");
    hipStream_t main = g.stream;
    g.stream = g.stream2;
    if (!rc && injection_on()) rc = run_injection((int)g.step, nullptr);    // Main_IFE_Test_2.f90:344
    if (!rc && g.mcc_enabled && g.mcc_ready) rc = run_mcc();
    g.stream = main;
    if (hipEventRecord(g.ev_side, g.stream2) != hipSuccess && !rc) rc = pic_fail(PIC_ERR_HIP, "side stream event");
    return rc;
}

int pic_step(int nsteps) {
    REQUIRE_DOMAIN();
    double acc[8] = {0};
    for (int k = 0; k < 6; ++k) g.kstat[k] = 0.0;
    for (int s = 0; s < PIC_MAX_SPECIES; ++s) g.ev_push_pending[s] = false;
    for (int it = 0; it < nsteps; ++it) {
        HIPCHK(hipEventRecord(g.ev[0], g.stream));
        for (int s = 0; s < g.nspecies; ++s) {
            if (!g.sp[s].defined) continue;
            TRY(pic_MoveAll(s, g.mesh.delta, 0));
        }
        HIPCHK(hipEventRecord(g.ev[1], g.stream));
        TRY(commit_locations());
        TRY(launch_finalize_deposit());
        HIPCHK(hipEventRecord(g.ev[2], g.stream));
        TRY(launch_rhs());
        HIPCHK(hipEventRecord(g.ev[3], g.stream));
        const bool ft = g.ft_every > 0 && g.solver_mode == 2 && learn_ready() && g.step > 0 &&
                        g.step % g.ft_every == 0;
        // injection and MCC go to the side stream at the solver's first host wait
        g.overlap_rc = PIC_OK;
        g.overlap_hook = side_inject_mcc;
        int rc_solve = PIC_OK;
        if (ft) {
            // online fine-tuning: this solve runs the plain inner CG and harvests pairs
            g.solver_mode = 1;
            rc_solve = learn_set_harvest(1, 0, 1);
            if (!rc_solve) rc_solve = run_pcg_mixed();
            g.solver_mode = 2;
            int rc2 = learn_set_harvest(0, 0, 1);
            if (!rc_solve) rc_solve = rc2;
        } else {
            rc_solve = run_solver();
        }
        overlap_point();                         // a solve that never waited: the side work now
        if (rc_solve) { (void)hipStreamWaitEvent(g.stream, g.ev_side, 0); return rc_solve; }
        HIPCHK(hipEventRecord(g.ev[4], g.stream));
        TRY(launch_ppr());
        HIPCHK(hipEventRecord(g.ev[5], g.stream));
        // the next step's push and the sort need the injected and collided particles
        HIPCHK(hipStreamWaitEvent(g.stream, g.ev_side, 0));
        HIPCHK(hipEventRecord(g.ev[6], g.stream));
        TRY(g.overlap_rc);
        if (ft) {
            // the training of the fine-tune, timed apart (host clock around the synchronous calls)
            HIPCHK(hipStreamSynchronize(g.stream));
            const auto t0 = std::chrono::steady_clock::now();
            double l0 = 0.0, l1 = 0.0;
            int32_t nset = 0, last_it = 0; int64_t nfb = 0; double lloss = 0.0;
            TRY(learn_info(&nset, &last_it, &nfb, &lloss));
            if (nset > 0)        // a solve that converged before its first snapshot harvests nothing
                TRY(learn_train(g.ft_train, g.ft_batch, g.ft_lr, g.ft_seed + (uint64_t)g.ft_count, &l0, &l1));
            g.ft_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            g.ft_count++;
            g.ft_last_loss = l1;
        }
        advance_step();
        if (g.sort_every > 0 && g.step % g.sort_every == 0) TRY(run_sort_all());
        HIPCHK(hipEventRecord(g.ev[7], g.stream));
        HIPCHK(hipEventSynchronize(g.ev[7]));
        float ms;
        for (int k = 0; k < 7; ++k) { hipEventElapsedTime(&ms, g.ev[k], g.ev[k + 1]); acc[k] += ms; }
        hipEventElapsedTime(&ms, g.ev[0], g.ev[7]); acc[7] += ms;
        for (int s = 0; s < PIC_MAX_SPECIES; ++s) {
            if (!g.ev_push_pending[s]) continue;
            // the push kernel alone (the wall kernel that follows it is timed apart)
            hipEventElapsedTime(&ms, g.ev_push[s][0], g.ev_push[s][1]);
            g.kstat[2] += ms; g.kstat[3] += 1.0;
            g.ev_push_pending[s] = false;
        }
        TRY(check_err_flags());
    }
    for (int k = 0; k < 8; ++k) g.phase_ms[k] = acc[k];
    return PIC_OK;
}

int pic_get_phase_times(double* ms, int n) {
    for (int k = 0; k < n && k < 8; ++k) ms[k] = g.phase_ms[k];
    for (int k = 8; k < n && k < 14; ++k) ms[k] = g.kstat[k - 8];
    return PIC_OK;
}

// ----------------------------------------------------------------------- fields
int pic_get_field(int which, double* dst, int64_t n) {
    REQUIRE_DOMAIN();
    int64_t nn = g.mesh.nnodes;
    const double* src = nullptr;
    int64_t cnt = 0;
    switch (which) {
    case PIC_F_RHO: src = g.d_rho; cnt = nn; break;
    case PIC_F_PHI: src = g.d_phi; cnt = nn; break;
    case PIC_F_EFX: src = g.d_efx; cnt = nn; break;
    case PIC_F_EFY: src = g.d_efy; cnt = nn; break;
    case PIC_F_COEF_PPR: src = g.d_coef; cnt = 6 * nn; break;
    case PIC_F_RHO_S: src = g.d_rho_s; cnt = nn * g.nspecies; break;
    case PIC_F_RHO_P: src = g.d_rho_p; cnt = nn * g.nspecies; break;
    case PIC_F_EK_TOT: src = g.d_ek; cnt = nn * g.nspecies; break;
    case PIC_F_EK_S: src = g.d_eks; cnt = nn * g.nspecies; break;
    case PIC_F_RHS: src = g.d_b; cnt = g.nunknowns; break;
    default: return pic_fail(PIC_ERR_ARG, "unknown field");
    }
    if (n < cnt) return pic_fail(PIC_ERR_ARG, "destination too small");
    HIPCHK(hipStreamSynchronize(g.stream));
    HIPCHK(hipMemcpy(dst, src, cnt * sizeof(double), hipMemcpyDeviceToHost));
    return PIC_OK;
}


int pic_set_field(int which, const double* src, int64_t n) {
    REQUIRE_DOMAIN();
    int64_t nn = g.mesh.nnodes;
    switch (which) {
    case PIC_F_RHO: if (n != nn) return pic_fail(PIC_ERR_ARG, "size"); HIPCHK(hipMemcpy(g.d_rho, src, nn * 8, hipMemcpyHostToDevice)); break;
    case PIC_F_PHI: if (n != nn) return pic_fail(PIC_ERR_ARG, "size"); HIPCHK(hipMemcpy(g.d_phi, src, nn * 8, hipMemcpyHostToDevice)); break;
    case PIC_F_EFX: if (n != nn) return pic_fail(PIC_ERR_ARG, "size"); HIPCHK(hipMemcpy(g.d_efx, src, nn * 8, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_d2f_n, dim3((int)((nn + 255) / 256)), dim3(256), 0, g.stream, nn, g.d_efx, g.d_efxf); break;
    case PIC_F_EFY: if (n != nn) return pic_fail(PIC_ERR_ARG, "size"); HIPCHK(hipMemcpy(g.d_efy, src, nn * 8, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_d2f_n, dim3((int)((nn + 255) / 256)), dim3(256), 0, g.stream, nn, g.d_efy, g.d_efyf); break;
    case PIC_F_COEF_PPR: if (n != 6 * nn) return pic_fail(PIC_ERR_ARG, "size");
        HIPCHK(hipMemcpy(g.d_coef, src, 6 * nn * 8, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_coef_split, dim3((int)((nn + 255) / 256)), dim3(256), 0, g.stream, (int)nn, g.d_coef, g.d_coef5, g.d_coeff); break;
    default: return pic_fail(PIC_ERR_ARG, "field not settable");
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(g.stream));
    return PIC_OK;
}

// ----------------------------------------------------------------------- drop-ins
// Host check of the reference's SPARSE array (allocates nothing): row pointers SROW(1..N+1)
// start at 1, never decrease and end within NZ + 1; every used JCOL lies in 1..N. Returns
// the stored-entry count SROW(N+1) - 1 in *nnz.
static int check_sparse(const pic_sparse* matrix, int NZ, int N, int* nnz) {
    if (matrix[0].SROW != 1 || matrix[N].SROW - 1 > NZ) return pic_fail(PIC_ERR_ARG, "SROW row pointers inconsistent");
    for (int i = 0; i < N; ++i)
        if (matrix[i + 1].SROW < matrix[i].SROW) return pic_fail(PIC_ERR_ARG, "SROW row pointers decrease");
    const int n = matrix[N].SROW - 1;
    for (int k = 0; k < n; ++k)
        if (matrix[k].JCOL < 1 || matrix[k].JCOL > N) return pic_fail(PIC_ERR_ARG, "JCOL out of range");
    *nnz = n;
    return PIC_OK;
}

int pic_MY_JPCG_Solver(const pic_sparse* matrix, double* R, double* X, const double* M, int NZ, int N, int* ITER) {
    REQUIRE_INIT();
    if (!matrix || !R || !X || !M || !ITER || N <= 0 || NZ < N + 1)
        return pic_fail(PIC_ERR_ARG, "bad MY_JPCG_Solver arguments");
    int nnz = 0;
    TRY(check_sparse(matrix, NZ, N, &nnz));
    std::vector<int> rowptr(N + 1);
    for (int i = 0; i <= N; ++i) rowptr[i] = matrix[i].SROW - 1;
    std::vector<int> col(nnz);
    std::vector<double> val(nnz);
    for (int k = 0; k < nnz; ++k) { col[k] = matrix[k].JCOL - 1; val[k] = matrix[k].K; }
    int *drp = nullptr, *dcol = nullptr;
    double *dval = nullptr, *ddiag = nullptr, *db = nullptr, *dx = nullptr, *dr = nullptr;
    HIPCHK(hipMalloc(&drp, (N + 1) * sizeof(int))); HIPCHK(hipMalloc(&dcol, nnz * sizeof(int)));
    HIPCHK(hipMalloc(&dval, nnz * sizeof(double))); HIPCHK(hipMalloc(&ddiag, N * sizeof(double)));
    HIPCHK(hipMalloc(&db, N * sizeof(double))); HIPCHK(hipMalloc(&dx, N * sizeof(double))); HIPCHK(hipMalloc(&dr, N * sizeof(double)));
    HIPCHK(hipMemcpy(drp, rowptr.data(), (N + 1) * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dcol, col.data(), nnz * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dval, val.data(), nnz * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ddiag, M, N * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(db, R, N * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dx, X, N * sizeof(double), hipMemcpyHostToDevice));
    // scratch for partial sums sized to this system
    double* save_partials = g.d_partials;
    double* part = nullptr;
    HIPCHK(hipMalloc(&part, std::max<size_t>(4 * ((size_t)N / 256 + 2) + 64, 4096) * sizeof(double)));
    g.d_partials = part;
    int iters = 0, ierr = 0;
    int rc = pcg_generic_fp64(N, drp, dcol, dval, ddiag, db, dx, g.pcg_tol, g.pcg_maxit, &iters, &ierr, dr);
    g.d_partials = save_partials;
    if (rc == PIC_OK) {
        HIPCHK(hipMemcpy(X, dx, N * sizeof(double), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(R, dr, N * sizeof(double), hipMemcpyDeviceToHost));
    }
    hipFree(part); hipFree(drp); hipFree(dcol); hipFree(dval); hipFree(ddiag); hipFree(db); hipFree(dx); hipFree(dr);
    if (rc) return rc;
    if (ierr == 1) return pic_fail(PIC_ERR_PCG_MAXITER, "MY_PCG IERR=1 (ITER left unset, as in the reference)");
    *ITER = iters - 1;     // MY_PCG.f90:129-133: ITER = i-1
    if (ierr == 2 || ierr == 3) return pic_fail(PIC_ERR_PCG_BREAKDOWN, "MY_PCG IERR=2/3");
    return PIC_OK;
}

int pic_MY_JPCG_Solver_noalloc(const pic_sparse* matrix, double* R, double* X, const double* M, int NZ, int N,
                               int* ITER) {
    REQUIRE_INIT();
    if (!matrix || !R || !X || !M || !ITER || N <= 0 || NZ < N + 1)
        return pic_fail(PIC_ERR_ARG, "bad MY_JPCG_Solver arguments");
    if (N > PIC_NOALLOC_MAX_N || NZ > PIC_NOALLOC_MAX_NZ) {
        char buf[200];
        snprintf(buf, sizeof buf, "MY_JPCG_Solver_noalloc: N = %d, NZ = %d exceed the scratch limits N <= %d, NZ <= %d",
                 N, NZ, PIC_NOALLOC_MAX_N, PIC_NOALLOC_MAX_NZ);
        return pic_fail(PIC_ERR_CAPACITY, buf);
    }
    int nnz = 0;
    TRY(check_sparse(matrix, NZ, N, &nnz));
    char* base = (char*)g_scratch_dev[g.device];
    if (!base) return pic_fail(PIC_ERR_STATE, "scratch address not set up (pic_init)");
    // carve the scratch: SPARSE array, diag, b, x, r, z, p, p2, q, partials, scalars
    pic_sparse* dmat = (pic_sparse*)base;
    double* v[8];
    for (int k = 0; k < 8; ++k) v[k] = (double*)(base + kNoallocMatBytes + (size_t)k * kNoallocVecBytes);
    double* part = (double*)(base + kNoallocMatBytes + 8 * kNoallocVecBytes);
    double* scal = (double*)(base + kNoallocMatBytes + 8 * kNoallocVecBytes + kNoallocPartBytes);
    HIPCHK(hipMemcpyAsync(dmat, matrix, (size_t)NZ * sizeof(pic_sparse), hipMemcpyHostToDevice, g.stream));
    HIPCHK(hipMemcpyAsync(v[0], M, (size_t)N * sizeof(double), hipMemcpyHostToDevice, g.stream));
    HIPCHK(hipMemcpyAsync(v[1], R, (size_t)N * sizeof(double), hipMemcpyHostToDevice, g.stream));
    HIPCHK(hipMemcpyAsync(v[2], X, (size_t)N * sizeof(double), hipMemcpyHostToDevice, g.stream));
    int iters = 0, ierr = 0;
    int rc = pcg_sparse_aos_fp64(N, dmat, v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7], part, PIC_NOALLOC_PSTRIDE,
                                 scal, g.pcg_tol, g.pcg_maxit, &iters, &ierr);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(X, v[2], (size_t)N * sizeof(double), hipMemcpyDeviceToHost, g.stream));
    HIPCHK(hipMemcpyAsync(R, v[3], (size_t)N * sizeof(double), hipMemcpyDeviceToHost, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    if (ierr == 1) return pic_fail(PIC_ERR_PCG_MAXITER, "MY_PCG IERR=1 (ITER left unset, as in the reference)");
    *ITER = iters - 1;     // MY_PCG.f90:129-133: ITER = i-1
    if (ierr == 2 || ierr == 3) return pic_fail(PIC_ERR_PCG_BREAKDOWN, "MY_PCG IERR=2/3");
    return PIC_OK;
}

int pic_Get_ParticleEField_PPR(int64_t n, const double* x, const double* y, const int32_t* EleIndex, double* Ex, double* Ey) {
    REQUIRE_DOMAIN();
    if (n <= 0) return PIC_OK;
    double* d; int* de;
    HIPCHK(hipMalloc(&d, 4 * (size_t)n * sizeof(double)));
    HIPCHK(hipMalloc(&de, (size_t)n * sizeof(int)));
    HIPCHK(hipMemcpy(d, x, n * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d + n, y, n * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(de, EleIndex, n * 4, hipMemcpyHostToDevice));
    TRY(launch_ppr_gather_points(n, d, d + n, de, d + 2 * n, d + 3 * n));
    HIPCHK(hipStreamSynchronize(g.stream));
    HIPCHK(hipMemcpy(Ex, d + 2 * n, n * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(Ey, d + 3 * n, n * 8, hipMemcpyDeviceToHost));
    hipFree(d); hipFree(de);
    return PIC_OK;
}

static int locate_common(int64_t n, const double* x, const double* y, int32_t* out, int mode) {
    REQUIRE_DOMAIN();
    if (n <= 0) return PIC_OK;
    double* d; int* de;
    HIPCHK(hipMalloc(&d, 2 * (size_t)n * sizeof(double)));
    HIPCHK(hipMalloc(&de, (size_t)n * sizeof(int)));
    HIPCHK(hipMemcpy(d, x, n * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d + n, y, n * 8, hipMemcpyHostToDevice));
    TRY(launch_locate_points(n, d, d + n, de, mode));
    HIPCHK(hipStreamSynchronize(g.stream));
    HIPCHK(hipMemcpy(out, de, n * 4, hipMemcpyDeviceToHost));
    hipFree(d); hipFree(de);
    return check_err_flags();
}

int pic_locate(int64_t n, const double* x, const double* y, int32_t* n_element) { return locate_common(n, x, y, n_element, 0); }
int pic_InjectPositioning(int64_t n, const double* x, const double* y, int32_t* n_element) { return locate_common(n, x, y, n_element, 1); }

// ParticlePositioning on the unrefined mesh (ParticlePositioning.f90:240-363): the final
// element is the given one; traverse_flag marks inner-edge points that are not corners.
int pic_ParticlePositioning(int64_t n, const double* x, const double* y, int32_t* n_element, int32_t* traverse_flag) {
    REQUIRE_DOMAIN();
    const MeshParams& m = g.mesh;
    const double sv = 1e-12;
    for (int64_t k = 0; k < n; ++k) {
        int e = n_element[k] - 1;
        traverse_flag[k] = 0;
        if (e < 0 || e >= m.nelem) return pic_fail(PIC_ERR_POSITION, "element out of range");
        int ci = e / m.ney, cj = e % m.ney;
        double l = x[k] - g.h_xs[ci], r = x[k] - g.h_xs[ci + 1], b = y[k] - g.h_ys[cj], t = y[k] - g.h_ys[cj + 1];
        bool L = std::fabs(l) < sv, Rr = std::fabs(r) < sv, B = std::fabs(b) < sv, T = std::fabs(t) < sv;
        if (!(L || Rr || B || T)) continue;
        if ((L && T) || (Rr && T) || (L && B) || (Rr && B)) continue;
        if (L) traverse_flag[k] = std::fabs(g.h_xs[ci] - m.xmin) < sv ? 0 : 1;
        else if (Rr) traverse_flag[k] = std::fabs(g.h_xs[ci + 1] - m.xmax) < sv ? 0 : 1;
        else if (B) traverse_flag[k] = std::fabs(g.h_ys[cj] - m.ymin) < sv ? 0 : 1;
        else if (T) traverse_flag[k] = std::fabs(g.h_ys[cj + 1] - m.ymax) < sv ? 0 : 1;
    }
    return PIC_OK;
}

// generate_Elementmap.f90:19-57 on the unrefined mesh (repeat_refinement must be 0).
int pic_generate_Elementmap(int32_t* Element_Map, int repeat_refinement) {
    REQUIRE_DOMAIN();
    if (repeat_refinement != 0) return pic_fail(PIC_ERR_UNSUPPORTED, "refined meshes (repeat_refinement > 0) are not built");
    const MeshParams& m = g.mesh;
    double dx = m.hx / 1.0, dy = m.hy / 1.0;
    for (int i = 1; i <= m.nex; ++i)
        for (int j = 1; j <= m.ney; ++j) {
            double x = m.xmin + (dx) * (2 * i - 1) / 2;
            double y = m.ymin + (dy) * (2 * j - 1) / 2;
            double rxp = (x - m.vox) / m.hx;
            int k = (int)rxp;
            double ryp = (y - m.voy) / m.hy;
            int l = (int)ryp;
            if (l == m.nny) l = l - 1;
            if (k == m.nnx) k = k - 1;
            int ne = l + (k - 1) * (m.nny - 1);
            Element_Map[(size_t)(i - 1) + (size_t)(j - 1) * m.nex] = ne;
        }
    return PIC_OK;
}

// ----------------------------------------------------------------------- learned preconditioner
int pic_learn_setup(int32_t max_levels, const int32_t* channels, int32_t nchannels, uint64_t seed) {
    REQUIRE_DOMAIN();
    return learn_setup(max_levels, channels, nchannels, seed);
}
int pic_learn_harvest(int32_t enable, int32_t capacity, int32_t every) {
    REQUIRE_DOMAIN();
    if (capacity < 1) return pic_fail(PIC_ERR_ARG, "capacity must be positive");
    return learn_set_harvest(enable, capacity, every);
}
int pic_learn_train(int32_t steps, int32_t batch, double lr, uint64_t seed, double* loss_first, double* loss_last) {
    REQUIRE_DOMAIN();
    return learn_train(steps, batch, lr, seed, loss_first, loss_last);
}
int pic_learn_eval(int32_t npairs, double* mean_loss) {
    REQUIRE_DOMAIN();
    if (!mean_loss) return pic_fail(PIC_ERR_ARG, "mean_loss is NULL");
    return learn_eval(npairs, mean_loss);
}
int pic_learn_data_mix(double harvested_share) {
    REQUIRE_DOMAIN();
    return learn_set_mix(harvested_share);
}
int pic_learn_finetune(int32_t every, int32_t train_steps, int32_t batch, double lr, uint64_t seed) {
    REQUIRE_DOMAIN();
    if (every < 0 || train_steps < 0 || batch < 1 || batch > 64 || !(lr > 0)) return pic_fail(PIC_ERR_ARG, "bad fine-tune parameters");
    if (every > 0 && !learn_ready()) return pic_fail(PIC_ERR_STATE, "pic_learn_setup not called");
    g.ft_every = every; g.ft_train = train_steps; g.ft_batch = batch; g.ft_lr = lr; g.ft_seed = seed;
    return PIC_OK;
}
int pic_learn_finetune_info(int64_t* count, double* train_ms, double* last_loss) {
    if (count) *count = g.ft_count;
    if (train_ms) *train_ms = g.ft_ms;
    if (last_loss) *last_loss = g.ft_last_loss;
    return PIC_OK;
}
int pic_learn_info(int32_t* dataset, int32_t* last_fcg_iterations, int64_t* fallbacks, double* last_loss) {
    return learn_info(dataset, last_fcg_iterations, fallbacks, last_loss);
}
int pic_learn_apply(const float* r, float* z, int64_t n) {
    REQUIRE_DOMAIN();
    if (!r || !z) return pic_fail(PIC_ERR_ARG, "NULL buffer");
    return learn_apply_host(r, z, n);
}
int pic_learn_get_params(float* dst, int64_t n, int64_t* nparams) {
    REQUIRE_DOMAIN();
    return learn_params(dst, nullptr, n, nparams);
}
int pic_learn_set_params(const float* src, int64_t n) {
    REQUIRE_DOMAIN();
    if (!src) return pic_fail(PIC_ERR_ARG, "NULL buffer");
    return learn_params(nullptr, src, n, nullptr);
}
int pic_learn_layout(int32_t* levels, int32_t* H, int32_t* W, int32_t* C, int64_t* offsets) {
    REQUIRE_DOMAIN();
    if (!levels || !H || !W || !C || !offsets) return pic_fail(PIC_ERR_ARG, "NULL buffer");
    return learn_layout(levels, H, W, C, offsets);
}
int pic_learn_set_pairs(int32_t npairs, const float* r, const float* e) {
    REQUIRE_DOMAIN();
    if (!r || !e) return pic_fail(PIC_ERR_ARG, "NULL buffer");
    return learn_set_pairs(npairs, r, e);
}
int pic_learn_get_pair(int32_t k, float* r, float* e, double* ea) {
    REQUIRE_DOMAIN();
    return learn_get_pair(k, r, e, ea);
}
int pic_learn_gradient(int32_t npairs, const int32_t* idx, double* loss, float* grad, int64_t n) {
    REQUIRE_DOMAIN();
    if (!idx || !loss || !grad) return pic_fail(PIC_ERR_ARG, "NULL buffer");
    return learn_gradient(npairs, idx, loss, grad, n);
}

int pic_warm_setup(int32_t levels, int32_t channels, uint64_t seed) {
    REQUIRE_DOMAIN();
    return warm_setup(levels, channels, seed);
}
int pic_warm_train(int32_t steps, int32_t batch, double lr, uint64_t seed, double* loss_first, double* loss_last) {
    REQUIRE_DOMAIN();
    return warm_train(steps, batch, lr, seed, loss_first, loss_last);
}
int pic_warm_enable(int32_t enable) {
    REQUIRE_DOMAIN();
    return warm_enable(enable);
}
int pic_warm_eval(int32_t npairs, double* mean_loss, double* mean_loss_scaled) {
    REQUIRE_DOMAIN();
    if (!mean_loss) return pic_fail(PIC_ERR_ARG, "NULL buffer");
    return warm_eval(npairs, mean_loss, mean_loss_scaled);
}
int pic_warm_apply(const float* r, float* y, int64_t n) {
    REQUIRE_DOMAIN();
    if (!r || !y) return pic_fail(PIC_ERR_ARG, "NULL buffer");
    return warm_apply_host(r, y, n);
}
int pic_warm_get_params(float* dst, int64_t n, int64_t* nparams) {
    REQUIRE_DOMAIN();
    return warm_params(dst, nullptr, n, nparams);
}
int pic_warm_set_params(const float* src, int64_t n) {
    REQUIRE_DOMAIN();
    if (!src) return pic_fail(PIC_ERR_ARG, "NULL buffer");
    return warm_params(nullptr, src, n, nullptr);
}
int pic_warm_layout(int32_t* levels, int32_t* H, int32_t* W, int32_t* channels) {
    REQUIRE_DOMAIN();
    if (!levels || !H || !W || !channels) return pic_fail(PIC_ERR_ARG, "NULL buffer");
    return warm_layout(levels, H, W, channels);
}
int pic_warm_gradient(int32_t npairs, const int32_t* idx, double* loss, float* grad, int64_t n) {
    REQUIRE_DOMAIN();
    if (!idx || !loss || !grad) return pic_fail(PIC_ERR_ARG, "NULL buffer");
    return warm_gradient(npairs, idx, loss, grad, n);
}
int pic_warm_info(int32_t* initial_pairs, int32_t* enabled, double* last_alpha, double* last_loss, int64_t* applied) {
    return warm_info(initial_pairs, enabled, last_alpha, last_loss, applied);
}

}  // extern "C"
