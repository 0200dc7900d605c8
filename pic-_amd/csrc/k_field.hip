// Field kernels: RHS assembly (rhs = B*rho - c), PPR gradient recovery, and the
// Jacobi-preconditioned conjugate gradient of MY_PCG (FP64), with an FP32 inner solve
// inside FP64 iterative refinement.
//
// PCG layout: CSR with int32 columns, one row per thread. Each iteration is two kernels:
//   k_pcg_spmv : p <- z + beta p (recomputed on the fly for the 9-point neighbourhood),
//                q <- A p, partial p.q
//   k_pcg_update: x += alpha p, r -= alpha q, z = r / diag, partial r.z and r.r
// Block partials are reduced by the last block to finish (fixed block order, so the
// scalars are deterministic); that block also forms alpha / beta and the convergence
// flag. Kernels of an iteration after convergence return at once, so the host checks
// the flag only every few iterations.
#include "pic_device.h"

#define PCG_BLOCK 256

// ---------------------------------------------------------------- RHS
__global__ void k_rhs(int n, const int* rowptr, const int* col, const double* val,
                      const double* rho, const double* c, double* b) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) s = s + val[k] * rho[col[k]];
    b[i] = s - c[i];
}

int launch_rhs() {
    int n = g.nunknowns;
    hipLaunchKernelGGL(k_rhs, dim3((n + 255) / 256), dim3(256), 0, g.stream, n, g.d_B_rowptr, g.d_B_col,
                       g.d_B_val, g.d_rho, g.d_c, g.d_b);
    HIPCHK(hipGetLastError());
    return PIC_OK;
}

// ---------------------------------------------------------------- PPR recovery
// GetEfield_SIDG_PPR.f90:153-241 with the geometry-only least-squares rows precomputed
// per patch class: temp = R * phi_patch; coef(1) = temp(1), coef(2:3) = temp(2:3)/h,
// coef(4:6) = temp(4:6)/(h*h); efx = -coef(2), efy = -coef(3) (:301-304).
struct PprArgs {
    int nnodes, nny;
    const int* cls;
    const PprClass* classes;
    const double* phi;
    double* coef6; double* coef5; float* coeff;
    double* efx; double* efy; float* efxf; float* efyf;
};

__global__ __launch_bounds__(256) void k_ppr(PprArgs a) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= a.nnodes) return;
    int c = a.cls[n];
    double t[6] = {0, 0, 0, 0, 0, 0};
    double co[6] = {0, 0, 0, 0, 0, 0};
    if (c >= 0) {
        const PprClass& P = a.classes[c];
        int i = n / a.nny, j = n - i * a.nny;
        for (int k = 0; k < P.n; ++k) {
            double ph = a.phi[(i + P.di[k]) * a.nny + (j + P.dj[k])];
#pragma unroll
            for (int r = 0; r < 6; ++r) t[r] = t[r] + P.R[r][k] * ph;
        }
        double h = P.h;
        co[0] = t[0];
        co[1] = t[1] / h; co[2] = t[2] / h;
        co[3] = t[3] / (h * h); co[4] = t[4] / (h * h); co[5] = t[5] / (h * h);
    }
#pragma unroll
    for (int r = 0; r < 6; ++r) a.coef6[(size_t)n * 6 + r] = co[r];
#pragma unroll
    for (int r = 0; r < 5; ++r) { a.coef5[(size_t)n * 5 + r] = co[r + 1]; a.coeff[(size_t)n * 5 + r] = (float)co[r + 1]; }
    double ex = -co[1], ey = -co[2];
    a.efx[n] = ex; a.efy[n] = ey;
    a.efxf[n] = (float)ex; a.efyf[n] = (float)ey;
}

int launch_ppr() {
    PprArgs a;
    a.nnodes = g.mesh.nnodes; a.nny = g.mesh.nny;
    a.cls = g.d_ppr_class; a.classes = g.d_ppr_classes; a.phi = g.d_phi;
    a.coef6 = g.d_coef; a.coef5 = g.d_coef5; a.coeff = g.d_coeff;
    a.efx = g.d_efx; a.efy = g.d_efy; a.efxf = g.d_efxf; a.efyf = g.d_efyf;
    hipLaunchKernelGGL(k_ppr, dim3((a.nnodes + 255) / 256), dim3(256), 0, g.stream, a);
    HIPCHK(hipGetLastError());
    return PIC_OK;
}

// ---------------------------------------------------------------- PCG
// scalar slots in d_scal
enum { S_RHO = 0, S_RHO_OLD, S_ALPHA, S_BETA, S_NORMRHS, S_NORMR, S_DONE, S_ITER, S_IERR, S_TOL2, S_MAXIT, S_PQ, S_NSCAL };

template <int NV>
__device__ inline void block_reduce_store(double v[NV], double* partials, int nblocks) {
    __shared__ double sh[NV][PCG_BLOCK / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        double s = v[k];
        for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
        if (lane == 0) sh[k][wid] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            double s = 0.0;
            for (int w = 0; w < PCG_BLOCK / 64; ++w) s += sh[k][w];
            partials[(size_t)k * nblocks + blockIdx.x] = s;
        }
    }
}

// Returns true in thread 0 of the last block to arrive (after all partials are visible).
__device__ inline bool last_block(unsigned int* counter) {
    __shared__ bool am_last;
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned int t = atomicAdd(counter, 1u);
        am_last = (t == gridDim.x - 1);
    }
    __syncthreads();
    return am_last;
}

__device__ inline double sum_partials(const double* partials, int nblocks) {
    double s = 0.0;
    for (int b = 0; b < nblocks; ++b) s += partials[b];
    return s;
}

// r = b - A x, z = r / diag, partials of r.z and b.b
template <typename T>
__global__ __launch_bounds__(PCG_BLOCK) void k_pcg_init(int n, const int* rowptr, const int* col, const T* val,
                                                        const T* diag, const T* b, const T* x, T* r, T* z,
                                                        double* scal, double* partials, unsigned int* counter) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double v[2] = {0.0, 0.0};
    if (i < n) {
        T s = 0;
        for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) s = s + val[k] * x[col[k]];
        T ri = b[i] - s;
        T zi = ri / diag[i];
        r[i] = ri; z[i] = zi;
        v[0] = (double)ri * (double)zi;
        v[1] = (double)b[i] * (double)b[i];
    }
    block_reduce_store<2>(v, partials, gridDim.x);
    if (last_block(counter) && threadIdx.x == 0) {
        scal[S_RHO] = sum_partials(partials, gridDim.x);
        scal[S_NORMRHS] = sum_partials(partials + gridDim.x, gridDim.x);
        scal[S_BETA] = 0.0;
        scal[S_DONE] = 0.0; scal[S_ITER] = 0.0; scal[S_IERR] = 0.0;
        if (scal[S_NORMRHS] == 0.0) scal[S_DONE] = 1.0;   // zero RHS: x = 0 is exact
        *counter = 0;
    }
}

template <typename T>
__global__ __launch_bounds__(PCG_BLOCK) void k_pcg_spmv(int n, const int* rowptr, const int* col, const T* val,
                                                        const T* z, const T* p_old, T* p_new, T* q,
                                                        double* scal, double* partials, unsigned int* counter) {
    if (scal[S_DONE] != 0.0) return;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    const T beta = (T)scal[S_BETA];
    const bool first = scal[S_ITER] == 0.0;
    double v[1] = {0.0};
    if (i < n) {
        T s = 0;
        for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) {
            int j = col[k];
            T pj = first ? z[j] : z[j] + beta * p_old[j];
            s = s + val[k] * pj;
        }
        T pi = first ? z[i] : z[i] + beta * p_old[i];
        p_new[i] = pi; q[i] = s;
        v[0] = (double)pi * (double)s;
    }
    block_reduce_store<1>(v, partials, gridDim.x);
    if (last_block(counter) && threadIdx.x == 0) {
        double pq = sum_partials(partials, gridDim.x);
        scal[S_PQ] = pq;
        if (pq == 0.0) { scal[S_IERR] = 3.0; scal[S_DONE] = 1.0; scal[S_ALPHA] = 0.0; }
        else scal[S_ALPHA] = scal[S_RHO] / pq;
        *counter = 0;
    }
}

template <typename T>
__global__ __launch_bounds__(PCG_BLOCK) void k_pcg_update(int n, const T* diag, const T* p, const T* q, T* x, T* r, T* z,
                                                          double* scal, double* partials, unsigned int* counter) {
    if (scal[S_DONE] != 0.0) return;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    const T alpha = (T)scal[S_ALPHA];
    double v[2] = {0.0, 0.0};
    if (i < n) {
        T xi = x[i] + alpha * p[i];
        T ri = r[i] - alpha * q[i];
        T zi = ri / diag[i];
        x[i] = xi; r[i] = ri; z[i] = zi;
        v[0] = (double)ri * (double)zi;
        v[1] = (double)ri * (double)ri;
    }
    block_reduce_store<2>(v, partials, gridDim.x);
    if (last_block(counter) && threadIdx.x == 0) {
        double rho_new = sum_partials(partials, gridDim.x);
        double normr = sum_partials(partials + gridDim.x, gridDim.x);
        double it = scal[S_ITER] + 1.0;
        scal[S_ITER] = it;
        scal[S_NORMR] = normr;
        double rho_old = scal[S_RHO];
        if (normr / scal[S_NORMRHS] <= scal[S_TOL2]) scal[S_DONE] = 1.0;
        else if (it >= scal[S_MAXIT]) { scal[S_DONE] = 1.0; scal[S_IERR] = 1.0; }
        else if (rho_old == 0.0) { scal[S_DONE] = 1.0; scal[S_IERR] = 2.0; }
        else { scal[S_BETA] = rho_new / rho_old; }
        scal[S_RHO_OLD] = rho_old;
        scal[S_RHO] = rho_new;
        *counter = 0;
    }
}

// true residual ||b - A x||^2 and ||b||^2 (FP64)
__global__ __launch_bounds__(PCG_BLOCK) void k_resid64(int n, const int* rowptr, const int* col, const double* val,
                                                       const double* b, const double* x, double* r,
                                                       double* out2, double* partials, unsigned int* counter) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double v[2] = {0.0, 0.0};
    if (i < n) {
        double s = 0.0;
        for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) s = s + val[k] * x[col[k]];
        double ri = b[i] - s;
        if (r) r[i] = ri;
        v[0] = ri * ri; v[1] = b[i] * b[i];
    }
    block_reduce_store<2>(v, partials, gridDim.x);
    if (last_block(counter) && threadIdx.x == 0) {
        out2[0] = sum_partials(partials, gridDim.x);
        out2[1] = sum_partials(partials + gridDim.x, gridDim.x);
        *counter = 0;
    }
}

__global__ void k_d2f(int n, const double* a, float* b) { int i = blockIdx.x * blockDim.x + threadIdx.x; if (i < n) b[i] = (float)a[i]; }
__global__ void k_axpy_f2d(int n, const float* a, double* b) { int i = blockIdx.x * blockDim.x + threadIdx.x; if (i < n) b[i] = b[i] + (double)a[i]; }
__global__ void k_zero_f(int n, float* a) { int i = blockIdx.x * blockDim.x + threadIdx.x; if (i < n) a[i] = 0.0f; }

// gather phi at unknowns / scatter back (IFE_Solve_2D.f90:80-84, 243-247, 514-516)
__global__ void k_phi_to_x(int n, const int* unknown_node, const double* phi, double* x) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = phi[unknown_node[i]];
}
__global__ void k_x_to_phi(int n, const int* unknown_node, const double* x, double* phi) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) phi[unknown_node[i]] = x[i];
}

// One Jacobi-PCG solve on device vectors (storage type T). Host loop in chunks.
template <typename T>
static int pcg_run(int n, const int* rowptr, const int* col, const T* val, const T* diag, const T* b, T* x,
                   T* r, T* z, T* p, T* p2, T* q, double tol, int maxit, int* iters, int* ierr) {
    int nb = (n + PCG_BLOCK - 1) / PCG_BLOCK;
    double* scal = g.d_scal;
    double h2[2] = {tol * tol, (double)maxit};
    HIPCHK(hipMemcpyAsync(scal + S_TOL2, h2, 2 * sizeof(double), hipMemcpyHostToDevice, g.stream));
    hipLaunchKernelGGL((k_pcg_init<T>), dim3(nb), dim3(PCG_BLOCK), 0, g.stream, n, rowptr, col, val, diag, b, x, r, z,
                       scal, g.d_partials, g.d_sync);
    HIPCHK(hipGetLastError());
    T* pa = p; T* pb = p2;
    const int CHUNK = 16;
    int done_iters = 0;
    for (;;) {
        for (int k = 0; k < CHUNK; ++k) {
            hipLaunchKernelGGL((k_pcg_spmv<T>), dim3(nb), dim3(PCG_BLOCK), 0, g.stream, n, rowptr, col, val, z, pa, pb, q,
                               scal, g.d_partials, g.d_sync);
            hipLaunchKernelGGL((k_pcg_update<T>), dim3(nb), dim3(PCG_BLOCK), 0, g.stream, n, diag, pb, q, x, r, z,
                               scal, g.d_partials, g.d_sync);
            T* t = pa; pa = pb; pb = t;
        }
        HIPCHK(hipGetLastError());
        done_iters += CHUNK;
        HIPCHK(hipMemcpyAsync(g.h_scal_pinned, scal, S_NSCAL * sizeof(double), hipMemcpyDeviceToHost, g.stream));
        HIPCHK(hipStreamSynchronize(g.stream));
        if (g.h_scal_pinned[S_DONE] != 0.0) break;
        if (done_iters > maxit + CHUNK) break;
    }
    // the buffer holding the last p is irrelevant to callers
    *iters = (int)g.h_scal_pinned[S_ITER];
    *ierr = (int)g.h_scal_pinned[S_IERR];
    return PIC_OK;
}

static int resid64(int n, const int* rowptr, const int* col, const double* val, const double* b, const double* x,
                   double* r, double* rr, double* bb) {
    int nb = (n + PCG_BLOCK - 1) / PCG_BLOCK;
    double* out = g.d_scal + S_NSCAL;
    hipLaunchKernelGGL(k_resid64, dim3(nb), dim3(PCG_BLOCK), 0, g.stream, n, rowptr, col, val, b, x, r, out,
                       g.d_partials, g.d_sync);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(g.h_scal_pinned + S_NSCAL, out, 2 * sizeof(double), hipMemcpyDeviceToHost, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    *rr = g.h_scal_pinned[S_NSCAL];
    *bb = g.h_scal_pinned[S_NSCAL + 1];
    return PIC_OK;
}

// Solve on the registered system: b = g.d_b, warm start from g.d_phi at unknowns.
int run_pcg_fp64(bool) {
    int n = g.nunknowns;
    int nb256 = (n + 255) / 256;
    hipLaunchKernelGGL(k_phi_to_x, dim3(nb256), dim3(256), 0, g.stream, n, g.d_unknown_node, g.d_phi, g.d_x);
    int iters = 0, ierr = 0;
    int rc = pcg_run<double>(n, g.d_A_rowptr, g.d_A_col, g.d_A_val, g.d_A_diag, g.d_b, g.d_x, g.d_r, g.d_z, g.d_p,
                             g.d_p2, g.d_q, g.pcg_tol, g.pcg_maxit, &iters, &ierr);
    if (rc) return rc;
    double rr, bb;
    rc = resid64(n, g.d_A_rowptr, g.d_A_col, g.d_A_val, g.d_b, g.d_x, nullptr, &rr, &bb);
    if (rc) return rc;
    hipLaunchKernelGGL(k_x_to_phi, dim3(nb256), dim3(256), 0, g.stream, n, g.d_unknown_node, g.d_x, g.d_phi);
    HIPCHK(hipGetLastError());
    g.last_iter = iters; g.last_outer = 1;
    g.last_relres = bb > 0 ? std::sqrt(rr / bb) : 0.0;
    if (ierr == 1) return pic_fail(PIC_ERR_PCG_MAXITER, "PCG reached PCG_MaxIT");
    if (ierr == 2 || ierr == 3) return pic_fail(PIC_ERR_PCG_BREAKDOWN, "PCG breakdown");
    return PIC_OK;
}

// FP32 Jacobi-PCG on the correction inside FP64 iterative refinement; the outer residual
// is recomputed in FP64 and decides convergence at pcg_tol.
int run_pcg_mixed() {
    int n = g.nunknowns;
    int nb256 = (n + 255) / 256;
    hipLaunchKernelGGL(k_phi_to_x, dim3(nb256), dim3(256), 0, g.stream, n, g.d_unknown_node, g.d_phi, g.d_x);
    int total = 0, outer = 0;
    double rr = 0, bb = 0;
    const double inner_tol = 1e-4;
    for (;;) {
        int rc = resid64(n, g.d_A_rowptr, g.d_A_col, g.d_A_val, g.d_b, g.d_x, g.d_r, &rr, &bb);
        if (rc) return rc;
        double rel = bb > 0 ? std::sqrt(rr / bb) : 0.0;
        if (rel <= g.pcg_tol || bb == 0.0) break;
        if (outer >= 30 || total >= g.pcg_maxit) {
            g.last_iter = total; g.last_outer = outer; g.last_relres = rel;
            return pic_fail(PIC_ERR_PCG_MAXITER, "mixed-precision PCG did not reach the tolerance");
        }
        hipLaunchKernelGGL(k_d2f, dim3(nb256), dim3(256), 0, g.stream, n, g.d_r, g.d_bf);
        hipLaunchKernelGGL(k_zero_f, dim3(nb256), dim3(256), 0, g.stream, n, g.d_xf);
        int it = 0, ierr = 0;
        rc = pcg_run<float>(n, g.d_A_rowptr, g.d_A_col, g.d_A_valf, g.d_A_diagf, g.d_bf, g.d_xf, g.d_rf, g.d_zf,
                            g.d_pf, g.d_p2f, g.d_qf, inner_tol, g.pcg_maxit - total, &it, &ierr);
        if (rc) return rc;
        hipLaunchKernelGGL(k_axpy_f2d, dim3(nb256), dim3(256), 0, g.stream, n, g.d_xf, g.d_x);
        total += it; outer++;
        if (ierr == 2 || ierr == 3) return pic_fail(PIC_ERR_PCG_BREAKDOWN, "inner PCG breakdown");
    }
    hipLaunchKernelGGL(k_x_to_phi, dim3(nb256), dim3(256), 0, g.stream, n, g.d_unknown_node, g.d_x, g.d_phi);
    HIPCHK(hipGetLastError());
    g.last_iter = total; g.last_outer = outer;
    g.last_relres = bb > 0 ? std::sqrt(rr / bb) : 0.0;
    return PIC_OK;
}

// Stand-alone FP64 PCG on caller-provided device CSR (MY_JPCG_Solver drop-in).
int pcg_generic_fp64(int n, const int* rowptr, const int* col, const double* val, const double* diag,
                     const double* b, double* x, double tol, int maxit, int* iters, int* ierr, double* r_out) {
    double *r, *z, *p, *p2, *q;
    HIPCHK(hipMalloc(&r, sizeof(double) * n)); HIPCHK(hipMalloc(&z, sizeof(double) * n));
    HIPCHK(hipMalloc(&p, sizeof(double) * n)); HIPCHK(hipMalloc(&p2, sizeof(double) * n));
    HIPCHK(hipMalloc(&q, sizeof(double) * n));
    int rc = pcg_run<double>(n, rowptr, col, val, diag, b, x, r, z, p, p2, q, tol, maxit, iters, ierr);
    if (rc == PIC_OK && r_out) HIPCHK(hipMemcpyAsync(r_out, r, sizeof(double) * n, hipMemcpyDeviceToDevice, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    hipFree(r); hipFree(z); hipFree(p); hipFree(p2); hipFree(q);
    return rc;
}
