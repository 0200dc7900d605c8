// Field kernels: RHS assembly (rhs = B*rho - c), PPR gradient recovery, and the
// Jacobi-preconditioned conjugate gradient of MY_PCG (FP64), with an FP32 inner solve
// inside FP64 iterative refinement.
//
// PCG layout: the nine stiffness diagonals in node space (or CSR for arbitrary matrices),
// one row per thread. Each iteration is two kernels:
//   k_pcg_spmv : p <- z + beta p (recomputed on the fly for the 9-point neighbourhood),
//                q <- A p, partial p.q
//   k_pcg_update: x += alpha p, r -= alpha q, z = r / diag, partial r.z and r.r
// Kernels of an iteration after convergence return at once, so the host checks the flag
// only every few iterations.
#include "pic_device.h"
#include <chrono>

// ---------------------------------------------------------------- RHS
// rhs_u = (B rho)_u - c_u for every unknown u, stored at its node (node-space PCG) and
// in unknown order (the IFE_Solve_2D rhs vector).
__global__ void k_rhs(int n, const int* rowptr, const int* col, const double* val, const double* rho,
                      const double* c, const int* unknown_node, double* b, double* bn) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) s = s + val[k] * rho[col[k]];
    double v = s - c[i];
    b[i] = v;
    bn[unknown_node[i]] = v;
}

int launch_rhs() {
    int n = g.nunknowns;
    hipLaunchKernelGGL(k_rhs, dim3((n + 255) / 256), dim3(256), 0, g.stream, n, g.d_B_rowptr, g.d_B_col,
                       g.d_B_val, g.d_rho, g.d_c, g.d_unknown_node, g.d_b, g.d_bn);
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
// No workgroup of a kernel waits for another: the block partial sums of one kernel are
// reduced in the prologue of the next kernel by every block (same operands, same fixed
// order, so every block and every run gets the same scalars). This avoids the device-scope
// fences and same-address atomics that a last-block reduction costs on a multi-XCD GPU.
// One row per thread, PCG_BLOCK-thread blocks; partial regions are g.pcg_pstride long.
#define PCG_BLOCK 256

// scalar slots in d_scal
enum { S_RHO0 = 0, S_RHO1, S_ALPHA, S_BETA, S_NORMRHS, S_NORMR, S_DONE, S_ITER, S_IERR, S_TOL2, S_MAXIT, S_PQ,
       S_GAM0, S_GAM1, S_ALP0, S_ALP1, S_NSCAL };
// partial-sum regions
enum { P_RZ = 0, P_RR = 1, P_BB = 2, P_PQ = 3, P_NREG = 4 };

// Operators: CSR (any matrix, the MY_JPCG_Solver drop-in) and the nine diagonals of the
// Q1 stiffness on the structured mesh in node numbering (offsets -nny-1 ... nny+1). Both
// sum a row's terms in ascending column order, so they round identically; the DIA
// operator adds the zero terms of Dirichlet neighbours (their vector entries are zero).
template <typename T> struct CsrOp {
    const int* rowptr; const int* col; const T* val;
    template <class F> __device__ inline T row(int i, F pj) const {
        T s = 0;
        for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) s = s + val[k] * pj(col[k]);
        return s;
    }
};
// The reference's TYPE SPARSE array read as is (IFE_MAIN_PARAM.f90:8-11, row pointers in SROW
// of entries 1..N+1, Sparse_Structure.f90:163-175; 1-based): the same terms in the same
// ascending order as CsrOp, so both round identically.
struct SparseAosOp {
    const pic_sparse* m;
    template <class F> __device__ inline double row(int i, F pj) const {
        double s = 0;
        const int k0 = m[i].SROW - 1, k1 = m[i + 1].SROW - 1;
        for (int k = k0; k < k1; ++k) s = s + m[k].K * pj(m[k].JCOL - 1);
        return s;
    }
};
template <typename T> struct DiaOp {
    const T* c[9];
    int off[9];
    template <class F> __device__ inline T row(int i, F pj) const {
        T s = 0;
#pragma unroll
        for (int k = 0; k < 9; ++k) s = s + c[k][i] * pj(i + off[k]);
        return s;
    }
};

template <int NV>
__device__ inline void block_partials(double v[NV], double* partials, int pstride, const int regs[NV]) {
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
            partials[(size_t)regs[k] * pstride + blockIdx.x] = s;
        }
    }
}

// Sum of the `nb` partials of region `reg` written by the previous kernel, in a fixed
// order; every thread of every block gets the same value.
__device__ inline double reduce_region(const double* partials, int pstride, int reg, int nb) {
    __shared__ double red[PCG_BLOCK];
    const double* pr = partials + (size_t)reg * pstride;
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += PCG_BLOCK) s += pr[b];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int off = PCG_BLOCK / 2; off > 0; off >>= 1) {
        if (threadIdx.x < (unsigned)off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    double t = red[0];
    __syncthreads();
    return t;
}

// r = b - A x, z = r / diag (0 where diag == 0: Dirichlet rows of the DIA operator);
// partials of r.z and b.b
template <typename T, class Op>
__global__ __launch_bounds__(PCG_BLOCK) void k_pcg_init(int n, Op A, const T* diag, const T* b, const T* x, T* r, T* z,
                                                        double* scal, double* partials, int pstride) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double v[2] = {0.0, 0.0};
    if (i < n) {
        T s = A.row(i, [&](int j) { return x[j]; });
        T ri = b[i] - s;
        T d = diag[i];
        T zi = d != (T)0 ? ri / d : (T)0;
        r[i] = ri; z[i] = zi;
        v[0] = (double)ri * (double)zi;
        v[1] = (double)b[i] * (double)b[i];
    }
    const int regs[2] = {P_RZ, P_BB};
    block_partials<2>(v, partials, pstride, regs);
    if (blockIdx.x == 0 && threadIdx.x == 0) { scal[S_DONE] = 0.0; scal[S_ITER] = 0.0; scal[S_IERR] = 0.0; }
}

// Iteration k (0-based) of MY_PCG.f90:73-127, first half: rho_k = r.z (previous kernel's
// partials), convergence test of iteration k-1 (||r||^2/||b||^2 <= tol^2 with b the RHS),
// p = z + (rho_k/rho_{k-1}) p (p = z when k == 0), q = A p, partials of p.q.
template <typename T, class Op>
__global__ __launch_bounds__(PCG_BLOCK) void k_pcg_spmv(int n, int k, int nb, Op A, const T* z, const T* p_old,
                                                        T* p_new, T* q, double* scal, double* partials, int pstride) {
    if (scal[S_DONE] != 0.0) return;
    double rho = reduce_region(partials, pstride, P_RZ, nb);
    double beta = 0.0;
    if (k == 0) {
        double bb = reduce_region(partials, pstride, P_BB, nb);
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            scal[S_NORMRHS] = bb;
            if (bb == 0.0) scal[S_DONE] = 1.0;      // zero RHS: x = 0 is exact
        }
        if (bb == 0.0) return;
    } else {
        double normr = reduce_region(partials, pstride, P_RR, nb);
        double rho_old = scal[(k - 1) & 1 ? S_RHO1 : S_RHO0];
        bool conv = normr / scal[S_NORMRHS] <= scal[S_TOL2];
        bool maxed = (double)k >= scal[S_MAXIT];
        if (conv || maxed || rho_old == 0.0) {
            if (blockIdx.x == 0 && threadIdx.x == 0) {
                scal[S_DONE] = 1.0; scal[S_ITER] = (double)k; scal[S_NORMR] = normr;
                scal[S_IERR] = maxed ? 1.0 : (conv ? 0.0 : 2.0);   // MY_PCG.f90:88-90, 129-131
            }
            return;
        }
        beta = rho / rho_old;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) { scal[k & 1 ? S_RHO1 : S_RHO0] = rho; scal[S_ITER] = (double)k; }
    const T tb = (T)beta;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double v[1] = {0.0};
    if (i < n) {
        T s;
        if (k == 0) s = A.row(i, [&](int j) { return z[j]; });
        else s = A.row(i, [&](int j) { return z[j] + tb * p_old[j]; });
        T pi = (k == 0) ? z[i] : z[i] + tb * p_old[i];
        p_new[i] = pi; q[i] = s;
        v[0] = (double)pi * (double)s;
    }
    const int regs[1] = {P_PQ};
    block_partials<1>(v, partials, pstride, regs);
}

// Second half: alpha = rho_k / p.q, x += alpha p, r -= alpha q, z = r / diag; partials of
// r.z and r.r for the next iteration.
template <typename T>
__global__ __launch_bounds__(PCG_BLOCK) void k_pcg_update(int n, int k, int nb, const T* diag, const T* p, const T* q,
                                                          T* x, T* r, T* z, double* scal, double* partials, int pstride) {
    if (scal[S_DONE] != 0.0) return;
    double pq = reduce_region(partials, pstride, P_PQ, nb);
    if (pq == 0.0) {
        if (blockIdx.x == 0 && threadIdx.x == 0) { scal[S_DONE] = 1.0; scal[S_IERR] = 3.0; scal[S_ITER] = (double)k; }
        return;
    }
    double alpha = scal[k & 1 ? S_RHO1 : S_RHO0] / pq;
    const T ta = (T)alpha;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double v[2] = {0.0, 0.0};
    if (i < n) {
        T xi = x[i] + ta * p[i];
        T ri = r[i] - ta * q[i];
        T d = diag[i];
        T zi = d != (T)0 ? ri / d : (T)0;
        x[i] = xi; r[i] = ri; z[i] = zi;
        v[0] = (double)ri * (double)zi;
        v[1] = (double)ri * (double)ri;
    }
    const int regs[2] = {P_RZ, P_RR};
    block_partials<2>(v, partials, pstride, regs);
}

// true residual ||b - A x||^2 and ||b||^2 (FP64): partials, then one block sums them
template <class Op>
__global__ __launch_bounds__(PCG_BLOCK) void k_resid64(int n, Op A, const double* b, const double* x, double* r,
                                                       double* partials, int pstride) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double v[2] = {0.0, 0.0};
    if (i < n) {
        double s = A.row(i, [&](int j) { return x[j]; });
        double ri = b[i] - s;
        if (r) r[i] = ri;
        v[0] = ri * ri; v[1] = b[i] * b[i];
    }
    const int regs[2] = {P_RR, P_BB};
    block_partials<2>(v, partials, pstride, regs);
}
__global__ __launch_bounds__(PCG_BLOCK) void k_sum2(int nb, const double* partials, int pstride, double* out2) {
    double a0 = reduce_region(partials, pstride, P_RR, nb);
    double a1 = reduce_region(partials, pstride, P_BB, nb);
    if (threadIdx.x == 0) { out2[0] = a0; out2[1] = a1; }
}

__global__ void k_d2f(int n, const double* a, float* b) { int i = blockIdx.x * blockDim.x + threadIdx.x; if (i < n) b[i] = (float)a[i]; }
__global__ void k_axpy_f2d(int n, const float* a, double* b) { int i = blockIdx.x * blockDim.x + threadIdx.x; if (i < n) b[i] = b[i] + (double)a[i]; }
__global__ void k_zero_f(int n, float* a) { int i = blockIdx.x * blockDim.x + threadIdx.x; if (i < n) a[i] = 0.0f; }

// phi at unknown nodes <-> node-space solution vector (IFE_Solve_2D.f90:80-84, 243-247, 514-516)
__global__ void k_phi_to_xn(int n, const int* nt2, const double* phi, double* x) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = nt2[i] > 0 ? phi[i] : 0.0;
}
__global__ void k_xn_to_phi(int n, const int* nt2, const double* x, double* phi) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && nt2[i] > 0) phi[i] = x[i];
}

// ---------------------------------------------------------------- one-kernel CG (FP32 inner solve)
// Jacobi-preconditioned CG on the symmetrically scaled system D^-1/2 A D^-1/2 y = D^-1/2 b
// (the same iterates as Jacobi-PCG on A x = b, with a unit diagonal, so the preconditioned
// residual is the residual itself), in the Chronopoulos-Gear arrangement: one reduction
// per iteration (gamma = r.r, delta = w.r with w = A r). Every iteration is ONE kernel:
//   beta = gamma/gamma_old, alpha = gamma/(delta - beta*gamma/alpha_old)   (prologue)
//   s = w + beta s, p = r + beta p, x += alpha p, r' = r - alpha s, w' = A r'
// where w' needs r' at the eight neighbours; it recomputes them from the previous
// iteration's r, w, s (double-buffered), with the same expression the owning row uses, so
// the values are identical. Partials of gamma' and delta' go to the other partial buffer.
#define CG_BLOCK 256
enum { CG_PG = 0, CG_PD = 1 };   // partial regions (x2 parities)

// Sum of two partial regions over nb blocks, fixed order, identical in every block.
__device__ inline void cg_reduce2(const double* pa, const double* pb, int nb, double& a, double& b) {
    __shared__ double sh[2][CG_BLOCK / 64];
    double sa = 0.0, sb = 0.0;
    for (int k = threadIdx.x; k < nb; k += CG_BLOCK) { sa += pa[k]; sb += pb[k]; }
    for (int off = 32; off > 0; off >>= 1) { sa += __shfl_xor(sa, off, 64); sb += __shfl_xor(sb, off, 64); }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) { sh[0][wid] = sa; sh[1][wid] = sb; }
    __syncthreads();
    a = 0.0; b = 0.0;
#pragma unroll
    for (int w = 0; w < CG_BLOCK / 64; ++w) { a += sh[0][w]; b += sh[1][w]; }
    __syncthreads();
}

__device__ inline void cg_partials2(double a, double b, double* pa, double* pb) {
    __shared__ double sh[2][CG_BLOCK / 64];
    for (int off = 32; off > 0; off >>= 1) { a += __shfl_xor(a, off, 64); b += __shfl_xor(b, off, 64); }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) { sh[0][wid] = a; sh[1][wid] = b; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double ta = 0.0, tb = 0.0;
#pragma unroll
        for (int w = 0; w < CG_BLOCK / 64; ++w) { ta += sh[0][w]; tb += sh[1][w]; }
        pa[blockIdx.x] = ta; pb[blockIdx.x] = tb;
    }
}

// r0 = D^-1/2 b64 (x0 = 0), w0 = A' r0; partials of r0.r0 and w0.r0 (parity 0).
template <int RPT>
__global__ __launch_bounds__(CG_BLOCK) void k_cg_init(int n, Dia8 A, const double* b64, const double* sinv, float* r,
                                                      float* w, float* x, double* scal, double* part, int pstride) {
    double g = 0.0, d = 0.0;
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
        int i = blockIdx.x * (CG_BLOCK * RPT) + q * CG_BLOCK + threadIdx.x;
        if (i < n) {
            float ri = (float)(sinv[i] * b64[i]);
            float wi = ri;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                int j = i + A.off[k];
                float rj = (float)(sinv[j] * b64[j]);
                wi = wi + A.c[k][i] * rj;
            }
            r[i] = ri; w[i] = wi; x[i] = 0.0f;
            g += (double)ri * (double)ri;
            d += (double)wi * (double)ri;
        }
    }
    cg_partials2(g, d, part + (size_t)(2 * 0 + CG_PG) * pstride, part + (size_t)(2 * 0 + CG_PD) * pstride);
    if (blockIdx.x == 0 && threadIdx.x == 0) { scal[S_DONE] = 0.0; scal[S_ITER] = 0.0; scal[S_IERR] = 0.0; }
}

template <int RPT>
__global__ __launch_bounds__(CG_BLOCK) void k_cg_iter(int n, int k, int nb, Dia8 A, const float* r_o, const float* w_o,
                                                      const float* s_o, float* r_n, float* w_n, float* s_n, float* p,
                                                      float* x, double* scal, double* part, int pstride) {
    __shared__ double sh[2][CG_BLOCK / 64];
    __shared__ double sh_done;
    const int par = k & 1;
    // Every load that does not depend on the scalars is issued first, so the latency of the
    // partial sums (written by all blocks of the previous kernel) overlaps the vector reads.
    const double* pa = part + (size_t)(2 * par + CG_PG) * pstride;
    const double* pb = part + (size_t)(2 * par + CG_PD) * pstride;
    double sa = 0.0, sb = 0.0;
    for (int q = threadIdx.x; q < nb; q += CG_BLOCK) { sa += pa[q]; sb += pb[q]; }
    if (threadIdx.x == 0) sh_done = scal[S_DONE];
    const double gam_old = scal[par ? S_GAM0 : S_GAM1], alp_old = scal[par ? S_ALP0 : S_ALP1];
    const double normrhs = scal[S_NORMRHS], tol2 = scal[S_TOL2], maxit = scal[S_MAXIT];
    float ro[RPT], wo[RPT], so[RPT], po[RPT], xo[RPT], cc[RPT][8], rj[RPT][8], wj[RPT][8], sj[RPT][8];
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
        int i = blockIdx.x * (CG_BLOCK * RPT) + q * CG_BLOCK + threadIdx.x;
        int ii = i < n ? i : n - 1;
        ro[q] = r_o[ii]; wo[q] = w_o[ii]; so[q] = s_o[ii]; po[q] = p[ii]; xo[q] = x[ii];
#pragma unroll
        for (int kk = 0; kk < 8; ++kk) {
            int j = ii + A.off[kk];
            cc[q][kk] = A.c[kk][ii]; rj[q][kk] = r_o[j]; wj[q][kk] = w_o[j]; sj[q][kk] = s_o[j];
        }
    }
    // block-uniform reduction of gamma = r.r and delta = w.r of the previous kernel
    for (int off = 32; off > 0; off >>= 1) { sa += __shfl_xor(sa, off, 64); sb += __shfl_xor(sb, off, 64); }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) { sh[0][wid] = sa; sh[1][wid] = sb; }
    __syncthreads();
    double gam = 0.0, del = 0.0;
#pragma unroll
    for (int w = 0; w < CG_BLOCK / 64; ++w) { gam += sh[0][w]; del += sh[1][w]; }
    if (sh_done != 0.0) return;
    __syncthreads();                               // sh is reused by the partials below
    double alpha, beta = 0.0;
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    if (k == 0) {
        if (lead) scal[S_NORMRHS] = gam;
        if (gam == 0.0) { if (lead) { scal[S_DONE] = 1.0; scal[S_ITER] = 0.0; } return; }
        alpha = gam / del;
    } else {
        bool conv = gam / normrhs <= tol2;
        bool maxed = (double)k >= maxit;
        if (conv || maxed) {
            if (lead) { scal[S_DONE] = 1.0; scal[S_ITER] = (double)k; scal[S_NORMR] = gam; scal[S_IERR] = maxed && !conv ? 1.0 : 0.0; }
            return;
        }
        beta = gam / gam_old;
        double den = del - beta * gam / alp_old;
        alpha = gam / den;
    }
    if (!(alpha > 0.0) || !(alpha < 1e300)) {      // breakdown (A not SPD in FP32, or NaN)
        if (lead) { scal[S_DONE] = 1.0; scal[S_IERR] = 2.0; scal[S_ITER] = (double)k; }
        return;
    }
    if (lead) { scal[par ? S_GAM1 : S_GAM0] = gam; scal[par ? S_ALP1 : S_ALP0] = alpha; scal[S_ITER] = (double)k; }
    const float fa = (float)alpha, fb = (float)beta;
    double g = 0.0, d = 0.0;
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
        int i = blockIdx.x * (CG_BLOCK * RPT) + q * CG_BLOCK + threadIdx.x;
        float si = (k == 0) ? wo[q] : wo[q] + fb * so[q];
        float pi = (k == 0) ? ro[q] : ro[q] + fb * po[q];
        float ri = ro[q] - fa * si;
        float wi = ri;
#pragma unroll
        for (int kk = 0; kk < 8; ++kk) {
            // the neighbour's own update, same expression and order as its row computes it
            float sjj = (k == 0) ? wj[q][kk] : wj[q][kk] + fb * sj[q][kk];
            float rjj = rj[q][kk] - fa * sjj;
            wi = wi + cc[q][kk] * rjj;
        }
        if (i < n) {
            x[i] = xo[q] + fa * pi;
            p[i] = pi; s_n[i] = si; r_n[i] = ri; w_n[i] = wi;
            g += (double)ri * (double)ri;
            d += (double)wi * (double)ri;
        }
    }
    const int np = par ^ 1;
    cg_partials2(g, d, part + (size_t)(2 * np + CG_PG) * pstride, part + (size_t)(2 * np + CG_PD) * pstride);
}

// x64 += D^-1/2 y
__global__ void k_cg_apply(int n, const double* sinv, const float* y, double* x) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = x[i] + sinv[i] * (double)y[i];
}

template <int RPT>
static int cg1_run_t(int n, const Dia8& A, const double* b64, double tol, int maxit, int* iters, int* ierr) {
    const int rows = CG_BLOCK * RPT;
    int nb = (n + rows - 1) / rows;
    if (nb > g.pcg_pstride) return pic_fail(PIC_ERR_STATE, "CG partial buffer too small");
    double* scal = g.d_scal;
    double* part = g.d_partials;
    int ps = g.pcg_pstride;
    double h2[2] = {tol * tol, (double)maxit};
    HIPCHK(hipMemcpyAsync(scal + S_TOL2, h2, 2 * sizeof(double), hipMemcpyHostToDevice, g.stream));
    float *r[2] = {g.d_cg_r[0], g.d_cg_r[1]}, *w[2] = {g.d_cg_w[0], g.d_cg_w[1]}, *s[2] = {g.d_cg_s[0], g.d_cg_s[1]};
    hipLaunchKernelGGL((k_cg_init<RPT>), dim3(nb), dim3(CG_BLOCK), 0, g.stream, n, A, b64, g.d_sinv, r[0], w[0],
                       g.d_cg_x, scal, part, ps);
    HIPCHK(hipGetLastError());
    const bool harvest = learn_harvesting();
    if (harvest) { int rc = learn_snapshot(g.d_cg_x, 0); if (rc) return rc; }
    const int CHUNK = 32;
    int k = 0, chunk = 0;
    for (;;) {
        HIPCHK(hipEventRecord(g.ev[10], g.stream));
        for (int c = 0; c < CHUNK; ++c, ++k) {
            int o = k & 1, nw = o ^ 1;
            hipLaunchKernelGGL((k_cg_iter<RPT>), dim3(nb), dim3(CG_BLOCK), 0, g.stream, n, k, nb, A, r[o], w[o], s[o],
                               r[nw], w[nw], s[nw], g.d_cg_p, g.d_cg_x, scal, part, ps);
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(g.ev[11], g.stream));
        HIPCHK(hipMemcpyAsync(g.h_scal_pinned, scal, S_NSCAL * sizeof(double), hipMemcpyDeviceToHost, g.stream));
        overlap_point();
        HIPCHK(hipStreamSynchronize(g.stream));
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, g.ev[10], g.ev[11]));
        g.kstat[0] += ms; g.kstat[1] += CHUNK;
        if (g.h_scal_pinned[S_DONE] != 0.0) break;
        if (k > maxit + CHUNK) { g.h_scal_pinned[S_IERR] = 1.0; g.h_scal_pinned[S_ITER] = k; break; }
        ++chunk;
        if (harvest) { int rc = learn_snapshot(g.d_cg_x, chunk); if (rc) return rc; }
    }
    if (harvest) { int rc = learn_harvest_finish(g.d_cg_x); if (rc) return rc; }
    *iters = (int)g.h_scal_pinned[S_ITER];
    *ierr = (int)g.h_scal_pinned[S_IERR];
    return PIC_OK;
}

static int cg1_run(int n, const double* b64, double tol, int maxit, int* iters, int* ierr) {
    Dia8 A;
    for (int k = 0; k < 8; ++k) { A.c[k] = g.d_cg_off[k]; A.off[k] = g.cg_off[k]; }
    switch (g.cg_rpt) {
    case 1: return cg1_run_t<1>(n, A, b64, tol, maxit, iters, ierr);
    case 2: return cg1_run_t<2>(n, A, b64, tol, maxit, iters, ierr);
    default: return cg1_run_t<4>(n, A, b64, tol, maxit, iters, ierr);
    }
}

static inline int pcg_blocks(int n) { return (n + PCG_BLOCK - 1) / PCG_BLOCK; }

void overlap_point() {
    if (!g.overlap_hook) return;
    int (*h)() = g.overlap_hook;
    g.overlap_hook = nullptr;
    g.overlap_rc = h();
}

// One Jacobi-PCG solve on device vectors (storage T, operator Op). The host launches the
// iterations in chunks and reads the done flag between chunks.
template <typename T, class Op>
static int pcg_run(int n, const Op& A, const T* diag, const T* b, T* x, T* r, T* z, T* p, T* p2, T* q, double tol,
                   int maxit, double* partials, int pstride, int* iters, int* ierr, double* scal = nullptr) {
    int nb = pcg_blocks(n);
    if (nb > pstride) return pic_fail(PIC_ERR_STATE, "PCG partial buffer too small");
    if (!scal) scal = g.d_scal;
    double h2[2] = {tol * tol, (double)maxit};
    HIPCHK(hipMemcpyAsync(scal + S_TOL2, h2, 2 * sizeof(double), hipMemcpyHostToDevice, g.stream));
    hipLaunchKernelGGL((k_pcg_init<T, Op>), dim3(nb), dim3(PCG_BLOCK), 0, g.stream, n, A, diag, b, x, r, z, scal,
                       partials, pstride);
    HIPCHK(hipGetLastError());
    T* pa = p; T* pb = p2;
    const int CHUNK = 32;
    int k = 0;
    for (;;) {
        for (int c = 0; c < CHUNK; ++c, ++k) {
            hipLaunchKernelGGL((k_pcg_spmv<T, Op>), dim3(nb), dim3(PCG_BLOCK), 0, g.stream, n, k, nb, A, z, pa, pb, q,
                               scal, partials, pstride);
            hipLaunchKernelGGL((k_pcg_update<T>), dim3(nb), dim3(PCG_BLOCK), 0, g.stream, n, k, nb, diag, pb, q, x, r, z,
                               scal, partials, pstride);
            T* t = pa; pa = pb; pb = t;
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(g.h_scal_pinned, scal, S_NSCAL * sizeof(double), hipMemcpyDeviceToHost, g.stream));
        overlap_point();
        HIPCHK(hipStreamSynchronize(g.stream));
        if (g.h_scal_pinned[S_DONE] != 0.0) break;
        if (k > maxit + CHUNK) { g.h_scal_pinned[S_IERR] = 1.0; g.h_scal_pinned[S_ITER] = k; break; }
    }
    *iters = (int)g.h_scal_pinned[S_ITER];
    *ierr = (int)g.h_scal_pinned[S_IERR];
    return PIC_OK;
}

template <class Op>
static int resid64(int n, const Op& A, const double* b, const double* x, double* r, double* rr, double* bb) {
    int nb = pcg_blocks(n);
    double* out = g.d_scal + S_NSCAL;
    hipLaunchKernelGGL((k_resid64<Op>), dim3(nb), dim3(PCG_BLOCK), 0, g.stream, n, A, b, x, r, g.d_partials, g.pcg_pstride);
    hipLaunchKernelGGL(k_sum2, dim3(1), dim3(PCG_BLOCK), 0, g.stream, nb, g.d_partials, g.pcg_pstride, out);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(g.h_scal_pinned + S_NSCAL, out, 2 * sizeof(double), hipMemcpyDeviceToHost, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    *rr = g.h_scal_pinned[S_NSCAL];
    *bb = g.h_scal_pinned[S_NSCAL + 1];
    return PIC_OK;
}

template <typename T>
static DiaOp<T> dia_op(int prec64) {
    DiaOp<T> A;
    for (int k = 0; k < 9; ++k) { A.c[k] = (const T*)(prec64 ? (void*)g.d_dia64[k] : (void*)g.d_dia32[k]); A.off[k] = g.dia_off[k]; }
    return A;
}

// Solve on the registered system in node space: b = g.d_bn, warm start from g.d_phi.
int run_pcg_fp64(bool) {
    int n = g.mesh.nnodes;
    int nb256 = (n + 255) / 256;
    hipLaunchKernelGGL(k_phi_to_xn, dim3(nb256), dim3(256), 0, g.stream, n, g.d_node_type2, g.d_phi, g.d_xn);
    DiaOp<double> A = dia_op<double>(1);
    int iters = 0, ierr = 0;
    int rc = pcg_run<double>(n, A, g.d_dia64[4], g.d_bn, g.d_xn, g.d_rn, g.d_zn, g.d_pn, g.d_p2n, g.d_qn, g.pcg_tol,
                             g.pcg_maxit, g.d_partials, g.pcg_pstride, &iters, &ierr);
    if (rc) return rc;
    double rr, bb;
    rc = resid64(n, A, g.d_bn, g.d_xn, nullptr, &rr, &bb);
    if (rc) return rc;
    hipLaunchKernelGGL(k_xn_to_phi, dim3(nb256), dim3(256), 0, g.stream, n, g.d_node_type2, g.d_xn, g.d_phi);
    HIPCHK(hipGetLastError());
    g.last_iter = iters; g.last_outer = 1;
    g.last_relres = bb > 0 ? std::sqrt(rr / bb) : 0.0;
    if (ierr == 1) return pic_fail(PIC_ERR_PCG_MAXITER, "PCG reached PCG_MaxIT");
    if (ierr == 2 || ierr == 3) return pic_fail(PIC_ERR_PCG_BREAKDOWN, "PCG breakdown");
    return PIC_OK;
}

// FP32 CG on the Jacobi-scaled correction equation inside FP64 iterative refinement; the
// outer residual is recomputed in FP64 with the unscaled operator and decides convergence
// at pcg_tol.
// FP64 residual r = b - A x with rr = ||r||^2, bb = ||b||^2 left at d_scal[S_NSCAL], [S_NSCAL+1]
template <typename Op>
static void resid64_launch(int n, const Op& A, const double* b, const double* x, double* r) {
    int nb = pcg_blocks(n);
    hipLaunchKernelGGL((k_resid64<Op>), dim3(nb), dim3(PCG_BLOCK), 0, g.stream, n, A, b, x, r, g.d_partials, g.pcg_pstride);
    hipLaunchKernelGGL(k_sum2, dim3(1), dim3(PCG_BLOCK), 0, g.stream, nb, g.d_partials, g.pcg_pstride, g.d_scal + S_NSCAL);
}

// Iterative refinement with the learned FCG as the inner solver. A batch of passes (as many
// as the last learned solve needed, each with the FCG iterations its counterpart needed, the
// last one with a spare while recent solves needed it) is launched with no host
// synchronisation in between: every pass tests the FP64 residual before it on the device and
// does nothing once rr <= tol^2 bb (or after a failed pass). One synchronisation per batch
// reads the records; a pass that ran out of iterations gets two more next time, and further
// batches of one pass follow until the tolerance is met.
// *fallback: the network failed (NaN, breakdown) or stagnated; the caller finishes with CG.
static int run_learned_refinement(int n, const DiaOp<double>& A64, int* total, int* outer, double* rr, double* bb,
                                  bool* fallback) {
    const double tol2 = g.pcg_tol * g.pcg_tol;
    const double inner = learn_inner_tol();
    const int cap = std::max(1, std::min(g.learn_maxit, 64));
    *fallback = false;
    int batch = std::min(std::max(g.lp_passes, 1), FCG_MAX_PASS);
    bool first = true;
    int done_passes = 0;
    for (;;) {
        const auto th0 = std::chrono::steady_clock::now();
        HIPCHK(hipEventRecord(g.ev[12], g.stream));
        if (first && warm_active()) {
            // warm start: xn += alpha D^-1/2 N(D^-1/2 (b - A xn)), then the passes start from it
            resid64_launch(n, A64, g.d_bn, g.d_xn, g.d_rn);
            int rc = warm_solve_start(n, g.d_rn, g.d_xn); if (rc) return rc;
        }
        resid64_launch(n, A64, g.d_bn, g.d_xn, g.d_rn);
        int K[FCG_MAX_PASS];
        for (int p = 0; p < batch; ++p) {
            const int idx = std::min(done_passes + p, FCG_MAX_PASS - 1);
            K[p] = std::min(cap, std::max(1, g.lp_iters[idx]) + (p == batch - 1 ? g.lp_spare : 0));
            int rc = learn_fcg_pass(n, g.d_rn, inner, K[p], tol2, g.d_scal + S_NSCAL, first && p == 0 ? 1 : 0, p,
                                    g.d_xn);
            if (rc) return rc;
            resid64_launch(n, A64, g.d_bn, g.d_xn, g.d_rn);
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(g.ev[13], g.stream));
        const double host_us = std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - th0).count();
        HIPCHK(hipMemcpyAsync(g.h_scal_pinned, g.d_scal, 64 * sizeof(double), hipMemcpyDeviceToHost, g.stream));
        overlap_point();
        HIPCHK(hipStreamSynchronize(g.stream));
        const double* h = g.h_scal_pinned;
        *rr = h[S_NSCAL]; *bb = h[S_NSCAL + 1];
        int its = 0;
        bool ran_out = false, spare_used = false;
        for (int p = 0; p < batch; ++p) {
            const double* rec = h + FCG_REC_BASE + FCG_RECW * p;
            if (rec[4] != 0.0) break;                       // skipped: converged before (or failed earlier)
            const int it = (int)rec[0];
            const int idx = std::min(done_passes, FCG_MAX_PASS - 1);
            if (p == batch - 1 && g.lp_spare && it >= K[p]) spare_used = true;
            its += it; *total += it; (*outer)++; done_passes++;
            g.last_learned_iters += it;
            if (learn_trace_on())
                fprintf(stderr, "learned pass %d: %d of %d iterations, done %g, err %g, rr/rr0 %.3e\n", done_passes, it,
                        K[p], rec[1], rec[2], rec[3]);
            if (rec[2] == 2.0 || rec[2] == 3.0) { *fallback = true; break; }
            if (rec[1] == 0.0) {
                // out of iterations: stagnation if the residual norm fell less than 10x
                if (K[p] >= 4 && !(rec[3] < 1e-2)) { *fallback = true; break; }
                g.lp_iters[idx] = K[p] + 2;
                ran_out = true;
            } else {
                g.lp_iters[idx] = std::max(1, it);
            }
        }
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, g.ev[12], g.ev[13]));
        g.kstat[4] += ms; g.kstat[5] += its;
        if (learn_trace_on())
            fprintf(stderr, "learned batch: %d passes, host enqueue %.1f us, device %.1f us\n", batch, host_us, 1e3 * ms);
        learn_note_iterations(its);
        if (ran_out) {
            g.lp_spare = 1; g.lp_calm = 0;
        } else if (g.lp_spare) {
            if (spare_used) g.lp_calm = 0;
            else if (++g.lp_calm >= 8) { g.lp_spare = 0; g.lp_calm = 0; }
        }
        if (*fallback) return PIC_OK;
        if (*bb == 0.0 || *rr <= tol2 * *bb) { g.lp_passes = std::max(1, done_passes); return PIC_OK; }
        if (*outer >= 30 || *total >= g.pcg_maxit) {
            g.last_iter = *total; g.last_outer = *outer; g.last_relres = std::sqrt(*rr / *bb);
            return pic_fail(PIC_ERR_PCG_MAXITER, "mixed-precision PCG did not reach the tolerance");
        }
        first = false;
        batch = 1;
    }
}

int run_pcg_mixed() {
    int n = g.mesh.nnodes;
    int nb256 = (n + 255) / 256;
    hipLaunchKernelGGL(k_phi_to_xn, dim3(nb256), dim3(256), 0, g.stream, n, g.d_node_type2, g.d_phi, g.d_xn);
    DiaOp<double> A64 = dia_op<double>(1);
    int total = 0, outer = 0;
    double rr = 0, bb = 0;
    const double inner_tol = 1e-4;
    g.last_learned_iters = 0; g.last_fallback = 0;
    // solver_mode 2: the learned FCG is the inner solver; when it fails or stagnates the plain
    // CG finishes this solve. The FP64 outer residual judges convergence either way.
    if (g.solver_mode == 2 && learn_ready()) {
        bool fallback = false;
        int rc = run_learned_refinement(n, A64, &total, &outer, &rr, &bb, &fallback);
        if (rc) return rc;
        if (!fallback) {
            hipLaunchKernelGGL(k_xn_to_phi, dim3(nb256), dim3(256), 0, g.stream, n, g.d_node_type2, g.d_xn, g.d_phi);
            HIPCHK(hipGetLastError());
            g.last_iter = total; g.last_outer = outer;
            g.last_relres = bb > 0 ? std::sqrt(rr / bb) : 0.0;
            return PIC_OK;
        }
        g.last_fallback = 1;
        learn_count_fallback();
    }
    for (;;) {
        int rc = resid64(n, A64, g.d_bn, g.d_xn, g.d_rn, &rr, &bb);
        if (rc) return rc;
        double rel = bb > 0 ? std::sqrt(rr / bb) : 0.0;
        if (rel <= g.pcg_tol || bb == 0.0) break;
        if (outer >= 30 || total >= g.pcg_maxit) {
            g.last_iter = total; g.last_outer = outer; g.last_relres = rel;
            return pic_fail(PIC_ERR_PCG_MAXITER, "mixed-precision PCG did not reach the tolerance");
        }
        int it = 0, ierr = 0;
        learn_set_solve_pass(outer);                 // the harvest tags a solve's initial residual
        rc = cg1_run(n, g.d_rn, inner_tol, g.pcg_maxit - total, &it, &ierr);
        if (rc) return rc;
        hipLaunchKernelGGL(k_cg_apply, dim3(nb256), dim3(256), 0, g.stream, n, g.d_sinv, g.d_cg_x, g.d_xn);
        total += it; outer++;
        if (ierr == 2 || ierr == 3) return pic_fail(PIC_ERR_PCG_BREAKDOWN, "inner PCG breakdown");
    }
    hipLaunchKernelGGL(k_xn_to_phi, dim3(nb256), dim3(256), 0, g.stream, n, g.d_node_type2, g.d_xn, g.d_phi);
    HIPCHK(hipGetLastError());
    g.last_iter = total; g.last_outer = outer;
    g.last_relres = bb > 0 ? std::sqrt(rr / bb) : 0.0;
    return PIC_OK;
}

// Stand-alone FP64 PCG on caller-provided device CSR (MY_JPCG_Solver drop-in).
int pcg_generic_fp64(int n, const int* rowptr, const int* col, const double* val, const double* diag,
                     const double* b, double* x, double tol, int maxit, int* iters, int* ierr, double* r_out) {
    double *r, *z, *p, *p2, *q, *part;
    int pstride = pcg_blocks(n) + 1;
    HIPCHK(hipMalloc(&r, sizeof(double) * n)); HIPCHK(hipMalloc(&z, sizeof(double) * n));
    HIPCHK(hipMalloc(&p, sizeof(double) * n)); HIPCHK(hipMalloc(&p2, sizeof(double) * n));
    HIPCHK(hipMalloc(&q, sizeof(double) * n)); HIPCHK(hipMalloc(&part, sizeof(double) * P_NREG * pstride));
    CsrOp<double> A; A.rowptr = rowptr; A.col = col; A.val = val;
    int rc = pcg_run<double>(n, A, diag, b, x, r, z, p, p2, q, tol, maxit, part, pstride, iters, ierr);
    if (rc == PIC_OK && r_out) HIPCHK(hipMemcpyAsync(r_out, r, sizeof(double) * n, hipMemcpyDeviceToDevice, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    (void)hipFree(r); (void)hipFree(z); (void)hipFree(p); (void)hipFree(p2); (void)hipFree(q); (void)hipFree(part);
    return rc;
}

// MY_PCG on the reference's SPARSE array already on the device (the allocation-free
// MY_JPCG_Solver): every buffer is the caller's (carved from the library's fixed scratch).
int pcg_sparse_aos_fp64(int n, const pic_sparse* dmat, const double* diag, const double* b, double* x, double* r,
                        double* z, double* p, double* p2, double* q, double* partials, int pstride, double* scal,
                        double tol, int maxit, int* iters, int* ierr) {
    SparseAosOp A; A.m = dmat;
    return pcg_run<double>(n, A, diag, b, x, r, z, p, p2, q, tol, maxit, partials, pstride, iters, ierr, scal);
}

// ------------------------------------------------------------------ IC(0)-PCG (solver_mode 3)
// A comparison solver: FP64 PCG with a multicolour incomplete Cholesky preconditioner on the
// node-space stiffness, to the same relative residual as MY_PCG. Colours by node parity,
// c = (i & 1) + 2 (j & 1) for node i * nny + j: nodes of one colour never share a nine-point
// stencil, so in the colour-major order each triangular solve is four parallel sweeps. The
// factor keeps the stencil's sparsity (IC(0), fill within the pattern updated): built once
// per domain on the host in FP64.
struct IcArgs {
    const double* L[8]; const double* U[8]; int off[8];
    const double* dinv;
    int nnx, nny;
};

__device__ inline int ic_node(const IcArgs& a, int color, int t, int* ok) {
    const int ci = color & 1, cj = color >> 1;
    const int wc = (a.nny - cj + 1) / 2, hc = (a.nnx - ci + 1) / 2;
    const int r = t / wc, c = t - r * wc;
    *ok = r < hc;
    return (ci + 2 * r) * a.nny + cj + 2 * c;
}

// forward (L y = r, colours ascending) or backward (L^T x = y, colours descending) sweep
// over one colour; reads only entries of other colours, written by earlier sweeps
__global__ __launch_bounds__(256) void k_ic_sweep(IcArgs a, int color, int backward, const double* in, double* out,
                                                  const double* scal) {
    if (scal[S_DONE] != 0.0) return;
    int ok = 0;
    const int i = ic_node(a, color, blockIdx.x * blockDim.x + threadIdx.x, &ok);
    if (!ok) return;
    const double d = a.dinv[i];
    double s = in[i];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const double c = backward ? a.U[q][i] : a.L[q][i];
        s -= c * out[i + a.off[q]];
    }
    out[i] = d != 0.0 ? s * d : 0.0;
}

// r = b - A x; partials of b.b (P_BB)
template <class Op>
__global__ __launch_bounds__(PCG_BLOCK) void k_ic_init(int n, Op A, const double* b, const double* x, double* r,
                                                       double* scal, double* partials, int pstride) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double v[1] = {0.0};
    if (i < n) {
        double ri = b[i] - A.row(i, [&](int j) { return x[j]; });
        r[i] = ri;
        v[0] = b[i] * b[i];
    }
    const int regs[1] = {P_BB};
    block_partials<1>(v, partials, pstride, regs);
    if (blockIdx.x == 0 && threadIdx.x == 0) { scal[S_DONE] = 0.0; scal[S_ITER] = 0.0; scal[S_IERR] = 0.0; }
}

// partials of r.z (P_RZ)
__global__ __launch_bounds__(PCG_BLOCK) void k_ic_rz(int n, const double* r, const double* z, const double* scal,
                                                     double* partials, int pstride) {
    if (scal[S_DONE] != 0.0) return;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double v[1] = {i < n ? r[i] * z[i] : 0.0};
    const int regs[1] = {P_RZ};
    block_partials<1>(v, partials, pstride, regs);
}

// alpha = rho_k / p.q, x += alpha p, r -= alpha q; partials of r.r (P_RR)
__global__ __launch_bounds__(PCG_BLOCK) void k_ic_update(int n, int k, int nb, const double* p, const double* q,
                                                         double* x, double* r, double* scal, double* partials,
                                                         int pstride) {
    if (scal[S_DONE] != 0.0) return;
    double pq = reduce_region(partials, pstride, P_PQ, nb);
    if (pq == 0.0) {
        if (blockIdx.x == 0 && threadIdx.x == 0) { scal[S_DONE] = 1.0; scal[S_IERR] = 3.0; scal[S_ITER] = (double)k; }
        return;
    }
    const double alpha = scal[k & 1 ? S_RHO1 : S_RHO0] / pq;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double v[1] = {0.0};
    if (i < n) {
        const double ri = r[i] - alpha * q[i];
        x[i] = x[i] + alpha * p[i];
        r[i] = ri;
        v[0] = ri * ri;
    }
    const int regs[1] = {P_RR};
    block_partials<1>(v, partials, pstride, regs);
}

// The factor, on the host: rows in colour order; for row i (unknown), each lower neighbour j
// (an unknown of a lower colour coupled to i) in ascending colour,
//   l_ij = (a_ij - sum_k l_ik l_jk) / l_jj    over k lower neighbours of both, below j
//   l_ii = sqrt(a_ii - sum_j l_ij^2)
// A pivot that is not positive (possible only off the M-matrix case) is replaced by a_ii.
static int ic_build() {
    const MeshParams& m = g.mesh;
    const int n = m.nnodes, nny = m.nny;
    std::vector<std::vector<double>> dia(9, std::vector<double>(n));
    for (int q = 0; q < 9; ++q)
        HIPCHK(hipMemcpy(dia[q].data(), g.d_dia64[q], n * sizeof(double), hipMemcpyDeviceToHost));
    auto color = [&](int i) { return (i / nny) % 2 + 2 * ((i % nny) % 2); };
    const int* off = g.dia_off;                        // off[4] = 0; off[8 - q] = -off[q]
    std::vector<double> L(8 * (size_t)n, 0.0), U(8 * (size_t)n, 0.0), lii(n, 0.0);
    auto lq = [](int q) { return q < 4 ? q : q - 1; };   // stencil direction -> factor slot
    // direction between two nodes (-1 if not stencil neighbours)
    auto dir = [&](int from, int to) {
        const int di = to / nny - from / nny, dj = to % nny - from % nny;
        if (di < -1 || di > 1 || dj < -1 || dj > 1) return -1;
        return (di + 1) * 3 + (dj + 1);
    };
    int shifted = 0;
    for (int c = 0; c < 4; ++c)
        for (int i = 0; i < n; ++i) {
            if (color(i) != c || !(dia[4][i] > 0.0)) continue;
            double sq = 0.0;
            for (int cj = 0; cj < c; ++cj)
                for (int q = 0; q < 9; ++q) {
                    if (q == 4 || dia[q][i] == 0.0) continue;
                    const int j = i + off[q];
                    if (j < 0 || j >= n || color(j) != cj || !(dia[4][j] > 0.0)) continue;
                    double a = dia[q][i];
                    // common lower neighbours k of i and j with colour below cj
                    for (int q2 = 0; q2 < 9; ++q2) {
                        if (q2 == 4) continue;
                        const int k = i + off[q2];
                        if (k < 0 || k >= n || color(k) >= cj) continue;
                        const double lik = L[(size_t)lq(q2) * n + i];
                        if (lik == 0.0) continue;
                        const int qj = dir(j, k);
                        if (qj < 0 || qj == 4) continue;
                        a -= lik * L[(size_t)lq(qj) * n + j];
                    }
                    const double l = a / lii[j];
                    L[(size_t)lq(q) * n + i] = l;
                    sq += l * l;
                }
            double p = dia[4][i] - sq;
            if (!(p > 1e-12 * dia[4][i])) { p = dia[4][i]; shifted++; }
            lii[i] = std::sqrt(p);
        }
    // U[q][i] = l_{j i} for j = i + off[q] of a higher colour (stored at L[8 - q][j])
    for (int i = 0; i < n; ++i)
        for (int q = 0; q < 9; ++q) {
            if (q == 4) continue;
            const int j = i + off[q];
            if (j < 0 || j >= n) continue;
            U[(size_t)lq(q) * n + i] = L[(size_t)lq(8 - q) * n + j];
        }
    std::vector<double> dinv(n, 0.0);
    for (int i = 0; i < n; ++i) dinv[i] = lii[i] > 0.0 ? 1.0 / lii[i] : 0.0;
    if (g.d_ic) { (void)hipFree(g.d_ic); g.d_ic = nullptr; }
    const size_t pad = ((size_t)nny + 2 + 63) / 64 * 64, len = (size_t)n + 2 * pad;
    const size_t bytes = (17 * (size_t)n + len) * sizeof(double);
    HIPCHK(hipMalloc(&g.d_ic, bytes));
    double* d = (double*)g.d_ic;
    HIPCHK(hipMemcpy(d, L.data(), 8 * (size_t)n * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d + 8 * (size_t)n, U.data(), 8 * (size_t)n * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d + 16 * (size_t)n, dinv.data(), (size_t)n * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemset(d + 17 * (size_t)n, 0, len * sizeof(double)));
    for (int q = 0; q < 8; ++q) { g.ic_L[q] = d + (size_t)q * n; g.ic_U[q] = d + (size_t)(8 + q) * n; }
    g.ic_dinv = d + 16 * (size_t)n;
    g.ic_y = d + 17 * (size_t)n + pad;
    g.ic_shifted = shifted;
    g.ic_epoch = g.domain_epoch;
    return PIC_OK;
}

static void ic_apply(const IcArgs& a, const double* r, double* z) {
    const int n = g.mesh.nnodes;
    for (int c = 0; c < 4; ++c) {
        const int nc = ((a.nnx - (c & 1) + 1) / 2) * ((a.nny - (c >> 1) + 1) / 2);
        hipLaunchKernelGGL(k_ic_sweep, dim3((nc + 255) / 256), dim3(256), 0, g.stream, a, c, 0, r, g.ic_y, g.d_scal);
    }
    for (int c = 3; c >= 0; --c) {
        const int nc = ((a.nnx - (c & 1) + 1) / 2) * ((a.nny - (c >> 1) + 1) / 2);
        hipLaunchKernelGGL(k_ic_sweep, dim3((nc + 255) / 256), dim3(256), 0, g.stream, a, c, 1, g.ic_y, z, g.d_scal);
    }
    (void)n;
}

int run_pcg_ic() {
    if (g.ic_epoch != g.domain_epoch) { int rc = ic_build(); if (rc) return rc; }
    const int n = g.mesh.nnodes, nb = pcg_blocks(n), nb256 = (n + 255) / 256;
    if (nb > g.pcg_pstride) return pic_fail(PIC_ERR_STATE, "PCG partial buffer too small");
    IcArgs a;
    for (int q = 0; q < 8; ++q) { a.L[q] = g.ic_L[q]; a.U[q] = g.ic_U[q]; a.off[q] = g.dia_off[q < 4 ? q : q + 1]; }
    a.dinv = g.ic_dinv; a.nnx = g.mesh.nnx; a.nny = g.mesh.nny;
    hipLaunchKernelGGL(k_phi_to_xn, dim3(nb256), dim3(256), 0, g.stream, n, g.d_node_type2, g.d_phi, g.d_xn);
    DiaOp<double> A = dia_op<double>(1);
    double* scal = g.d_scal;
    double h2[2] = {g.pcg_tol * g.pcg_tol, (double)g.pcg_maxit};
    HIPCHK(hipMemcpyAsync(scal + S_TOL2, h2, 2 * sizeof(double), hipMemcpyHostToDevice, g.stream));
    hipLaunchKernelGGL((k_ic_init<DiaOp<double>>), dim3(nb), dim3(PCG_BLOCK), 0, g.stream, n, A, g.d_bn, g.d_xn, g.d_rn,
                       scal, g.d_partials, g.pcg_pstride);
    ic_apply(a, g.d_rn, g.d_zn);
    hipLaunchKernelGGL(k_ic_rz, dim3(nb), dim3(PCG_BLOCK), 0, g.stream, n, g.d_rn, g.d_zn, scal, g.d_partials,
                       g.pcg_pstride);
    HIPCHK(hipGetLastError());
    double* pa = g.d_pn; double* pb = g.d_p2n;
    const int CHUNK = 16;
    int k = 0;
    for (;;) {
        for (int c = 0; c < CHUNK; ++c, ++k) {
            hipLaunchKernelGGL((k_pcg_spmv<double, DiaOp<double>>), dim3(nb), dim3(PCG_BLOCK), 0, g.stream, n, k, nb, A,
                               g.d_zn, pa, pb, g.d_qn, scal, g.d_partials, g.pcg_pstride);
            hipLaunchKernelGGL(k_ic_update, dim3(nb), dim3(PCG_BLOCK), 0, g.stream, n, k, nb, pb, g.d_qn, g.d_xn, g.d_rn,
                               scal, g.d_partials, g.pcg_pstride);
            ic_apply(a, g.d_rn, g.d_zn);
            hipLaunchKernelGGL(k_ic_rz, dim3(nb), dim3(PCG_BLOCK), 0, g.stream, n, g.d_rn, g.d_zn, scal, g.d_partials,
                               g.pcg_pstride);
            double* t = pa; pa = pb; pb = t;
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(g.h_scal_pinned, scal, S_NSCAL * sizeof(double), hipMemcpyDeviceToHost, g.stream));
        overlap_point();
        HIPCHK(hipStreamSynchronize(g.stream));
        if (g.h_scal_pinned[S_DONE] != 0.0) break;
        if (k > g.pcg_maxit + CHUNK) { g.h_scal_pinned[S_IERR] = 1.0; g.h_scal_pinned[S_ITER] = k; break; }
    }
    const int iters = (int)g.h_scal_pinned[S_ITER], ierr = (int)g.h_scal_pinned[S_IERR];
    double rr, bb;
    { int rc = resid64(n, A, g.d_bn, g.d_xn, nullptr, &rr, &bb); if (rc) return rc; }
    hipLaunchKernelGGL(k_xn_to_phi, dim3(nb256), dim3(256), 0, g.stream, n, g.d_node_type2, g.d_xn, g.d_phi);
    HIPCHK(hipGetLastError());
    g.last_iter = iters; g.last_outer = 1;
    g.last_relres = bb > 0 ? std::sqrt(rr / bb) : 0.0;
    if (ierr == 1) return pic_fail(PIC_ERR_PCG_MAXITER, "IC(0)-PCG reached PCG_MaxIT");
    if (ierr == 2 || ierr == 3) return pic_fail(PIC_ERR_PCG_BREAKDOWN, "IC(0)-PCG breakdown");
    return PIC_OK;
}
