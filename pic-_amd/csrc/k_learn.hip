// Learned preconditioner for the field solve: a multigrid V-cycle on the node grid whose
// smoothers on the finest levels are learned convolutions. It is trained on the GPU with its
// own forward and backward kernels on residual/error pairs harvested from the plain solver,
// and applied as the preconditioner of a flexible CG (FCG(1)) on the Jacobi-scaled system.
//
// Hierarchy (set up once per domain, FP64 on the host, stored FP32): level 0 is the
// Jacobi-scaled node operator A' (unit diagonal on unknowns); level l+1 holds the nodes
// (I, J) <-> fine (2I, 2J); P_l is bilinear interpolation with the rows of fine Dirichlet
// nodes removed, R_l = P_l^T, A_{l+1} = P_l^T A_l P_l (nine diagonals per level); the
// coarsening stops at the first level with at most 33 x 33 nodes (once every learned level
// has a coarser one), and that coarsest level is solved exactly with a dense inverse.
//
// Network = symmetric V(1,1)-cycle (encoder: smooth, residual, restrict; decoder:
// interpolate, residual, smooth):
//   x  = K_l r;  rc = R_l (r - A_l x);  ec = V_{l+1}(rc);  x2 = x + P_l ec;
//   y  = x2 + K_l (r - A_l x2)
// Smoothers: on the finest levels l < Ls a learned convolution
//   K_l = Dh In_l^T S_l In_l Dh    (Dh = diag(A_l)^-1/2, In_l: 3x3 conv 1 -> C_l channels with
//                                   weights W_l[t][c], S_l = Q_l Q_l^T, so K_l is symmetric
//                                   positive semi-definite)
// initialised to damped Jacobi (In = identity tap on channel 0, S = omega e0 e0^T); damped
// Jacobi omega/diag on the coarser levels. M is symmetric; the untrained M is the classical
// Jacobi V-cycle.
//
// The smoother is linear with one input channel, so In^T S In collapses to 81 tap-pair
// weights (a symmetric 5x5 stencil away from the grid edge), recomputed from the FP32 master
// weights after every update; the V-cycle (solve and training forward) applies it in FP32.
// Training minimises sum (M r - e)^T A' (M r - e) / (e^T A' e) over pairs r = A' e. With
// g = dL/dy, the gradient of a learned smoother is dL/dK = s2(g) (x) r_l(r) + r_l(g) (x) s2(r)
// (inputs of its pre-/post-smoothing in the V-cycles run on r and on g). In the factored
// form, with U = In Dh u and V = In Dh v, dS = sum_p U V^T and dIn = corr(S V, Dh u) +
// corr(S U, Dh v); both are linear in the 3x3 neighbourhoods of Dh u and Dh v, so they
// contract exactly to the tap-pair Gram matrix of the two scalar fields (k_tapgram below):
// the gradient needs one pass over four scalar fields per level, no lifted C-channel tensors
// and no matrix-core work (there is no GEMM-shaped operand: the input has one channel).
#include "pic_device.h"
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#define LN_MAX_LEVELS 12
#define LN_MAX_LEARNED 4
#define LN_KW 128                                    // smoother weights per level: k81 then k25
#define LN_K25 96
#define LN_OMEGA 0.8f
#define LN_MAX_COARSE 2048
#define LN_DENSE_NODES 1089                          // 33 x 33: coarsest level of the hierarchy

// ---------------------------------------------------------------- multigrid transfer kernels
// ---------------------------------------------------------------- FCG(1) on the scaled system
// scal slots (shared with the CG): NORMRHS, DONE, ITER, IERR, TOL2, MAXIT; pq in GAM0/GAM1
enum { F_S_NORMRHS = 4, F_S_NORMR = 5, F_S_DONE = 6, F_S_ITER = 7, F_S_IERR = 8, F_S_TOL2 = 9, F_S_MAXIT = 10,
       F_S_PQ0 = 12, F_S_PQ1 = 13, F_S_FAIL = 22 };
// The iteration index k of a pass is a launch argument of the first V-cycle kernel, which
// runs the convergence test (FcgCheck), and of the update. A refinement pass leaves a
// record at FCG_REC_BASE + FCG_RECW * slot: iterations, done, error code, last ||r||^2/||r0||^2,
// skipped (the outer residual met its tolerance before the pass, or an earlier pass failed).
// FAIL is set by a pass that ends in a NaN or a breakdown.
#define FCG_BLOCK 256
// tightest relative inner tolerance, squared: (3.2e-4)^2. An FP32 pass cannot cut the FP64
// outer residual by much more (the FP32 operator's coefficient rounding times the condition
// number), so iterations past it would be lost
#define FCG_INNER_FLOOR2 1e-7

__device__ inline void fcg_block_partials(int nv, double* v, double* part, int pstride) {
    __shared__ double sh[4][FCG_BLOCK / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int k = 0; k < nv; ++k) {
        double s = v[k];
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
        if (lane == 0) sh[k][wid] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0)
        for (int k = 0; k < nv; ++k) part[(size_t)k * pstride + blockIdx.x] = sh[k][0] + sh[k][1] + sh[k][2] + sh[k][3];
}
__device__ inline void fcg_reduce(int nv, const double* part, int pstride, int nb, double* out) {
    __shared__ double sh[4][FCG_BLOCK / 64];
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    for (int q = threadIdx.x; q < nb; q += FCG_BLOCK)
        for (int k = 0; k < nv; ++k) s[k] += part[(size_t)k * pstride + q];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int k = 0; k < nv; ++k) {
        for (int off = 32; off > 0; off >>= 1) s[k] += __shfl_xor(s[k], off, 64);
        if (lane == 0) sh[k][wid] = s[k];
    }
    __syncthreads();
    for (int k = 0; k < nv; ++k) out[k] = sh[k][0] + sh[k][1] + sh[k][2] + sh[k][3];
    __syncthreads();
}


// Convergence test of the current r before FCG iteration k, run by every block of the first
// V-cycle kernel in its prologue (the partial sums of r.r from the previous update or the
// pass start are reduced in the same fixed order by every block, so all blocks take the same
// decision; block 0 publishes it): done when ||r||^2 / ||r_0||^2 <= tol^2 (k = 0 sets
// ||r_0||^2), at k >= maxit or on a NaN. The later kernels of the V-cycle and the FCG vector
// kernels read DONE. part == nullptr: no test (training and evaluation passes).
struct FcgCheck { const double* part; int pstride, nb, k; double* scal; };

__device__ inline bool fcg_check_block(const FcgCheck& c) {
    const double pre = c.scal[F_S_DONE];
    double rr;
    fcg_reduce(1, c.part, c.pstride, c.nb, &rr);
    if (pre != 0.0) return true;
    const double nb2 = c.k == 0 ? rr : c.scal[F_S_NORMRHS];
    const bool conv = nb2 == 0.0 || rr / nb2 <= c.scal[F_S_TOL2];
    const bool stop = conv || (double)c.k >= c.scal[F_S_MAXIT] || !(rr == rr);
    if (blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0 && threadIdx.x == 0) {
        if (c.k == 0) c.scal[F_S_NORMRHS] = rr;
        c.scal[F_S_NORMR] = rr;
        if (stop) {
            c.scal[F_S_ITER] = (double)c.k;
            if (!conv) c.scal[F_S_IERR] = rr == rr ? 1.0 : 2.0;
            c.scal[F_S_DONE] = 1.0;
        }
    }
    return stop;
}

struct MgArgs {
    int Hf, Wf, Hc, Wc;
    const float* dia;            // fine operator, nine diagonals [9][Hf*Wf], tap t <-> (t/3-1, t%3-1)
    const float* dinv;           // fine 1/diag (0 on Dirichlet nodes)
    const float* r;              // fine residual (input of this level)
    float* x;                    // fine pre-smoothed correction
    float* rc;                   // coarse residual (down)
    const float* ec;             // coarse correction (up)
    float* y;                    // fine output (up)
    float* s2;                   // fine post-smoothing residual (up, learned smoother)
    int64_t fs, cs;              // batch strides of fine / coarse vectors (r, x, y, s2 / rc, ec)
    int64_t r_bs, y_bs;          // batch strides of r and y (level 0 uses the caller's vectors)
    float omega;
    int jacobi;                  // 1: x = omega/diag r here (down) and the post-smoothing too (up)
    const double* done;
    const float* dh;             // learned levels: diag^-1/2 and the 81 tap-pair smoother weights
    const float* k81;
    FcgCheck chk;                // first kernel of an FCG V-cycle: the convergence test (else part == nullptr)
};

// Down, tiled: a block owns CI x CJ coarse nodes. It stages x on the fine nodes
// [2I0-2, 2I0+2CI] x [2J0-2, 2J0+2CJ] (Jacobi levels: omega/diag r, computed here and written
// back for the owned fine nodes; learned levels: the smoother's output) and r on the
// restriction footprint in LDS, forms the fine residual s = r - A x on the footprint, then
// rc = R s.
#define MGD_CI 8
#define MGD_CJ 32
#define MGD_XI (2 * MGD_CI + 3)
#define MGD_XJ (2 * MGD_CJ + 3)
#define MGD_SI (2 * MGD_CI + 1)
#define MGD_SJ (2 * MGD_CJ + 1)
#define MGD_NX (MGD_XI * MGD_XJ)
#define MGD_NS (MGD_SI * MGD_SJ)
#define MGD_XIT ((MGD_NX + 255) / 256)
#define MGD_SIT ((MGD_NS + 255) / 256)
// Every global load of the kernel (operator rows of the residual nodes, r and x / 1/diag on
// the staged nodes) is issued before the first barrier, so the block waits for memory once.
template <bool JAC>
__device__ inline void mg_down_body(const MgArgs& a, float* xs, float* rx, float* ss) {
    const int b = blockIdx.z;
    const int I0 = blockIdx.y * MGD_CI, J0 = blockIdx.x * MGD_CJ;
    const int nf = a.Hf * a.Wf;
    const float* r = a.r + (int64_t)b * a.r_bs;
    float* x = a.x + (int64_t)b * a.fs;
    const int fi0 = 2 * I0 - 2, fj0 = 2 * J0 - 2;
    float dv[MGD_SIT][9];
#pragma unroll
    for (int it = 0; it < MGD_SIT; ++it) {
        const int k = threadIdx.x + it * 256;
        const int u = k / MGD_SJ, v = k - u * MGD_SJ;
        const int fi = fi0 + 1 + u, fj = fj0 + 1 + v;
        const bool in = k < MGD_NS && fi >= 0 && fi < a.Hf && fj >= 0 && fj < a.Wf;
        const int f = in ? fi * a.Wf + fj : 0;
#pragma unroll
        for (int t = 0; t < 9; ++t) dv[it][t] = a.dia[(int64_t)t * nf + f];
        if (!in) dv[it][4] = 0.0f;                      // no residual outside the grid
    }
    float xv[MGD_XIT], rv[MGD_XIT];
#pragma unroll
    for (int it = 0; it < MGD_XIT; ++it) {
        const int k = threadIdx.x + it * 256;
        const int u = k / MGD_XJ, v = k - u * MGD_XJ;
        const int fi = fi0 + u, fj = fj0 + v;
        const bool in = k < MGD_NX && fi >= 0 && fi < a.Hf && fj >= 0 && fj < a.Wf;
        const int f = in ? fi * a.Wf + fj : 0;
        rv[it] = r[f];
        xv[it] = JAC ? a.dinv[f] : x[f];
        if (!in) { rv[it] = 0.0f; xv[it] = 0.0f; }
    }
    if (a.chk.part ? fcg_check_block(a.chk) : (a.done && *a.done != 0.0)) return;
#pragma unroll
    for (int it = 0; it < MGD_XIT; ++it) {
        const int k = threadIdx.x + it * 256;
        if (k < MGD_NX) {
            const float val = JAC ? a.omega * xv[it] * rv[it] : xv[it];
            xs[k] = val;
            rx[k] = rv[it];
            if (JAC) {                                   // owned fine nodes [2I0, 2I0+2CI) x [2J0, 2J0+2CJ)
                const int u = k / MGD_XJ - 2, v = k - (k / MGD_XJ) * MGD_XJ - 2;
                const int fi = 2 * I0 + u, fj = 2 * J0 + v;
                if (u >= 0 && u < 2 * MGD_CI && v >= 0 && v < 2 * MGD_CJ && fi < a.Hf && fj < a.Wf)
                    x[fi * a.Wf + fj] = val;
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < MGD_SIT; ++it) {
        const int k = threadIdx.x + it * 256;
        if (k < MGD_NS) {
            const int u = k / MGD_SJ, v = k - u * MGD_SJ;
            float sv = 0.0f;
            if (dv[it][4] != 0.0f) {
                sv = rx[(u + 1) * MGD_XJ + (v + 1)];
#pragma unroll
                for (int t = 0; t < 9; ++t) sv = sv - dv[it][t] * xs[(u + t / 3) * MGD_XJ + (v + t % 3)];
            }
            ss[k] = sv;
        }
    }
    __syncthreads();
    const int ti = threadIdx.x / MGD_CJ, tj = threadIdx.x - (threadIdx.x / MGD_CJ) * MGD_CJ;
    const int I = I0 + ti, J = J0 + tj;
    if (I < a.Hc && J < a.Wc) {
        float acc = 0.0f;
#pragma unroll
        for (int u = -1; u <= 1; ++u)
#pragma unroll
            for (int v = -1; v <= 1; ++v)
                acc = acc + (u == 0 ? 1.0f : 0.5f) * (v == 0 ? 1.0f : 0.5f) * ss[(2 * ti + 1 + u) * MGD_SJ + (2 * tj + 1 + v)];
        a.rc[(int64_t)b * a.cs + I * a.Wc + J] = acc;
    }
}

__global__ __launch_bounds__(256) void k_mg_down(MgArgs a) {
    __shared__ float xs[MGD_NX];
    __shared__ float rx[MGD_NX];
    __shared__ float ss[MGD_NS];
    if (a.jacobi) mg_down_body<true>(a, xs, rx, ss);
    else mg_down_body<false>(a, xs, rx, ss);
}

// Down on a learned level with the pre-smoothing fused in: the block computes the learned
// smoother x = Dh K Dh r (81 tap pairs, the arithmetic of k_smooth_k81) on the staged fine
// nodes from Dh r on a further halo of 2, writes x on its owned fine nodes, then forms the
// residual and restricts it exactly as k_mg_down does.
#define MGF_UI (MGD_XI + 4)
#define MGF_UJ (MGD_XJ + 4)
#define MGF_NU (MGF_UI * MGF_UJ)
#define MGF_UIT ((MGF_NU + 255) / 256)
__global__ __launch_bounds__(256) void k_mg_smooth_down(MgArgs a) {
    __shared__ float us[MGF_NU];                 // Dh r on the U region (0 outside the grid)
    __shared__ float ru[MGF_NU];                 // r on the U region
    __shared__ float du[MGF_NU];                 // Dh on the U region
    __shared__ float xs[MGD_NX];
    __shared__ float ss[MGD_NS];
    __shared__ float kw[LN_KW];
    const int b = blockIdx.z;
    const int I0 = blockIdx.y * MGD_CI, J0 = blockIdx.x * MGD_CJ;
    const int nf = a.Hf * a.Wf;
    const float* r = a.r + (int64_t)b * a.r_bs;
    float* x = a.x + (int64_t)b * a.fs;
    const int fi0 = 2 * I0 - 2, fj0 = 2 * J0 - 2;          // X region origin
    const int ui0 = fi0 - 2, uj0 = fj0 - 2;                // U region origin
    // ---- all global loads first
    float dv[MGD_SIT][9];
#pragma unroll
    for (int it = 0; it < MGD_SIT; ++it) {
        const int k = threadIdx.x + it * 256;
        const int u = k / MGD_SJ, v = k - u * MGD_SJ;
        const int fi = fi0 + 1 + u, fj = fj0 + 1 + v;
        const bool in = k < MGD_NS && fi >= 0 && fi < a.Hf && fj >= 0 && fj < a.Wf;
        const int f = in ? fi * a.Wf + fj : 0;
#pragma unroll
        for (int t = 0; t < 9; ++t) dv[it][t] = a.dia[(int64_t)t * nf + f];
        if (!in) dv[it][4] = 0.0f;
    }
    float hr[MGF_UIT], hd[MGF_UIT];
#pragma unroll
    for (int it = 0; it < MGF_UIT; ++it) {
        const int k = threadIdx.x + it * 256;
        const int u = k / MGF_UJ, v = k - u * MGF_UJ;
        const int gi = ui0 + u, gj = uj0 + v;
        const bool in = k < MGF_NU && gi >= 0 && gi < a.Hf && gj >= 0 && gj < a.Wf;
        const int g = in ? gi * a.Wf + gj : 0;
        hr[it] = r[g];
        hd[it] = a.dh[g];
        if (!in) { hr[it] = 0.0f; hd[it] = 0.0f; }
    }
    const float kv = a.k81[threadIdx.x < LN_KW ? threadIdx.x : 0];
    if (a.chk.part ? fcg_check_block(a.chk) : (a.done && *a.done != 0.0)) return;
#pragma unroll
    for (int it = 0; it < MGF_UIT; ++it) {
        const int k = threadIdx.x + it * 256;
        if (k < MGF_NU) { us[k] = hd[it] * hr[it]; ru[k] = hr[it]; du[k] = hd[it]; }
    }
    if (threadIdx.x < LN_KW) kw[threadIdx.x] = kv;
    __syncthreads();
    // ---- learned smoother on the X region
#pragma unroll
    for (int it = 0; it < MGD_XIT; ++it) {
        const int k = threadIdx.x + it * 256;
        if (k < MGD_NX) {
            const int u = k / MGD_XJ, v = k - u * MGD_XJ;
            const int i = fi0 + u, j = fj0 + v;
            float val = 0.0f;
            if (i >= 0 && i < a.Hf && j >= 0 && j < a.Wf) {
                const float dp = du[(u + 2) * MGF_UJ + (v + 2)];
                float sm = 0.0f;
                if (dp != 0.0f) {
                    if (i >= 1 && i <= a.Hf - 2 && j >= 1 && j <= a.Wf - 2) {
#pragma unroll
                        for (int d = 0; d < 25; ++d)
                            sm = sm + kw[LN_K25 + d] * us[(u + d / 5) * MGF_UJ + (v + d % 5)];
                    } else {
#pragma unroll
                        for (int t = 0; t < 9; ++t) {
                            const int qi = i - (t / 3 - 1), qj = j - (t % 3 - 1);   // intermediate pixel p - d_t
                            if (qi < 0 || qi >= a.Hf || qj < 0 || qj >= a.Wf) continue;
#pragma unroll
                            for (int t2 = 0; t2 < 9; ++t2)
                                sm = sm + kw[9 * t + t2] * us[(u + 2 - (t / 3 - 1) + (t2 / 3 - 1)) * MGF_UJ + (v + 2 - (t % 3 - 1) + (t2 % 3 - 1))];
                        }
                    }
                    sm = dp * sm;
                }
                val = sm;
                const int ou = u - 2, ov = v - 2;                  // owned fine nodes
                if (ou >= 0 && ou < 2 * MGD_CI && ov >= 0 && ov < 2 * MGD_CJ) x[i * a.Wf + j] = val;
            }
            xs[k] = val;
        }
    }
    __syncthreads();
    // ---- residual on the restriction footprint, then restriction (as k_mg_down)
#pragma unroll
    for (int it = 0; it < MGD_SIT; ++it) {
        const int k = threadIdx.x + it * 256;
        if (k < MGD_NS) {
            const int u = k / MGD_SJ, v = k - u * MGD_SJ;
            float sv = 0.0f;
            if (dv[it][4] != 0.0f) {
                sv = ru[(u + 3) * MGF_UJ + (v + 3)];
#pragma unroll
                for (int t = 0; t < 9; ++t) sv = sv - dv[it][t] * xs[(u + t / 3) * MGD_XJ + (v + t % 3)];
            }
            ss[k] = sv;
        }
    }
    __syncthreads();
    const int ti = threadIdx.x / MGD_CJ, tj = threadIdx.x - (threadIdx.x / MGD_CJ) * MGD_CJ;
    const int I = I0 + ti, J = J0 + tj;
    if (I < a.Hc && J < a.Wc) {
        float acc = 0.0f;
#pragma unroll
        for (int u = -1; u <= 1; ++u)
#pragma unroll
            for (int v = -1; v <= 1; ++v)
                acc = acc + (u == 0 ? 1.0f : 0.5f) * (v == 0 ? 1.0f : 0.5f) * ss[(2 * ti + 1 + u) * MGD_SJ + (2 * tj + 1 + v)];
        a.rc[(int64_t)b * a.cs + I * a.Wc + J] = acc;
    }
}

// Up, tiled: a block owns FI x FJ fine nodes; it stages ec on the covering coarse nodes and x
// on the fine halo in LDS, forms x2 = x + P ec on the halo, then (Jacobi levels)
// y = x2 + omega/diag (r - A x2), or (learned levels) y = x2 and s2 = r - A x2.
#define MGU_FI 4
#define MGU_FJ 64
#define MGU_CI (MGU_FI / 2 + 2)
#define MGU_CJ (MGU_FJ / 2 + 2)
#define MGU_NE (MGU_CI * MGU_CJ)
#define MGU_NX ((MGU_FI + 2) * (MGU_FJ + 2))
#define MGU_XIT ((MGU_NX + 255) / 256)
__global__ __launch_bounds__(256) void k_mg_up(MgArgs a) {
    static_assert(MGU_NE <= 256 && MGU_FI * MGU_FJ == 256, "k_mg_up tile");
    __shared__ float es[MGU_NE];
    __shared__ float x2s[MGU_NX];
    const int b = blockIdx.z;
    const int fi0 = blockIdx.y * MGU_FI, fj0 = blockIdx.x * MGU_FJ;
    const int nf = a.Hf * a.Wf;
    const float* r = a.r + (int64_t)b * a.r_bs;
    const float* x = a.x + (int64_t)b * a.fs;
    const float* ec = a.ec + (int64_t)b * a.cs;
    float* y = a.y + (int64_t)b * a.y_bs;
    const int ci0 = (fi0 - 1) >> 1, cj0 = (fj0 - 1) >> 1;    // floor
    // all global loads first: coarse correction, halo x and its Dirichlet marks, the own row
    float ev;
    {
        const int k = threadIdx.x;
        const int u = k / MGU_CJ, v = k - u * MGU_CJ;
        const int I = ci0 + u, J = cj0 + v;
        const bool in = k < MGU_NE && I >= 0 && I < a.Hc && J >= 0 && J < a.Wc;
        ev = ec[in ? I * a.Wc + J : 0];
        if (!in) ev = 0.0f;
    }
    float hx[MGU_XIT], hd[MGU_XIT];
#pragma unroll
    for (int it = 0; it < MGU_XIT; ++it) {
        const int k = threadIdx.x + it * 256;
        const int u = k / (MGU_FJ + 2), v = k - u * (MGU_FJ + 2);
        const int gi = fi0 - 1 + u, gj = fj0 - 1 + v;
        const bool in = k < MGU_NX && gi >= 0 && gi < a.Hf && gj >= 0 && gj < a.Wf;
        const int g = in ? gi * a.Wf + gj : 0;
        hd[it] = a.dia[4 * (int64_t)nf + g];
        hx[it] = x[g];
        if (!in) hd[it] = 0.0f;
    }
    const int ti = threadIdx.x / MGU_FJ, tj = threadIdx.x - (threadIdx.x / MGU_FJ) * MGU_FJ;
    const int fi = fi0 + ti, fj = fj0 + tj;
    const bool own = fi < a.Hf && fj < a.Wf;
    const int f = own ? fi * a.Wf + fj : 0;
    float dv[9];
#pragma unroll
    for (int t = 0; t < 9; ++t) dv[t] = a.dia[(int64_t)t * nf + f];
    const float rf = r[f];
    const float di = a.jacobi ? a.dinv[f] : 0.0f;
    if (a.done && *a.done != 0.0) return;
    if (threadIdx.x < MGU_NE) es[threadIdx.x] = ev;
    __syncthreads();
#pragma unroll
    for (int it = 0; it < MGU_XIT; ++it) {
        const int k = threadIdx.x + it * 256;
        if (k < MGU_NX) {
            const int u = k / (MGU_FJ + 2), v = k - u * (MGU_FJ + 2);
            const int gi = fi0 - 1 + u, gj = fj0 - 1 + v;
            float val = 0.0f;
            if (hd[it] != 0.0f) {
                const int i0 = gi >> 1, j0 = gj >> 1;
                const bool oi = gi & 1, oj = gj & 1;
                float pe = 0.0f;
                for (int p = 0; p < (oi ? 2 : 1); ++p)
                    for (int q = 0; q < (oj ? 2 : 1); ++q)
                        pe = pe + (oi ? 0.5f : 1.0f) * (oj ? 0.5f : 1.0f) * es[(i0 + p - ci0) * MGU_CJ + (j0 + q - cj0)];
                val = hx[it] + pe;
            }
            x2s[k] = val;
        }
    }
    __syncthreads();
    if (!own) return;
    const float xc = x2s[(ti + 1) * (MGU_FJ + 2) + (tj + 1)];
    if (dv[4] == 0.0f) {
        y[f] = 0.0f;
        if (!a.jacobi) a.s2[(int64_t)b * a.fs + f] = 0.0f;
        return;
    }
    float sres = rf;
#pragma unroll
    for (int t = 0; t < 9; ++t) sres = sres - dv[t] * x2s[(ti + t / 3) * (MGU_FJ + 2) + (tj + t % 3)];
    if (a.jacobi) {
        y[f] = xc + a.omega * di * sres;
    } else {
        y[f] = xc;
        a.s2[(int64_t)b * a.fs + f] = sres;
    }
}

// Up on a learned level with the post-smoothing fused in: y = x2 + Dh K Dh s2, s2 = r - A x2,
// x2 = x + P ec, with the arithmetic of k_mg_up followed by k_smooth_k81 (add), so the result is
// the same bits. A block owns MGU_FI x MGU_FJ fine nodes; it stages ec on the covering coarse
// nodes and x on the fine halo 3 (x2 there), forms s2 on the halo 2 (the smoother's footprint:
// every operator row of those nodes is loaded up front) and writes y and s2 (the training
// gradient's site) on its own nodes. Replaces two launches by one at 2.1x the operator rows.
#define MGS_HI (MGU_FI + 6)                          // x2: fine rows fi0-3 .. fi0+FI+2
#define MGS_HJ (MGU_FJ + 6)
#define MGS_NX (MGS_HI * MGS_HJ)
#define MGS_XIT ((MGS_NX + 255) / 256)
#define MGS_SI (MGU_FI + 4)                          // s2: halo 2
#define MGS_SJ (MGU_FJ + 4)
#define MGS_NS (MGS_SI * MGS_SJ)
#define MGS_SIT ((MGS_NS + 255) / 256)
#define MGS_CJ (MGU_FJ / 2 + 6)                       // coarse columns staged (fj0/2-2 .. fj0/2+FJ/2+3)
#define MGS_NE ((MGU_FI / 2 + 4) * MGS_CJ)
__global__ __launch_bounds__(256) void k_mg_up_smooth(MgArgs a) {
    static_assert(MGU_FI * MGU_FJ == 256 && MGS_NE <= 256, "k_mg_up_smooth tile");
    __shared__ float es[MGS_NE];
    __shared__ float x2s[MGS_NX];
    __shared__ float us[MGS_NS];
    __shared__ float s2s[MGS_NS];
    __shared__ float kw[LN_KW];
    const int b = blockIdx.z;
    const int fi0 = blockIdx.y * MGU_FI, fj0 = blockIdx.x * MGU_FJ;
    const int H = a.Hf, W = a.Wf;
    const int nf = H * W;
    const float* r = a.r + (int64_t)b * a.r_bs;
    const float* x = a.x + (int64_t)b * a.fs;
    const float* ec = a.ec + (int64_t)b * a.cs;
    float* y = a.y + (int64_t)b * a.y_bs;
    constexpr int ECI = MGU_FI / 2 + 4;                 // coarse rows staged (fi0/2-2 .. fi0/2+FI/2+1)
    const int ci0 = (fi0 - 3) >> 1, cj0 = (fj0 - 3) >> 1;    // floor
    // ---- every global load first
    const float kv = a.k81[threadIdx.x < LN_KW ? threadIdx.x : 0];
    float ev;
    {
        const int k = threadIdx.x;
        const int u = k / MGS_CJ, v = k - u * MGS_CJ;
        const int I = ci0 + u, J = cj0 + v;
        const bool in = k < MGS_NE && u < ECI && I >= 0 && I < a.Hc && J >= 0 && J < a.Wc;
        ev = ec[in ? I * a.Wc + J : 0];
        if (!in) ev = 0.0f;
    }
    float hx[MGS_XIT], hd[MGS_XIT];
#pragma unroll
    for (int it = 0; it < MGS_XIT; ++it) {
        const int k = threadIdx.x + it * 256;
        const int u = k / MGS_HJ, v = k - u * MGS_HJ;
        const int gi = fi0 - 3 + u, gj = fj0 - 3 + v;
        const bool in = k < MGS_NX && gi >= 0 && gi < H && gj >= 0 && gj < W;
        const int g = in ? gi * W + gj : 0;
        hd[it] = a.dia[4 * (int64_t)nf + g];
        hx[it] = x[g];
        if (!in) hd[it] = 0.0f;
    }
    float sr[MGS_SIT], sdh[MGS_SIT], sdv[MGS_SIT][9];
#pragma unroll
    for (int it = 0; it < MGS_SIT; ++it) {
        const int k = threadIdx.x + it * 256;
        const int u = k / MGS_SJ, v = k - u * MGS_SJ;
        const int gi = fi0 - 2 + u, gj = fj0 - 2 + v;
        const bool in = k < MGS_NS && gi >= 0 && gi < H && gj >= 0 && gj < W;
        const int g = in ? gi * W + gj : 0;
        sr[it] = r[g];
        sdh[it] = a.dh[g];
#pragma unroll
        for (int t = 0; t < 9; ++t) sdv[it][t] = a.dia[(int64_t)t * nf + g];
        if (!in) { sdh[it] = 0.0f; sdv[it][4] = 0.0f; }
    }
    if (a.done && *a.done != 0.0) return;
    if (threadIdx.x < MGS_NE) es[threadIdx.x] = ev;
    if (threadIdx.x < LN_KW) kw[threadIdx.x] = kv;
    __syncthreads();
    // ---- x2 = x + P ec on the halo 3 (0 on Dirichlet nodes and off the grid)
#pragma unroll
    for (int it = 0; it < MGS_XIT; ++it) {
        const int k = threadIdx.x + it * 256;
        if (k < MGS_NX) {
            const int u = k / MGS_HJ, v = k - u * MGS_HJ;
            const int gi = fi0 - 3 + u, gj = fj0 - 3 + v;
            float val = 0.0f;
            if (hd[it] != 0.0f) {
                const int i0 = gi >> 1, j0 = gj >> 1;
                const bool oi = gi & 1, oj = gj & 1;
                float pe = 0.0f;
                for (int p = 0; p < (oi ? 2 : 1); ++p)
                    for (int q = 0; q < (oj ? 2 : 1); ++q)
                        pe = pe + (oi ? 0.5f : 1.0f) * (oj ? 0.5f : 1.0f) * es[(i0 + p - ci0) * MGS_CJ + (j0 + q - cj0)];
                val = hx[it] + pe;
            }
            x2s[k] = val;
        }
    }
    __syncthreads();
    // ---- s2 = r - A x2 on the halo 2 (0 on Dirichlet nodes and off the grid), and Dh s2
#pragma unroll
    for (int it = 0; it < MGS_SIT; ++it) {
        const int k = threadIdx.x + it * 256;
        if (k < MGS_NS) {
            const int u = k / MGS_SJ, v = k - u * MGS_SJ;
            float sres = 0.0f;
            if (sdv[it][4] != 0.0f) {
                sres = sr[it];
#pragma unroll
                for (int t = 0; t < 9; ++t) sres = sres - sdv[it][t] * x2s[(u + t / 3) * MGS_HJ + (v + t % 3)];
            }
            s2s[k] = sres;
            us[k] = sdh[it] * sres;
        }
    }
    __syncthreads();
    const int ti = threadIdx.x / MGU_FJ, tj = threadIdx.x - (threadIdx.x / MGU_FJ) * MGU_FJ;
    const int i = fi0 + ti, j = fj0 + tj;
    if (i >= H || j >= W) return;
    const int f = i * W + j;
    const float dp = a.dh[f];
    // y before the smoothing: x2 (0 on a Dirichlet node, where x2 is 0)
    const float o0 = x2s[(ti + 3) * MGS_HJ + (tj + 3)];
    float s = 0.0f;
    if (dp != 0.0f) {
        if (i >= 1 && i <= H - 2 && j >= 1 && j <= W - 2) {
#pragma unroll
            for (int d = 0; d < 25; ++d) s = s + kw[LN_K25 + d] * us[(ti + d / 5) * MGS_SJ + (tj + d % 5)];
        } else {
#pragma unroll
            for (int t = 0; t < 9; ++t) {
                const int qi = i - (t / 3 - 1), qj = j - (t % 3 - 1);
                if (qi < 0 || qi >= H || qj < 0 || qj >= W) continue;
#pragma unroll
                for (int t2 = 0; t2 < 9; ++t2)
                    s = s + kw[9 * t + t2] * us[(ti + 2 - (t / 3 - 1) + (t2 / 3 - 1)) * MGS_SJ + (tj + 2 - (t % 3 - 1) + (t2 % 3 - 1))];
            }
        }
        s = dp * s;
    }
    y[f] = o0 + s;
    a.s2[(int64_t)b * a.fs + f] = s2s[(ti + 2) * MGS_SJ + (tj + 2)];
}

// Coarsest level: y = inv r with the dense inverse of the level's operator over its unknown
// nodes (rows and columns of Dirichlet nodes are zero; n <= LN_MAX_COARSE nodes). One wave per
// row: the lanes read the row with unit stride (the inverse, 4.7 MB at 33 x 33, stays in the
// last-level cache between applications) against r staged in LDS, then a wave reduction in a
// fixed order. MGX_ROWS rows per block, so the rows spread over the whole GPU; the batch entry
// is blockIdx.y. Replaces the Jacobi levels below 33 x 33: one launch instead of a chain of
// small phases, and an exact coarse correction (fewer FCG iterations).
#define MGX_ROWS 4
#define MGX_RIT (LN_MAX_COARSE / 64)
#define MGX_SIT (LN_MAX_COARSE / (64 * MGX_ROWS))
__global__ __launch_bounds__(64 * MGX_ROWS) void k_mg_dense(int n, const float* inv, const float* r, int64_t bs, float* y,
                                                            const double* done) {
    __shared__ float rs[LN_MAX_COARSE];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int row = blockIdx.x * MGX_ROWS + w;
    const float* rb = r + (int64_t)blockIdx.y * bs;
    const float* ir = inv + (int64_t)(row < n ? row : 0) * n;
    // every global load first: the row, then the block's share of r
    float iv[MGX_RIT];
#pragma unroll
    for (int c = 0; c < MGX_RIT; ++c) {
        const int j = lane + 64 * c;
        iv[c] = j < n ? ir[j] : 0.0f;
    }
    float rv[MGX_SIT];
#pragma unroll
    for (int q = 0; q < MGX_SIT; ++q) {
        const int k = threadIdx.x + 64 * MGX_ROWS * q;
        rv[q] = k < n ? rb[k] : 0.0f;
    }
    if (done && *done != 0.0) return;
#pragma unroll
    for (int q = 0; q < MGX_SIT; ++q) rs[threadIdx.x + 64 * MGX_ROWS * q] = rv[q];
    __syncthreads();
    float acc = 0.0f;
#pragma unroll
    for (int c = 0; c < MGX_RIT; ++c) acc = fmaf(iv[c], rs[lane + 64 * c], acc);
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if (lane == 0 && row < n) y[(int64_t)blockIdx.y * bs + row] = acc;
}

// Learned smoother, collapsed: out = [out +] Dh (In^T S In) (Dh v) on one level with the
// 5 x 5 stencil k25 on interior nodes and the 81 tap-pair weights k81 on the outer ring (zero
// outside the grid, as the factored form). A block owns
// SMI x SMJ nodes and stages Dh v on the tile with a halo of 2 in LDS.
#define SMI 4
#define SMJ 64
#define SM_NU ((SMI + 4) * (SMJ + 4))
#define SM_UIT ((SM_NU + 255) / 256)
__global__ __launch_bounds__(256) void k_smooth_k81(int H, int W, const float* k81, const float* dh, const float* v,
                                                    int64_t v_bs, float* out, int64_t o_bs, int add, const double* done,
                                                    FcgCheck chk) {
    static_assert(SMI * SMJ == 256, "k_smooth_k81 tile");
    __shared__ float us[SM_NU];
    __shared__ float kw[LN_KW];
    const float* vb = v + (int64_t)blockIdx.z * v_bs;
    float* ob = out + (int64_t)blockIdx.z * o_bs;
    const int i0 = blockIdx.y * SMI, j0 = blockIdx.x * SMJ;
    // all global loads first: weights, the scaled halo, the own node's scale and output
    const float kv = k81[threadIdx.x < LN_KW ? threadIdx.x : 0];
    float hd[SM_UIT], hv[SM_UIT];
#pragma unroll
    for (int it = 0; it < SM_UIT; ++it) {
        const int k = threadIdx.x + it * 256;
        const int u = k / (SMJ + 4), w = k - u * (SMJ + 4);
        const int ii = i0 - 2 + u, jj = j0 - 2 + w;
        const bool in = k < SM_NU && ii >= 0 && ii < H && jj >= 0 && jj < W;
        const int q = in ? ii * W + jj : 0;
        hd[it] = dh[q];
        hv[it] = vb[q];
        if (!in) { hd[it] = 0.0f; hv[it] = 0.0f; }
    }
    const int ti = threadIdx.x / SMJ, tj = threadIdx.x - (threadIdx.x / SMJ) * SMJ;
    const int i = i0 + ti, j = j0 + tj;
    const bool own = i < H && j < W;
    const int p = own ? i * W + j : 0;
    const float dp = dh[p];
    const float o0 = add ? ob[p] : 0.0f;
    if (chk.part ? fcg_check_block(chk) : (done && *done != 0.0)) return;
    if (threadIdx.x < LN_KW) kw[threadIdx.x] = kv;
#pragma unroll
    for (int it = 0; it < SM_UIT; ++it) {
        const int k = threadIdx.x + it * 256;
        if (k < SM_NU) us[k] = hd[it] * hv[it];
    }
    __syncthreads();
    if (!own) return;
    float s = 0.0f;
    if (dp != 0.0f) {
        if (i >= 1 && i <= H - 2 && j >= 1 && j <= W - 2) {
#pragma unroll
            for (int d = 0; d < 25; ++d) s = s + kw[LN_K25 + d] * us[(ti + d / 5) * (SMJ + 4) + (tj + d % 5)];
        } else {
#pragma unroll
            for (int t = 0; t < 9; ++t) {
                const int qi = i - (t / 3 - 1), qj = j - (t % 3 - 1);   // intermediate pixel p - d_t
                if (qi < 0 || qi >= H || qj < 0 || qj >= W) continue;
#pragma unroll
                for (int t2 = 0; t2 < 9; ++t2)
                    s = s + kw[9 * t + t2] * us[(ti + 2 - (t / 3 - 1) + (t2 / 3 - 1)) * (SMJ + 4) + (tj + 2 - (t % 3 - 1) + (t2 % 3 - 1))];
            }
        }
        s = dp * s;
    }
    ob[p] = add ? o0 + s : s;
}

// ---------------------------------------------------------------- weight gradient: tap-pair Gram matrix
// G[t][t2] = sum_{b, p on the grid} (u0~_b(p + d_t) v0~_b(p + d_t2) + u1~_b(p + d_t) v1~_b(p + d_t2))
// with x~ = Dh x (zero off the grid), over the two pairings (u0, v0), (u1, v1) of one level;
// d_t = (t / 3 - 1, t % 3 - 1) in (row, column). With Gs = G + G^T and S = Q Q^T the factored
// gradient contracts exactly to
//   dIn[t][c] = (Gs W S)[t][c],   dS + dS^T = W^T Gs W,   dQ = (W^T Gs W) Q
// (corr(S V, u~) + corr(S U, v~) and sum_p U V^T with U = In u~, V = In v~, summed over the
// pairings; U and V are linear in the 3 x 3 neighbourhoods of u~ and v~).
// A block owns TG_I x TG_J nodes of one batch entry (blockIdx.z). It stages the four scaled
// fields on the tile with a halo of 1 in LDS (each field read once), then each thread walks a
// column of TG_R nodes with the 3-row window of every field in registers (12 LDS reads per
// node for 162 FMAs). The 81 sums are reduced over the block (wave shuffles, then the four
// waves in a fixed order) into one FP64 row of partials per block.
#define TG_J 64
#define TG_R 8
#define TG_I (4 * TG_R)
struct TapGramArgs {
    int H, W;
    const float* dh;
    const float* f[4];                               // u0, v0, u1, v1
    int64_t bs[4];                                   // batch strides (floats)
    double* part;                                    // [block][81]
};

__global__ __launch_bounds__(256) void k_tapgram(TapGramArgs a) {
    constexpr int SH = TG_I + 2, SW = TG_J + 2;
    __shared__ float sf[4][SH][SW];
    __shared__ float red[4][81];
    const int tid = threadIdx.x;
    const int i0 = blockIdx.y * TG_I, j0 = blockIdx.x * TG_J, b = blockIdx.z;
    const float* fb[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) fb[q] = a.f[q] + (int64_t)b * a.bs[q];
    for (int k = tid; k < SH * SW; k += 256) {
        const int li = k / SW, lj = k - li * SW;     // constant divisor
        const int gi = i0 + li - 1, gj = j0 + lj - 1;
        float v[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        if (gi >= 0 && gi < a.H && gj >= 0 && gj < a.W) {
            const int64_t n = (int64_t)gi * a.W + gj;
            const float d = a.dh[n];
#pragma unroll
            for (int q = 0; q < 4; ++q) v[q] = fb[q][n] * d;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) sf[q][li][lj] = v[q];
    }
    __syncthreads();
    const int tx = tid & 63, ty = tid >> 6;
    const int r0 = ty * TG_R;                        // first LDS row of the window (node row r0 + 1)
    const bool colok = j0 + tx < a.W;
    float G[81];
#pragma unroll
    for (int k = 0; k < 81; ++k) G[k] = 0.0f;
    float w[4][3][3];                                // [field][window row][column tx .. tx + 2]
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int rr = 0; rr < 2; ++rr)
#pragma unroll
            for (int c = 0; c < 3; ++c) w[q][rr + 1][c] = sf[q][r0 + rr][tx + c];
#pragma unroll
    for (int r = 0; r < TG_R; ++r) {
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                w[q][0][c] = w[q][1][c];
                w[q][1][c] = w[q][2][c];
                w[q][2][c] = sf[q][r0 + r + 2][tx + c];
            }
        if (colok && i0 + r0 + r < a.H) {
#pragma unroll
            for (int t = 0; t < 9; ++t) {
                const float u0 = w[0][t / 3][t % 3], u1 = w[2][t / 3][t % 3];
#pragma unroll
                for (int t2 = 0; t2 < 9; ++t2)
                    G[9 * t + t2] = fmaf(u0, w[1][t2 / 3][t2 % 3], fmaf(u1, w[3][t2 / 3][t2 % 3], G[9 * t + t2]));
            }
        }
    }
    const int lane = tid & 63;
#pragma unroll
    for (int k = 0; k < 81; ++k) {
        float v = G[k];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        if (lane == 0) red[ty][k] = v;
    }
    __syncthreads();
    if (tid < 81) {
        const int64_t blk = ((int64_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
        a.part[blk * 81 + tid] = (((double)red[0][tid] + (double)red[1][tid]) + (double)red[2][tid]) + (double)red[3][tid];
    }
}

// G[k] = sum over blocks of part[blk][k] (one workgroup per entry, fixed order)
__global__ __launch_bounds__(256) void k_tapgram_sum(int nblk, const double* part, double* G) {
    __shared__ double sh[256];
    const int k = blockIdx.x, tid = threadIdx.x;
    double s = 0.0;
    for (int b = tid; b < nblk; b += 256) s += part[(int64_t)b * 81 + k];
    sh[tid] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (tid < o) sh[tid] += sh[tid + o];
        __syncthreads();
    }
    if (tid == 0) G[k] = sh[0];
}

// One learned level's gradient from its Gram matrix (one workgroup, C <= 64):
// T = Gs W (9 x C), dIn = T S, dQ = (W^T T) Q.
__global__ __launch_bounds__(256) void k_tapgrad(int C, const double* G, const float* Wt, const float* Q, float* dW,
                                                 float* dQ) {
    __shared__ float gs[81], w[9 * 64], T[9 * 64], S[64 * 64], M[64 * 64];
    const int tid = threadIdx.x;
    const int CC = C * C;
    if (tid < 81) gs[tid] = (float)(G[tid] + G[(tid % 9) * 9 + tid / 9]);
    for (int k = tid; k < 9 * C; k += 256) w[k] = Wt[k];
    for (int k = tid; k < CC; k += 256) {
        const int i = k / C, j = k - i * C;
        float acc = 0.0f;
        for (int m = 0; m < C; ++m) acc = fmaf(Q[i * C + m], Q[j * C + m], acc);
        S[k] = acc;
    }
    __syncthreads();
    for (int k = tid; k < 9 * C; k += 256) {
        const int t = k / C, c = k - t * C;
        float acc = 0.0f;
        for (int t2 = 0; t2 < 9; ++t2) acc = fmaf(gs[9 * t + t2], w[t2 * C + c], acc);
        T[k] = acc;
    }
    __syncthreads();
    for (int k = tid; k < 9 * C; k += 256) {
        const int t = k / C, c = k - t * C;
        float acc = 0.0f;
        for (int m = 0; m < C; ++m) acc = fmaf(T[t * C + m], S[m * C + c], acc);
        dW[k] = acc;
    }
    for (int k = tid; k < CC; k += 256) {
        const int i = k / C, j = k - i * C;
        float acc = 0.0f;
        for (int t = 0; t < 9; ++t) acc = fmaf(w[t * C + i], T[t * C + j], acc);
        M[k] = acc;
    }
    __syncthreads();
    for (int k = tid; k < CC; k += 256) {
        const int i = k / C, j = k - i * C;
        float acc = 0.0f;
        for (int m = 0; m < C; ++m) acc = fmaf(M[i * C + m], Q[m * C + j], acc);
        dQ[k] = acc;
    }
}

__global__ void k_adam(int64_t n, float* p, const float* gr, float* m, float* v, float lr, float b1, float b2,
                       float c1, float c2, float eps, float gscale) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    float g = gr[k] * gscale;
    float mk = b1 * m[k] + (1.0f - b1) * g;
    float vk = b2 * v[k] + (1.0f - b2) * g * g;
    m[k] = mk; v[k] = vk;
    p[k] = p[k] - lr * (mk / c1) / (sqrtf(vk / c2) + eps);
}

// Collapsed smoother: In^T S In through the channels is, tap pair by tap pair,
//   k81[t][t2] = sum_{c, c2} W[t][c] S[c][c2] W[t2][c2]      (FP32 from the master weights)
// applied as y(p) = sum_t [p - d_t on the grid] sum_t2 k81[t][t2] u(p - d_t + d_t2), which is
// exactly the factored form (its intermediate channels exist on the grid only). Where every
// p - d_t is on the grid (all but the outer ring of nodes) this is the 5 x 5 stencil
//   k25[delta] = sum_{d_t2 - d_t = delta} k81[t][t2]           (summed in FP64, rounded once)
// One block per level writes both: kw[0..80] = k81, kw[LN_K25 .. LN_K25+24] = k25.
__global__ __launch_bounds__(128) void k_pack_k81(int C, const float* W, const float* Q, float* kw) {
    __shared__ float S[64 * 64], WS[9 * 64], k81s[81];
    const int d = threadIdx.x;
    for (int k = d; k < C * C; k += 128) {           // S = Q Q^T
        const int i = k / C, j = k - i * C;
        float acc = 0.0f;
        for (int m = 0; m < C; ++m) acc += Q[i * C + m] * Q[j * C + m];
        S[k] = acc;
    }
    __syncthreads();
    for (int k = d; k < 9 * C; k += 128) {           // W S (9 x C)
        const int t = k / C, c = k - t * C;
        float acc = 0.0f;
        for (int m = 0; m < C; ++m) acc += W[t * C + m] * S[m * C + c];
        WS[k] = acc;
    }
    __syncthreads();
    if (d < 81) {
        const int t = d / 9, t2 = d - t * 9;
        float acc = 0.0f;
        for (int c = 0; c < C; ++c) acc += WS[t * C + c] * W[t2 * C + c];
        kw[d] = acc;
        k81s[d] = acc;
    }
    __syncthreads();
    if (d < 25) {
        const int di = d / 5 - 2, dj = d % 5 - 2;
        double acc = 0.0;
        for (int t = 0; t < 9; ++t)
            for (int t2 = 0; t2 < 9; ++t2)
                if (t2 / 3 - t / 3 == di && t2 % 3 - t % 3 == dj) acc += (double)k81s[9 * t + t2];
        kw[LN_K25 + d] = (float)acc;
    }
}

// The scaled operator A' (unit diagonal on unknown rows, eight scaled off-diagonals, Dia8)
// on unpadded vectors of length n reads out-of-range neighbours as zero.

// loss and its gradient: res = y - e, g = 2 m . A' res / ea, loss partial res.A'res / ea
__global__ __launch_bounds__(256) void k_loss_grad(int n, int batch, Dia8 A, const double* sinv, const float* y,
                                                   const float* e, const double* ea, float* g,
                                                   double* lpart) {
    __shared__ double sh[4];
    double acc = 0.0;
    for (int b = 0; b < batch; ++b) {
        int i = blockIdx.x * blockDim.x + threadIdx.x;
        const float* yb = y + (int64_t)b * n;
        const float* eb = e + (int64_t)b * n;
        double inv = 1.0 / ea[b];
        if (i < n) {
            float ri = yb[i] - eb[i];
            float ar = ri;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                int j = i + A.off[k];
                float rj = (j >= 0 && j < n) ? yb[j] - eb[j] : 0.0f;
                ar = ar + A.c[k][i] * rj;
            }
            float gi = sinv[i] > 0.0 ? (float)(2.0 * inv) * ar : 0.0f;
            g[(int64_t)b * n + i] = gi;
            acc += (double)ri * (double)ar * inv;
        }
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) lpart[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}
__global__ void k_sum_partials(int nb, const double* part, double* out) {
    __shared__ double sh[256];
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += 256) s += part[b];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < (unsigned)off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0];
}

__global__ __launch_bounds__(256) void k_dot_f(int n, const float* a, const float* b, double* part) {
    __shared__ double sh[4];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double acc = i < n ? (double)a[i] * (double)b[i] : 0.0;
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

// ---------------------------------------------------------------- synthetic pairs
// Rough training errors complement the harvested (mostly smooth) ones so that every
// frequency is in the loss: e = mask * uniform(-sqrt 3, sqrt 3) from the counter-based stream
// (seed; node, step, tag), then `s` damped-Jacobi sweeps e <- e - A'e/2 shift it to smoother modes.
__global__ void k_rough_e(int n, uint64_t seed, uint32_t step, uint32_t tag, const double* sinv, float* e) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    pic_stream rs;
    pic_stream_init(&rs, seed, (uint32_t)i, step, tag);
    double u = pic_stream_u01(&rs);
    e[i] = sinv[i] > 0.0 ? (float)((2.0 * u - 1.0) * 1.7320508075688772) : 0.0f;
}
__device__ inline float apply_scaled(int n, const Dia8& A, const float* v, int i) {
    float s = v[i];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        int j = i + A.off[k];
        if (j >= 0 && j < n) s = s + A.c[k][i] * v[j];
    }
    return s;
}
__global__ void k_smooth_e(int n, Dia8 A, const float* e, float* e_out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) e_out[i] = e[i] - 0.5f * apply_scaled(n, A, e, i);
}
// r = A' e, partial e.r
__global__ __launch_bounds__(256) void k_pair_from_e(int n, Dia8 A, const float* e, float* r, double* part) {
    __shared__ double sh[4];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double acc = 0.0;
    if (i < n) {
        float ri = apply_scaled(n, A, e, i);
        r[i] = ri;
        acc = (double)e[i] * (double)ri;
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}
__global__ void k_gather_ea(int nb, const int* idx, const double* ds_ea, double* eab) {
    int b = threadIdx.x;
    if (b < nb) eab[b] = ds_ea[idx[b]];
}

// ---------------------------------------------------------------- harvest
// e = x_final - x_snap, r = A' e (padded node vectors in, unpadded dataset rows out), partial e.r
__global__ __launch_bounds__(256) void k_harvest(int n, Dia8 A, const float* xf, const float* xs, float* e_out,
                                                 float* r_out, double* part) {
    __shared__ double sh[4];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double acc = 0.0;
    if (i < n) {
        float ei = xf[i] - xs[i];
        float ri = ei;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            int j = i + A.off[k];
            ri = ri + A.c[k][i] * (xf[j] - xs[j]);
        }
        e_out[i] = ei; r_out[i] = ri;
        acc = (double)ei * (double)ri;
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

// ---------------------------------------------------------------- FCG(1) on the scaled system
// (scalar slots, block reductions and the merged convergence test are defined with the V-cycle)

// r = D^-1/2 b64, x = 0; partial r.r (region 0). The pass is skipped (DONE at once) when the
// FP64 residual before it meets the outer tolerance, rr <= otol2 * bb (outer_rrbb = rr, bb),
// or when an earlier pass of this solve failed; first = 1 starts a solve (clears FAIL).
__global__ __launch_bounds__(FCG_BLOCK) void k_fcg_init(int n, const double* b64, const double* sinv, float* r, float* x,
                                                        double* scal, double* part, int pstride, double tol2,
                                                        double maxit, double otol2, const double* outer_rrbb,
                                                        int first, int slot) {
    int i = blockIdx.x * FCG_BLOCK + threadIdx.x;
    double v[1] = {0.0};
    if (i < n) {
        float ri = (float)(sinv[i] * b64[i]);
        r[i] = ri; x[i] = 0.0f;
        v[0] = (double)ri * (double)ri;
    }
    fcg_block_partials(1, v, part, pstride);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const double orr = outer_rrbb[0], obb = outer_rrbb[1];
        const bool failed = !first && scal[F_S_FAIL] != 0.0;
        const bool skip = failed || obb == 0.0 || orr <= otol2 * obb;
        if (first) scal[F_S_FAIL] = 0.0;
        scal[F_S_DONE] = skip ? 1.0 : 0.0; scal[F_S_ITER] = 0.0; scal[F_S_IERR] = 0.0;
        // the pass aims at the reduction the outer residual still needs (half of it in norm, the
        // inner residual being in the scaled system), at least a halving and no tighter than
        // what the FP32 iteration reaches reliably; tol2 when there is no outer residual yet
        double t2 = tol2;
        if (orr > 0.0) t2 = fmin(0.25, fmax(FCG_INNER_FLOOR2, 0.25 * otol2 * obb / orr));
        scal[F_S_TOL2] = t2; scal[F_S_MAXIT] = maxit;
        scal[F_S_NORMR] = 0.0; scal[F_S_NORMRHS] = 0.0;
        scal[FCG_REC_BASE + FCG_RECW * slot + 4] = skip ? 1.0 : 0.0;
    }
}

// end of a pass: xn += D^-1/2 x (zero when skipped); block 0 writes the pass record. A pass
// that used its whole budget is still recorded as done when the residual of its last update
// (region-0 partials) meets the inner tolerance.
__global__ __launch_bounds__(FCG_BLOCK) void k_fcg_finish(int n, int nb, const double* sinv, const float* x, double* xn,
                                                          double* scal, const double* part, int pstride, int slot) {
    int i = blockIdx.x * FCG_BLOCK + threadIdx.x;
    if (i < n) xn[i] = xn[i] + sinv[i] * (double)x[i];
    if (blockIdx.x != 0) return;
    double rr;
    fcg_reduce(1, part, pstride, nb, &rr);
    if (threadIdx.x == 0) {
        double* rec = scal + FCG_REC_BASE + FCG_RECW * slot;
        const double ierr = scal[F_S_IERR];
        const double rr0 = scal[F_S_NORMRHS];
        double done = scal[F_S_DONE];
        double ratio = rr0 > 0.0 ? scal[F_S_NORMR] / rr0 : 0.0;
        if (done == 0.0 && rr0 > 0.0) {
            ratio = rr / rr0;
            if (ratio <= scal[F_S_TOL2]) done = 1.0;
        }
        rec[0] = scal[F_S_ITER]; rec[1] = done; rec[2] = ierr; rec[3] = ratio;
        if (ierr == 2.0 || ierr == 3.0) scal[F_S_FAIL] = 1.0;
    }
}

// w = A' z; partials (z, r), (z, w), (z, q_prev) in regions 1, 2, 3
__global__ __launch_bounds__(FCG_BLOCK) void k_fcg_aw(int n, Dia8 A, const float* z, const float* r, const float* q,
                                                      float* w, const double* scal, double* part, int pstride) {
    // every global load first (the pass-done flag with them), so the block waits for memory once
    const int i = blockIdx.x * FCG_BLOCK + threadIdx.x;
    const int ii = i < n ? i : 0;
    const float zi = z[ii], ri = r[ii], qi = q[ii];
    float cz[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) cz[k] = A.c[k][ii] * z[ii + A.off[k]];
    if (scal[F_S_DONE] != 0.0) return;
    double v[3] = {0.0, 0.0, 0.0};
    if (i < n) {
        float wi = zi;
#pragma unroll
        for (int k = 0; k < 8; ++k) wi = wi + cz[k];
        w[i] = wi;
        v[0] = (double)zi * (double)ri;
        v[1] = (double)zi * (double)wi;
        v[2] = (double)zi * (double)qi;
    }
    fcg_block_partials(3, v, part + (size_t)pstride, pstride);
}

// beta = -(z, q_prev)/(p_prev, q_prev); p = z + beta p; q = w + beta q;
// pq = (z,w) + 2 beta (z,q_prev) + beta^2 pq_prev; alpha = (z, r)/pq; x += alpha p; r -= alpha q;
// partial r.r (region 0)
__global__ __launch_bounds__(FCG_BLOCK) void k_fcg_update(int n, int nb, const float* z, const float* w, float* p,
                                                          float* q, float* x, float* r, double* scal, double* part,
                                                          int pstride, int kk) {
    __shared__ double s_done;
    // the vectors are loaded before the partials are reduced, so both memory round trips overlap
    const int i0 = blockIdx.x * FCG_BLOCK + threadIdx.x;
    const int ic = i0 < n ? i0 : 0;
    const float z0 = z[ic], w0 = w[ic], p0 = p[ic], q0 = q[ic], x0 = x[ic], r0 = r[ic];
    if (threadIdx.x == 0) s_done = scal[F_S_DONE];
    double d[3];
    fcg_reduce(3, part + (size_t)pstride, pstride, nb, d);     // contains a barrier after s_done
    if (s_done != 0.0) return;
    const double k = (double)kk;
    const int par = kk & 1;
    double zr = d[0], zw = d[1], zq = d[2];
    double beta = 0.0, pq;
    if (k == 0.0) {
        pq = zw;
    } else {
        double pq_prev = scal[par ? F_S_PQ0 : F_S_PQ1];
        beta = -zq / pq_prev;
        pq = zw + 2.0 * beta * zq + beta * beta * pq_prev;
    }
    double alpha = zr / pq;
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    if (!(pq > 0.0) || !(alpha == alpha)) {
        if (lead) { scal[F_S_DONE] = 1.0; scal[F_S_IERR] = 3.0; scal[F_S_ITER] = k; }
        return;
    }
    const float fb = (float)beta, fa = (float)alpha;
    const int i = i0;
    double v[1] = {0.0};
    if (i < n) {
        float pi = (k == 0.0) ? z0 : z0 + fb * p0;
        float qi = (k == 0.0) ? w0 : w0 + fb * q0;
        float xi = x0 + fa * pi;
        float ri = r0 - fa * qi;
        p[i] = pi; q[i] = qi; x[i] = xi; r[i] = ri;
        v[0] = (double)ri * (double)ri;
    }
    fcg_block_partials(1, v, part, pstride);
    if (lead) { scal[par ? F_S_PQ1 : F_S_PQ0] = pq; scal[F_S_ITER] = k + 1.0; }
}

// ================================================================ host side
struct MgLevel {
    int H = 0, W = 0; int64_t n = 0;
    float* dia = nullptr;        // [9][n]
    float* dinv = nullptr;       // [n]
    float* dh = nullptr;         // [n]  diag^-1/2
    // per run (0: input r, 1: loss gradient g) vectors of bcap samples
    float *r[2] = {nullptr, nullptr}, *x[2] = {nullptr, nullptr}, *y[2] = {nullptr, nullptr}, *s2[2] = {nullptr, nullptr};
};

struct LearnNet {
    bool ready = false;
    int L = 0, Ls = 0;                               // levels, learned smoother levels
    MgLevel lv[LN_MAX_LEVELS];
    int C[LN_MAX_LEARNED] = {0};
    float* d_inv = nullptr; int ncoarse = 0;         // dense inverse of the coarsest operator [n][n]
    void* d_ops = nullptr;                           // one allocation: dia, dinv, dh of all levels
    // parameters (FP32, flat): per learned level W_in [9][C] then Q [C][C]
    int64_t off_in[LN_MAX_LEARNED] = {0}, off_q[LN_MAX_LEARNED] = {0}, nparams = 0;
    float *d_p = nullptr, *d_m = nullptr, *d_v = nullptr, *d_g = nullptr;
    double* d_G = nullptr;                           // tap-pair Gram matrices [Ls][81]
    float* d_k81 = nullptr;                          // collapsed smoother weights [Ls][LN_KW] (k81, k25)
    int adam_t = 0;
    // activations for a batch of bcap samples
    int bcap = 0;
    void* d_act = nullptr;
    float *y0 = nullptr, *gl = nullptr, *rb = nullptr, *eb = nullptr;   // [bcap][N]
    double* d_gpart = nullptr;                       // Gram partials [blocks][81] (in d_act)
    double* d_lpart = nullptr; double* d_lsum = nullptr;
    int* d_idx = nullptr;
    double* d_eab = nullptr;                         // e^T A' e of the batch entries
    uint32_t rough_step = 0;                         // counter of synthetic pairs drawn
    // data set of harvested pairs (ring)
    float *ds_r = nullptr, *ds_e = nullptr; double* ds_ea = nullptr;
    int ds_cap = 0, ds_n = 0, ds_head = 0;
    bool harvest = false;
    int harvest_every = 1;                           // snapshot every k-th CG chunk
    float* d_snap = nullptr; int snap_cap = 0, nsnap = 0;
    size_t snap_len = 0, snap_pad = 0;
    // per ring slot: 1 when the pair holds the initial residual of a solve (first refinement
    // pass, chunk 0: e is the whole correction of the solve); per snapshot its chunk
    std::vector<int> ds_initial;
    int snap_chunk[64] = {0};
    int solve_pass = 0;                              // refinement pass of the running inner CG
    double* d_hpart = nullptr;
    // FCG work vectors (padded node space)
    void* d_fvec = nullptr; float *z = nullptr, *w = nullptr, *q = nullptr;
    // statistics; PIC_FCG_TRACE=1 prints each learned solve's residual history
    int last_iters = 0; long long fallbacks = 0;
    double last_loss = 0.0;
    bool trace = false;
    double harvest_share = 0.25;                     // share of harvested pairs in a training minibatch
    double inner_tol = 1e-3;                         // FCG tolerance per refinement pass (the FP64 outer
                                                     // residual decides convergence)
};
static LearnNet ln;

static int64_t ln_npix0() { return ln.lv[0].n; }

static void free_learn() {
    void* ptrs[] = {ln.d_k81, ln.d_inv, ln.d_ops, ln.d_p, ln.d_m, ln.d_v, ln.d_g, ln.d_G, ln.d_act,
                    ln.d_lpart, ln.d_lsum, ln.d_idx, ln.d_eab, ln.ds_r, ln.ds_e, ln.ds_ea, ln.d_snap, ln.d_hpart, ln.d_fvec};
    for (void* p : ptrs) if (p) (void)hipFree(p);
    ln = LearnNet();
}

static MgArgs mg_args(int l, int run, const float* r, int64_t r_bs, float* y, int64_t y_bs, const double* done) {
    const MgLevel& f = ln.lv[l];
    const MgLevel& c = ln.lv[l + 1];
    MgArgs a;
    std::memset(&a, 0, sizeof(a));
    a.Hf = f.H; a.Wf = f.W; a.Hc = c.H; a.Wc = c.W;
    a.dia = f.dia; a.dinv = f.dinv; a.r = r; a.x = f.x[run];
    a.rc = c.r[run]; a.ec = c.y[run]; a.y = y; a.s2 = f.s2[run];
    a.fs = f.n; a.cs = c.n; a.r_bs = r_bs; a.y_bs = y_bs;
    a.omega = LN_OMEGA; a.jacobi = l >= ln.Ls; a.done = done;
    return a;
}

static void smooth5(int l, int batch, const float* v, int64_t v_bs, float* out, int64_t o_bs, bool add, const double* done,
                    const FcgCheck& chk) {
    const MgLevel& lv = ln.lv[l];
    dim3 grid((unsigned)((lv.W + SMJ - 1) / SMJ), (unsigned)((lv.H + SMI - 1) / SMI), batch);
    hipLaunchKernelGGL(k_smooth_k81, grid, dim3(256), 0, g.stream, lv.H, lv.W, ln.d_k81 + LN_KW * l, lv.dh, v, v_bs, out,
                       o_bs, add ? 1 : 0, done, chk);
}

// y = M r for `batch` inputs (level-0 vectors with batch stride bs); run selects the work set.
// chk (FCG solve only): the convergence test, run by the first kernel of the V-cycle.
static int net_apply(int batch, const float* r0, float* y0, int64_t bs, int run, const double* done,
                     const FcgCheck* chk = nullptr) {
    const int L = ln.L;
    const FcgCheck none = {nullptr, 0, 0, 0, nullptr};
    for (int l = 0; l < L - 1; ++l) {
        const FcgCheck& c0 = (l == 0 && chk) ? *chk : none;
        const float* r = l == 0 ? r0 : ln.lv[l].r[run];
        int64_t r_bs = l == 0 ? bs : ln.lv[l].n;
        MgArgs a = mg_args(l, run, r, r_bs, nullptr, 0, done);
        a.chk = c0;
        const MgLevel& c = ln.lv[l + 1];
        dim3 grid((unsigned)((c.W + MGD_CJ - 1) / MGD_CJ), (unsigned)((c.H + MGD_CI - 1) / MGD_CI), batch);
        // the fused smoother + down pays where its grid fills the GPU; on smaller levels the
        // smoother's own launch spreads the 25-tap work over more blocks
        const bool fuse = l < ln.Ls && (int64_t)grid.x * grid.y * batch >= 256;
        if (l < ln.Ls && !fuse) {
            smooth5(l, batch, r, r_bs, ln.lv[l].x[run], ln.lv[l].n, false, done, c0);
            a.chk = none;                            // tested by the smoother's launch
        }
        if (fuse) {
            a.dh = ln.lv[l].dh; a.k81 = ln.d_k81 + LN_KW * l;
            hipLaunchKernelGGL(k_mg_smooth_down, grid, dim3(256), 0, g.stream, a);
        } else {
            hipLaunchKernelGGL(k_mg_down, grid, dim3(256), 0, g.stream, a);
        }
    }
    {
        const MgLevel& v = ln.lv[L - 1];
        const int n = (int)v.n;
        hipLaunchKernelGGL(k_mg_dense, dim3((unsigned)((n + MGX_ROWS - 1) / MGX_ROWS), batch), dim3(64 * MGX_ROWS), 0,
                           g.stream, n, ln.d_inv, v.r[run], v.n, v.y[run], done);
    }
    for (int l = L - 2; l >= 0; --l) {
        const float* r = l == 0 ? r0 : ln.lv[l].r[run];
        int64_t r_bs = l == 0 ? bs : ln.lv[l].n;
        float* y = l == 0 ? y0 : ln.lv[l].y[run];
        int64_t y_bs = l == 0 ? bs : ln.lv[l].n;
        MgArgs a = mg_args(l, run, r, r_bs, y, y_bs, done);
        const MgLevel& f = ln.lv[l];
        dim3 grid((unsigned)((f.W + MGU_FJ - 1) / MGU_FJ), (unsigned)((f.H + MGU_FI - 1) / MGU_FI), batch);
        if (l < ln.Ls) {
            // learned level: interpolation, residual and post-smoothing in one launch
            a.dh = ln.lv[l].dh; a.k81 = ln.d_k81 + LN_KW * l;
            hipLaunchKernelGGL(k_mg_up_smooth, grid, dim3(256), 0, g.stream, a);
        } else {
            hipLaunchKernelGGL(k_mg_up, grid, dim3(256), 0, g.stream, a);
        }
    }
    HIPCHK(hipGetLastError());
    return PIC_OK;
}

static int pack_weights() {
    for (int l = 0; l < ln.Ls; ++l) {
        hipLaunchKernelGGL(k_pack_k81, dim3(1), dim3(128), 0, g.stream, ln.C[l], ln.d_p + ln.off_in[l], ln.d_p + ln.off_q[l],
                           ln.d_k81 + LN_KW * l);
    }
    HIPCHK(hipGetLastError());
    return PIC_OK;
}

// blocks of k_tapgram on level l for a batch (level 0 is the largest)
static int64_t tapgram_blocks(int l, int batch) {
    const MgLevel& v = ln.lv[l];
    return (int64_t)((v.W + TG_J - 1) / TG_J) * ((v.H + TG_I - 1) / TG_I) * batch;
}

static int ensure_batch(int batch) {
    if (batch <= ln.bcap) return PIC_OK;
    if (ln.d_act) { (void)hipFree(ln.d_act); ln.d_act = nullptr; }
    auto al = [](size_t b) { return (b + 255) / 256 * 256; };
    size_t bytes = 0;
    for (int l = 0; l < ln.L; ++l) bytes += 8 * al((size_t)ln.lv[l].n * sizeof(float) * batch);
    const size_t N = (size_t)ln_npix0();
    bytes += 4 * al(N * sizeof(float) * batch);
    bytes += al((size_t)tapgram_blocks(0, batch) * 81 * sizeof(double));
    HIPCHK(hipMalloc(&ln.d_act, bytes));
    HIPCHK(hipMemset(ln.d_act, 0, bytes));
    char* p = (char*)ln.d_act;
    for (int l = 0; l < ln.L; ++l) {
        MgLevel& v = ln.lv[l];
        float** vs[8] = {&v.r[0], &v.r[1], &v.x[0], &v.x[1], &v.y[0], &v.y[1], &v.s2[0], &v.s2[1]};
        for (int k = 0; k < 8; ++k) { *vs[k] = (float*)p; p += al((size_t)v.n * sizeof(float) * batch); }
    }
    float** fs[4] = {&ln.y0, &ln.gl, &ln.rb, &ln.eb};
    for (int s = 0; s < 4; ++s) { *fs[s] = (float*)p; p += al(N * sizeof(float) * batch); }
    ln.d_gpart = (double*)p;
    ln.bcap = batch;
    return PIC_OK;
}

// Dia8 view of the scaled level-0 operator (off-diagonals of the one-kernel CG)
static Dia8 scaled_op() {
    Dia8 A;
    for (int k = 0; k < 8; ++k) { A.c[k] = g.d_cg_off[k]; A.off[k] = g.cg_off[k]; }
    return A;
}

// ---- hierarchy on the host (FP64)
static void galerkin(int Hf, int Wf, const std::vector<double>& af, int Hc, int Wc, std::vector<double>& ac) {
    const int64_t nf = (int64_t)Hf * Wf, nc = (int64_t)Hc * Wc;
    ac.assign(9 * nc, 0.0);
    auto mask = [&](int i, int j) { return af[4 * nf + (int64_t)i * Wf + j] != 0.0; };
    for (int I = 0; I < Hc; ++I)
        for (int J = 0; J < Wc; ++J) {
            const int64_t c = (int64_t)I * Wc + J;
            for (int u = -1; u <= 1; ++u)
                for (int v = -1; v <= 1; ++v) {
                    int f1i = 2 * I + u, f1j = 2 * J + v;
                    if (f1i < 0 || f1i >= Hf || f1j < 0 || f1j >= Wf || !mask(f1i, f1j)) continue;
                    double w1 = (u == 0 ? 1.0 : 0.5) * (v == 0 ? 1.0 : 0.5);
                    int64_t f1 = (int64_t)f1i * Wf + f1j;
                    for (int t = 0; t < 9; ++t) {
                        double a12 = af[t * nf + f1];
                        if (a12 == 0.0) continue;
                        int f2i = f1i + t / 3 - 1, f2j = f1j + t % 3 - 1;
                        if (f2i < 0 || f2i >= Hf || f2j < 0 || f2j >= Wf || !mask(f2i, f2j)) continue;
                        // parents of f2
                        int i0 = f2i >> 1, j0 = f2j >> 1;
                        bool oi = f2i & 1, oj = f2j & 1;
                        for (int a = 0; a < (oi ? 2 : 1); ++a)
                            for (int b = 0; b < (oj ? 2 : 1); ++b) {
                                int I2 = i0 + a, J2 = j0 + b;
                                if (I2 >= Hc || J2 >= Wc) continue;
                                double w2 = (oi ? 0.5 : 1.0) * (oj ? 0.5 : 1.0);
                                int dI = I2 - I, dJ = J2 - J;
                                ac[(int64_t)(3 * (dI + 1) + (dJ + 1)) * nc + c] += w1 * a12 * w2;
                            }
                    }
                }
        }
}

static bool dense_inverse(std::vector<double>& a, int n) {
    std::vector<double> inv(n * (size_t)n, 0.0);
    for (int i = 0; i < n; ++i) inv[(size_t)i * n + i] = 1.0;
    for (int c = 0; c < n; ++c) {
        int piv = c;
        for (int r = c + 1; r < n; ++r) if (std::fabs(a[(size_t)r * n + c]) > std::fabs(a[(size_t)piv * n + c])) piv = r;
        if (a[(size_t)piv * n + c] == 0.0) return false;
        if (piv != c)
            for (int k = 0; k < n; ++k) { std::swap(a[(size_t)c * n + k], a[(size_t)piv * n + k]); std::swap(inv[(size_t)c * n + k], inv[(size_t)piv * n + k]); }
        double d = 1.0 / a[(size_t)c * n + c];
        for (int k = 0; k < n; ++k) { a[(size_t)c * n + k] *= d; inv[(size_t)c * n + k] *= d; }
        for (int r = 0; r < n; ++r) {
            if (r == c) continue;
            double f = a[(size_t)r * n + c];
            if (f == 0.0) continue;
            for (int k = 0; k < n; ++k) { a[(size_t)r * n + k] -= f * a[(size_t)c * n + k]; inv[(size_t)r * n + k] -= f * inv[(size_t)c * n + k]; }
        }
    }
    a.swap(inv);
    return true;
}

// Initial parameters: In = identity tap on channel 0 (+ small random taps on the others),
// Q = sqrt(omega) e0 e0^T (+ small random), so the untrained smoother is omega/diag.
static void init_params(std::vector<float>& P, uint64_t seed) {
    uint64_t st = seed * 0x9E3779B97F4A7C15ull + 1;
    auto rnd = [&]() {                                // uniform in [-1, 1)
        st ^= st << 13; st ^= st >> 7; st ^= st << 17;
        return (float)((double)(st >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0);
    };
    const float small = 0.01f;
    for (int l = 0; l < ln.Ls; ++l) {
        const int C = ln.C[l];
        for (int t = 0; t < 9; ++t)
            for (int co = 0; co < C; ++co)
                P[ln.off_in[l] + (int64_t)t * C + co] = (co == 0) ? (t == 4 ? 1.0f : 0.0f) : small * rnd();
        for (int i = 0; i < C; ++i)
            for (int j = 0; j < C; ++j)
                P[ln.off_q[l] + i * C + j] = (i == 0 && j == 0) ? std::sqrt(LN_OMEGA) : small * rnd();
    }
}

int learn_setup(int max_levels, const int32_t* channels, int nch, uint64_t seed) {
    if (!g.have_domain) return pic_fail(PIC_ERR_STATE, "pic_setup_domain not called");
    if (max_levels < 2 || max_levels > LN_MAX_LEVELS) return pic_fail(PIC_ERR_ARG, "levels must be 2..12");
    if (nch < 0 || nch > LN_MAX_LEARNED) return pic_fail(PIC_ERR_ARG, "0..4 learned smoother levels");
    free_learn();
    g.lp_passes = 2;                                 // launch budget of the first learned solve
    for (int k = 0; k < FCG_MAX_PASS; ++k) g.lp_iters[k] = 8;
    g.lp_spare = 1; g.lp_calm = 0;
    // level sizes: coarse node I <-> fine node 2I, down to about 3 x 3
    int H[LN_MAX_LEVELS], W[LN_MAX_LEVELS];
    H[0] = g.mesh.nnx; W[0] = g.mesh.nny;
    // coarsen until the coarsest level has at most LN_DENSE_NODES nodes and every learned
    // level has a coarser one (at most max_levels levels, down to about 3 x 3 nodes)
    int L = 1;
    const int min_levels = std::max(2, nch + 1);
    while (L < max_levels && std::min(H[L - 1], W[L - 1]) > 3 &&
           ((int64_t)H[L - 1] * W[L - 1] > LN_DENSE_NODES || L < min_levels)) {
        H[L] = (H[L - 1] + 1) / 2; W[L] = (W[L - 1] + 1) / 2; L++;
    }
    if (L < 2) return pic_fail(PIC_ERR_UNSUPPORTED, "mesh too small for a multigrid hierarchy");
    if ((int64_t)H[L - 1] * W[L - 1] > LN_MAX_COARSE) return pic_fail(PIC_ERR_UNSUPPORTED, "coarsest level too large");
    ln.L = L;
    ln.Ls = std::min(nch, L - 1);
    for (int l = 0; l < ln.Ls; ++l) {
        int c = channels[l];
        if (c < 1 || c > 64) return pic_fail(PIC_ERR_ARG, "channels must be 1..64");
        ln.C[l] = c;
    }
    // level-0 operator: A' = D^-1/2 A D^-1/2 on unknowns, nine node-space diagonals
    const int64_t N = (int64_t)H[0] * W[0];
    std::vector<std::vector<double>> ops(L);
    ops[0].assign(9 * N, 0.0);
    {
        const int nu = (int)g.h_A_diag.size();
        for (int u = 0; u < nu; ++u) {
            int fu = g.h_unknown_node[u] - 1;
            double su = 1.0 / std::sqrt(g.h_A_diag[u]);
            for (int k = g.h_A_rowptr[u]; k < g.h_A_rowptr[u + 1]; ++k) {
                int v = g.h_A_col[k];
                int fv = g.h_unknown_node[v] - 1;
                int di = fv / W[0] - fu / W[0], dj = fv % W[0] - fu % W[0];
                if (di < -1 || di > 1 || dj < -1 || dj > 1) return pic_fail(PIC_ERR_UNSUPPORTED, "stiffness off the 9-point stencil");
                ops[0][(int64_t)(3 * (di + 1) + (dj + 1)) * N + fu] = g.h_A_val[k] * su / std::sqrt(g.h_A_diag[v]);
            }
        }
    }
    for (int l = 1; l < L; ++l) galerkin(H[l - 1], W[l - 1], ops[l - 1], H[l], W[l], ops[l]);
    // coarsest dense inverse over its unknowns
    const int nc = H[L - 1] * W[L - 1];
    std::vector<float> invf((size_t)nc * nc, 0.0f);
    {
        const std::vector<double>& a = ops[L - 1];
        std::vector<int> idx;
        for (int i = 0; i < nc; ++i) if (a[4 * (int64_t)nc + i] != 0.0) idx.push_back(i);
        const int m = (int)idx.size();
        if (m > 0) {
            std::vector<double> d((size_t)m * m, 0.0);
            for (int r = 0; r < m; ++r) {
                int i = idx[r], ii = i / W[L - 1], jj = i % W[L - 1];
                for (int t = 0; t < 9; ++t) {
                    int ni = ii + t / 3 - 1, nj = jj + t % 3 - 1;
                    if (ni < 0 || ni >= H[L - 1] || nj < 0 || nj >= W[L - 1]) continue;
                    int j = ni * W[L - 1] + nj;
                    auto it = std::lower_bound(idx.begin(), idx.end(), j);
                    if (it == idx.end() || *it != j) continue;
                    d[(size_t)r * m + (it - idx.begin())] += a[(int64_t)t * nc + i];
                }
            }
            if (!dense_inverse(d, m)) return pic_fail(PIC_ERR_STATE, "singular coarsest operator");
            for (int r = 0; r < m; ++r)
                for (int c = 0; c < m; ++c) invf[(size_t)idx[r] * nc + idx[c]] = (float)d[(size_t)r * m + c];
        }
    }
    // upload operators: [dia 9n | dinv n | dh n] per level
    size_t tot = 0;
    for (int l = 0; l < L; ++l) tot += 11 * (size_t)H[l] * W[l];
    HIPCHK(hipMalloc(&ln.d_ops, tot * sizeof(float)));
    {
        std::vector<float> host(tot);
        float* hp = host.data();
        float* dp = (float*)ln.d_ops;
        for (int l = 0; l < L; ++l) {
            MgLevel& v = ln.lv[l];
            v.H = H[l]; v.W = W[l]; v.n = (int64_t)H[l] * W[l];
            for (int64_t k = 0; k < 9 * v.n; ++k) hp[k] = (float)ops[l][k];
            for (int64_t i = 0; i < v.n; ++i) {
                double d = ops[l][4 * v.n + i];
                hp[9 * v.n + i] = d != 0.0 ? (float)(1.0 / d) : 0.0f;
                hp[10 * v.n + i] = d > 0.0 ? (float)(1.0 / std::sqrt(d)) : 0.0f;
            }
            v.dia = dp; v.dinv = dp + 9 * v.n; v.dh = dp + 10 * v.n;
            hp += 11 * v.n; dp += 11 * v.n;
        }
        HIPCHK(hipMemcpy(ln.d_ops, host.data(), tot * sizeof(float), hipMemcpyHostToDevice));
    }
    HIPCHK(hipMalloc(&ln.d_inv, invf.size() * sizeof(float)));
    HIPCHK(hipMemcpy(ln.d_inv, invf.data(), invf.size() * sizeof(float), hipMemcpyHostToDevice));
    ln.ncoarse = nc;
    // parameters
    int64_t o = 0;
    for (int l = 0; l < ln.Ls; ++l) { ln.off_in[l] = o; o += 9LL * ln.C[l]; }
    for (int l = 0; l < ln.Ls; ++l) { ln.off_q[l] = o; o += (int64_t)ln.C[l] * ln.C[l]; }
    ln.nparams = o;
    std::vector<float> P(std::max<int64_t>(o, 1), 0.0f);
    init_params(P, seed);
    const int64_t no = std::max<int64_t>(o, 1);
    HIPCHK(hipMalloc(&ln.d_p, no * sizeof(float))); HIPCHK(hipMalloc(&ln.d_m, no * sizeof(float)));
    HIPCHK(hipMalloc(&ln.d_v, no * sizeof(float))); HIPCHK(hipMalloc(&ln.d_g, no * sizeof(float)));
    HIPCHK(hipMalloc(&ln.d_G, 81 * LN_MAX_LEARNED * sizeof(double)));
    HIPCHK(hipMemcpy(ln.d_p, P.data(), no * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(hipMemset(ln.d_m, 0, no * sizeof(float))); HIPCHK(hipMemset(ln.d_v, 0, no * sizeof(float)));
    HIPCHK(hipMemset(ln.d_g, 0, no * sizeof(float)));
    HIPCHK(hipMalloc(&ln.d_k81, LN_KW * LN_MAX_LEARNED * sizeof(float)));
    HIPCHK(hipMalloc(&ln.d_lpart, ((N + 255) / 256 + 64) * sizeof(double)));
    HIPCHK(hipMalloc(&ln.d_lsum, 64 * sizeof(double)));
    HIPCHK(hipMalloc(&ln.d_idx, 1024 * sizeof(int)));
    HIPCHK(hipMalloc(&ln.d_eab, 1024 * sizeof(double)));
    HIPCHK(hipMalloc(&ln.d_hpart, ((N + 255) / 256 + 64) * sizeof(double)));
    // FCG vectors with the padding of the CG vectors
    size_t pad = ((size_t)g.mesh.nny + 2 + 63) / 64 * 64, len = (size_t)N + 2 * pad;
    HIPCHK(hipMalloc(&ln.d_fvec, 3 * len * sizeof(float)));
    HIPCHK(hipMemset(ln.d_fvec, 0, 3 * len * sizeof(float)));
    ln.z = (float*)ln.d_fvec + pad; ln.w = ln.z + len; ln.q = ln.w + len;
    ln.snap_pad = pad; ln.snap_len = len;
    int rc = pack_weights(); if (rc) return rc;
    rc = ensure_batch(1); if (rc) return rc;
    HIPCHK(hipStreamSynchronize(g.stream));
    const char* tr = std::getenv("PIC_FCG_TRACE");
    ln.trace = tr && tr[0] == '1';
    if (const char* e = std::getenv("PIC_LEARN_INNER_TOL")) { double v = std::atof(e); if (v > 0 && v < 1) ln.inner_tol = v; }
    ln.ready = true;
    return PIC_OK;
}

// --------------------------------------------------------------- data set
int learn_set_harvest(int enable, int capacity, int every) {
    if (!ln.ready) return pic_fail(PIC_ERR_STATE, "pic_learn_setup not called");
    if (capacity == 0) capacity = ln.ds_cap > 0 ? ln.ds_cap : 64;      // keep the current ring
    if (capacity < 1 || every < 1) return pic_fail(PIC_ERR_ARG, "capacity and every must be positive");
    const size_t N = (size_t)ln_npix0();
    if (capacity != ln.ds_cap) {
        if (ln.ds_r) (void)hipFree(ln.ds_r);
        if (ln.ds_e) (void)hipFree(ln.ds_e);
        if (ln.ds_ea) (void)hipFree(ln.ds_ea);
        ln.ds_r = ln.ds_e = nullptr; ln.ds_ea = nullptr;
        HIPCHK(hipMalloc(&ln.ds_r, N * capacity * sizeof(float)));
        HIPCHK(hipMalloc(&ln.ds_e, N * capacity * sizeof(float)));
        HIPCHK(hipMalloc(&ln.ds_ea, capacity * sizeof(double)));
        ln.ds_cap = capacity; ln.ds_n = 0; ln.ds_head = 0;
        ln.ds_initial.assign(capacity, 0);
    }
    if (!ln.d_snap) {
        ln.snap_cap = 64;
        HIPCHK(hipMalloc(&ln.d_snap, ln.snap_len * ln.snap_cap * sizeof(float)));
        HIPCHK(hipMemset(ln.d_snap, 0, ln.snap_len * ln.snap_cap * sizeof(float)));
    }
    ln.harvest = enable != 0;
    ln.harvest_every = every;
    return PIC_OK;
}

bool learn_harvesting() { return ln.ready && ln.harvest; }

// snapshot of the inner CG iterate (padded node vector) at a chunk boundary
int learn_snapshot(const float* x, int chunk) {
    if (!learn_harvesting()) return PIC_OK;
    if (chunk == 0) ln.nsnap = 0;
    if (chunk % ln.harvest_every != 0 || ln.nsnap >= ln.snap_cap) return PIC_OK;
    float* dst = ln.d_snap + ln.nsnap * ln.snap_len;
    HIPCHK(hipMemcpyAsync(dst, x - ln.snap_pad, ln.snap_len * sizeof(float), hipMemcpyDeviceToDevice, g.stream));
    ln.snap_chunk[ln.nsnap] = chunk;
    ln.nsnap++;
    return PIC_OK;
}

// pairs (r = A' e, e = x_final - x_snap) of the finished inner solve into the ring
int learn_harvest_finish(const float* xfinal) {
    if (!learn_harvesting()) return PIC_OK;
    const int n = (int)ln_npix0();
    const int nb = (n + 255) / 256;
    Dia8 A = scaled_op();
    for (int s = 0; s < ln.nsnap; ++s) {
        const float* xs = ln.d_snap + s * ln.snap_len + ln.snap_pad;
        int slot = ln.ds_head;
        float* e_out = ln.ds_e + (size_t)slot * n;
        float* r_out = ln.ds_r + (size_t)slot * n;
        hipLaunchKernelGGL(k_harvest, dim3(nb), dim3(256), 0, g.stream, n, A, xfinal, xs, e_out, r_out, ln.d_hpart);
        hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, g.stream, nb, ln.d_hpart, ln.ds_ea + slot);
        HIPCHK(hipGetLastError());
        double ea = 0.0;
        HIPCHK(hipMemcpyAsync(&ea, ln.ds_ea + slot, sizeof(double), hipMemcpyDeviceToHost, g.stream));
        HIPCHK(hipStreamSynchronize(g.stream));
        if (!(ea > 0.0)) continue;                      // converged snapshot: no information
        ln.ds_initial[slot] = (ln.solve_pass == 0 && ln.snap_chunk[s] == 0) ? 1 : 0;
        ln.ds_head = (ln.ds_head + 1) % ln.ds_cap;
        ln.ds_n = std::min(ln.ds_n + 1, ln.ds_cap);
    }
    ln.nsnap = 0;
    return PIC_OK;
}

// --------------------------------------------------------------- training
// Loss and gradient for `batch` pairs already in ln.rb / ln.eb (indices in d_idx), gradient
// summed over the batch into ln.d_g. Returns the summed loss in *loss (nullptr: no read-back,
// no host synchronisation).
static int loss_and_gradient(int batch, double* loss) {
    const int n = (int)ln_npix0();
    const int64_t N = n;
    int rc = net_apply(batch, ln.rb, ln.y0, N, 0, nullptr); if (rc) return rc;
    const int nb = (n + 255) / 256;
    hipLaunchKernelGGL(k_loss_grad, dim3(nb), dim3(256), 0, g.stream, n, batch, scaled_op(), g.d_sinv, ln.y0, ln.eb,
                       ln.d_eab, ln.gl, ln.d_lpart);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, g.stream, nb, ln.d_lpart, ln.d_lsum);
    HIPCHK(hipGetLastError());
    // the same V-cycle on g (its level-0 output is not needed; it lands in ln.y0's twin below)
    rc = net_apply(batch, ln.gl, ln.lv[0].y[1], N, 1, nullptr); if (rc) return rc;
    for (int l = 0; l < ln.Ls; ++l) {
        const MgLevel& v = ln.lv[l];
        const float* r_r = l == 0 ? ln.rb : v.r[0];
        const float* r_g = l == 0 ? ln.gl : v.r[1];
        const int64_t rbs = l == 0 ? N : v.n;
        // pairings (u, v): (s2(g), r(r)) and (r(g), s2(r))
        TapGramArgs ta;
        ta.H = v.H; ta.W = v.W; ta.dh = v.dh; ta.part = ln.d_gpart;
        ta.f[0] = v.s2[1]; ta.bs[0] = v.n; ta.f[1] = r_r; ta.bs[1] = rbs;
        ta.f[2] = r_g; ta.bs[2] = rbs; ta.f[3] = v.s2[0]; ta.bs[3] = v.n;
        dim3 grid((unsigned)((v.W + TG_J - 1) / TG_J), (unsigned)((v.H + TG_I - 1) / TG_I), batch);
        hipLaunchKernelGGL(k_tapgram, grid, dim3(256), 0, g.stream, ta);
        hipLaunchKernelGGL(k_tapgram_sum, dim3(81), dim3(256), 0, g.stream, (int)tapgram_blocks(l, batch), ln.d_gpart,
                           ln.d_G + 81 * l);
        hipLaunchKernelGGL(k_tapgrad, dim3(1), dim3(256), 0, g.stream, ln.C[l], ln.d_G + 81 * l, ln.d_p + ln.off_in[l],
                           ln.d_p + ln.off_q[l], ln.d_g + ln.off_in[l], ln.d_g + ln.off_q[l]);
    }
    HIPCHK(hipGetLastError());
    if (loss) {
        HIPCHK(hipMemcpyAsync(loss, ln.d_lsum, sizeof(double), hipMemcpyDeviceToHost, g.stream));
        HIPCHK(hipStreamSynchronize(g.stream));
    }
    return PIC_OK;
}

// batch entries [0, nh) from the ring (indices idx), [nh, batch) synthetic rough pairs
static int gather_batch(int batch, int nh, const int* idx) {
    const size_t N = (size_t)ln_npix0();
    const int n = (int)N, nbk = (n + 255) / 256;
    for (int b = 0; b < nh; ++b) {
        HIPCHK(hipMemcpyAsync(ln.rb + b * N, ln.ds_r + (size_t)idx[b] * N, N * sizeof(float), hipMemcpyDeviceToDevice, g.stream));
        HIPCHK(hipMemcpyAsync(ln.eb + b * N, ln.ds_e + (size_t)idx[b] * N, N * sizeof(float), hipMemcpyDeviceToDevice, g.stream));
    }
    if (nh > 0) {
        HIPCHK(hipMemcpyAsync(ln.d_idx, idx, nh * sizeof(int), hipMemcpyHostToDevice, g.stream));
        hipLaunchKernelGGL(k_gather_ea, dim3(1), dim3(64), 0, g.stream, nh, ln.d_idx, ln.ds_ea, ln.d_eab);
    }
    Dia8 A = scaled_op();
    for (int b = nh; b < batch; ++b) {
        const uint32_t step = ln.rough_step++;
        float* e = ln.eb + b * N;
        float* tmp = ln.rb + b * N;
        hipLaunchKernelGGL(k_rough_e, dim3(nbk), dim3(256), 0, g.stream, n, (uint64_t)0x5EEDull, step, 0x10000u,
                           g.d_sinv, e);
        const int sweeps = (int)(step % 9);            // 0..8 smoothing sweeps: a spread of frequencies
        for (int s = 0; s < sweeps; ++s) {
            hipLaunchKernelGGL(k_smooth_e, dim3(nbk), dim3(256), 0, g.stream, n, A, e, tmp);
            HIPCHK(hipMemcpyAsync(e, tmp, N * sizeof(float), hipMemcpyDeviceToDevice, g.stream));
        }
        hipLaunchKernelGGL(k_pair_from_e, dim3(nbk), dim3(256), 0, g.stream, n, A, e, ln.rb + b * N, ln.d_hpart);
        hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, g.stream, nbk, ln.d_hpart, ln.d_eab + b);
    }
    HIPCHK(hipGetLastError());
    return PIC_OK;
}

int learn_train(int steps, int batch, double lr, uint64_t seed, double* loss_first, double* loss_last) {
    if (!ln.ready) return pic_fail(PIC_ERR_STATE, "pic_learn_setup not called");
    if (ln.nparams == 0) return pic_fail(PIC_ERR_STATE, "learned network: no learned smoother levels");
    if (batch < 1 || batch > 64 || steps < 0) return pic_fail(PIC_ERR_ARG, "batch must be 1..64");
    // harvested pairs in a minibatch (rounded, at least one for a positive share); the rest are
    // synthetic pairs of every smoothness
    const int nh = ln.harvest_share > 0.0 ? std::max(1, (int)std::lround(batch * ln.harvest_share)) : 0;
    if (nh > 0 && ln.ds_n < 1) return pic_fail(PIC_ERR_STATE, "learned network: empty data set");
    int rc = ensure_batch(batch); if (rc) return rc;
    uint64_t st = seed * 0x9E3779B97F4A7C15ull + 7;
    std::vector<int> idx(batch);
    const float b1 = 0.9f, b2 = 0.999f;
    for (int s = 0; s < steps; ++s) {
        for (int b = 0; b < nh; ++b) {
            st ^= st << 13; st ^= st >> 7; st ^= st << 17;
            idx[b] = (int)(st % (uint64_t)ln.ds_n);
        }
        rc = gather_batch(batch, nh, idx.data()); if (rc) return rc;
        // the loss is read back on the first and last step only (no host wait in between)
        const bool want = s == 0 || s == steps - 1;
        double loss = 0.0;
        rc = loss_and_gradient(batch, want ? &loss : nullptr); if (rc) return rc;
        if (want) {
            loss /= batch;
            if (s == 0 && loss_first) *loss_first = loss;
            if (loss_last) *loss_last = loss;
            ln.last_loss = loss;
        }
        ln.adam_t++;
        float c1 = 1.0f - std::pow(b1, (float)ln.adam_t), c2 = 1.0f - std::pow(b2, (float)ln.adam_t);
        hipLaunchKernelGGL(k_adam, dim3((unsigned)((ln.nparams + 255) / 256)), dim3(256), 0, g.stream, ln.nparams, ln.d_p,
                           ln.d_g, ln.d_m, ln.d_v, (float)lr, b1, b2, c1, c2, 1e-8f, 1.0f / batch);
        HIPCHK(hipGetLastError());
        rc = pack_weights(); if (rc) return rc;
    }
    HIPCHK(hipStreamSynchronize(g.stream));
    return PIC_OK;
}

// mean loss over pairs [0, npairs) of the data set (evaluation, no update)
int learn_eval(int npairs, double* mean_loss) {
    if (!ln.ready) return pic_fail(PIC_ERR_STATE, "pic_learn_setup not called");
    if (npairs < 1 || npairs > ln.ds_n) return pic_fail(PIC_ERR_ARG, "npairs out of range");
    int rc = ensure_batch(1); if (rc) return rc;
    const int n = (int)ln_npix0();
    const int nb = (n + 255) / 256;
    double tot = 0.0;
    for (int k = 0; k < npairs; ++k) {
        rc = gather_batch(1, 1, &k); if (rc) return rc;
        rc = net_apply(1, ln.rb, ln.y0, n, 0, nullptr); if (rc) return rc;
        hipLaunchKernelGGL(k_loss_grad, dim3(nb), dim3(256), 0, g.stream, n, 1, scaled_op(), g.d_sinv, ln.y0, ln.eb,
                           ln.d_eab, ln.gl, ln.d_lpart);
        hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, g.stream, nb, ln.d_lpart, ln.d_lsum);
        double l = 0.0;
        HIPCHK(hipMemcpyAsync(&l, ln.d_lsum, sizeof(double), hipMemcpyDeviceToHost, g.stream));
        HIPCHK(hipStreamSynchronize(g.stream));
        tot += l;
    }
    *mean_loss = tot / npairs;
    return PIC_OK;
}

// --------------------------------------------------------------- FCG solve
// One refinement pass of FCG(1) on the scaled correction equation A' y = D^-1/2 b64 with the
// network as preconditioner, launched without any host synchronisation: init (outer test on
// outer_rrbb against otol2), K iterations (each: convergence test, V-cycle, A z + dots,
// update; all no-ops once the pass is done), then xn += D^-1/2 y and the pass record in
// record slot `slot`. The inner tolerance is inner_tol on the residual norm.
int learn_fcg_pass(int n, const double* b64, double inner_tol, int K, double otol2, const double* outer_rrbb,
                   int first, int slot, double* xn) {
    const int nb = (n + FCG_BLOCK - 1) / FCG_BLOCK;
    if (nb > g.pcg_pstride) return pic_fail(PIC_ERR_STATE, "FCG partial buffer too small");
    if (slot < 0 || slot >= FCG_MAX_PASS) return pic_fail(PIC_ERR_ARG, "FCG pass record slot out of range");
    double* scal = g.d_scal;
    double* part = g.d_partials;
    const int ps = g.pcg_pstride;
    float* r = g.d_cg_r[0];
    float* x = g.d_cg_x;
    float* p = g.d_cg_p;
    hipLaunchKernelGGL(k_fcg_init, dim3(nb), dim3(FCG_BLOCK), 0, g.stream, n, b64, g.d_sinv, r, x, scal, part, ps,
                       inner_tol * inner_tol, (double)K + 1.0, otol2, outer_rrbb, first, slot);
    const Dia8 A = scaled_op();
    for (int c = 0; c < K; ++c) {
        const FcgCheck chk = {part, ps, nb, c, scal};     // the V-cycle's first kernel tests convergence
        int rc = net_apply(1, r, ln.z, n, 0, scal + F_S_DONE, &chk); if (rc) return rc;
        hipLaunchKernelGGL(k_fcg_aw, dim3(nb), dim3(FCG_BLOCK), 0, g.stream, n, A, ln.z, r, ln.q, ln.w, scal, part, ps);
        hipLaunchKernelGGL(k_fcg_update, dim3(nb), dim3(FCG_BLOCK), 0, g.stream, n, nb, ln.z, ln.w, p, ln.q, x, r,
                           scal, part, ps, c);
    }
    hipLaunchKernelGGL(k_fcg_finish, dim3(nb), dim3(FCG_BLOCK), 0, g.stream, n, nb, g.d_sinv, x, xn, scal, part, ps,
                       slot);
    HIPCHK(hipGetLastError());
    return PIC_OK;
}

bool learn_trace_on() { return ln.trace; }
void learn_note_iterations(int it) { ln.last_iters = it; }

bool learn_ready() { return ln.ready; }

int learn_set_mix(double harvested_share) {
    if (!ln.ready) return pic_fail(PIC_ERR_STATE, "pic_learn_setup not called");
    if (!(harvested_share >= 0.0 && harvested_share <= 1.0)) return pic_fail(PIC_ERR_ARG, "share must be in [0, 1]");
    ln.harvest_share = harvested_share;
    return PIC_OK;
}
double learn_inner_tol() { return ln.inner_tol; }

// --------------------------------------------------------------- test/diagnostic access
int learn_apply_host(const float* r, float* z, int64_t n) {
    if (!ln.ready) return pic_fail(PIC_ERR_STATE, "pic_learn_setup not called");
    if (n != ln_npix0()) return pic_fail(PIC_ERR_ARG, "n must equal the node count");
    int rc = ensure_batch(1); if (rc) return rc;
    HIPCHK(hipMemcpyAsync(ln.rb, r, n * sizeof(float), hipMemcpyHostToDevice, g.stream));
    rc = net_apply(1, ln.rb, ln.y0, n, 0, nullptr); if (rc) return rc;
    HIPCHK(hipMemcpyAsync(z, ln.y0, n * sizeof(float), hipMemcpyDeviceToHost, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    return PIC_OK;
}

int learn_params(float* dst, const float* src, int64_t n, int64_t* nparams) {
    if (!ln.ready) return pic_fail(PIC_ERR_STATE, "pic_learn_setup not called");
    if (nparams) *nparams = ln.nparams;
    if (dst) {
        if (n < ln.nparams) return pic_fail(PIC_ERR_ARG, "buffer too small for the parameters");
        HIPCHK(hipMemcpy(dst, ln.d_p, ln.nparams * sizeof(float), hipMemcpyDeviceToHost));
    }
    if (src) {
        if (n != ln.nparams) return pic_fail(PIC_ERR_ARG, "parameter count mismatch");
        // new parameters start a new optimisation: the Adam moments are cleared
        HIPCHK(hipStreamSynchronize(g.stream));
        HIPCHK(hipMemcpy(ln.d_p, src, n * sizeof(float), hipMemcpyHostToDevice));
        HIPCHK(hipMemset(ln.d_m, 0, n * sizeof(float)));
        HIPCHK(hipMemset(ln.d_v, 0, n * sizeof(float)));
        ln.adam_t = 0;
        int rc = pack_weights(); if (rc) return rc;
        HIPCHK(hipStreamSynchronize(g.stream));
    }
    return PIC_OK;
}

int learn_layout(int32_t* levels, int32_t* H, int32_t* W, int32_t* C, int64_t* offsets) {
    if (!ln.ready) return pic_fail(PIC_ERR_STATE, "pic_learn_setup not called");
    // levels[0] = L, levels[1] = learned smoother levels Ls; H, W per level; C per learned level;
    // offsets: [0..Ls) W_in, [Ls..2Ls) Q, [2Ls] total
    levels[0] = ln.L; levels[1] = ln.Ls;
    for (int l = 0; l < ln.L; ++l) { H[l] = ln.lv[l].H; W[l] = ln.lv[l].W; }
    for (int l = 0; l < ln.Ls; ++l) { C[l] = ln.C[l]; offsets[l] = ln.off_in[l]; offsets[ln.Ls + l] = ln.off_q[l]; }
    offsets[2 * ln.Ls] = ln.nparams;
    return PIC_OK;
}

int learn_set_pairs(int npairs, const float* r, const float* e) {
    if (!ln.ready) return pic_fail(PIC_ERR_STATE, "pic_learn_setup not called");
    if (npairs < 1) return pic_fail(PIC_ERR_ARG, "npairs must be positive");
    const size_t N = (size_t)ln_npix0();
    if (ln.ds_cap < npairs) { int rc = learn_set_harvest(0, npairs, 1); if (rc) return rc; }
    HIPCHK(hipMemcpy(ln.ds_r, r, N * npairs * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ln.ds_e, e, N * npairs * sizeof(float), hipMemcpyHostToDevice));
    // e^T A' e = e . r on the device (r = A' e is given)
    const int n = (int)N, nb = (n + 255) / 256;
    for (int k = 0; k < npairs; ++k) {
        hipLaunchKernelGGL(k_dot_f, dim3(nb), dim3(256), 0, g.stream, n, ln.ds_e + k * N, ln.ds_r + k * N, ln.d_hpart);
        hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, g.stream, nb, ln.d_hpart, ln.ds_ea + k);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(g.stream));
    ln.ds_n = npairs; ln.ds_head = npairs % ln.ds_cap;
    for (int k = 0; k < ln.ds_cap; ++k) ln.ds_initial[k] = k < npairs ? 1 : 0;   // given pairs: solves' initial residuals
    return PIC_OK;
}

int learn_get_pair(int k, float* r, float* e, double* ea) {
    if (!ln.ready) return pic_fail(PIC_ERR_STATE, "pic_learn_setup not called");
    if (k < 0 || k >= ln.ds_n) return pic_fail(PIC_ERR_ARG, "pair index out of range");
    const size_t N = (size_t)ln_npix0();
    if (r) HIPCHK(hipMemcpy(r, ln.ds_r + k * N, N * sizeof(float), hipMemcpyDeviceToHost));
    if (e) HIPCHK(hipMemcpy(e, ln.ds_e + k * N, N * sizeof(float), hipMemcpyDeviceToHost));
    if (ea) HIPCHK(hipMemcpy(ea, ln.ds_ea + k, sizeof(double), hipMemcpyDeviceToHost));
    return PIC_OK;
}

// loss and gradient (no update) for pairs idx[0..np) of the data set
int learn_gradient(int np, const int32_t* idx, double* loss, float* grad, int64_t n) {
    if (!ln.ready) return pic_fail(PIC_ERR_STATE, "pic_learn_setup not called");
    if (np < 1 || np > 64) return pic_fail(PIC_ERR_ARG, "1..64 pairs");
    if (ln.nparams == 0) return pic_fail(PIC_ERR_STATE, "learned network: no learned smoother levels");
    if (n < ln.nparams) return pic_fail(PIC_ERR_ARG, "gradient buffer too small");
    for (int k = 0; k < np; ++k) if (idx[k] < 0 || idx[k] >= ln.ds_n) return pic_fail(PIC_ERR_ARG, "pair index out of range");
    int rc = ensure_batch(np); if (rc) return rc;
    rc = gather_batch(np, np, idx); if (rc) return rc;
    rc = loss_and_gradient(np, loss); if (rc) return rc;
    HIPCHK(hipMemcpy(grad, ln.d_g, ln.nparams * sizeof(float), hipMemcpyDeviceToHost));
    return PIC_OK;
}

int learn_info(int32_t* dataset, int32_t* last_iters, int64_t* fallbacks, double* last_loss) {
    if (dataset) *dataset = ln.ds_n;
    if (last_iters) *last_iters = ln.last_iters;
    if (fallbacks) *fallbacks = ln.fallbacks;
    if (last_loss) *last_loss = ln.last_loss;
    return PIC_OK;
}
void learn_count_fallback() { ln.fallbacks++; }
void learn_free() { free_learn(); }

// --------------------------------------------------------------- access for the warm start
// (k_warm.hip trains on the same ring and minibatch buffers)
void learn_set_solve_pass(int pass) { ln.solve_pass = pass; }

void learn_ring_initial(std::vector<int>& idx) {
    idx.clear();
    for (int k = 0; k < ln.ds_n; ++k) if (ln.ds_initial[k]) idx.push_back(k);
}

Dia8 learn_scaled_op() { return scaled_op(); }

// pairs idx[0..batch) of the ring into the minibatch buffers
int learn_batch_pairs(int batch, const int* idx, LearnBatch* out) {
    if (!ln.ready) return pic_fail(PIC_ERR_STATE, "pic_learn_setup not called");
    for (int k = 0; k < batch; ++k) if (idx[k] < 0 || idx[k] >= ln.ds_n) return pic_fail(PIC_ERR_ARG, "pair index out of range");
    int rc = ensure_batch(batch); if (rc) return rc;
    rc = gather_batch(batch, batch, idx); if (rc) return rc;
    out->r = ln.rb; out->e = ln.eb; out->y = ln.y0; out->gy = ln.gl;
    return PIC_OK;
}

// energy-norm loss of the minibatch outputs y against e and its gradient into gy; the summed
// loss is read back when loss != nullptr
int learn_batch_loss(int batch, double* loss) {
    const int n = (int)ln_npix0();
    const int nb = (n + 255) / 256;
    hipLaunchKernelGGL(k_loss_grad, dim3(nb), dim3(256), 0, g.stream, n, batch, scaled_op(), g.d_sinv, ln.y0, ln.eb,
                       ln.d_eab, ln.gl, ln.d_lpart);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, g.stream, nb, ln.d_lpart, ln.d_lsum);
    HIPCHK(hipGetLastError());
    if (loss) {
        HIPCHK(hipMemcpyAsync(loss, ln.d_lsum, sizeof(double), hipMemcpyDeviceToHost, g.stream));
        HIPCHK(hipStreamSynchronize(g.stream));
    }
    return PIC_OK;
}

// loss of minibatch entry 0 (output y) and the loss of its A'-optimal multiple alpha y:
// with L e'Ae = y'Ay - 2 y'r + e'r, the scaled loss is 1 - (y'r)^2 / (y'Ay e'r)
int learn_batch_loss_scaled(const float* y, double* loss, double* scaled) {
    const int n = (int)ln_npix0();
    const int nb = (n + 255) / 256;
    int rc = learn_batch_loss(1, loss); if (rc) return rc;
    hipLaunchKernelGGL(k_dot_f, dim3(nb), dim3(256), 0, g.stream, n, y, ln.rb, ln.d_hpart);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, g.stream, nb, ln.d_hpart, ln.d_lsum + 1);
    HIPCHK(hipGetLastError());
    double h[2] = {0.0, 0.0};
    HIPCHK(hipMemcpyAsync(h, ln.d_lsum + 1, sizeof(double), hipMemcpyDeviceToHost, g.stream));
    HIPCHK(hipMemcpyAsync(h + 1, ln.d_eab, sizeof(double), hipMemcpyDeviceToHost, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    const double yr = h[0], ea = h[1];
    const double yay = *loss * ea + 2.0 * yr - ea;
    *scaled = (yay > 0.0 && ea > 0.0) ? 1.0 - yr * yr / (yay * ea) : *loss;
    return PIC_OK;
}

void learn_adam_step(int64_t n, float* p, const float* gr, float* m, float* v, float lr, int t, float gscale) {
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - std::pow(b1, (float)t), c2 = 1.0f - std::pow(b2, (float)t);
    hipLaunchKernelGGL(k_adam, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, g.stream, n, p, gr, m, v, lr, b1, b2, c1, c2,
                       1e-8f, gscale);
}
