// Learned preconditioner for the field solve: a convolutional encoder-decoder on the node
// grid, BF16 MFMA with FP32 accumulation, trained on the GPU with its own forward and
// backward kernels on residual/error pairs harvested from the plain solver, and applied as
// the preconditioner of a flexible CG (FCG(1)) on the Jacobi-scaled system.
//
// Network (linear, tied weights, symmetric positive semi-definite by construction):
//   E_0 = In r                      3x3 conv, 1 -> C_0 channels, node grid H_0 x W_0
//   E_l = Down_l E_{l-1}            3x3 stride-2 conv (coarse node I <-> fine node 2I)
//   D_{L-1} = S_{L-1} E_{L-1}
//   D_l = S_l E_l + Down_{l+1}^T D_{l+1}
//   M r = m . In^T D_0              m = 1 on unknown nodes, 0 elsewhere
// with S_l = Q_l Q_l^T (1x1 channel mixing), i.e. M = sum_l P_l^T S_l P_l, P_l = Down_l..Down_1 In.
// Initialised as a multilevel (BPX-type) preconditioner: channel 0 of In is the identity,
// channel 0 of every Down is full weighting, S_l = s_l e_0 e_0^T; the other channels start
// small and random. Training minimises the energy-norm error of M on pairs (r, e), r = A' e:
//   loss = sum_pairs (M r - e)^T A' (M r - e) / (e^T A' e)
// Gradients (dL/dy = g = 2 m . A'(y - e) / e^T A' e):
//   G_0 = In g, G_l = Down_l G_{l-1};  H_l = S_l G_l + Down_{l+1}^T H_{l+1}
//   dIn   = corr(H_0, r) + corr(D_0, g)
//   dDown_l = corr(H_l, E_{l-1}) + corr(D_l, G_{l-1})
//   dS_l = sum_p G_l E_l^T,  dQ_l = (dS_l + dS_l^T) Q_l
// where corr(o, x)[t][co][ci] = sum_p o[co](p) x[ci](src(p, t)).
//
// Activations are NHWC bf16 (pixel-major, channels contiguous, C a multiple of 16), one
// pixel per node (node index i*nny + j of the node numbering). MFMA 16x16x16 bf16: a wave
// computes 16 output pixels x 16 output channels per instruction, K = 16 input channels of
// one tap.
#include "pic_device.h"
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

typedef uint16_t bf16;
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

#define LN_MAX_LEVELS 10

__host__ __device__ inline bf16 f2bf(float f) {
    // round to nearest even (finite inputs)
#ifdef __HIP_DEVICE_COMPILE__
    uint32_t u = __float_as_uint(f);
#else
    uint32_t u; std::memcpy(&u, &f, 4);
#endif
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (bf16)(u >> 16);
}
__device__ inline float bf2f(bf16 h) { return __uint_as_float((uint32_t)h << 16); }

enum { CM_SAME = 0, CM_SAME_T = 1, CM_DOWN = 2, CM_UP = 3, CM_NONE = 4, CM_CENTER = 5 };

// source pixel of tap (di, dj) for output pixel (pi, pj); false if it does not exist
__device__ inline bool conv_src(int mode, int pi, int pj, int di, int dj, int Hin, int Win, int& si, int& sj) {
    switch (mode) {
    case CM_SAME: si = pi + di; sj = pj + dj; break;
    case CM_SAME_T: si = pi - di; sj = pj - dj; break;
    case CM_DOWN: si = 2 * pi + di; sj = 2 * pj + dj; break;
    case CM_UP: {
        int a = pi - di, b = pj - dj;
        if ((a & 1) || (b & 1)) return false;
        si = a >> 1; sj = b >> 1; break;
    }
    default: si = pi; sj = pj; break;
    }
    return si >= 0 && si < Hin && sj >= 0 && sj < Win;
}

// ---------------------------------------------------------------- convolution (MFMA)
struct ConvArgs {
    const bf16* in; const float* in_f32; int Cin, Hin, Win; int64_t in_bs;
    const bf16* w;                                   // [9][Cout][Cin]
    const bf16* in2; int Cin2; int64_t in2_bs; const bf16* w2;   // 1x1 term at the output pixel: [Cout][Cin2]
    bf16* out; float* out_f32; int Cout, Hout, Wout; int64_t out_bs;
    const double* mask;                              // out_f32: zero where mask[p] <= 0
    int mode;
    const double* done;                              // skip the launch when *done != 0
};

template <int NT>
__global__ __launch_bounds__(256) void k_conv(ConvArgs a) {
    if (a.done && *a.done != 0.0) return;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int b = blockIdx.y;
    const int npix = a.Hout * a.Wout;
    const int tile = (blockIdx.x * 4 + wv) * 16;
    if (tile >= npix) return;                        // no barrier in this kernel
    const int row = lane & 15, kq = lane >> 4;
    const int pa = tile + row;
    const bool va = pa < npix;
    const int pi = va ? pa / a.Wout : 0, pj = va ? pa - (pa / a.Wout) * a.Wout : 0;
    f32x4 acc[NT];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) acc[nt] = f32x4{0.f, 0.f, 0.f, 0.f};
    const bf16x4 zero4 = {0, 0, 0, 0};
    if (a.mode != CM_NONE) {
        const bf16* inb = a.in ? a.in + (int64_t)b * a.in_bs : nullptr;
        const float* inf = a.in_f32 ? a.in_f32 + (int64_t)b * a.in_bs : nullptr;
        for (int t = 0; t < 9; ++t) {
            const int di = t / 3 - 1, dj = t % 3 - 1;
            int si = 0, sj = 0;
            bool ok = va && conv_src(a.mode, pi, pj, di, dj, a.Hin, a.Win, si, sj);
            const int64_t sp = (int64_t)si * a.Win + sj;
            for (int cb = 0; cb < a.Cin; cb += 16) {
                bf16x4 av;
                if (inf) {                             // one FP32 input channel (channel 0)
                    float v = ok ? inf[sp] : 0.0f;
                    av = zero4;
                    if (kq == 0 && cb == 0) av[0] = (short)f2bf(v);
                } else {
                    av = ok ? *(const bf16x4*)(inb + sp * a.Cin + cb + 4 * kq) : zero4;
                }
                const bf16* wt = a.w + (int64_t)t * a.Cout * a.Cin + cb + 4 * kq;
#pragma unroll
                for (int nt = 0; nt < NT; ++nt) {
                    bf16x4 bv = *(const bf16x4*)(wt + (int64_t)(nt * 16 + row) * a.Cin);
                    acc[nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[nt], 0, 0, 0);
                }
            }
        }
    }
    if (a.in2) {
        const bf16* in2b = a.in2 + (int64_t)b * a.in2_bs;
        for (int cb = 0; cb < a.Cin2; cb += 16) {
            bf16x4 av = va ? *(const bf16x4*)(in2b + (int64_t)pa * a.Cin2 + cb + 4 * kq) : zero4;
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
                bf16x4 bv = *(const bf16x4*)(a.w2 + (int64_t)(nt * 16 + row) * a.Cin2 + cb + 4 * kq);
                acc[nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[nt], 0, 0, 0);
            }
        }
    }
    // lane holds output rows (pixels) tile + 4*kq + r, column (channel) nt*16 + row
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        int p = tile + 4 * kq + r;
        if (p >= npix) continue;
        if (a.out_f32) {
            if (row == 0) {
                float v = acc[0][r];
                if (a.mask && !(a.mask[p] > 0.0)) v = 0.0f;
                a.out_f32[(int64_t)b * a.out_bs + p] = v;
            }
        } else {
            bf16* o = a.out + (int64_t)b * a.out_bs + (int64_t)p * a.Cout + row;
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) o[nt * 16] = f2bf(acc[nt][r]);
        }
    }
}

// ---------------------------------------------------------------- weight-gradient correlation (MFMA)
// part[blk][t][co][ci] = sum_{b, p in block} og[b][p][co] * in[b][src(p, t)][ci]
struct CorrArgs {
    const bf16* og; int Cout, Hout, Wout; int64_t og_bs;
    const bf16* in; const float* in_f32; int Cin, Hin, Win; int64_t in_bs;
    int mode;                                        // CM_SAME, CM_DOWN or CM_CENTER (1 tap)
    int batch, ppb;                                  // batch size, pixels per block (multiple of 16)
    float* part;
};

template <int MT, int NT>
__global__ __launch_bounds__(256) void k_corr(CorrArgs a) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int row = lane & 15, kq = lane >> 4;
    const int ntaps = a.mode == CM_CENTER ? 1 : 9;
    const int npix = a.Hout * a.Wout;
    const int p0 = blockIdx.x * a.ppb, p1 = min(npix, p0 + a.ppb);
    const int64_t tsize = (int64_t)a.Cout * a.Cin;
    float* part = a.part + (int64_t)blockIdx.x * ntaps * tsize;
    for (int t = wv; t < ntaps; t += 4) {
        const int di = a.mode == CM_CENTER ? 0 : t / 3 - 1, dj = a.mode == CM_CENTER ? 0 : t % 3 - 1;
        const int smode = a.mode == CM_CENTER ? CM_SAME : a.mode;
        f32x4 acc[MT][NT];
#pragma unroll
        for (int i = 0; i < MT; ++i)
#pragma unroll
            for (int j = 0; j < NT; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
        for (int b = 0; b < a.batch; ++b) {
            const bf16* ogb = a.og + (int64_t)b * a.og_bs;
            const bf16* inb = a.in ? a.in + (int64_t)b * a.in_bs : nullptr;
            const float* inf = a.in_f32 ? a.in_f32 + (int64_t)b * a.in_bs : nullptr;
            for (int pc = p0; pc < p1; pc += 16) {
                // K index k = 4*kq + j  <->  pixel pc + k
                int64_t sp[4]; bool okp[4]; bool vp[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    int p = pc + 4 * kq + j;
                    vp[j] = p < p1;
                    int pi = vp[j] ? p / a.Wout : 0, pj = vp[j] ? p - (p / a.Wout) * a.Wout : 0;
                    int si = 0, sj = 0;
                    okp[j] = vp[j] && conv_src(smode, pi, pj, di, dj, a.Hin, a.Win, si, sj);
                    sp[j] = (int64_t)si * a.Win + sj;
                }
                bf16x4 av[MT], bv[NT];
#pragma unroll
                for (int i = 0; i < MT; ++i) {
                    int co = i * 16 + row;
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        int p = pc + 4 * kq + j;
                        av[i][j] = vp[j] ? (short)ogb[(int64_t)p * a.Cout + co] : (short)0;
                    }
                }
#pragma unroll
                for (int i = 0; i < NT; ++i) {
                    int ci = i * 16 + row;
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        short v = 0;
                        if (okp[j]) {
                            if (inf) v = (ci == 0) ? (short)f2bf(inf[sp[j]]) : (short)0;
                            else v = (short)inb[sp[j] * a.Cin + ci];
                        }
                        bv[i][j] = v;
                    }
                }
#pragma unroll
                for (int i = 0; i < MT; ++i)
#pragma unroll
                    for (int j = 0; j < NT; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av[i], bv[j], acc[i][j], 0, 0, 0);
            }
        }
        // D rows = co (4*kq + r + 16 i), columns = ci (row + 16 j)
        float* pt = part + (int64_t)t * tsize;
#pragma unroll
        for (int i = 0; i < MT; ++i)
#pragma unroll
            for (int j = 0; j < NT; ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    int co = i * 16 + 4 * kq + r, ci = j * 16 + row;
                    pt[(int64_t)co * a.Cin + ci] = acc[i][j][r];
                }
    }
}

// grad[k] = sum over regions, over blocks, of part[region + blk*size + k] (fixed order)
struct GradRegion { int64_t off; int nblk; };
__global__ void k_reduce_grad(int64_t size, const float* part, GradRegion r0, GradRegion r1, int nreg, float* grad) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= size) return;
    float s = 0.0f;
    for (int b = 0; b < r0.nblk; ++b) s += part[r0.off + (int64_t)b * size + k];
    if (nreg > 1)
        for (int b = 0; b < r1.nblk; ++b) s += part[r1.off + (int64_t)b * size + k];
    grad[k] = s;
}

// dQ = (dS + dS^T) Q   (C x C)
__global__ void k_qgrad(int C, const float* dS, const float* Q, float* dQ) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= C * C) return;
    int i = k / C, j = k - i * C;
    float s = 0.0f;
    for (int m = 0; m < C; ++m) s += (dS[i * C + m] + dS[m * C + i]) * Q[m * C + j];
    dQ[k] = s;
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

// packed bf16 weights: W [9][Co][Ci] -> Wb (same) and WbT [9][Ci][Co]
__global__ void k_pack_conv(int Co, int Ci, const float* W, bf16* Wb, bf16* WbT) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= 9 * Co * Ci) return;
    int t = k / (Co * Ci), r = k - t * Co * Ci, co = r / Ci, ci = r - co * Ci;
    bf16 h = f2bf(W[k]);
    Wb[k] = h;
    WbT[(int64_t)t * Co * Ci + (int64_t)ci * Co + co] = h;
}
// S = Q Q^T -> bf16 [C][C]
__global__ void k_pack_s(int C, const float* Q, bf16* S) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= C * C) return;
    int i = k / C, j = k - i * C;
    float s = 0.0f;
    for (int m = 0; m < C; ++m) s += Q[i * C + m] * Q[j * C + m];
    S[k] = f2bf(s);
}

// The scaled operator A' (unit diagonal on unknown rows, eight scaled off-diagonals, Dia8)
// on unpadded vectors of length n reads out-of-range neighbours as zero.

// loss and its gradient: res = y - e, g = 2 m . A' res / ea, loss partial res.A'res / ea
__global__ __launch_bounds__(256) void k_loss_grad(int n, int batch, Dia8 A, const double* sinv, const float* y,
                                                   const float* e, const double* ea, const int* idx, float* g,
                                                   double* lpart) {
    __shared__ double sh[4];
    double acc = 0.0;
    for (int b = 0; b < batch; ++b) {
        int i = blockIdx.x * blockDim.x + threadIdx.x;
        const float* yb = y + (int64_t)b * n;
        const float* eb = e + (int64_t)b * n;
        double inv = 1.0 / ea[idx[b]];
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
// scal slots (shared with the CG): NORMRHS, DONE, ITER, IERR, TOL2, MAXIT; pq in GAM0/GAM1
enum { F_S_NORMRHS = 4, F_S_NORMR = 5, F_S_DONE = 6, F_S_ITER = 7, F_S_IERR = 8, F_S_TOL2 = 9, F_S_MAXIT = 10,
       F_S_PQ0 = 12, F_S_PQ1 = 13 };
#define FCG_BLOCK 256

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

// r = D^-1/2 b64, x = 0; partial r.r (region 0)
__global__ __launch_bounds__(FCG_BLOCK) void k_fcg_init(int n, const double* b64, const double* sinv, float* r, float* x,
                                                        double* scal, double* part, int pstride) {
    int i = blockIdx.x * FCG_BLOCK + threadIdx.x;
    double v[1] = {0.0};
    if (i < n) {
        float ri = (float)(sinv[i] * b64[i]);
        r[i] = ri; x[i] = 0.0f;
        v[0] = (double)ri * (double)ri;
    }
    fcg_block_partials(1, v, part, pstride);
    if (blockIdx.x == 0 && threadIdx.x == 0) { scal[F_S_DONE] = 0.0; scal[F_S_ITER] = 0.0; scal[F_S_IERR] = 0.0; }
}

// convergence test of the current r before iteration k (one block): done if
// ||r||^2/||b||^2 <= tol^2 or k >= maxit
__global__ __launch_bounds__(FCG_BLOCK) void k_fcg_check(int nb, int k, double* scal, const double* part, int pstride) {
    if (scal[F_S_DONE] != 0.0) return;
    double rr;
    fcg_reduce(1, part, pstride, nb, &rr);
    if (threadIdx.x != 0) return;
    if (k == 0) scal[F_S_NORMRHS] = rr;
    double nb2 = (k == 0) ? rr : scal[F_S_NORMRHS];
    scal[F_S_NORMR] = rr;
    if (nb2 == 0.0 || rr / nb2 <= scal[F_S_TOL2]) { scal[F_S_DONE] = 1.0; scal[F_S_ITER] = (double)k; return; }
    if ((double)k >= scal[F_S_MAXIT] || !(rr == rr)) {
        scal[F_S_DONE] = 1.0; scal[F_S_ITER] = (double)k; scal[F_S_IERR] = rr == rr ? 1.0 : 2.0;
    }
}

// w = A' z; partials (z, r), (z, w), (z, q_prev) in regions 1, 2, 3
__global__ __launch_bounds__(FCG_BLOCK) void k_fcg_aw(int n, Dia8 A, const float* z, const float* r, const float* q,
                                                      float* w, const double* scal, double* part, int pstride) {
    if (scal[F_S_DONE] != 0.0) return;
    int i = blockIdx.x * FCG_BLOCK + threadIdx.x;
    double v[3] = {0.0, 0.0, 0.0};
    if (i < n) {
        float zi = z[i];
        float wi = zi;
#pragma unroll
        for (int k = 0; k < 8; ++k) wi = wi + A.c[k][i] * z[i + A.off[k]];
        w[i] = wi;
        v[0] = (double)zi * (double)r[i];
        v[1] = (double)zi * (double)wi;
        v[2] = (double)zi * (double)q[i];
    }
    fcg_block_partials(3, v, part + (size_t)pstride, pstride);
}

// beta = -(z, q_prev)/(p_prev, q_prev); p = z + beta p; q = w + beta q;
// pq = (z,w) + 2 beta (z,q_prev) + beta^2 pq_prev; alpha = (z, r)/pq; x += alpha p; r -= alpha q;
// partial r.r (region 0)
__global__ __launch_bounds__(FCG_BLOCK) void k_fcg_update(int n, int nb, int kk, const float* z, const float* w, float* p,
                                                          float* q, float* x, float* r, double* scal, double* part,
                                                          int pstride) {
    __shared__ double s_done;
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
    int i = blockIdx.x * FCG_BLOCK + threadIdx.x;
    double v[1] = {0.0};
    if (i < n) {
        float pi = (k == 0.0) ? z[i] : z[i] + fb * p[i];
        float qi = (k == 0.0) ? w[i] : w[i] + fb * q[i];
        float xi = x[i] + fa * pi;
        float ri = r[i] - fa * qi;
        p[i] = pi; q[i] = qi; x[i] = xi; r[i] = ri;
        v[0] = (double)ri * (double)ri;
    }
    fcg_block_partials(1, v, part, pstride);
    if (lead) { scal[par ? F_S_PQ1 : F_S_PQ0] = pq; scal[F_S_ITER] = k + 1.0; }
}

// ================================================================ host side
struct LearnNet {
    bool ready = false;
    int L = 0;                                       // levels 0..L-1
    int H[LN_MAX_LEVELS] = {0}, W[LN_MAX_LEVELS] = {0}, C[LN_MAX_LEVELS] = {0};
    int64_t npix[LN_MAX_LEVELS] = {0};
    // parameters (FP32, flat): In [9][C0][16]; Down_l [9][C_l][C_{l-1}] (l >= 1); Q_l [C_l][C_l]
    int64_t off_in = 0, off_dn[LN_MAX_LEVELS] = {0}, off_q[LN_MAX_LEVELS] = {0}, nparams = 0;
    float *d_p = nullptr, *d_m = nullptr, *d_v = nullptr, *d_g = nullptr, *d_dS = nullptr;
    bf16 *d_pk = nullptr;                            // packed bf16 weights (one allocation)
    bf16 *pin = nullptr, *pinT = nullptr, *pdn[LN_MAX_LEVELS] = {nullptr}, *pdnT[LN_MAX_LEVELS] = {nullptr},
         *pS[LN_MAX_LEVELS] = {nullptr};
    int adam_t = 0;
    // activations for a batch of bcap samples
    int bcap = 0;
    void* d_act = nullptr;
    bf16 *E[LN_MAX_LEVELS] = {nullptr}, *D[LN_MAX_LEVELS] = {nullptr}, *G[LN_MAX_LEVELS] = {nullptr},
         *Hh[LN_MAX_LEVELS] = {nullptr};
    float *y = nullptr, *gl = nullptr, *rb = nullptr, *eb = nullptr;   // [bcap][N]
    float* d_part = nullptr; int64_t part_cap = 0;
    double* d_lpart = nullptr; double* d_lsum = nullptr;
    int* d_idx = nullptr;
    // data set of harvested pairs (ring)
    float *ds_r = nullptr, *ds_e = nullptr; double* ds_ea = nullptr;
    int ds_cap = 0, ds_n = 0, ds_head = 0;
    bool harvest = false;
    int harvest_every = 1;                           // snapshot every k-th CG chunk
    float* d_snap = nullptr; int snap_cap = 0, nsnap = 0;
    size_t snap_len = 0, snap_pad = 0;
    double* d_hpart = nullptr;
    // FCG work vectors (padded node space)
    void* d_fvec = nullptr; float *z = nullptr, *w = nullptr, *q = nullptr;
    // statistics of the last learned solve
    int last_iters = 0, last_fallback = 0; long long fallbacks = 0;
    double last_loss = 0.0;
};
static LearnNet ln;

static int64_t ln_npix0() { return (int64_t)ln.H[0] * ln.W[0]; }

static void free_learn() {
    void* ptrs[] = {ln.d_p, ln.d_m, ln.d_v, ln.d_g, ln.d_dS, ln.d_pk, ln.d_act, ln.d_part, ln.d_lpart, ln.d_lsum,
                    ln.d_idx, ln.ds_r, ln.ds_e, ln.ds_ea, ln.d_snap, ln.d_hpart, ln.d_fvec};
    for (void* p : ptrs) if (p) (void)hipFree(p);
    ln = LearnNet();
}

static int launch_conv(const ConvArgs& a, int batch) {
    int npix = a.Hout * a.Wout;
    dim3 grid((npix + 63) / 64, batch);
    switch (a.Cout / 16) {
    case 1: hipLaunchKernelGGL(k_conv<1>, grid, dim3(256), 0, g.stream, a); break;
    case 2: hipLaunchKernelGGL(k_conv<2>, grid, dim3(256), 0, g.stream, a); break;
    case 4: hipLaunchKernelGGL(k_conv<4>, grid, dim3(256), 0, g.stream, a); break;
    default: return pic_fail(PIC_ERR_ARG, "learned network: channel count must be 16, 32 or 64");
    }
    HIPCHK(hipGetLastError());
    return PIC_OK;
}

// Forward of the network on `batch` inputs in_f32 (node-space, stride in_bs) into out_f32
// (stride out_bs; may be null: then only the encoder/decoder tensors E, D are produced).
static int net_forward(int batch, const float* in_f32, int64_t in_bs, float* out_f32, int64_t out_bs, bf16** E,
                       bf16** D, const double* done) {
    const int L = ln.L;
    ConvArgs a;
    // E_0 = In r
    std::memset(&a, 0, sizeof(a));
    a.in_f32 = in_f32; a.Cin = 16; a.Hin = ln.H[0]; a.Win = ln.W[0]; a.in_bs = in_bs;
    a.w = ln.pin; a.out = E[0]; a.Cout = ln.C[0]; a.Hout = ln.H[0]; a.Wout = ln.W[0];
    a.out_bs = ln.npix[0] * ln.C[0]; a.mode = CM_SAME; a.done = done;
    int rc = launch_conv(a, batch); if (rc) return rc;
    for (int l = 1; l < L; ++l) {
        std::memset(&a, 0, sizeof(a));
        a.in = E[l - 1]; a.Cin = ln.C[l - 1]; a.Hin = ln.H[l - 1]; a.Win = ln.W[l - 1]; a.in_bs = ln.npix[l - 1] * ln.C[l - 1];
        a.w = ln.pdn[l]; a.out = E[l]; a.Cout = ln.C[l]; a.Hout = ln.H[l]; a.Wout = ln.W[l];
        a.out_bs = ln.npix[l] * ln.C[l]; a.mode = CM_DOWN; a.done = done;
        rc = launch_conv(a, batch); if (rc) return rc;
    }
    for (int l = L - 1; l >= 0; --l) {
        std::memset(&a, 0, sizeof(a));
        if (l == L - 1) {
            a.mode = CM_NONE;
        } else {
            a.mode = CM_UP;
            a.in = D[l + 1]; a.Cin = ln.C[l + 1]; a.Hin = ln.H[l + 1]; a.Win = ln.W[l + 1];
            a.in_bs = ln.npix[l + 1] * ln.C[l + 1]; a.w = ln.pdnT[l + 1];
        }
        a.in2 = E[l]; a.Cin2 = ln.C[l]; a.in2_bs = ln.npix[l] * ln.C[l]; a.w2 = ln.pS[l];
        a.out = D[l]; a.Cout = ln.C[l]; a.Hout = ln.H[l]; a.Wout = ln.W[l]; a.out_bs = ln.npix[l] * ln.C[l];
        a.done = done;
        rc = launch_conv(a, batch); if (rc) return rc;
    }
    if (out_f32) {
        std::memset(&a, 0, sizeof(a));
        a.in = D[0]; a.Cin = ln.C[0]; a.Hin = ln.H[0]; a.Win = ln.W[0]; a.in_bs = ln.npix[0] * ln.C[0];
        a.w = ln.pinT; a.out_f32 = out_f32; a.Cout = 16; a.Hout = ln.H[0]; a.Wout = ln.W[0]; a.out_bs = out_bs;
        a.mask = g.d_sinv; a.mode = CM_SAME_T; a.done = done;
        rc = launch_conv(a, batch); if (rc) return rc;
    }
    return PIC_OK;
}

static int pack_weights() {
    const int L = ln.L;
    int n = 9 * ln.C[0] * 16;
    hipLaunchKernelGGL(k_pack_conv, dim3((n + 255) / 256), dim3(256), 0, g.stream, ln.C[0], 16, ln.d_p + ln.off_in,
                       ln.pin, ln.pinT);
    for (int l = 1; l < L; ++l) {
        n = 9 * ln.C[l] * ln.C[l - 1];
        hipLaunchKernelGGL(k_pack_conv, dim3((n + 255) / 256), dim3(256), 0, g.stream, ln.C[l], ln.C[l - 1],
                           ln.d_p + ln.off_dn[l], ln.pdn[l], ln.pdnT[l]);
    }
    for (int l = 0; l < L; ++l) {
        n = ln.C[l] * ln.C[l];
        hipLaunchKernelGGL(k_pack_s, dim3((n + 255) / 256), dim3(256), 0, g.stream, ln.C[l], ln.d_p + ln.off_q[l], ln.pS[l]);
    }
    HIPCHK(hipGetLastError());
    return PIC_OK;
}

static int ensure_batch(int batch) {
    if (batch <= ln.bcap) return PIC_OK;
    if (ln.d_act) { (void)hipFree(ln.d_act); ln.d_act = nullptr; }
    size_t per = 0;
    for (int l = 0; l < ln.L; ++l) per += (size_t)ln.npix[l] * ln.C[l] * sizeof(bf16);
    const size_t N = (size_t)ln_npix0();
    size_t bytes = 4 * per * batch + 4 * N * sizeof(float) * batch + 256 * 64;
    HIPCHK(hipMalloc(&ln.d_act, bytes));
    HIPCHK(hipMemset(ln.d_act, 0, bytes));
    char* p = (char*)ln.d_act;
    bf16** sets[4] = {ln.E, ln.D, ln.G, ln.Hh};
    for (int s = 0; s < 4; ++s)
        for (int l = 0; l < ln.L; ++l) {
            sets[s][l] = (bf16*)p;
            p += ((size_t)ln.npix[l] * ln.C[l] * sizeof(bf16) * batch + 255) / 256 * 256;
        }
    float** fs[4] = {&ln.y, &ln.gl, &ln.rb, &ln.eb};
    for (int s = 0; s < 4; ++s) { *fs[s] = (float*)p; p += (N * sizeof(float) * batch + 255) / 256 * 256; }
    ln.bcap = batch;
    return PIC_OK;
}

// Dia8 view of the scaled operator on node space
static Dia8 scaled_op() {
    Dia8 A;
    for (int k = 0; k < 8; ++k) { A.c[k] = g.d_cg_off[k]; A.off[k] = g.cg_off[k]; }
    return A;
}

// Initial parameters: channel 0 is a multilevel preconditioner (identity lift, full
// weighting restrictions, S_l = e_0 e_0^T), the other channels small and random.
static void init_params(std::vector<float>& P, uint64_t seed) {
    uint64_t st = seed * 0x9E3779B97F4A7C15ull + 1;
    auto rnd = [&]() {                                // uniform in [-1, 1)
        st ^= st << 13; st ^= st >> 7; st ^= st << 17;
        return (float)((double)(st >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0);
    };
    const float fw[9] = {0.25f, 0.5f, 0.25f, 0.5f, 1.0f, 0.5f, 0.25f, 0.5f, 0.25f};
    const float small = 0.01f;
    // In [9][C0][16]: only input channel 0 exists
    for (int t = 0; t < 9; ++t)
        for (int co = 0; co < ln.C[0]; ++co)
            P[ln.off_in + (t * ln.C[0] + co) * 16 + 0] = (co == 0) ? (t == 4 ? 1.0f : 0.0f) : small * rnd();
    for (int l = 1; l < ln.L; ++l)
        for (int t = 0; t < 9; ++t)
            for (int co = 0; co < ln.C[l]; ++co)
                for (int ci = 0; ci < ln.C[l - 1]; ++ci)
                    P[ln.off_dn[l] + ((int64_t)t * ln.C[l] + co) * ln.C[l - 1] + ci] =
                        (co == 0 && ci == 0) ? fw[t] : small * rnd();
    for (int l = 0; l < ln.L; ++l)
        for (int i = 0; i < ln.C[l]; ++i)
            for (int j = 0; j < ln.C[l]; ++j)
                P[ln.off_q[l] + i * ln.C[l] + j] = (i == 0 && j == 0) ? 1.0f : small * rnd();
}

int learn_setup(int max_levels, const int32_t* channels, int nch, uint64_t seed) {
    if (!g.have_domain) return pic_fail(PIC_ERR_STATE, "pic_setup_domain not called");
    if (max_levels < 1 || max_levels > LN_MAX_LEVELS) return pic_fail(PIC_ERR_ARG, "levels must be 1..10");
    free_learn();
    // level sizes: coarse node I <-> fine node 2I, down to about 3 x 3
    ln.H[0] = g.mesh.nnx; ln.W[0] = g.mesh.nny;
    int L = 1;
    while (L < max_levels && std::min(ln.H[L - 1], ln.W[L - 1]) > 3) {
        ln.H[L] = (ln.H[L - 1] + 1) / 2; ln.W[L] = (ln.W[L - 1] + 1) / 2; L++;
    }
    ln.L = L;
    for (int l = 0; l < L; ++l) {
        int c = channels && nch > 0 ? channels[std::min(l, nch - 1)] : (l < 2 ? 16 : 32);
        if (c != 16 && c != 32 && c != 64) return pic_fail(PIC_ERR_ARG, "channels must be 16, 32 or 64");
        ln.C[l] = c;
        ln.npix[l] = (int64_t)ln.H[l] * ln.W[l];
    }
    int64_t o = 0;
    ln.off_in = o; o += 9 * ln.C[0] * 16;
    for (int l = 1; l < L; ++l) { ln.off_dn[l] = o; o += 9LL * ln.C[l] * ln.C[l - 1]; }
    for (int l = 0; l < L; ++l) { ln.off_q[l] = o; o += (int64_t)ln.C[l] * ln.C[l]; }
    ln.nparams = o;
    std::vector<float> P(o, 0.0f);
    init_params(P, seed);
    HIPCHK(hipMalloc(&ln.d_p, o * sizeof(float))); HIPCHK(hipMalloc(&ln.d_m, o * sizeof(float)));
    HIPCHK(hipMalloc(&ln.d_v, o * sizeof(float))); HIPCHK(hipMalloc(&ln.d_g, o * sizeof(float)));
    HIPCHK(hipMalloc(&ln.d_dS, o * sizeof(float)));
    HIPCHK(hipMemcpy(ln.d_p, P.data(), o * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(hipMemset(ln.d_m, 0, o * sizeof(float))); HIPCHK(hipMemset(ln.d_v, 0, o * sizeof(float)));
    HIPCHK(hipMemset(ln.d_g, 0, o * sizeof(float))); HIPCHK(hipMemset(ln.d_dS, 0, o * sizeof(float)));
    // packed weights: In, In^T, Down_l, Down_l^T, S_l
    size_t nb = 2 * 9 * ln.C[0] * 16;
    for (int l = 1; l < L; ++l) nb += 2 * 9 * (size_t)ln.C[l] * ln.C[l - 1];
    for (int l = 0; l < L; ++l) nb += (size_t)ln.C[l] * ln.C[l];
    HIPCHK(hipMalloc(&ln.d_pk, nb * sizeof(bf16)));
    bf16* q = ln.d_pk;
    ln.pin = q; q += 9 * ln.C[0] * 16; ln.pinT = q; q += 9 * ln.C[0] * 16;
    for (int l = 1; l < L; ++l) {
        size_t s = 9 * (size_t)ln.C[l] * ln.C[l - 1];
        ln.pdn[l] = q; q += s; ln.pdnT[l] = q; q += s;
    }
    for (int l = 0; l < L; ++l) { ln.pS[l] = q; q += (size_t)ln.C[l] * ln.C[l]; }
    // gradient partials: two correlations per conv tensor, <= 256 blocks each
    int64_t pc = 2LL * 256 * 9 * ln.C[0] * 16;
    for (int l = 1; l < L; ++l) pc += 2LL * 256 * 9 * ln.C[l] * ln.C[l - 1];
    for (int l = 0; l < L; ++l) pc += 256LL * ln.C[l] * ln.C[l];
    ln.part_cap = pc;
    HIPCHK(hipMalloc(&ln.d_part, pc * sizeof(float)));
    int64_t N = ln_npix0();
    HIPCHK(hipMalloc(&ln.d_lpart, ((N + 255) / 256 + 64) * sizeof(double)));
    HIPCHK(hipMalloc(&ln.d_lsum, 64 * sizeof(double)));
    HIPCHK(hipMalloc(&ln.d_idx, 1024 * sizeof(int)));
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
    ln.ready = true;
    return PIC_OK;
}

// --------------------------------------------------------------- data set
int learn_set_harvest(int enable, int capacity, int every) {
    if (!ln.ready) return pic_fail(PIC_ERR_STATE, "pic_learn_setup not called");
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
        ln.ds_head = (ln.ds_head + 1) % ln.ds_cap;
        ln.ds_n = std::min(ln.ds_n + 1, ln.ds_cap);
    }
    ln.nsnap = 0;
    return PIC_OK;
}

// --------------------------------------------------------------- training
struct CorrPlan { int64_t off; int nblk; };

static int launch_corr(CorrArgs a, int64_t& part_off, CorrPlan& plan) {
    const int npix = a.Hout * a.Wout;
    int nblk = std::min(256, std::max(1, (npix + 63) / 64));
    int ppb = ((npix + nblk - 1) / nblk + 15) / 16 * 16;
    nblk = (npix + ppb - 1) / ppb;
    const int ntaps = a.mode == CM_CENTER ? 1 : 9;
    const int64_t size = (int64_t)ntaps * a.Cout * a.Cin;
    if (part_off + (int64_t)nblk * size > ln.part_cap) return pic_fail(PIC_ERR_STATE, "gradient partial buffer too small");
    a.ppb = ppb;
    a.part = ln.d_part + part_off;
    plan.off = part_off; plan.nblk = nblk;
    part_off += (int64_t)nblk * size;
    const int mt = a.Cout / 16, nt = a.Cin / 16;
    dim3 grid(nblk), blk(256);
#define CORR_CASE(M_, N_) if (mt == M_ && nt == N_) { hipLaunchKernelGGL((k_corr<M_, N_>), grid, blk, 0, g.stream, a); HIPCHK(hipGetLastError()); return PIC_OK; }
    CORR_CASE(1, 1) CORR_CASE(1, 2) CORR_CASE(1, 4) CORR_CASE(2, 1) CORR_CASE(2, 2) CORR_CASE(2, 4)
    CORR_CASE(4, 1) CORR_CASE(4, 2) CORR_CASE(4, 4)
#undef CORR_CASE
    return pic_fail(PIC_ERR_ARG, "learned network: unsupported channel pair");
}

static CorrArgs corr_args(const bf16* og, int lo, const bf16* in, const float* in_f32, int li, int Cin, int mode, int batch) {
    CorrArgs a;
    std::memset(&a, 0, sizeof(a));
    a.og = og; a.Cout = ln.C[lo]; a.Hout = ln.H[lo]; a.Wout = ln.W[lo]; a.og_bs = ln.npix[lo] * ln.C[lo];
    a.in = in; a.in_f32 = in_f32; a.Cin = Cin; a.Hin = ln.H[li]; a.Win = ln.W[li];
    a.in_bs = in_f32 ? ln.npix[li] : ln.npix[li] * Cin;
    a.mode = mode; a.batch = batch;
    return a;
}

// Loss and gradient for `batch` pairs already in ln.rb / ln.eb (indices idx into ds_ea),
// gradient written to ln.d_g (sum over the batch). Returns the summed loss in *loss.
static int loss_and_gradient(int batch, double* loss) {
    const int n = (int)ln_npix0();
    const int L = ln.L;
    const int64_t N = n;
    int rc = net_forward(batch, ln.rb, N, ln.y, N, ln.E, ln.D, nullptr); if (rc) return rc;
    const int nb = (n + 255) / 256;
    hipLaunchKernelGGL(k_loss_grad, dim3(nb), dim3(256), 0, g.stream, n, batch, scaled_op(), g.d_sinv, ln.y, ln.eb,
                       ln.ds_ea, ln.d_idx, ln.gl, ln.d_lpart);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, g.stream, nb, ln.d_lpart, ln.d_lsum);
    HIPCHK(hipGetLastError());
    // backward: G = encoder(g), H = decoder(G)
    rc = net_forward(batch, ln.gl, N, nullptr, 0, ln.G, ln.Hh, nullptr); if (rc) return rc;
    // correlations
    int64_t po = 0;
    CorrPlan pin1, pin2, pdn1[LN_MAX_LEVELS], pdn2[LN_MAX_LEVELS], pq[LN_MAX_LEVELS];
    rc = launch_corr(corr_args(ln.Hh[0], 0, nullptr, ln.rb, 0, 16, CM_SAME, batch), po, pin1); if (rc) return rc;
    rc = launch_corr(corr_args(ln.D[0], 0, nullptr, ln.gl, 0, 16, CM_SAME, batch), po, pin2); if (rc) return rc;
    for (int l = 1; l < L; ++l) {
        rc = launch_corr(corr_args(ln.Hh[l], l, ln.E[l - 1], nullptr, l - 1, ln.C[l - 1], CM_DOWN, batch), po, pdn1[l]);
        if (rc) return rc;
        rc = launch_corr(corr_args(ln.D[l], l, ln.G[l - 1], nullptr, l - 1, ln.C[l - 1], CM_DOWN, batch), po, pdn2[l]);
        if (rc) return rc;
    }
    for (int l = 0; l < L; ++l) {
        rc = launch_corr(corr_args(ln.G[l], l, ln.E[l], nullptr, l, ln.C[l], CM_CENTER, batch), po, pq[l]);
        if (rc) return rc;
    }
    // reductions (fixed order)
    auto reduce = [&](int64_t size, CorrPlan a, CorrPlan b, int nreg, float* dst) {
        GradRegion r0{a.off, a.nblk}, r1{b.off, b.nblk};
        hipLaunchKernelGGL(k_reduce_grad, dim3((unsigned)((size + 255) / 256)), dim3(256), 0, g.stream, size, ln.d_part,
                           r0, r1, nreg, dst);
    };
    reduce(9LL * ln.C[0] * 16, pin1, pin2, 2, ln.d_g + ln.off_in);
    for (int l = 1; l < L; ++l) reduce(9LL * ln.C[l] * ln.C[l - 1], pdn1[l], pdn2[l], 2, ln.d_g + ln.off_dn[l]);
    for (int l = 0; l < L; ++l) {
        int C = ln.C[l];
        reduce((int64_t)C * C, pq[l], pq[l], 1, ln.d_dS + ln.off_q[l]);
        hipLaunchKernelGGL(k_qgrad, dim3((C * C + 255) / 256), dim3(256), 0, g.stream, C, ln.d_dS + ln.off_q[l],
                           ln.d_p + ln.off_q[l], ln.d_g + ln.off_q[l]);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(loss, ln.d_lsum, sizeof(double), hipMemcpyDeviceToHost, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    return PIC_OK;
}

static int gather_batch(int batch, const int* idx) {
    const size_t N = (size_t)ln_npix0();
    for (int b = 0; b < batch; ++b) {
        HIPCHK(hipMemcpyAsync(ln.rb + b * N, ln.ds_r + (size_t)idx[b] * N, N * sizeof(float), hipMemcpyDeviceToDevice, g.stream));
        HIPCHK(hipMemcpyAsync(ln.eb + b * N, ln.ds_e + (size_t)idx[b] * N, N * sizeof(float), hipMemcpyDeviceToDevice, g.stream));
    }
    HIPCHK(hipMemcpyAsync(ln.d_idx, idx, batch * sizeof(int), hipMemcpyHostToDevice, g.stream));
    return PIC_OK;
}

int learn_train(int steps, int batch, double lr, uint64_t seed, double* loss_first, double* loss_last) {
    if (!ln.ready) return pic_fail(PIC_ERR_STATE, "pic_learn_setup not called");
    if (ln.ds_n < 1) return pic_fail(PIC_ERR_STATE, "learned network: empty data set");
    if (batch < 1 || batch > 64 || steps < 0) return pic_fail(PIC_ERR_ARG, "batch must be 1..64");
    int rc = ensure_batch(batch); if (rc) return rc;
    uint64_t st = seed * 0x9E3779B97F4A7C15ull + 7;
    std::vector<int> idx(batch);
    const float b1 = 0.9f, b2 = 0.999f;
    for (int s = 0; s < steps; ++s) {
        for (int b = 0; b < batch; ++b) {
            st ^= st << 13; st ^= st >> 7; st ^= st << 17;
            idx[b] = (int)(st % (uint64_t)ln.ds_n);
        }
        rc = gather_batch(batch, idx.data()); if (rc) return rc;
        double loss = 0.0;
        rc = loss_and_gradient(batch, &loss); if (rc) return rc;
        loss /= batch;
        if (s == 0 && loss_first) *loss_first = loss;
        if (loss_last) *loss_last = loss;
        ln.last_loss = loss;
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
        rc = gather_batch(1, &k); if (rc) return rc;
        rc = net_forward(1, ln.rb, n, ln.y, n, ln.E, ln.D, nullptr); if (rc) return rc;
        hipLaunchKernelGGL(k_loss_grad, dim3(nb), dim3(256), 0, g.stream, n, 1, scaled_op(), g.d_sinv, ln.y, ln.eb,
                           ln.ds_ea, ln.d_idx, ln.gl, ln.d_lpart);
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
// FCG(1) on the scaled correction equation A' y = D^-1/2 b64, network preconditioner.
// Returns PIC_OK with *stagnated set when the residual stops decreasing (the caller then
// continues with the plain CG).
int learn_fcg_run(int n, const double* b64, double tol, int maxit, int* iters, int* ierr, int* stagnated) {
    *stagnated = 0;
    const int nb = (n + FCG_BLOCK - 1) / FCG_BLOCK;
    if (nb > g.pcg_pstride) return pic_fail(PIC_ERR_STATE, "FCG partial buffer too small");
    double* scal = g.d_scal;
    double* part = g.d_partials;
    const int ps = g.pcg_pstride;
    double h2[2] = {tol * tol, (double)maxit};
    HIPCHK(hipMemcpyAsync(scal + F_S_TOL2, h2, 2 * sizeof(double), hipMemcpyHostToDevice, g.stream));
    float* r = g.d_cg_r[0];
    float* x = g.d_cg_x;
    float* p = g.d_cg_p;
    hipLaunchKernelGGL(k_fcg_init, dim3(nb), dim3(FCG_BLOCK), 0, g.stream, n, b64, g.d_sinv, r, x, scal, part, ps);
    HIPCHK(hipGetLastError());
    Dia8 A = scaled_op();
    const int CH = 4;
    int k = 0;
    double prev_rr = -1.0;
    int slow = 0;
    for (;;) {
        HIPCHK(hipEventRecord(g.ev[12], g.stream));
        for (int c = 0; c < CH; ++c, ++k) {
            hipLaunchKernelGGL(k_fcg_check, dim3(1), dim3(FCG_BLOCK), 0, g.stream, nb, k, scal, part, ps);
            int rc = net_forward(1, r, n, ln.z, n, ln.E, ln.D, scal + F_S_DONE); if (rc) return rc;
            hipLaunchKernelGGL(k_fcg_aw, dim3(nb), dim3(FCG_BLOCK), 0, g.stream, n, A, ln.z, r, ln.q, ln.w, scal, part, ps);
            hipLaunchKernelGGL(k_fcg_update, dim3(nb), dim3(FCG_BLOCK), 0, g.stream, n, nb, k, ln.z, ln.w, p, ln.q, x, r,
                               scal, part, ps);
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(g.ev[13], g.stream));
        HIPCHK(hipMemcpyAsync(g.h_scal_pinned, scal, 16 * sizeof(double), hipMemcpyDeviceToHost, g.stream));
        HIPCHK(hipStreamSynchronize(g.stream));
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, g.ev[12], g.ev[13]));
        g.kstat[4] += ms; g.kstat[5] += CH;
        if (g.h_scal_pinned[F_S_DONE] != 0.0) break;
        double rr = g.h_scal_pinned[F_S_NORMR];
        // stagnation: less than a 2x residual-norm reduction over a chunk, three times
        if (prev_rr > 0.0 && !(rr < 0.25 * prev_rr)) slow++;
        prev_rr = rr;
        if (slow >= 3) { *stagnated = 1; g.h_scal_pinned[F_S_ITER] = k; g.h_scal_pinned[F_S_IERR] = 0.0; break; }
        if (k > maxit + CH) { g.h_scal_pinned[F_S_IERR] = 1.0; g.h_scal_pinned[F_S_ITER] = k; break; }
    }
    *iters = (int)g.h_scal_pinned[F_S_ITER];
    *ierr = (int)g.h_scal_pinned[F_S_IERR];
    if (*ierr == 2 || *ierr == 3) *stagnated = 1;        // NaN or breakdown: hand over to the plain CG
    ln.last_iters = *iters;
    return PIC_OK;
}

bool learn_ready() { return ln.ready; }

// --------------------------------------------------------------- test/diagnostic access
int learn_apply_host(const float* r, float* z, int64_t n) {
    if (!ln.ready) return pic_fail(PIC_ERR_STATE, "pic_learn_setup not called");
    if (n != ln_npix0()) return pic_fail(PIC_ERR_ARG, "n must equal the node count");
    int rc = ensure_batch(1); if (rc) return rc;
    HIPCHK(hipMemcpyAsync(ln.rb, r, n * sizeof(float), hipMemcpyHostToDevice, g.stream));
    rc = net_forward(1, ln.rb, n, ln.y, n, ln.E, ln.D, nullptr); if (rc) return rc;
    HIPCHK(hipMemcpyAsync(z, ln.y, n * sizeof(float), hipMemcpyDeviceToHost, g.stream));
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
        HIPCHK(hipMemcpy(ln.d_p, src, n * sizeof(float), hipMemcpyHostToDevice));
        int rc = pack_weights(); if (rc) return rc;
        HIPCHK(hipStreamSynchronize(g.stream));
    }
    return PIC_OK;
}

int learn_layout(int32_t* levels, int32_t* H, int32_t* W, int32_t* C, int64_t* offsets) {
    if (!ln.ready) return pic_fail(PIC_ERR_STATE, "pic_learn_setup not called");
    *levels = ln.L;
    for (int l = 0; l < ln.L; ++l) { H[l] = ln.H[l]; W[l] = ln.W[l]; C[l] = ln.C[l]; }
    // offsets: [0] In, [1..L-1] Down_l, [L..2L-1] Q_l, [2L] total
    offsets[0] = ln.off_in;
    for (int l = 1; l < ln.L; ++l) offsets[l] = ln.off_dn[l];
    for (int l = 0; l < ln.L; ++l) offsets[ln.L + l] = ln.off_q[l];
    offsets[2 * ln.L] = ln.nparams;
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
    if (n < ln.nparams) return pic_fail(PIC_ERR_ARG, "gradient buffer too small");
    for (int k = 0; k < np; ++k) if (idx[k] < 0 || idx[k] >= ln.ds_n) return pic_fail(PIC_ERR_ARG, "pair index out of range");
    int rc = ensure_batch(np); if (rc) return rc;
    rc = gather_batch(np, idx); if (rc) return rc;
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
