// Warm start of the field solve: a convolutional encoder-decoder on the node grid that maps
// the Jacobi-scaled residual of the solve's starting potential to a correction of it. For a
// cold start (potential = the Dirichlet values) that residual is the right-hand side B rho - c
// of the deposited charge, so the network maps the charge to a first potential; inside the
// time loop it maps the change of the charge since the last solve to the change of the
// potential. The flexible CG with the learned V-cycle then starts from that guess; the plain
// CG stays the fallback and the FP64 outer residual the judge of convergence.
//
// Network (linear, no biases, so N(a r) = a N(r) for every a: the warm start does not depend
// on the units or the sign of the charge, as A'^-1 does not), on the levels of the node grid
// (level l+1 holds the nodes (I, J) <-> (2I, 2J) of level l, L levels, C channels each):
//   h_0 = m_0 . lift(r)                           3x3, 1 -> C
//   h_l = m_l . conv(R h_{l-1}; E_l)              3x3, C -> C, l = 1..L-1   (encoder)
//   g_{L-1} = h_{L-1}
//   g_l = m_l . conv(P g_{l+1} + h_l; D_l)        3x3, C -> C, l = L-2..0   (decoder, skips)
//   y   = m_0 . out(g_0)                          3x3, C -> 1
// m_l: unknown-node masks (m_{l+1}(I, J) = m_l(2I, 2J)); P: bilinear interpolation from the
// coarse nodes inside the grid; R = P^T / 4 (full weighting 1/4, 1/8, 1/16).
//
// Matrix cores: the C -> C convolutions (forward and the backward data pass, which is the
// same convolution with flipped taps and transposed channels) are implicit GEMMs on
// v_mfma_f32_16x16x32_bf16 (A = packed BF16 weights, output channels x (tap, input channel);
// B = BF16 activations staged in LDS, (tap, input channel) x 16 nodes of a row; FP32
// accumulation), and so are the weight gradients (input channels x nodes times nodes x output
// channels, 32 nodes per instruction). The restriction of the encoder and the interpolation +
// skip of the decoder are computed while staging the convolution's input tile, so the
// forward pass is 2L launches. Activations are stored BF16, node-major with the C channels
// contiguous (32 B per node at C = 16); the 1-channel lift and output convolutions run in
// FP32 on the vector units.
//
// Training: Adam on the energy-norm error of the raw output against the pair's solution,
// (y - e)^T A' (y - e) / (e^T A' e), on the (r, e) pairs of the harvest ring that hold the
// initial residual of a solve. Weight gradients go through fixed-order block partials, so a
// gradient is the same bits on every run.
#include "pic_device.h"
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#define WS_MAX_LEVELS 8
#define WS_TI 8

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 ws_bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float ws_f32x4;
typedef uint32_t ws_u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t ws_u32x2 __attribute__((ext_vector_type(2)));

enum { WS_PLAIN = 0, WS_RESTRICT = 1, WS_INTERP = 2 };

template <int C> struct WsShape {
    static constexpr int TJ = C == 16 ? 64 : 32;          // conv tile: WS_TI x TJ nodes
    static constexpr int G = C / 8;                        // 8-channel groups (16 B)
    static constexpr int S = (9 * C + 31) / 32;            // K steps of 32 (tap, channel) pairs
    static constexpr int MT = C / 16;                      // 16-channel tiles
};

__device__ inline uint16_t ws_f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
__device__ inline float ws_bf2f(uint32_t b) { return __uint_as_float(b << 16); }
__device__ inline uint32_t ws_pack2(float a, float b) { return (uint32_t)ws_f2bf(a) | ((uint32_t)ws_f2bf(b) << 16); }
__device__ inline void ws_unpack8(ws_u32x4 v, float f[8]) {
#pragma unroll
    for (int k = 0; k < 4; ++k) { f[2 * k] = __uint_as_float(v[k] << 16); f[2 * k + 1] = __uint_as_float(v[k] & 0xFFFF0000u); }
}
__device__ inline ws_u32x4 ws_pack8(const float f[8]) {
    ws_u32x4 v;
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = ws_pack2(f[2 * k], f[2 * k + 1]);
    return v;
}

// Input of a convolution on the (H, W) grid, read from a BF16 tensor:
//   PLAIN    x = src                                   (src on (H, W))
//   RESTRICT x = scale R src                            (src on the finer grid (Hs, Ws))
//   INTERP   x = scale P src + add                      (src on the coarser grid (Hs, Ws), add on (H, W))
struct WsIn {
    int H, W, Hs, Ws;
    const uint16_t* src; int64_t src_bs;
    const uint16_t* add; int64_t add_bs;
    float scale;
};

// channels [8 gi, 8 gi + 8) of the input at node (i, j) of the (H, W) grid, FP32
template <int C, int MODE>
__device__ inline void ws_input8(const WsIn& in, int b, int i, int j, int gi, float f[8]) {
    const uint16_t* src = in.src + (int64_t)b * in.src_bs + 8 * gi;
    if (MODE == WS_PLAIN) {
        ws_unpack8(*(const ws_u32x4*)(src + ((int64_t)i * in.W + j) * C), f);
        return;
    }
    float v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) f[k] = 0.0f;
    if (MODE == WS_RESTRICT) {
#pragma unroll
        for (int du = -1; du <= 1; ++du) {
            const int fi = 2 * i + du;
            if (fi < 0 || fi >= in.Hs) continue;
#pragma unroll
            for (int dv = -1; dv <= 1; ++dv) {
                const int fj = 2 * j + dv;
                if (fj < 0 || fj >= in.Ws) continue;
                const float w = (du == 0 ? 0.5f : 0.25f) * (dv == 0 ? 0.5f : 0.25f);
                ws_unpack8(*(const ws_u32x4*)(src + ((int64_t)fi * in.Ws + fj) * C), v);
#pragma unroll
                for (int k = 0; k < 8; ++k) f[k] = f[k] + w * v[k];
            }
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) f[k] = in.scale * f[k];
    } else {
        const int I0 = i >> 1, J0 = j >> 1, oi = i & 1, oj = j & 1;
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            if (a > oi || I0 + a >= in.Hs) continue;
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                if (c > oj || J0 + c >= in.Ws) continue;
                const float w = (oi ? 0.5f : 1.0f) * (oj ? 0.5f : 1.0f);
                ws_unpack8(*(const ws_u32x4*)(src + ((int64_t)(I0 + a) * in.Ws + (J0 + c)) * C), v);
#pragma unroll
                for (int k = 0; k < 8; ++k) f[k] = f[k] + w * v[k];
            }
        }
        if (in.add) {
            ws_unpack8(*(const ws_u32x4*)(in.add + (int64_t)b * in.add_bs + ((int64_t)i * in.W + j) * C + 8 * gi), v);
#pragma unroll
            for (int k = 0; k < 8; ++k) f[k] = in.scale * f[k] + v[k];
        } else {
#pragma unroll
            for (int k = 0; k < 8; ++k) f[k] = in.scale * f[k];
        }
    }
}

// Stage the input on the nodes [i0 - 1, i0 + TI] x [j0 - 1, j0 + TJ] (zero outside the grid)
// into xs[node][group] (16 B per entry).
template <int C, int MODE, int TI, int TJ>
__device__ inline void ws_stage(const WsIn& in, int b, int i0, int j0, ws_u32x4* xs) {
    constexpr int G = C / 8, HI = TI + 2, HJ = TJ + 2;
    for (int k = threadIdx.x; k < HI * HJ * G; k += 256) {
        const int node = k / G, gi = k - node * G;
        const int u = node / HJ, v = node - u * HJ;
        const int i = i0 - 1 + u, j = j0 - 1 + v;
        ws_u32x4 val = {0u, 0u, 0u, 0u};
        if (i >= 0 && i < in.H && j >= 0 && j < in.W) {
            if (MODE == WS_PLAIN) {
                val = *(const ws_u32x4*)(in.src + (int64_t)b * in.src_bs + ((int64_t)i * in.W + j) * C + 8 * gi);
            } else {
                float f[8];
                ws_input8<C, MODE>(in, b, i, j, gi, f);
                val = ws_pack8(f);
            }
        }
        xs[k] = val;
    }
}

// out = mask . conv(x; W) on the (H, W) grid, C -> C channels, batch entry blockIdx.z.
// wp: packed BF16 weights [MT][S][64 lanes] of 8 values: lane l of tile mt, step s holds
// W[tap][ci0 .. ci0 + 7][16 mt + l % 16] with 32 s + 8 (l / 16) = C tap + ci0 (zero for tap 9).
// A workgroup owns WS_CTI x WS_CTJ = 4 x 16 nodes, one 16-node row segment per wave: every
// level is a latency-bound pass (its input was written by the previous launch), so the tile
// is small enough that the input (6 x 18 nodes) is staged in one round of loads per thread
// and the grid has many workgroups even on the coarse levels.
#define WS_CTI 4
#define WS_CTJ 16
template <int C, int MODE>
__global__ __launch_bounds__(256) void k_ws_conv(WsIn in, const ws_u32x4* __restrict__ wp,
                                                 const float* __restrict__ mask, uint16_t* __restrict__ out,
                                                 int64_t out_bs) {
    using Sh = WsShape<C>;
    constexpr int TI = WS_CTI, TJ = WS_CTJ, G = Sh::G, S = Sh::S, MT = Sh::MT, HJ = TJ + 2;
    __shared__ ws_u32x4 xs[(TI + 2) * HJ * G];
    const int b = blockIdx.z;
    const int i0 = blockIdx.y * TI, j0 = blockIdx.x * TJ;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    ws_u32x4 wr[MT][S];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int s = 0; s < S; ++s) wr[mt][s] = wp[(mt * S + s) * 64 + lane];
    ws_stage<C, MODE, TI, TJ>(in, b, i0, j0, xs);
    __syncthreads();
    const int q = lane >> 4, pc = lane & 15;
    constexpr int SEGR = TJ / 16;                       // 16-node segments per tile row
    for (int seg = wid; seg < TI * SEGR; seg += 4) {
        const int ri = seg / SEGR, cj = (seg - ri * SEGR) * 16 + pc;
        ws_f32x4 acc[MT];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) acc[mt] = (ws_f32x4){0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int k0 = 32 * s + 8 * q;
            const int tap = k0 / C, ci0 = k0 - tap * C;
            ws_u32x4 bv = {0u, 0u, 0u, 0u};
            if (tap < 9) bv = xs[((ri + tap / 3) * HJ + (cj + tap % 3)) * G + ci0 / 8];
#pragma unroll
            for (int mt = 0; mt < MT; ++mt)
                acc[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(ws_bf16x8, wr[mt][s]),
                                                                  __builtin_bit_cast(ws_bf16x8, bv), acc[mt], 0, 0, 0);
        }
        const int i = i0 + ri, j = j0 + cj;
        if (i < in.H && j < in.W) {
            const int64_t p = (int64_t)i * in.W + j;
            const float m = mask ? mask[p] : 1.0f;
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
                ws_u32x2 o;
                o[0] = ws_pack2(acc[mt][0] * m, acc[mt][1] * m);
                o[1] = ws_pack2(acc[mt][2] * m, acc[mt][3] * m);
                *(ws_u32x2*)(out + (int64_t)b * out_bs + p * C + 16 * mt + 4 * q) = o;
            }
        }
    }
}

// Weight gradient dW[tap][ci][co] = sum_{b, p} x(p + d_tap)[ci] z(p)[co] over the whole batch:
// tiles (batch folded in) strided over the grid of blocks; per wave, chunks of 32 nodes of a
// tile row are the K of one MFMA (A = x at the shifted nodes, input channels x nodes; B = z,
// nodes x output channels). Each wave writes its partial sums to part[(block * 4 + wave)][9 C C];
// k_ws_reduce adds them in a fixed order.
template <int C, int MODE>
__global__ __launch_bounds__(256) void k_ws_wgrad(WsIn in, const uint16_t* __restrict__ z, int64_t z_bs, int ntj,
                                                  int nti, int batch, float* __restrict__ part) {
    using Sh = WsShape<C>;
    constexpr int TJ = Sh::TJ, G = Sh::G, MT = Sh::MT, HJ = TJ + 2;
    __shared__ ws_u32x4 xs[(WS_TI + 2) * HJ * G];
    __shared__ ws_u32x4 zs[WS_TI * TJ * G];
    const uint16_t* xh = (const uint16_t*)xs;
    const uint16_t* zh = (const uint16_t*)zs;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int q = lane >> 4, ch = lane & 15;
    ws_f32x4 acc[9][MT][MT];
#pragma unroll
    for (int t = 0; t < 9; ++t)
#pragma unroll
        for (int a = 0; a < MT; ++a)
#pragma unroll
            for (int c = 0; c < MT; ++c) acc[t][a][c] = (ws_f32x4){0.0f, 0.0f, 0.0f, 0.0f};
    const int ntiles = ntj * nti * batch;
    constexpr int CPR = TJ / 32;                        // 32-node chunks per tile row
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int b = tile / (ntj * nti), rem = tile - b * ntj * nti;
        const int i0 = (rem / ntj) * WS_TI, j0 = (rem - (rem / ntj) * ntj) * TJ;
        __syncthreads();                                 // the previous tile's reads are done
        ws_stage<C, MODE, WS_TI, TJ>(in, b, i0, j0, xs);
        for (int k = threadIdx.x; k < WS_TI * TJ * G; k += 256) {
            const int node = k / G, gi = k - node * G;
            const int u = node / TJ, v = node - u * TJ;
            const int i = i0 + u, j = j0 + v;
            ws_u32x4 val = {0u, 0u, 0u, 0u};
            if (i < in.H && j < in.W) val = *(const ws_u32x4*)(z + (int64_t)b * z_bs + ((int64_t)i * in.W + j) * C + 8 * gi);
            zs[k] = val;
        }
        __syncthreads();
        for (int chk = wid; chk < WS_TI * CPR; chk += 4) {
            const int ri = chk / CPR, cj0 = (chk - ri * CPR) * 32 + 8 * q;
            ws_bf16x8 bz[MT];
#pragma unroll
            for (int nt = 0; nt < MT; ++nt) {
                uint16_t tmp[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) tmp[j] = zh[(ri * TJ + cj0 + j) * C + 16 * nt + ch];
                bz[nt] = __builtin_bit_cast(ws_bf16x8, tmp);
            }
#pragma unroll
            for (int t = 0; t < 9; ++t) {
#pragma unroll
                for (int mt = 0; mt < MT; ++mt) {
                    uint16_t tmp[8];
#pragma unroll
                    for (int j = 0; j < 8; ++j) tmp[j] = xh[((ri + t / 3) * HJ + (cj0 + j + t % 3)) * C + 16 * mt + ch];
                    const ws_bf16x8 ax = __builtin_bit_cast(ws_bf16x8, tmp);
#pragma unroll
                    for (int nt = 0; nt < MT; ++nt)
                        acc[t][mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ax, bz[nt], acc[t][mt][nt], 0, 0, 0);
                }
            }
        }
    }
    float* pw = part + (int64_t)(blockIdx.x * 4 + wid) * 9 * C * C;
#pragma unroll
    for (int t = 0; t < 9; ++t)
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
            for (int nt = 0; nt < MT; ++nt)
#pragma unroll
                for (int r = 0; r < 4; ++r) pw[t * C * C + (16 * mt + 4 * q + r) * C + 16 * nt + ch] = acc[t][mt][nt][r];
}

// Fixed-order sums of partials: out[y][k] = sum of part[p][k] over the parts p of chunk y
// (chunk y: parts [y * chunk, min((y + 1) * chunk, nparts)), 4 part-lanes of a block summed in
// FP64 and combined in order). One pass with one chunk gives the final sums (FP32); a first
// pass over many chunks writes FP64 chunk sums for the second.
template <typename TI, typename TO>
__global__ __launch_bounds__(256) void k_ws_reduce(int nparts, int chunk, int K, const TI* __restrict__ part,
                                                   TO* __restrict__ out) {
    __shared__ double sh[4][64];
    const int kk = blockIdx.x * 64 + (threadIdx.x & 63), pl = threadIdx.x >> 6;
    const int p0 = blockIdx.y * chunk, p1 = min(nparts, p0 + chunk);
    double s = 0.0;
    if (kk < K)
        for (int p = p0 + pl; p < p1; p += 4) s += (double)part[(int64_t)p * K + kk];
    sh[pl][threadIdx.x & 63] = s;
    __syncthreads();
    if (pl == 0 && kk < K)
        out[(int64_t)blockIdx.y * K + kk] =
            (TO)(((sh[0][threadIdx.x] + sh[1][threadIdx.x]) + sh[2][threadIdx.x]) + sh[3][threadIdx.x]);
}

// 1 -> C lift: out(p)[c] = mask(p) sum_tap w[tap'][c] s(p + d_tap), tap' = flip ? 8 - tap : tap
// (flip: the adjoint of the C -> 1 output convolution). One thread per node and all C
// channels: the weight reads are the same for every lane (scalar loads).
template <int C>
__global__ __launch_bounds__(256) void k_ws_lift(int H, int W, const float* __restrict__ s, int64_t s_bs,
                                                 const float* __restrict__ w, int flip, const float* __restrict__ mask,
                                                 uint16_t* __restrict__ out, int64_t o_bs) {
    constexpr int G = C / 8;
    const int b = blockIdx.y;
    const int p = blockIdx.x * 256 + threadIdx.x;           // node index (a level has < 2^31 nodes)
    if (p >= H * W) return;
    const int i = p / W, j = p - (p / W) * W;
    const float* sb = s + (int64_t)b * s_bs;
    float sv[9];
#pragma unroll
    for (int t = 0; t < 9; ++t) {
        const int ii = i + t / 3 - 1, jj = j + t % 3 - 1;
        sv[t] = (ii < 0 || ii >= H || jj < 0 || jj >= W) ? 0.0f : sb[(int64_t)ii * W + jj];
    }
    const float m = mask[p];
#pragma unroll
    for (int gi = 0; gi < G; ++gi) {
        float f[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) f[c] = 0.0f;
#pragma unroll
        for (int t = 0; t < 9; ++t) {
            const int tt = flip ? 8 - t : t;
#pragma unroll
            for (int c = 0; c < 8; ++c) f[c] = f[c] + w[tt * C + 8 * gi + c] * sv[t];
        }
#pragma unroll
        for (int c = 0; c < 8; ++c) f[c] = f[c] * m;
        *(ws_u32x4*)(out + (int64_t)b * o_bs + (int64_t)p * C + 8 * gi) = ws_pack8(f);
    }
}

// C -> 1 output: y(p) = mask(p) sum_tap sum_c w[tap][c] v(p + d_tap)[c], FP32
template <int C>
__global__ __launch_bounds__(256) void k_ws_out(int H, int W, const uint16_t* __restrict__ v, int64_t v_bs,
                                                const float* __restrict__ w, const float* __restrict__ mask,
                                                float* __restrict__ y, int64_t y_bs) {
    constexpr int G = C / 8;
    const int b = blockIdx.y;
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= H * W) return;
    const int i = p / W, j = p - (p / W) * W;
    const uint16_t* vb = v + (int64_t)b * v_bs;
    float acc = 0.0f;
#pragma unroll
    for (int t = 0; t < 9; ++t) {
        const int ii = i + t / 3 - 1, jj = j + t % 3 - 1;
        if (ii < 0 || ii >= H || jj < 0 || jj >= W) continue;
#pragma unroll
        for (int gi = 0; gi < G; ++gi) {
            float f[8];
            ws_unpack8(*(const ws_u32x4*)(vb + ((int64_t)ii * W + jj) * C + 8 * gi), f);
#pragma unroll
            for (int c = 0; c < 8; ++c) acc = acc + w[t * C + 8 * gi + c] * f[c];
        }
    }
    y[(int64_t)b * y_bs + p] = acc * mask[p];
}

// Weight gradients of the 1-channel convolutions, summed over the batch:
//   LIFT: dW[tap][c] = sum s(p + d_tap) v(p)[c]     OUT: dW[tap][c] = sum v(p + d_tap)[c] s(p)
// A block owns WS_CORR_ROWS rows of one batch entry (blockIdx.y); threads: c = tid % C, node
// lane tid / C stepping along the row; per block 9 C partial sums (fixed order).
#define WS_CORR_ROWS 2
template <int C, bool LIFT>
__global__ __launch_bounds__(256) void k_ws_corr(int H, int W, const float* __restrict__ s, int64_t s_bs,
                                                 const uint16_t* __restrict__ v, int64_t v_bs, float* __restrict__ part) {
    __shared__ float red[256 * 9];
    constexpr int NL = 256 / C;
    const int c = threadIdx.x % C, nl = threadIdx.x / C;
    const int b = blockIdx.y;
    const int i0 = blockIdx.x * WS_CORR_ROWS, i1 = min(H, i0 + WS_CORR_ROWS);
    const float* sb = s + (int64_t)b * s_bs;
    const uint16_t* vb = v + (int64_t)b * v_bs;
    float acc[9];
#pragma unroll
    for (int t = 0; t < 9; ++t) acc[t] = 0.0f;
    for (int i = i0; i < i1; ++i) {
        for (int j = nl; j < W; j += NL) {
            const int p = i * W + j;
            if (LIFT) {
                const float vv = ws_bf2f(vb[(int64_t)p * C + c]);
#pragma unroll
                for (int t = 0; t < 9; ++t) {
                    const int ii = i + t / 3 - 1, jj = j + t % 3 - 1;
                    if (ii >= 0 && ii < H && jj >= 0 && jj < W) acc[t] = acc[t] + sb[ii * W + jj] * vv;
                }
            } else {
                const float sv = sb[p];
#pragma unroll
                for (int t = 0; t < 9; ++t) {
                    const int ii = i + t / 3 - 1, jj = j + t % 3 - 1;
                    if (ii >= 0 && ii < H && jj >= 0 && jj < W) acc[t] = acc[t] + ws_bf2f(vb[(int64_t)(ii * W + jj) * C + c]) * sv;
                }
            }
        }
    }
#pragma unroll
    for (int t = 0; t < 9; ++t) red[threadIdx.x * 9 + t] = acc[t];
    __syncthreads();
    const int64_t blk = (int64_t)blockIdx.y * gridDim.x + blockIdx.x;
    for (int k = threadIdx.x; k < 9 * C; k += 256) {
        const int t = k / C, cc = k - t * C;
        float sum = 0.0f;
        for (int l = 0; l < NL; ++l) sum += red[(l * C + cc) * 9 + t];
        part[blk * 9 * C + t * C + cc] = sum;
    }
}

// Transfer operators of the backward pass (one thread per node and 8 channels):
//   RESTRICT: out (coarse) = mask . scale R src (fine)      INTERP: out (fine) = mask . (scale P src + add)
template <int C, int MODE>
__global__ __launch_bounds__(256) void k_ws_transfer(WsIn in, const float* __restrict__ mask, uint16_t* __restrict__ out,
                                                     int64_t o_bs) {
    constexpr int G = C / 8;
    const int b = blockIdx.y;
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= in.H * in.W * G) return;
    const int p = k / G;
    const int gi = k - p * G;
    const int i = p / in.W, j = p - (p / in.W) * in.W;
    float f[8];
    ws_input8<C, MODE>(in, b, i, j, gi, f);
    const float m = mask ? mask[p] : 1.0f;
#pragma unroll
    for (int c = 0; c < 8; ++c) f[c] = f[c] * m;
    *(ws_u32x4*)(out + (int64_t)b * o_bs + (int64_t)p * C + 8 * gi) = ws_pack8(f);
}

// Packs the FP32 master weights of every C -> C convolution into the MFMA A-operand order,
// forward (dir 0) and adjoint (dir 1: W'[tap][co][ci] = W[8 - tap][ci][co]).
struct WsPackArgs { int64_t off[2 * WS_MAX_LEVELS]; int nconv; };
template <int C>
__global__ __launch_bounds__(256) void k_ws_pack(WsPackArgs a, const float* __restrict__ p, ws_u32x4* __restrict__ wp) {
    using Sh = WsShape<C>;
    constexpr int S = Sh::S, MT = Sh::MT, per = MT * S * 64;
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= a.nconv * 2 * per) return;
    const int cv = k / (2 * per), r1 = k - cv * 2 * per;
    const int dir = r1 / per, r2 = r1 - dir * per;
    const int mt = r2 / (S * 64), r3 = r2 - mt * S * 64;
    const int s = r3 / 64, lane = r3 - s * 64;
    const float* w = p + a.off[cv];
    const int m = 16 * mt + (lane & 15);
    const int k0 = 32 * s + 8 * (lane >> 4);
    const int tap = k0 / C, ci0 = k0 - tap * C;
    float f[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        float v = 0.0f;
        if (tap < 9) v = dir == 0 ? w[(tap * C + ci0 + j) * C + m] : w[((8 - tap) * C + m) * C + ci0 + j];
        f[j] = v;
    }
    wp[k] = ws_pack8(f);
}

// Line search of the warm start along y (scaled system): dots (y, r') and (y, A' y) with
// r' = D^-1/2 rn, per block (regions 0 and 1 of part, stride nb).
__global__ __launch_bounds__(256) void k_ws_dots(int n, Dia8 A, const double* __restrict__ sinv,
                                                 const double* __restrict__ rn, const float* __restrict__ y,
                                                 double* __restrict__ part, int nb) {
    __shared__ double sh[2][4];
    const int i = blockIdx.x * 256 + threadIdx.x;
    double yr = 0.0, yay = 0.0;
    if (i < n) {
        float ay = y[i];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int j = i + A.off[k];
            const float yj = (j >= 0 && j < n) ? y[j] : 0.0f;
            ay = ay + A.c[k][i] * yj;
        }
        if (sinv[i] > 0.0) {
            yr = (double)y[i] * (sinv[i] * rn[i]);
            yay = (double)y[i] * (double)ay;
        }
    }
    for (int off = 32; off > 0; off >>= 1) { yr += __shfl_xor(yr, off, 64); yay += __shfl_xor(yay, off, 64); }
    if ((threadIdx.x & 63) == 0) { sh[0][threadIdx.x >> 6] = yr; sh[1][threadIdx.x >> 6] = yay; }
    __syncthreads();
    if (threadIdx.x == 0) {
        part[blockIdx.x] = ((sh[0][0] + sh[0][1]) + sh[0][2]) + sh[0][3];
        part[nb + blockIdx.x] = ((sh[1][0] + sh[1][1]) + sh[1][2]) + sh[1][3];
    }
}

// xn += alpha D^-1/2 y with alpha = (y, r') / (y, A' y), the A'-norm-optimal multiple of y (0
// when the curvature is not positive or not finite); every block reduces the partials in the
// same order, block 0 records alpha.
__global__ __launch_bounds__(256) void k_ws_update(int n, int nb, const double* __restrict__ part,
                                                   const double* __restrict__ sinv, const float* __restrict__ y,
                                                   double* __restrict__ xn, double* __restrict__ alpha_out) {
    __shared__ double sh[2][4];
    double yr = 0.0, yay = 0.0;
    for (int k = threadIdx.x; k < nb; k += 256) { yr += part[k]; yay += part[nb + k]; }
    for (int off = 32; off > 0; off >>= 1) { yr += __shfl_xor(yr, off, 64); yay += __shfl_xor(yay, off, 64); }
    if ((threadIdx.x & 63) == 0) { sh[0][threadIdx.x >> 6] = yr; sh[1][threadIdx.x >> 6] = yay; }
    __syncthreads();
    yr = ((sh[0][0] + sh[0][1]) + sh[0][2]) + sh[0][3];
    yay = ((sh[1][0] + sh[1][1]) + sh[1][2]) + sh[1][3];
    double alpha = yay > 0.0 ? yr / yay : 0.0;
    if (!(alpha == alpha) || fabs(alpha) > 1e30) alpha = 0.0;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n && alpha != 0.0) xn[i] = xn[i] + alpha * sinv[i] * (double)y[i];
    if (blockIdx.x == 0 && threadIdx.x == 0) *alpha_out = alpha;
}

// r' = D^-1/2 rn in FP32 (the network's input in the solve)
__global__ void k_ws_scale_in(int n, const double* __restrict__ sinv, const double* __restrict__ rn, float* __restrict__ r) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) r[i] = (float)(sinv[i] * rn[i]);
}

// ================================================================ host side
struct WarmNet {
    bool ready = false, enabled = false;
    int L = 0, C = 16;
    int H[WS_MAX_LEVELS] = {0}, W[WS_MAX_LEVELS] = {0};
    int64_t n[WS_MAX_LEVELS] = {0};
    void* d_mask = nullptr; const float* mask[WS_MAX_LEVELS] = {nullptr};
    // FP32 parameters: lift [9][C], out [9][C], enc_l [9][C][C] (l = 1..L-1), dec_l [9][C][C] (l = 0..L-2)
    int64_t off_lift = 0, off_out = 0, off_enc[WS_MAX_LEVELS] = {0}, off_dec[WS_MAX_LEVELS] = {0}, nparams = 0;
    float *d_p = nullptr, *d_m = nullptr, *d_v = nullptr, *d_g = nullptr;
    ws_u32x4* d_wp = nullptr; int64_t wp_per = 0;      // packed conv weights: conv k, dir -> d_wp + (2k + dir) wp_per
    int adam_t = 0;
    int bcap = 0; void* d_act = nullptr;
    uint16_t *h[WS_MAX_LEVELS] = {nullptr}, *g[WS_MAX_LEVELS] = {nullptr};
    uint16_t *gb[WS_MAX_LEVELS] = {nullptr}, *ub[WS_MAX_LEVELS] = {nullptr}, *hb[WS_MAX_LEVELS] = {nullptr},
             *tb[WS_MAX_LEVELS] = {nullptr};
    float* d_part = nullptr; int64_t part_cap = 0;      // floats
    double* d_part2 = nullptr;                          // FP64 chunk sums of the partials
    float *d_in = nullptr, *d_y = nullptr;              // solve: r' and y (one node vector each)
    double *d_dots = nullptr, *d_alpha = nullptr;
    double last_alpha = 0.0, last_loss = 0.0;
    int64_t applied = 0;
    // the solve's application as one HIP graph (scale, network, line search), captured for the
    // buffers it was recorded with and captured again when any of them moves
    hipGraph_t graph = nullptr; hipGraphExec_t graph_exec = nullptr;
    const double* graph_rn = nullptr; double* graph_xn = nullptr; void* graph_act = nullptr; int graph_n = 0;
    int mode = 0;                                       // 0 off, 1 graph, 2 eager launches
};
static WarmNet wn;

static void ws_drop_graph() {
    if (wn.graph_exec) (void)hipGraphExecDestroy(wn.graph_exec);
    if (wn.graph) (void)hipGraphDestroy(wn.graph);
    wn.graph_exec = nullptr; wn.graph = nullptr;
}

static void ws_free() {
    ws_drop_graph();
    void* ptrs[] = {wn.d_mask, wn.d_p, wn.d_m, wn.d_v, wn.d_g, wn.d_wp, wn.d_act, wn.d_part, wn.d_part2, wn.d_in, wn.d_y,
                    wn.d_dots, wn.d_alpha};
    for (void* p : ptrs) if (p) (void)hipFree(p);
    wn = WarmNet();
}

static int conv_index_enc(int l) { return l - 1; }
static int conv_index_dec(int l) { return wn.L - 1 + l; }
static const ws_u32x4* ws_weights(int conv, int dir) { return wn.d_wp + (int64_t)(2 * conv + dir) * wn.wp_per; }

template <int C> static int ws_pack_t() {
    WsPackArgs a;
    std::memset(&a, 0, sizeof(a));
    a.nconv = 2 * (wn.L - 1);
    for (int l = 1; l < wn.L; ++l) a.off[conv_index_enc(l)] = wn.off_enc[l];
    for (int l = 0; l + 1 < wn.L; ++l) a.off[conv_index_dec(l)] = wn.off_dec[l];
    const int64_t tot = (int64_t)a.nconv * 2 * wn.wp_per;
    hipLaunchKernelGGL(k_ws_pack<C>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, g.stream, a, wn.d_p, wn.d_wp);
    HIPCHK(hipGetLastError());
    return PIC_OK;
}
static int ws_pack() { return wn.C == 16 ? ws_pack_t<16>() : ws_pack_t<32>(); }

template <int C, int MODE>
static void ws_conv_launch(const WsIn& in, int batch, const ws_u32x4* w, const float* mask, uint16_t* out, int64_t obs) {
    dim3 grid((unsigned)((in.W + WS_CTJ - 1) / WS_CTJ), (unsigned)((in.H + WS_CTI - 1) / WS_CTI), batch);
    hipLaunchKernelGGL((k_ws_conv<C, MODE>), grid, dim3(256), 0, g.stream, in, w, mask, out, obs);
}

static WsIn ws_in(int l, int ls, const uint16_t* src, int64_t sbs, const uint16_t* add, int64_t abs_, float scale) {
    WsIn in;
    in.H = wn.H[l]; in.W = wn.W[l]; in.Hs = wn.H[ls]; in.Ws = wn.W[ls];
    in.src = src; in.src_bs = sbs; in.add = add; in.add_bs = abs_; in.scale = scale;
    return in;
}

static int64_t act_bs(int l) { return wn.n[l] * wn.C; }        // batch stride of a level-l activation
static uint16_t* g_or_h(int l) { return l == wn.L - 1 ? wn.h[l] : wn.g[l]; }

// y = N(r) for `batch` inputs (FP32 node vectors, batch strides rbs / ybs); keeps h_l, g_l
template <int C>
static int ws_forward_t(int batch, const float* r, int64_t rbs, float* y, int64_t ybs) {
    const int L = wn.L;
    const int64_t n0 = wn.n[0];
    hipLaunchKernelGGL(k_ws_lift<C>, dim3((unsigned)((n0 + 255) / 256), batch), dim3(256), 0, g.stream,
                       wn.H[0], wn.W[0], r, rbs, wn.d_p + wn.off_lift, 0, wn.mask[0], wn.h[0], act_bs(0));
    for (int l = 1; l < L; ++l)
        ws_conv_launch<C, WS_RESTRICT>(ws_in(l, l - 1, wn.h[l - 1], act_bs(l - 1), nullptr, 0, 1.0f), batch,
                                       ws_weights(conv_index_enc(l), 0), wn.mask[l], wn.h[l], act_bs(l));
    for (int l = L - 2; l >= 0; --l)
        ws_conv_launch<C, WS_INTERP>(ws_in(l, l + 1, g_or_h(l + 1), act_bs(l + 1), wn.h[l], act_bs(l), 1.0f), batch,
                                     ws_weights(conv_index_dec(l), 0), wn.mask[l], wn.g[l], act_bs(l));
    hipLaunchKernelGGL(k_ws_out<C>, dim3((unsigned)((n0 + 255) / 256), batch), dim3(256), 0, g.stream, wn.H[0], wn.W[0],
                       wn.g[0], act_bs(0), wn.d_p + wn.off_out, wn.mask[0], y, ybs);
    HIPCHK(hipGetLastError());
    return PIC_OK;
}

static int ws_blocks_for(int64_t work) { return (int)std::max<int64_t>(1, std::min<int64_t>(work, 512)); }

// grad[k] = sum of the nparts partial sums part[p][k] (two fixed-order passes: chunks of 64
// parts to FP64 chunk sums, then the chunk sums)
static void ws_reduce_launch(int nparts, int K, float* grad) {
    const int chunk = 64, nch = (nparts + chunk - 1) / chunk;
    const unsigned kb = (unsigned)((K + 63) / 64);
    hipLaunchKernelGGL((k_ws_reduce<float, double>), dim3(kb, nch), dim3(256), 0, g.stream, nparts, chunk, K,
                       (const float*)wn.d_part, wn.d_part2);
    hipLaunchKernelGGL((k_ws_reduce<double, float>), dim3(kb, 1), dim3(256), 0, g.stream, nch, nch, K,
                       (const double*)wn.d_part2, grad);
}

template <int C, int MODE>
static int ws_wgrad_launch(const WsIn& in, int batch, const uint16_t* z, int64_t zbs, float* grad) {
    constexpr int TJ = WsShape<C>::TJ;
    const int ntj = (in.W + TJ - 1) / TJ, nti = (in.H + WS_TI - 1) / WS_TI;
    const int nblk = ws_blocks_for((int64_t)ntj * nti * batch);
    const int K = 9 * C * C;
    if ((int64_t)nblk * 4 * K > wn.part_cap) return pic_fail(PIC_ERR_STATE, "warm-start partial buffer too small");
    hipLaunchKernelGGL((k_ws_wgrad<C, MODE>), dim3(nblk), dim3(256), 0, g.stream, in, z, zbs, ntj, nti, batch, wn.d_part);
    ws_reduce_launch(nblk * 4, K, grad);
    return PIC_OK;
}

static int64_t ws_corr_blocks(int batch) { return (int64_t)((wn.H[0] + WS_CORR_ROWS - 1) / WS_CORR_ROWS) * batch; }

template <int C, bool LIFT>
static int ws_corr_launch(int batch, const float* s, int64_t sbs, const uint16_t* v, int64_t vbs, float* grad) {
    const int nbx = (wn.H[0] + WS_CORR_ROWS - 1) / WS_CORR_ROWS;
    const int64_t nblk = ws_corr_blocks(batch);
    if (nblk * 9 * C > wn.part_cap) return pic_fail(PIC_ERR_STATE, "warm-start partial buffer too small");
    hipLaunchKernelGGL((k_ws_corr<C, LIFT>), dim3((unsigned)nbx, batch), dim3(256), 0, g.stream, wn.H[0], wn.W[0], s, sbs,
                       v, vbs, wn.d_part);
    ws_reduce_launch((int)nblk, 9 * C, grad);
    return PIC_OK;
}

// gradient of sum_b <gy_b, N(r_b)> into wn.d_g (needs the activations of ws_forward_t)
template <int C>
static int ws_backward_t(int batch, const float* r, int64_t rbs, const float* gy, int64_t gbs) {
    const int L = wn.L;
    const int64_t n0 = wn.n[0];
    int rc = ws_corr_launch<C, false>(batch, gy, gbs, wn.g[0], act_bs(0), wn.d_g + wn.off_out); if (rc) return rc;
    hipLaunchKernelGGL(k_ws_lift<C>, dim3((unsigned)((n0 + 255) / 256), batch), dim3(256), 0, g.stream,
                       wn.H[0], wn.W[0], gy, gbs, wn.d_p + wn.off_out, 1, wn.mask[0], wn.gb[0], act_bs(0));
    // decoder, finest first: z_l = gb[l] (masked by its producer)
    for (int l = 0; l + 1 < L; ++l) {
        rc = ws_wgrad_launch<C, WS_INTERP>(ws_in(l, l + 1, g_or_h(l + 1), act_bs(l + 1), wn.h[l], act_bs(l), 1.0f), batch,
                                           wn.gb[l], act_bs(l), wn.d_g + wn.off_dec[l]);
        if (rc) return rc;
        ws_conv_launch<C, WS_PLAIN>(ws_in(l, l, wn.gb[l], act_bs(l), nullptr, 0, 1.0f), batch,
                                    ws_weights(conv_index_dec(l), 1), nullptr, wn.ub[l], act_bs(l));
        // d g_{l+1} = P^T u_bar = 4 R u_bar, masked (on the top level this is d h_{L-1})
        WsIn tr = ws_in(l + 1, l, wn.ub[l], act_bs(l), nullptr, 0, 4.0f);
        hipLaunchKernelGGL((k_ws_transfer<C, WS_RESTRICT>), dim3((unsigned)((wn.n[l + 1] * (C / 8) + 255) / 256), batch),
                           dim3(256), 0, g.stream, tr, wn.mask[l + 1], wn.gb[l + 1], act_bs(l + 1));
    }
    // encoder, coarsest first: z_l = d h_l
    for (int l = L - 1; l >= 1; --l) {
        const uint16_t* zl = l == L - 1 ? wn.gb[L - 1] : wn.hb[l];
        rc = ws_wgrad_launch<C, WS_RESTRICT>(ws_in(l, l - 1, wn.h[l - 1], act_bs(l - 1), nullptr, 0, 1.0f), batch, zl,
                                             act_bs(l), wn.d_g + wn.off_enc[l]);
        if (rc) return rc;
        ws_conv_launch<C, WS_PLAIN>(ws_in(l, l, zl, act_bs(l), nullptr, 0, 1.0f), batch, ws_weights(conv_index_enc(l), 1),
                                    nullptr, wn.tb[l], act_bs(l));
        // d h_{l-1} = mask . (u_bar_{l-1} + R^T t_bar_l), R^T = P / 4
        WsIn tr = ws_in(l - 1, l, wn.tb[l], act_bs(l), wn.ub[l - 1], act_bs(l - 1), 0.25f);
        hipLaunchKernelGGL((k_ws_transfer<C, WS_INTERP>), dim3((unsigned)((wn.n[l - 1] * (C / 8) + 255) / 256), batch),
                           dim3(256), 0, g.stream, tr, wn.mask[l - 1], wn.hb[l - 1], act_bs(l - 1));
    }
    rc = ws_corr_launch<C, true>(batch, r, rbs, wn.hb[0], act_bs(0), wn.d_g + wn.off_lift); if (rc) return rc;
    HIPCHK(hipGetLastError());
    return PIC_OK;
}

static int ws_forward(int batch, const float* r, int64_t rbs, float* y, int64_t ybs) {
    return wn.C == 16 ? ws_forward_t<16>(batch, r, rbs, y, ybs) : ws_forward_t<32>(batch, r, rbs, y, ybs);
}
static int ws_backward(int batch, const float* r, int64_t rbs, const float* gy, int64_t gbs) {
    return wn.C == 16 ? ws_backward_t<16>(batch, r, rbs, gy, gbs) : ws_backward_t<32>(batch, r, rbs, gy, gbs);
}

static int ws_ensure_batch(int batch) {
    if (batch <= wn.bcap) return PIC_OK;
    if (wn.d_act) { (void)hipFree(wn.d_act); wn.d_act = nullptr; }
    if (wn.d_part) { (void)hipFree(wn.d_part); wn.d_part = nullptr; }
    auto al = [](size_t b) { return (b + 255) / 256 * 256; };
    size_t bytes = 0;
    for (int l = 0; l < wn.L; ++l) bytes += 6 * al((size_t)act_bs(l) * sizeof(uint16_t) * batch);
    HIPCHK(hipMalloc(&wn.d_act, bytes));
    HIPCHK(hipMemset(wn.d_act, 0, bytes));
    char* p = (char*)wn.d_act;
    for (int l = 0; l < wn.L; ++l) {
        uint16_t** vs[6] = {&wn.h[l], &wn.g[l], &wn.gb[l], &wn.ub[l], &wn.hb[l], &wn.tb[l]};
        for (int k = 0; k < 6; ++k) { *vs[k] = (uint16_t*)p; p += al((size_t)act_bs(l) * sizeof(uint16_t) * batch); }
    }
    // partial sums: the largest weight-gradient launch (512 blocks x 4 waves x 9 C C) or 1-channel
    // one (a block per WS_CORR_ROWS rows of a batch entry); chunk sums of 64 parts each in FP64
    const int64_t corr_blocks = ws_corr_blocks(batch);
    wn.part_cap = std::max<int64_t>(512LL * 4 * 9 * wn.C * wn.C, corr_blocks * 9 * wn.C);
    HIPCHK(hipMalloc(&wn.d_part, wn.part_cap * sizeof(float)));
    if (wn.d_part2) { (void)hipFree(wn.d_part2); wn.d_part2 = nullptr; }
    const int64_t nch_max = std::max<int64_t>((512LL * 4 + 63) / 64, (corr_blocks + 63) / 64);
    HIPCHK(hipMalloc(&wn.d_part2, nch_max * 9 * wn.C * wn.C * sizeof(double)));
    wn.bcap = batch;
    return PIC_OK;
}

// ---- parameters
static void ws_init_params(std::vector<float>& P, uint64_t seed) {
    uint64_t st = seed * 0x9E3779B97F4A7C15ull + 3;
    auto rnd = [&]() {                                 // uniform in [-1, 1)
        st ^= st << 13; st ^= st >> 7; st ^= st << 17;
        return (float)((double)(st >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0);
    };
    const int C = wn.C;
    // uniform with variance 1 / fan_in: half-width sqrt(3 / fan_in)
    const float a1 = std::sqrt(3.0f / 9.0f), aC = std::sqrt(3.0f / (9.0f * C));
    for (int k = 0; k < 9 * C; ++k) P[wn.off_lift + k] = a1 * rnd();
    for (int k = 0; k < 9 * C; ++k) P[wn.off_out + k] = aC * rnd();
    for (int l = 1; l < wn.L; ++l)
        for (int k = 0; k < 9 * C * C; ++k) P[wn.off_enc[l] + k] = aC * rnd();
    for (int l = 0; l + 1 < wn.L; ++l)
        for (int k = 0; k < 9 * C * C; ++k) P[wn.off_dec[l] + k] = aC * rnd();
}

int warm_setup(int levels, int channels, uint64_t seed) {
    if (!g.have_domain) return pic_fail(PIC_ERR_STATE, "pic_setup_domain not called");
    if (!learn_ready()) return pic_fail(PIC_ERR_STATE, "pic_learn_setup not called (the warm start trains on its ring)");
    if (channels != 16 && channels != 32) return pic_fail(PIC_ERR_ARG, "channels must be 16 or 32");
    if (levels < 2 || levels > WS_MAX_LEVELS) return pic_fail(PIC_ERR_ARG, "levels must be 2..8");
    ws_free();
    wn.C = channels;
    int L = 1;
    wn.H[0] = g.mesh.nnx; wn.W[0] = g.mesh.nny;
    while (L < levels && std::min(wn.H[L - 1], wn.W[L - 1]) > 5) {
        wn.H[L] = (wn.H[L - 1] + 1) / 2; wn.W[L] = (wn.W[L - 1] + 1) / 2; ++L;
    }
    if (L < 2) return pic_fail(PIC_ERR_UNSUPPORTED, "mesh too small for the warm-start network");
    wn.L = L;
    for (int l = 0; l < L; ++l) wn.n[l] = (int64_t)wn.H[l] * wn.W[l];
    // masks: unknown nodes, m_{l+1}(I, J) = m_l(2I, 2J)
    size_t tot = 0;
    for (int l = 0; l < L; ++l) tot += (size_t)wn.n[l];
    std::vector<float> hm(tot, 0.0f);
    for (int u : g.h_unknown_node) hm[(size_t)(u - 1)] = 1.0f;
    size_t o = 0;
    for (int l = 1; l < L; ++l) {
        const size_t of = o;
        o += (size_t)wn.n[l - 1];
        for (int I = 0; I < wn.H[l]; ++I)
            for (int J = 0; J < wn.W[l]; ++J) {
                const int fi = 2 * I, fj = 2 * J;
                hm[o + (size_t)I * wn.W[l] + J] =
                    (fi < wn.H[l - 1] && fj < wn.W[l - 1]) ? hm[of + (size_t)fi * wn.W[l - 1] + fj] : 0.0f;
            }
    }
    HIPCHK(hipMalloc(&wn.d_mask, tot * sizeof(float)));
    HIPCHK(hipMemcpy(wn.d_mask, hm.data(), tot * sizeof(float), hipMemcpyHostToDevice));
    o = 0;
    for (int l = 0; l < L; ++l) { wn.mask[l] = (const float*)wn.d_mask + o; o += (size_t)wn.n[l]; }
    // parameters
    const int C = channels;
    int64_t off = 0;
    wn.off_lift = off; off += 9 * C;
    wn.off_out = off; off += 9 * C;
    for (int l = 1; l < L; ++l) { wn.off_enc[l] = off; off += 9LL * C * C; }
    for (int l = 0; l + 1 < L; ++l) { wn.off_dec[l] = off; off += 9LL * C * C; }
    wn.nparams = off;
    std::vector<float> P(off, 0.0f);
    ws_init_params(P, seed);
    HIPCHK(hipMalloc(&wn.d_p, off * sizeof(float))); HIPCHK(hipMalloc(&wn.d_m, off * sizeof(float)));
    HIPCHK(hipMalloc(&wn.d_v, off * sizeof(float))); HIPCHK(hipMalloc(&wn.d_g, off * sizeof(float)));
    HIPCHK(hipMemcpy(wn.d_p, P.data(), off * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(hipMemset(wn.d_m, 0, off * sizeof(float))); HIPCHK(hipMemset(wn.d_v, 0, off * sizeof(float)));
    HIPCHK(hipMemset(wn.d_g, 0, off * sizeof(float)));
    wn.wp_per = (C / 16) * ((9 * C + 31) / 32) * 64;
    HIPCHK(hipMalloc(&wn.d_wp, (size_t)2 * 2 * (L - 1) * wn.wp_per * sizeof(ws_u32x4)));
    const int64_t n0 = wn.n[0];
    HIPCHK(hipMalloc(&wn.d_in, n0 * sizeof(float)));
    HIPCHK(hipMalloc(&wn.d_y, n0 * sizeof(float)));
    HIPCHK(hipMalloc(&wn.d_dots, 2 * ((n0 + 255) / 256 + 1) * sizeof(double)));
    HIPCHK(hipMalloc(&wn.d_alpha, sizeof(double)));
    HIPCHK(hipMemset(wn.d_alpha, 0, sizeof(double)));
    int rc = ws_pack(); if (rc) return rc;
    rc = ws_ensure_batch(1); if (rc) return rc;
    HIPCHK(hipStreamSynchronize(g.stream));
    wn.ready = true;
    return PIC_OK;
}

bool warm_active() { return wn.ready && wn.enabled && learn_ready(); }

int warm_enable(int enable) {
    if (enable < 0 || enable > 2) return pic_fail(PIC_ERR_ARG, "enable must be 0, 1 or 2");
    if (enable && !wn.ready) return pic_fail(PIC_ERR_STATE, "pic_warm_setup not called");
    wn.enabled = enable != 0;
    wn.mode = enable;
    return PIC_OK;
}

static int ws_enqueue_solve_start(int n, const double* rn, double* xn) {
    const int nb = (n + 255) / 256;
    hipLaunchKernelGGL(k_ws_scale_in, dim3(nb), dim3(256), 0, g.stream, n, g.d_sinv, rn, wn.d_in);
    int rc = ws_forward(1, wn.d_in, n, wn.d_y, n); if (rc) return rc;
    hipLaunchKernelGGL(k_ws_dots, dim3(nb), dim3(256), 0, g.stream, n, learn_scaled_op(), g.d_sinv, rn, wn.d_y, wn.d_dots, nb);
    hipLaunchKernelGGL(k_ws_update, dim3(nb), dim3(256), 0, g.stream, n, nb, wn.d_dots, g.d_sinv, wn.d_y, xn, wn.d_alpha);
    HIPCHK(hipGetLastError());
    return PIC_OK;
}

// In the solve: xn += alpha D^-1/2 N(D^-1/2 rn), rn the FP64 residual of xn (no host wait).
// Mode 1 replays a captured graph of the same launches (no memset nodes: every buffer is
// written by a kernel), mode 2 launches them one by one.
int warm_solve_start(int n, const double* rn, double* xn) {
    if (!warm_active()) return PIC_OK;
    if ((int64_t)n != wn.n[0]) return pic_fail(PIC_ERR_STATE, "warm-start network built for another mesh");
    if (wn.mode == 2) {
        int rc = ws_enqueue_solve_start(n, rn, xn); if (rc) return rc;
    } else {
        if (!wn.graph_exec || wn.graph_rn != rn || wn.graph_xn != xn || wn.graph_act != wn.d_act || wn.graph_n != n) {
            ws_drop_graph();
            HIPCHK(hipStreamBeginCapture(g.stream, hipStreamCaptureModeThreadLocal));
            int rc = ws_enqueue_solve_start(n, rn, xn);
            hipGraph_t gr = nullptr;
            hipError_t e = hipStreamEndCapture(g.stream, &gr);
            if (rc) { if (gr) (void)hipGraphDestroy(gr); return rc; }
            if (e != hipSuccess) return pic_fail(PIC_ERR_HIP, std::string("warm-start graph capture: ") + hipGetErrorString(e));
            wn.graph = gr;
            HIPCHK(hipGraphInstantiate(&wn.graph_exec, wn.graph, nullptr, nullptr, 0));
            wn.graph_rn = rn; wn.graph_xn = xn; wn.graph_act = wn.d_act; wn.graph_n = n;
        }
        HIPCHK(hipGraphLaunch(wn.graph_exec, g.stream));
    }
    wn.applied++;
    return PIC_OK;
}

// ---- training on the ring pairs that hold the initial residual of a solve
static int ws_loss_grad(int batch, const int* idx, double* loss) {
    LearnBatch lb;
    int rc = learn_batch_pairs(batch, idx, &lb); if (rc) return rc;
    const int64_t N = wn.n[0];
    rc = ws_forward(batch, lb.r, N, lb.y, N); if (rc) return rc;
    rc = learn_batch_loss(batch, loss); if (rc) return rc;
    return ws_backward(batch, lb.r, N, lb.gy, N);
}

int warm_train(int steps, int batch, double lr, uint64_t seed, double* loss_first, double* loss_last) {
    if (!wn.ready) return pic_fail(PIC_ERR_STATE, "pic_warm_setup not called");
    if (batch < 1 || batch > 64 || steps < 0) return pic_fail(PIC_ERR_ARG, "batch must be 1..64");
    std::vector<int> pool;
    learn_ring_initial(pool);
    if (pool.empty()) return pic_fail(PIC_ERR_STATE, "warm start: no harvested initial-residual pairs");
    int rc = ws_ensure_batch(batch); if (rc) return rc;
    uint64_t st = seed * 0x9E3779B97F4A7C15ull + 11;
    std::vector<int> idx(batch);
    for (int s = 0; s < steps; ++s) {
        for (int b = 0; b < batch; ++b) {
            st ^= st << 13; st ^= st >> 7; st ^= st << 17;
            idx[b] = pool[(size_t)(st % (uint64_t)pool.size())];
        }
        const bool want = s == 0 || s == steps - 1;
        double loss = 0.0;
        rc = ws_loss_grad(batch, idx.data(), want ? &loss : nullptr); if (rc) return rc;
        if (want) {
            loss /= batch;
            if (s == 0 && loss_first) *loss_first = loss;
            if (loss_last) *loss_last = loss;
            wn.last_loss = loss;
        }
        wn.adam_t++;
        learn_adam_step(wn.nparams, wn.d_p, wn.d_g, wn.d_m, wn.d_v, (float)lr, wn.adam_t, 1.0f / batch);
        rc = ws_pack(); if (rc) return rc;
    }
    HIPCHK(hipStreamSynchronize(g.stream));
    return PIC_OK;
}

// mean loss (energy-norm error of the raw output, relative) over the first npairs
// initial-residual pairs of the ring; *scaled: the same after the line search
int warm_eval(int npairs, double* mean_loss, double* mean_scaled) {
    if (!wn.ready) return pic_fail(PIC_ERR_STATE, "pic_warm_setup not called");
    std::vector<int> pool;
    learn_ring_initial(pool);
    if (npairs < 1 || npairs > (int)pool.size()) return pic_fail(PIC_ERR_ARG, "npairs out of range");
    double tot = 0.0, tot_s = 0.0;
    for (int k = 0; k < npairs; ++k) {
        LearnBatch lb;
        int rc = learn_batch_pairs(1, &pool[k], &lb); if (rc) return rc;
        const int64_t N = wn.n[0];
        rc = ws_forward(1, lb.r, N, lb.y, N); if (rc) return rc;
        double l = 0.0, ls = 0.0;
        rc = learn_batch_loss_scaled(lb.y, &l, &ls); if (rc) return rc;
        tot += l; tot_s += ls;
    }
    *mean_loss = tot / npairs;
    if (mean_scaled) *mean_scaled = tot_s / npairs;
    return PIC_OK;
}

int warm_apply_host(const float* r, float* y, int64_t n) {
    if (!wn.ready) return pic_fail(PIC_ERR_STATE, "pic_warm_setup not called");
    if (n != wn.n[0]) return pic_fail(PIC_ERR_ARG, "n must equal the node count");
    HIPCHK(hipMemcpyAsync(wn.d_in, r, n * sizeof(float), hipMemcpyHostToDevice, g.stream));
    int rc = ws_forward(1, wn.d_in, n, wn.d_y, n); if (rc) return rc;
    HIPCHK(hipMemcpyAsync(y, wn.d_y, n * sizeof(float), hipMemcpyDeviceToHost, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    return PIC_OK;
}

int warm_params(float* dst, const float* src, int64_t n, int64_t* nparams) {
    if (!wn.ready) return pic_fail(PIC_ERR_STATE, "pic_warm_setup not called");
    if (nparams) *nparams = wn.nparams;
    if (dst) {
        if (n < wn.nparams) return pic_fail(PIC_ERR_ARG, "buffer too small for the parameters");
        HIPCHK(hipMemcpy(dst, wn.d_p, wn.nparams * sizeof(float), hipMemcpyDeviceToHost));
    }
    if (src) {
        if (n != wn.nparams) return pic_fail(PIC_ERR_ARG, "parameter count mismatch");
        HIPCHK(hipStreamSynchronize(g.stream));
        HIPCHK(hipMemcpy(wn.d_p, src, n * sizeof(float), hipMemcpyHostToDevice));
        HIPCHK(hipMemset(wn.d_m, 0, n * sizeof(float)));
        HIPCHK(hipMemset(wn.d_v, 0, n * sizeof(float)));
        wn.adam_t = 0;
        int rc = ws_pack(); if (rc) return rc;
        HIPCHK(hipStreamSynchronize(g.stream));
    }
    return PIC_OK;
}

int warm_layout(int32_t* levels, int32_t* H, int32_t* W, int32_t* channels) {
    if (!wn.ready) return pic_fail(PIC_ERR_STATE, "pic_warm_setup not called");
    *levels = wn.L; *channels = wn.C;
    for (int l = 0; l < wn.L; ++l) { H[l] = wn.H[l]; W[l] = wn.W[l]; }
    return PIC_OK;
}

int warm_gradient(int np, const int32_t* idx, double* loss, float* grad, int64_t n) {
    if (!wn.ready) return pic_fail(PIC_ERR_STATE, "pic_warm_setup not called");
    if (np < 1 || np > 64) return pic_fail(PIC_ERR_ARG, "1..64 pairs");
    if (n < wn.nparams) return pic_fail(PIC_ERR_ARG, "gradient buffer too small");
    int rc = ws_ensure_batch(np); if (rc) return rc;
    rc = ws_loss_grad(np, idx, loss); if (rc) return rc;
    HIPCHK(hipMemcpy(grad, wn.d_g, wn.nparams * sizeof(float), hipMemcpyDeviceToHost));
    return PIC_OK;
}

int warm_info(int32_t* initial_pairs, int32_t* enabled, double* last_alpha, double* last_loss, int64_t* applied) {
    std::vector<int> pool;
    if (learn_ready()) learn_ring_initial(pool);
    if (initial_pairs) *initial_pairs = (int32_t)pool.size();
    if (enabled) *enabled = warm_active() ? 1 : 0;
    if (last_alpha) {
        double a = 0.0;
        if (wn.ready) HIPCHK(hipMemcpy(&a, wn.d_alpha, sizeof(double), hipMemcpyDeviceToHost));
        *last_alpha = a;
    }
    if (last_loss) *last_loss = wn.last_loss;
    if (applied) *applied = wn.applied;
    return PIC_OK;
}

void warm_free() { ws_free(); }
