// Device helpers shared by the kernel translation units: element lookup, fixed-point
// quantisation and wave-level segmented reduction.
#pragma once
#include "pic_state.h"
#include "philox.h"

// Element lookup of GetParChg_2D.f90:137-163 on the unrefined mesh, FP64 throughout:
//   estimate i = INT((x - Vert_o) * hxi), j likewise, clamped at nx-1 / ny-1;
//   if Mod(rxp, Mapxmin) > 1e-10 and Mod(ryp, Mapymin) > 1e-10 the InjectPositioning
//   branch (InjectPositioning.f90:13-22: i = INT((x - dxmin)/dx) + 1, Element_Map is the
//   identity without refinement); otherwise ParticlePositioning, which on an unrefined
//   mesh without DG edges returns the estimate (ParticlePositioning.f90:240-363 with
//   empty EdgeParent -> part_edge_count(1) == 0 -> Finalindex, GetParChg_2D.f90:776-778).
// MOD of reals is taken as the exact remainder (fmod), which is what a == INT(a/p)*p
// rounds to whenever p is a power of two. Returns the 0-based element, sets *bad when
// the result falls outside the mesh.
PIC_HD int pic_locate_ij(double X, double Y, const MeshParams& m, int* bad, int* ci, int* cj) {
    double rxp = (X - m.vox) * m.hxi;
    int i = (int)rxp;
    double ryp = (Y - m.voy) * m.hyi;
    int j = (int)ryp;
    if (j == m.nny) j = j - 1;
    if (i == m.nnx) i = i - 1;
    double modx, mody;
    if (m.pow2_map) {
        // exact: a power-of-two divisor (1/mapx is exact and x*(1/mapx) == x/mapx)
        modx = rxp - m.mapx * trunc(rxp * m.imapx);
        mody = ryp - m.mapy * trunc(ryp * m.imapy);
    } else {
        modx = fmod(rxp, m.mapx);
        mody = fmod(ryp, m.mapy);
    }
    if (modx > 1e-10 && mody > 1e-10) {
        double rx = m.pow2_map ? (X - m.xmin) * m.imapx : (X - m.xmin) / m.mapx;
        i = (int)rx + 1;
        double ry = m.pow2_map ? (Y - m.ymin) * m.imapy : (Y - m.ymin) / m.mapy;
        j = (int)ry + 1;
    }
    if (i < 1 || i > m.nex || j < 1 || j > m.ney || !(X == X) || !(Y == Y)) {
        *bad = 1;
        i = i < 1 ? 1 : (i > m.nex ? m.nex : i);
        j = j < 1 ? 1 : (j > m.ney ? m.ney : j);
    }
    *ci = i - 1; *cj = j - 1;
    return (j - 1) + (i - 1) * m.ney;
}

// pic_locate_ij for an FP32 coordinate pair inside the domain (xmin <= X, ymin <= Y) on a mesh
// where it reduces exactly (locate_f32: power-of-two spacing and map, origin a multiple of the
// spacing): X - Vert_o and X - xmin are then exact in FP64 and the scaling by 1/h is exact, so
// INT((X - xmin) / mapx) + 1 equals INT(rxp) and the grid-line test only chooses between equal
// values. Same element, same clamps and range test, without the grid-line arithmetic.
PIC_HD int pic_locate_inside_f32(float Xf, float Yf, const MeshParams& m, int* bad, int* ci, int* cj) {
    const double X = (double)Xf, Y = (double)Yf;
    int i = (int)((X - m.vox) * m.hxi);
    int j = (int)((Y - m.voy) * m.hyi);
    if (j == m.nny) j = j - 1;
    if (i == m.nnx) i = i - 1;
    if (i < 1 || i > m.nex || j < 1 || j > m.ney || !(X == X) || !(Y == Y)) {
        *bad = 1;
        i = i < 1 ? 1 : (i > m.nex ? m.nex : i);
        j = j < 1 ? 1 : (j > m.ney ? m.ney : j);
    }
    *ci = i - 1; *cj = j - 1;
    return (j - 1) + (i - 1) * m.ney;
}

PIC_HD int pic_locate_elem(double X, double Y, const MeshParams& m, int* bad) {
    int ci, cj;
    return pic_locate_ij(X, Y, m, bad, &ci, &cj);
}

// Column and row of element e (ci = e / ney, cj = e mod ney), 0 <= e < nelem. Below 2^24
// elements the FP32 product e * RN(1/ney) is within 1 of e / ney (e is exact in FP32, two
// roundings give a relative error under 2^-23 + 2^-48, and e / ney < nex < 2^23 when ney > 1;
// ney = 1 is exact), so truncation is off by at most one, which the remainder test corrects.
PIC_HD void pic_elem_ij(int e, const MeshParams& m, int* ci, int* cj) {
    int q, r;
    if (m.ij_exact_div) {
        q = e / m.ney; r = e - q * m.ney;
    } else {
        q = (int)((float)e * m.ney_rcp);
        r = e - q * m.ney;
        q += (r >= m.ney) - (r < 0);
        r = e - q * m.ney;
    }
    *ci = q; *cj = r;
}

// Loads and stores at base[i] with a 32-bit byte offset (i * sizeof(T) < 2^32): the address
// is the wave-uniform base plus a zero-extended 32-bit lane offset, one instruction shorter
// than 64-bit index arithmetic per access.
template <typename T>
PIC_HD T pic_ld(const T* base, unsigned int i) {
    return *(const T*)((const char*)base + (size_t)(i * (unsigned int)sizeof(T)));
}
template <typename T>
PIC_HD void pic_st(T* base, unsigned int i, T v) {
    *(T*)((char*)base + (size_t)(i * (unsigned int)sizeof(T))) = v;
}

// Slot layout after a cell sort: within every full group of 64 K slots the sorted position
// K l + k (l = 0..63, k = 0..K-1) sits at slot 64 k + l, so lane l of a wave that walks the
// group 64 slots at a time meets K consecutive sorted particles (a partial last group keeps
// the sorted order). K is the push kernel's slots per thread (pic_push_ppt): 8 for FP32
// particles on a Cartesian mesh of the common layout, else 4. Any order is valid input for
// the push; the layout only sets how often a thread's consecutive slots share an element.
#define PIC_PPT_NARROW 4
#define PIC_PPT_WIDE 8
PIC_HD int64_t pic_sorted_position(int64_t slot, int64_t nfull, int ppt) {
    if (slot >= nfull) return slot;
    const int64_t g = slot & ~(int64_t)(64 * ppt - 1);
    const int r = (int)(slot - g);
    return g + (int64_t)((r & 63) * ppt + (r >> 6));
}

// Sort tiles: SORT_TILE x SORT_TILE elements, tiles visited column by column with the
// direction alternating (serpentine), elements inside a tile column-major.
#define SORT_TILE 16
PIC_HD unsigned int pic_sort_key(int e, int ney, int nty) {
    int ci = e / ney, cj = e - ci * ney;
    int ti = ci / SORT_TILE, tj = cj / SORT_TILE;
    int tpos = (ti & 1) ? (nty - 1 - tj) : tj;
    unsigned int tile = (unsigned int)(ti * nty + tpos);
    return tile * (SORT_TILE * SORT_TILE) + (unsigned int)((ci % SORT_TILE) * SORT_TILE + (cj % SORT_TILE));
}

// Device error bits.
enum {
    PIC_EBIT_POSITION = 1u,
    PIC_EBIT_BOUNDARY = 2u,
    PIC_EBIT_NONFINITE = 4u,
    PIC_EBIT_WRONGCROSS = 8u,
    PIC_EBIT_EKRHO = 16u,
    PIC_EBIT_OVERFLOW = 32u,
    PIC_EBIT_CAPACITY = 64u,
    PIC_EBIT_MCCLOOP = 128u,
    PIC_EBIT_SAMPLER = 256u
};

// Fixed-point quantum of deposition: value * 2^40 rounded to nearest (even), as int64.
// Integer sums are exact and independent of order, so the node moments are
// bit-identical whatever the particle order, atomic interleaving or kernel fusion.
// For |x| < 2^51, x + 1.5*2^52 rounds x to an integer (nearest even) in the low mantissa
// bits, so the bit pattern minus that of 1.5*2^52 is the rounded value: one add and one
// integer subtract instead of the generic double -> int64 conversion sequence.
#if defined(__HIPCC__)
__device__ inline long long pic_fix(double v) {
    const double x = v * PIC_FIX_SCALE;
    if (fabs(x) < 0x1p51) {
        const double t = x + 0x1.8p52;
        return __double_as_longlong(t) - __double_as_longlong(0x1.8p52);
    }
    return __double2ll_rn(x);
}

// Range of the fixed-point deposition. Each node keeps three int64 sums per species: the
// weights W and the weight x energy quanta W E at scale 2^40, and the whole-eV parts of the
// rare quanta with W E >= 2^11 eV (a kappa re-emission draws energies up to ~1e16 eV), whose
// fraction stays in the 2^40 sum. The node's energy moment is the exact integer sum of all
// quanta, held as lo + 2^40 hi, so it does not depend on the order of the additions and no
// single particle can wrap the 2^40 sum. Out of range (PIC_EBIT_OVERFLOW): a non-finite
// quantum, W E >= 2^62 eV, or a node sum that comes out negative (every quantum is >= 0).
#define PIC_FIX_SPLIT 2048.0          // 2^11: W E at or above it is split into whole eV + fraction
#define PIC_FIX_HI_LIMIT 0x1p62

// round(v 2^40) as lo + 2^40 hi: hi = 0 below 2^11, otherwise hi = floor(v) (exact integer) and
// lo = round((v - floor(v)) 2^40) (v - floor(v) is exact: floor(v) lies within [v/2, v]).
// Returns false when v is out of range (both parts 0 then).
__device__ inline bool pic_fix_split(double v, long long* lo, long long* hi) {
    if (fabs(v) < PIC_FIX_SPLIT) { *hi = 0; *lo = pic_fix(v); return true; }
    if (!(fabs(v) < PIC_FIX_HI_LIMIT)) { *hi = 0; *lo = 0; return false; }
    const double ip = floor(v);
    *hi = (long long)ip;
    *lo = pic_fix(v - ip);
    return true;
}

// The quanta of one particle's deposit on element e: q[k] = round(W[k] 2^40) and q[4 + k] +
// 2^40 hi = round(W[k] energy 2^40) for the four corners (LL, LR, UR, UL). When |W| < 2 and
// |energy| < 1000 every value v has |v| < 2^11, so v + 1.5*2^12 lies in [2^12, 2^13) where the
// spacing is 2^-40: the add rounds v*2^40 to the nearest (even) integer, held in the low
// mantissa bits (the same rounding as pic_fix, one FP64 add and one 32-bit subtract on the
// high word per value), and every hi part is 0. The test is one uniform branch per particle
// instead of one per value. On the other path a nonzero hi part goes straight to acc_h of its
// corner node (a global atomic: rare), and a value out of range sets PIC_EBIT_OVERFLOW.
__device__ inline void pic_fix8(const double W[4], double energy, long long q[8], int e, const MeshParams& m,
                                long long* acc_h, unsigned int* err) {
    const double wmax = fmax(fmax(fabs(W[0]), fabs(W[1])), fmax(fabs(W[2]), fabs(W[3])));
    if (wmax < 2.0 && fabs(energy) < 1000.0) {
        const long long off = __double_as_longlong(0x1.8p12);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            q[k] = __double_as_longlong(__dadd_rn(W[k], 0x1.8p12)) - off;
            q[4 + k] = __double_as_longlong(__dadd_rn(__dmul_rn(W[k], energy), 0x1.8p12)) - off;   // never fused
        }
        return;
    }
    int ci, cj;
    pic_elem_ij(e, m, &ci, &cj);
    const int n1 = ci * m.nny + cj;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        long long hi;
        q[k] = pic_fix(W[k]);
        if (!pic_fix_split(W[k] * energy, &q[4 + k], &hi)) atomicOr(err, PIC_EBIT_OVERFLOW);
        const int node = n1 + (k == 1 || k == 2 ? m.nny : 0) + (k >= 2 ? 1 : 0);
        if (hi != 0) atomicAdd((unsigned long long*)&acc_h[node], (unsigned long long)hi);
    }
}
#endif


// First failing particle, recorded for the error message (first writer wins).
struct PicErrInfo {
    unsigned int taken;
    unsigned int bit;
    long long slot;
    double v[8];     // x, y, vx, vy, boundary iterations, old element, new element, species
};
#if defined(__HIPCC__)
__device__ inline void pic_record(PicErrInfo* ei, unsigned int bit, long long slot, double x, double y, double vx,
                                  double vy, double it, double e_old, double e_new, double sp) {
    if (ei && atomicCAS(&ei->taken, 0u, 1u) == 0u) {
        ei->bit = bit; ei->slot = slot;
        ei->v[0] = x; ei->v[1] = y; ei->v[2] = vx; ei->v[3] = vy; ei->v[4] = it; ei->v[5] = e_old;
        ei->v[6] = e_new; ei->v[7] = sp;
    }
}
#endif

// Counter slots of SpeciesHost::counters.
enum {
    PIC_C_EMIT = 0, PIC_C_RIGHT = 1, PIC_C_OBJECT = 2, PIC_C_REFLECT = 3,
    PIC_C_MCC_SELECTED = 4, PIC_C_MCC_REAL = 5, PIC_C_MCC_CREATED = 6, PIC_C_DELETED = 7
};

// Jacobi-scaled node-space operator: eight scaled off-diagonals (the unit diagonal is
// implicit), offsets -nny-1 ... nny+1 without 0.
struct Dia8 { const float* c[8]; int off[8]; };
