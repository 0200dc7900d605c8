// Particle kernels: the fused mover (MoveAll/MoveOne/Get_ParticleEField_PPR/
// AdjustParticleBoundary) with the element lookup and deposition of GetParChg_2D, the
// stand-alone lookup+deposition, the deposition finalisation (corner -> node sums and
// normalisation) and the batched drop-ins (point lookup, PPR point gather).
//
// Particles are structure-of-arrays slots kept in tile-major element order (periodic
// cell sort). A wave owns a group of 64 PPT slots and a thread PPT of them (4, or 8 for FP32
// particles on a Cartesian mesh of the common layout), 64 slots apart, so every access is lane-contiguous; the sort places consecutive sorted
// particles on a thread's slots. Its
// deposits are summed in registers while consecutive particles share an element; the
// per-thread runs are merged by one segmented wave reduction keyed by element, and run
// totals go to an LDS node window of the block (flushed once with global atomics) or,
// outside the window, straight to global atomics. All sums are int64 fixed point, so
// the node moments are identical whatever path or order a deposit takes.
#include "pic_device.h"
#include "pic_trig.h"

#define PUSH_BLOCK 256
// LDS deposition window of a push block: element columns ci - M .. ci + 1 + M around the
// element of the block's middle particle, and the rows of its sort tile +- M, with M =
// PUSH_WIN_MARGIN cells of drift since the last sort. A block's PUSH_BLOCK x PPT sorted particles
// cover one or two element columns of one tile; deposits outside the window go to global
// atomics.
#ifndef PUSH_WIN_MARGIN
#define PUSH_WIN_MARGIN 16
#endif
#define WIN_NX (2 * PUSH_WIN_MARGIN + 3)
#define WIN_NY (SORT_TILE + 2 * PUSH_WIN_MARGIN + 1)
#define WIN_NODES (WIN_NX * WIN_NY)

template <typename R> struct PushArgs {
    R *x, *y, *vx, *vy, *vz;
    R* th;                    // azimuth (delta = 1), else null
    const int* cell;
    int* cell_next;
    int64_t n;
    MeshParams m;
    const DevObject* objs;    // device copy of the objects (keeps them out of SGPRs)
    const R* efx; const R* efy;
    const R* coef5;           // c2..c6 per node, node-major (5 per node)
    const R* xsR; const R* ysR;
    R qm, dt;
    R xmin, xmax, ymin, ymax, emit_eps;   // domain and the 10E-6 offset of the reference
    R hxR, hxiR, hyR, hyiR;               // mesh spacing and reciprocal in R (exact if pow2)
    R x0R, y0R;                           // xmin, ymin in R (node lines, node_x / node_y)
    double vref_inv;          // 1/v_ref  (VelRes(1.0/PB%VFactor))
    double kappa_c;           // (Kappa-1.5)/beta, beta = m/(2 kB T)   (SI)
    double ek_c;              // 0.5 Mass v_ref^2 / JtoeV: kinetic energy (eV) = |v|^2 ek_c
    uint64_t seed; uint32_t step; uint32_t tag_emit;
    long long* acc_w; long long* acc_e;
    long long* acc_h;         // whole-eV parts of the rare W E quanta >= 2^11 eV (pic_fix_split)
    unsigned long long* counters;
    unsigned int* err;
    PicErrInfo* errinfo;
    int species;
    int* wall_list; unsigned int* wall_count;   // particles whose first pass reaches a wall
};

__device__ inline long long shfl_xor_ll(long long v, int d) {
    int lo = (int)(unsigned long long)v, hi = (int)((unsigned long long)v >> 32);
    lo = __shfl_xor(lo, d, 64); hi = __shfl_xor(hi, d, 64);
    return (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

__device__ inline long long shfl_up_ll(long long v, int d) {
    int lo = (int)(unsigned long long)v, hi = (int)((unsigned long long)v >> 32);
    lo = __shfl_up(lo, d, 64); hi = __shfl_up(hi, d, 64);
    return (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

// Wave total of a per-lane count, added to a global counter by one lane (rare events).
__device__ inline void wave_add(unsigned long long* ctr, int c) {
    if (__ballot(c != 0) == 0ull) return;
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(ctr, (unsigned long long)c);
}

// LDS deposition window of a block: node columns x0..x0+nx-1, rows y0..y0+ny-1.
struct DepWindow { int x0, y0, nx, ny; };

// 64-bit add to an LDS word through an LDS-typed pointer: the compiler cannot merge it with
// the global atomic of the other branch into one flat atomic (which would count on both the
// vector-memory and the LDS wait counters).
typedef __attribute__((address_space(3))) unsigned long long lds_u64;
__device__ inline void lds_add(long long* p, long long v) {
    __hip_atomic_fetch_add((lds_u64*)p, (unsigned long long)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Add the eight moments of element `key` to the window (when its four nodes lie in it)
// or to the global node accumulators.
__device__ inline void add_element(int key, const long long v[8], const MeshParams& m, long long* acc_w,
                                   long long* acc_e, const DepWindow& w, long long* winW, long long* winE) {
    int ci, cj;
    pic_elem_ij(key, m, &ci, &cj);
    int li = ci - w.x0, lj = cj - w.y0;
    if (li >= 0 && lj >= 0 && li + 1 < w.nx && lj + 1 < w.ny) {
        int l1 = li * w.ny + lj, l2 = l1 + w.ny, l3 = l2 + 1, l4 = l1 + 1;
        lds_add(&winW[l1], v[0]); lds_add(&winW[l2], v[1]); lds_add(&winW[l3], v[2]); lds_add(&winW[l4], v[3]);
        lds_add(&winE[l1], v[4]); lds_add(&winE[l2], v[5]); lds_add(&winE[l3], v[6]); lds_add(&winE[l4], v[7]);
    } else {
        int n1 = ci * m.nny + cj, n2 = n1 + m.nny, n3 = n2 + 1, n4 = n1 + 1;
        atomicAdd((unsigned long long*)&acc_w[n1], (unsigned long long)v[0]);
        atomicAdd((unsigned long long*)&acc_w[n2], (unsigned long long)v[1]);
        atomicAdd((unsigned long long*)&acc_w[n3], (unsigned long long)v[2]);
        atomicAdd((unsigned long long*)&acc_w[n4], (unsigned long long)v[3]);
        atomicAdd((unsigned long long*)&acc_e[n1], (unsigned long long)v[4]);
        atomicAdd((unsigned long long*)&acc_e[n2], (unsigned long long)v[5]);
        atomicAdd((unsigned long long*)&acc_e[n3], (unsigned long long)v[6]);
        atomicAdd((unsigned long long*)&acc_e[n4], (unsigned long long)v[7]);
    }
}

// Segmented inclusive sum over the wave of eight int64 values keyed by `key` (runs of
// equal keys in lane order); the lane at the tail of each run adds the run total.
__device__ inline void wave_deposit(int key, const MeshParams& m, long long v[8], long long* acc_w, long long* acc_e,
                                    const DepWindow& w, long long* winW, long long* winE) {
    const int lane = threadIdx.x & 63;
    int kprev = __shfl_up(key, 1, 64);
    int knext = __shfl_down(key, 1, 64);
    bool head = (lane == 0) || (kprev != key);
    bool tail = (lane == 63) || (knext != key);
    unsigned long long heads = __ballot(head);
    unsigned long long below = heads & ((lane == 63) ? ~0ull : ((2ull << lane) - 1ull));
    int seg_start = 63 - __clzll((long long)below);
    if (heads != 1ull) {
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            bool take = (lane - off) >= seg_start;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                long long u = shfl_up_ll(v[k], off);
                if (take) v[k] += u;
            }
        }
    } else {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
            for (int k = 0; k < 8; ++k) v[k] += shfl_xor_ll(v[k], off);
        }
    }
    if (tail && key >= 0) add_element(key, v, m, acc_w, acc_e, w, winW, winE);
}

// Specular folding of y into [ymin, ymax] (reflections off both walls), Vy reversed.
template <typename R>
__device__ inline void fold_y(R& Y, R& VY, R ymin, R ymax) {
    R L = ymax - ymin;
    R d = Y - ymin;
    if (!(d >= (R)0 && d <= L)) {
        R per = (R)2 * L;
        d = fmod(d, per);
        if (d < (R)0) d = d + per;
        if (d > L) d = per - d;
        Y = ymin + d;
    }
    VY = -VY;
}

// Deposition weights of GetParChg_2D.f90:788-821 on element `e`, FP64. With a
// power-of-two spacing the division is the (bit-identical) multiplication by 1/h.
// The node lines are xmin + hx i, ymin + hy j (the set-up's formula, same roundings).
__device__ inline void deposit_weights_ij(double X, double Y, int ci, int cj, const MeshParams& m, double W[4]) {
    if (m.unif_r64 && m.delta == 0) {
        // every element's spacing is hx (hy) exactly: the same arithmetic as below without the
        // far node lines
        const double x1 = m.xmin + m.hx * (double)ci, y1 = m.ymin + m.hy * (double)cj;
        const double dx = (X - x1) * m.hxi, dy = (Y - y1) * m.hyi;
        const double xm = 1.0 - dx, ym = 1.0 - dy;
        W[0] = xm * ym; W[1] = dx * ym; W[2] = dx * dy; W[3] = xm * dy;
        return;
    }
    const double x1 = m.xmin + m.hx * (double)ci, x2 = m.xmin + m.hx * (double)(ci + 1);
    const double y1 = m.ymin + m.hy * (double)cj, y2 = m.ymin + m.hy * (double)(cj + 1);
    double hxp = x2 - x1, hyp = y2 - y1;
    double dx = (m.pow2_map && hxp == m.hx) ? (X - x1) * m.hxi : (X - x1) / hxp;
    double dy = (m.pow2_map && hyp == m.hy) ? (Y - y1) * m.hyi : (Y - y1) / hyp;
    double xm = 1.0 - dx, ym = 1.0 - dy;
    if (m.delta == 0) {
        W[0] = xm * ym; W[1] = dx * ym; W[2] = dx * dy; W[3] = xm * dy;
    } else {
        double beta = (fabs(y1 - m.ymin) < 1e-12) ? 0.75 : 1.0;
        double R1 = m.ymin + y1, R2 = m.ymin + y2, Rr = Y, den = R2 * R2 - R1 * R1;
        W[0] = xm * (R2 * R2 - Rr * Rr) / den * beta;
        W[1] = dx * (R2 * R2 - Rr * Rr) / den * beta;
        W[2] = dx * (Rr * Rr - R1 * R1) / den;
        W[3] = xm * (Rr * Rr - R1 * R1) / den;
    }
}

__device__ inline void deposit_weights(double X, double Y, int e, const MeshParams& m, double W[4]) {
    int ci, cj;
    pic_elem_ij(e, m, &ci, &cj);
    deposit_weights_ij(X, Y, ci, cj, m, W);
}

struct PushEvents { int emit, right, object, reflect, deleted; };

// Node coordinate lines in R, recomputed (no loads in the hot loop). The FP64 lines are
// xmin + hx i by construction (the set-up's formula, evaluated with the same two roundings);
// the FP32 lines are their FP32 roundings, evaluated in FP32 when the set-up found that exact
// for every node (nodes_f32), else as the FP64 formula rounded once to FP32 (the value the
// set-up stores in its FP32 node table).
template <typename R>
__device__ inline R node_x(const PushArgs<R>& a, int i) {
    if (sizeof(R) == 8 || !a.m.nodes_f32) return (R)(a.m.xmin + a.m.hx * (double)i);
    return a.x0R + a.hxR * (R)i;
}
template <typename R>
__device__ inline R node_y(const PushArgs<R>& a, int j) {
    if (sizeof(R) == 8 || !a.m.nodes_f32) return (R)(a.m.ymin + a.m.hy * (double)j);
    return a.y0R + a.hyR * (R)j;
}

// The loads of a particle's first pass that depend only on its element: the neighbourhood
// flags and the four corner fields. Issued for every slot of a thread before any slot is
// computed, so a thread waits for its gathers once.
template <typename R> struct HotIn { unsigned int nbr; R ex[4], ey[4]; };
template <typename R>
__device__ inline void hot_gather(const PushArgs<R>& a, int e, HotIn<R>& h) {
    const MeshParams& m = a.m;
    int ci, cj;
    pic_elem_ij(e, m, &ci, &cj);
    h.nbr = pic_ld(m.elem_nbr, e);
    // a 24-bit product (mesh sides < 2^24): the generic 32-bit multiply-add is a 64-bit
    // v_mad_u64_u32 whose addend pair can take the register of a load still in flight
    const int n1 = (int)__umul24((unsigned)ci, (unsigned)m.nny) + cj, n2 = n1 + m.nny, n3 = n2 + 1, n4 = n1 + 1;
    h.ex[0] = pic_ld(a.efx, n1); h.ex[1] = pic_ld(a.efx, n2); h.ex[2] = pic_ld(a.efx, n3); h.ex[3] = pic_ld(a.efx, n4);
    h.ey[0] = pic_ld(a.efy, n1); h.ey[1] = pic_ld(a.efy, n2); h.ey[2] = pic_ld(a.efy, n3); h.ey[3] = pic_ld(a.efy, n4);
}

// The whole-eV parts of a particle's energy quanta (pic_fix8 returned fx != 0): global
// atomics on the element's corner nodes (rare: W E >= 2^11 eV), and the range flag.
// The nodes come from the element index (not from registers the common path would have to
// keep live).
__device__ inline void deposit_hi(int fx, int e, const MeshParams& m, const long long hv[4], long long* acc_h,
                                  unsigned int* err) {
    if (fx & 2) atomicOr(err, PIC_EBIT_OVERFLOW);
    int ci, cj;
    pic_elem_ij(e, m, &ci, &cj);
    const int n1 = ci * m.nny + cj, nd[4] = {n1, n1 + m.nny, n1 + m.nny + 1, n1 + 1};
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (hv[k] != 0) atomicAdd((unsigned long long*)&acc_h[nd[k]], (unsigned long long)hv[k]);
}

// MoveAll for one particle (MCCInterface.f90:591-615 with MoveOne + AdjustParticleBoundary):
// state in registers, returns false when the particle is deleted.
template <typename R>
__device__ inline bool move_one(const PushArgs<R>& a, int64_t p, int e_old, R& X, R& Y, R& VX, R& VY, R& VZ, R& TH,
                                bool& vz_changed, PushEvents& ev) {
    const MeshParams& m = a.m;
    R T = a.dt;
    int ivel = 1, ipos = 1, it = 0;
    pic_stream rs; bool rs_init = false;
    bool emitted = false, absorbed_wall = false, absorbed_obj = false, reflected = false;
    while (ivel == 1 || ipos == 1) {
        if (++it > 64) {
            atomicOr(a.err, PIC_EBIT_BOUNDARY);
            pic_record(a.errinfo, PIC_EBIT_BOUNDARY, p, X, Y, VX, VY, it, e_old, -1, a.species);
            break;
        }
        // ---------------- MoveOne: gather + push (MoveOne.f90:77-146, 213-219)
        if (ivel == 1) {
            int n = e_old;
            int ci = n / m.ney, cj = n - ci * m.ney;
            R x1 = a.xsR[ci], x2 = a.xsR[ci + 1], y1 = a.ysR[cj], y2 = a.ysR[cj + 1];
            R hxp = x2 - x1, hyp = y2 - y1;
            R dx = (m.pow2_map && hxp == a.hxR) ? (X - x1) * a.hxiR : (X - x1) / hxp;
            R dy = (m.pow2_map && hyp == a.hyR) ? (Y - y1) * a.hyiR : (Y - y1) / hyp;
            R xm = (R)1 - dx, ym = (R)1 - dy;
            R W1, W2, W3, W4;
            if (m.delta == 0) {
                W1 = xm * ym; W2 = dx * ym; W3 = dx * dy; W4 = xm * dy;
            } else {
                // MoveOne.f90:94-104: R1, R2 = dymin + node radii, no axis factor here
                const R R1 = (R)m.ymin + y1, R2 = (R)m.ymin + y2, Rp = Y;
                const R den = R2 * R2 - R1 * R1;
                W1 = xm * (R2 * R2 - Rp * Rp) / den;
                W2 = dx * (R2 * R2 - Rp * Rp) / den;
                W3 = dx * (Rp * Rp - R1 * R1) / den;
                W4 = xm * (Rp * Rp - R1 * R1) / den;
            }
            int n1 = ci * m.nny + cj, n2 = n1 + m.nny, n3 = n2 + 1, n4 = n1 + 1;
            R Ex, Ey;
            if (m.element_index[n] > 0) {
                // Get_ParticleEField_PPR.f90:19-33: vacuum-side nodes only
                R ex = 0, ey = 0;
                int count = 0;
                const int nd[4] = {n1, n2, n3, n4};
                const R nx_[4] = {x1, x2, x2, x1}, ny_[4] = {y1, y1, y2, y2};
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    if (m.node_index[nd[k]] == -1) {
                        const R* c = a.coef5 + (size_t)nd[k] * 5;
                        count++;
                        R xr = X - nx_[k], yr = Y - ny_[k];
                        ex = ex + c[0] + (R)2 * c[2] * xr + c[3] * yr;
                        ey = ey + c[1] + (R)2 * c[4] * yr + c[3] * xr;
                    }
                }
                ex = ex / (R)count;
                ey = ey / (R)count;
                Ex = -ex; Ey = -ey;
            } else {
                Ex = W1 * a.efx[n1] + W2 * a.efx[n2] + W3 * a.efx[n3] + W4 * a.efx[n4];
                Ey = W1 * a.efy[n1] + W2 * a.efy[n2] + W3 * a.efy[n3] + W4 * a.efy[n4];
            }
            if (m.delta == 0) {
                VX = VX + a.qm * Ex * T;
                VY = VY + a.qm * Ey * T;
            } else {
                // MoveOne.f90:221-236: E_r turned into the fixed Cartesian frame by the
                // azimuth; the tangential field Efield(3) = efz is zero in the electrostatic model
                double sn, cs;
                pic_sincos((double)TH, &sn, &cs);
                const R xe = Ey * (R)cs - (R)0 * (R)sn;
                const R ye = Ey * (R)sn + (R)0 * (R)cs;
                VX = VX + a.qm * Ex * T;
                VY = VY + a.qm * xe * T;
                VZ = VZ + a.qm * ye * T;
                vz_changed = true;
            }
            ivel = 0;
        }
        if (m.delta == 0) {
            X = X + T * VX;
            Y = Y + T * VY;
        } else {
            // MoveOne.f90:246-263: move in Cartesian coordinates (in FP64), back to (r, th)
            double sn, cs;
            pic_sincos((double)TH, &sn, &cs);
            double xc = (double)Y * cs, yc = (double)Y * sn;
            xc = xc + (double)T * (double)VY;
            if (xc == 0.0) xc = m.hy * (double)1.0e-5f;      // hx(2)*1.0E-5 (a default-real literal)
            yc = yc + (double)T * (double)VZ;
            X = X + T * VX;
            const double rr = sqrt(xc * xc + yc * yc);
            double tt = pic_atan(yc / xc);
            if (xc <= 0.0) tt = tt + 3.14159265358979;         // PI_LY
            Y = (R)rr; TH = (R)tt;
        }
        ipos = 0;
        // ---------------- AdjustParticleBoundary (MCCInterface.f90:361-499)
        // Back-up times lie in [0, T] whenever the reference's arithmetic is well posed;
        // outside [-T/2, 2T] (possible only after the re-emission has divided Vy by v_ref
        // again) the position is folded into the domain instead.
        if (X <= a.xmin) {
            ipos = 1;
            R TR = (X - a.xmin) / VX;
            bool ok = TR >= (R)-0.5 * T && TR <= (R)2 * T;
            if (ok) { T = TR; Y = Y - VY * TR; }
            X = a.xmin + a.emit_eps;
            if (!rs_init) { pic_stream_init(&rs, a.seed, (uint32_t)p, a.step, a.tag_emit); rs_init = true; }
            double R1 = pic_stream_u01(&rs);
            // Kappa = 2: ranf**(-1/(Kappa-1)) == 1/ranf
            double vxs = sqrt(a.kappa_c * (1.0 / R1 - 1.0));
            VX = (R)(vxs * a.vref_inv);
            VY = (R)((double)VY * a.vref_inv);
            VZ = (R)((double)VZ * a.vref_inv);
            emitted = true;
            vz_changed = true;
            if (!ok) { T = 0; ipos = 0; fold_y(Y, VY, a.ymin, a.ymax); }
        } else if (X > a.xmax) {
            X = (R)-2000;
            absorbed_wall = true;
        } else if (Y < a.ymin) {
            if (m.delta == 0) {
                ipos = 1;
                R TR = (Y - a.ymin) / VY;
                reflected = true;
                if (TR >= (R)-0.5 * T && TR <= (R)2 * T) {
                    T = TR;
                    X = X - VX * TR;
                    VY = -VY;
                    Y = a.ymin + a.emit_eps;
                    if (Y > a.ymax) atomicOr(a.err, PIC_EBIT_BOUNDARY);
                } else {
                    T = 0; ipos = 0; fold_y(Y, VY, a.ymin, a.ymax);
                }
            } else {
                atomicOr(a.err, PIC_EBIT_WRONGCROSS);
            }
        } else if (Y > a.ymax) {
            if (m.delta == 0) {
                ipos = 1;
                R TR = (Y - a.ymax) / VY;
                reflected = true;
                if (TR >= (R)-0.5 * T && TR <= (R)2 * T) {
                    T = TR;
                    X = X - VX * TR;
                    VY = -VY;
                    Y = a.ymax - a.emit_eps;
                    if (Y < a.ymin) atomicOr(a.err, PIC_EBIT_BOUNDARY);
                } else {
                    T = 0; ipos = 0; fold_y(Y, VY, a.ymin, a.ymax);
                }
            } else {
                // axisymmetric outer radius: the reference has no branch (the particle would
                // leave the mesh); absorbed like the x > xmax wall
                X = (R)-2000;
            }
        }
        for (int k = 0; k < m.n_objects; ++k) {
            const DevObject o = a.objs[k];
            double xd = (double)X, yd = (double)Y;
            bool in = false;
            if (o.shape == 3) in = xd > o.x0 && xd < o.x1 && yd > o.y0 && yd < o.y1;
            else if (o.shape == 1) in = (xd - o.x0) * (xd - o.x0) + (yd - o.y0) * (yd - o.y0) < o.r2;
            if (in) { X = (R)-2000; absorbed_obj = true; }
        }
    }
    bool live = !(X <= a.xmin);          // MoveAll deletion test, MCCInterface.f90:617-620
    if (!(X == X) || !(Y == Y) || !(VX == VX) || !(VY == VY)) { atomicOr(a.err, PIC_EBIT_NONFINITE); live = false; }
    ev.emit += emitted; ev.reflect += reflected;
    ev.right += (absorbed_wall && !live); ev.object += (absorbed_obj && !live); ev.deleted += !live;
    return live;
}

// First pass of MoveAll's loop for one particle (gather + push, MoveOne.f90:77-146, 213-219,
// 242-245) and the element lookup of the new position, for the common case only. Returns
// true (state untouched) when the particle needs any other branch of MoveOne /
// AdjustParticleBoundary: its element is cut by or near an object (the object-side field
// of Get_ParticleEField_PPR.f90:19-33), the pass ends outside [xmin, xmax] x [ymin, ymax]
// (walls), or it ends in an element near an object (the object test,
// MCCInterface.f90:486-499, could fire) or the lookup fails. The wall kernel then runs the
// whole loop from the original state with the same arithmetic. Otherwise the reference's
// loop ends after this pass with no boundary or object event, and *e_new is the element.
template <typename R>
__device__ inline bool push_hot(const PushArgs<R>& a, int e_old, const HotIn<R>& h, R& X, R& Y, R& VX, R& VY,
                                int* e_new, int* ci_new, int* cj_new) {
    const MeshParams& m = a.m;
    const R T = a.dt;
    int ci, cj;
    pic_elem_ij(e_old, m, &ci, &cj);
    const unsigned int nbr = h.nbr;                           // flags of the 3x3 neighbourhood
    const R x1 = node_x(a, ci), y1 = node_y(a, cj);
    const R e1x = h.ex[0], e2x = h.ex[1], e3x = h.ex[2], e4x = h.ex[3];
    const R e1y = h.ey[0], e2y = h.ey[1], e3y = h.ey[2], e4y = h.ey[3];
    R dx, dy;
    if (sizeof(R) == 4 ? m.unif_r32 : m.unif_r64) {
        // every element's spacing is hxR (hyR) exactly: the multiplication by the reciprocal
        // below, without the far node lines
        dx = (X - x1) * a.hxiR;
        dy = (Y - y1) * a.hyiR;
    } else {
        const R x2 = node_x(a, ci + 1), y2 = node_y(a, cj + 1);
        R hxp = x2 - x1, hyp = y2 - y1;
        dx = (m.pow2_map && hxp == a.hxR) ? (X - x1) * a.hxiR : (X - x1) / hxp;
        dy = (m.pow2_map && hyp == a.hyR) ? (Y - y1) * a.hyiR : (Y - y1) / hyp;
    }
    R xm = (R)1 - dx, ym = (R)1 - dy;
    R W1 = xm * ym, W2 = dx * ym, W3 = dx * dy, W4 = xm * dy;
    R Ex = W1 * e1x + W2 * e2x + W3 * e3x + W4 * e4x;
    R Ey = W1 * e1y + W2 * e2y + W3 * e3y + W4 * e4y;
    R VX1 = VX + a.qm * Ex * T;
    R VY1 = VY + a.qm * Ey * T;
    R X1 = X + T * VX1;
    R Y1 = Y + T * VY1;
    // the pass is computed for every particle and the element flag tested with the walls,
    // so no branch separates the flag load from the field loads
    if (!(nbr & 16u) || X1 <= a.xmin || X1 > a.xmax || Y1 < a.ymin || Y1 > a.ymax) return true;
    int bad = 0;
    int en = (sizeof(R) == 4 && m.locate_f32) ? pic_locate_inside_f32((float)X1, (float)Y1, m, &bad, ci_new, cj_new)
                                               : pic_locate_ij((double)X1, (double)Y1, m, &bad, ci_new, cj_new);
    // the new element's flag from the neighbourhood mask (no second load); a particle that
    // moved further than one element is handed over like a flagged one
    const int di = *ci_new - ci + 1, dj = *cj_new - cj + 1;
    if (bad || (unsigned int)di > 2u || (unsigned int)dj > 2u || !((nbr >> (3 * di + dj)) & 1u)) return true;
    X = X1; Y = Y1; VX = VX1; VY = VY1;
    *e_new = en;
    return false;
}

// The axisymmetric first pass (MoveOne.f90:94-104 gather weights, :221-236 the radial field
// turned by the azimuth, :246-263 the move in Cartesian coordinates and back to (r, th)) for
// the common case, with the same arithmetic as move_one so that both paths give the same
// bits; the conditions for handing over are those of push_hot (r < ymin is the reference's
// wrong-crossing error, r > ymax an absorption: both go to the wall kernel).
template <typename R>
__device__ inline bool push_hot_rz(const PushArgs<R>& a, int e_old, const HotIn<R>& h, R& X, R& Y, R& VX, R& VY,
                                   R& VZ, R& TH, int* e_new, int* ci_new, int* cj_new) {
    const MeshParams& m = a.m;
    const R T = a.dt;
    int ci, cj;
    pic_elem_ij(e_old, m, &ci, &cj);
    const unsigned int nbr = h.nbr;
    const R x1 = node_x(a, ci), y1 = node_y(a, cj), y2 = node_y(a, cj + 1);
    const R e1x = h.ex[0], e2x = h.ex[1], e3x = h.ex[2], e4x = h.ex[3];
    const R e1y = h.ey[0], e2y = h.ey[1], e3y = h.ey[2], e4y = h.ey[3];
    R dx;
    if (sizeof(R) == 4 ? m.unif_r32 : m.unif_r64) {
        dx = (X - x1) * a.hxiR;
    } else {
        const R hxp = node_x(a, ci + 1) - x1;
        dx = (m.pow2_map && hxp == a.hxR) ? (X - x1) * a.hxiR : (X - x1) / hxp;
    }
    R xm = (R)1 - dx;
    const R R1 = (R)m.ymin + y1, R2 = (R)m.ymin + y2, Rp = Y;
    const R den = R2 * R2 - R1 * R1;
    const R W1 = xm * (R2 * R2 - Rp * Rp) / den;
    const R W2 = dx * (R2 * R2 - Rp * Rp) / den;
    const R W3 = dx * (Rp * Rp - R1 * R1) / den;
    const R W4 = xm * (Rp * Rp - R1 * R1) / den;
    const R Ex = W1 * e1x + W2 * e2x + W3 * e3x + W4 * e4x;
    const R Ey = W1 * e1y + W2 * e2y + W3 * e3y + W4 * e4y;
    double sn, cs;
    pic_sincos((double)TH, &sn, &cs);
    const R xe = Ey * (R)cs - (R)0 * (R)sn;
    const R ye = Ey * (R)sn + (R)0 * (R)cs;
    const R VX1 = VX + a.qm * Ex * T;
    const R VY1 = VY + a.qm * xe * T;
    const R VZ1 = VZ + a.qm * ye * T;
    double xc = (double)Y * cs, yc = (double)Y * sn;
    xc = xc + (double)T * (double)VY1;
    if (xc == 0.0) xc = m.hy * (double)1.0e-5f;
    yc = yc + (double)T * (double)VZ1;
    const R X1 = X + T * VX1;
    const double rr = sqrt(xc * xc + yc * yc);
    double tt = pic_atan(yc / xc);
    if (xc <= 0.0) tt = tt + 3.14159265358979;
    const R Y1 = (R)rr;
    if (!(nbr & 16u) || X1 <= a.xmin || X1 > a.xmax || Y1 < a.ymin || Y1 > a.ymax) return true;
    int bad = 0;
    int en = (sizeof(R) == 4 && m.locate_f32) ? pic_locate_inside_f32((float)X1, (float)Y1, m, &bad, ci_new, cj_new)
                                               : pic_locate_ij((double)X1, (double)Y1, m, &bad, ci_new, cj_new);
    const int di = *ci_new - ci + 1, dj = *cj_new - cj + 1;
    if (bad || (unsigned int)di > 2u || (unsigned int)dj > 2u || !((nbr >> (3 * di + dj)) & 1u)) return true;
    X = X1; Y = Y1; VX = VX1; VY = VY1; VZ = VZ1; TH = (R)tt;
    *e_new = en;
    return false;
}

// The particles the push kernel handed over: the whole MoveOne/AdjustParticleBoundary loop,
// lookup, write-back, and the deposit merged per wave (segmented by element, the list is
// in slot order within each wave's append, so neighbours mostly share an element) before
// the global atomics. Grid-stride over the list, every lane of a wave in every round.
template <typename R>
__global__ __launch_bounds__(256) void k_push_walls(PushArgs<R> a) {
    const MeshParams& m = a.m;
    const unsigned int nw = *a.wall_count;
    const DepWindow none = {0, 0, 0, 0};
    const unsigned int lane = threadIdx.x & 63;
    const unsigned int wave0 = (blockIdx.x * blockDim.x + threadIdx.x) - lane;
    const unsigned int stride = gridDim.x * blockDim.x;
    PushEvents ev = {0, 0, 0, 0, 0};
    for (unsigned int base = wave0; base < nw; base += stride) {
        const unsigned int i = base + lane;
        int key = -1;
        long long v[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) v[q] = 0;
        const int64_t p = i < nw ? (int64_t)a.wall_list[i] : 0;
        const int e_old = i < nw ? a.cell[p] : -1;
        if (e_old >= 0) {
            R X = a.x[p], Y = a.y[p], VX = a.vx[p], VY = a.vy[p], VZ = a.vz[p];
            R TH = a.th ? a.th[p] : (R)0;
            bool vzc = false;
            bool live = move_one<R>(a, p, e_old, X, Y, VX, VY, VZ, TH, vzc, ev);
            a.x[p] = X; a.y[p] = Y; a.vx[p] = VX; a.vy[p] = VY;
            if (vzc) a.vz[p] = VZ;
            if (a.th) a.th[p] = TH;
            if (live) {
                int bad = 0;
                double Xd = (double)X, Yd = (double)Y;
                key = pic_locate_elem(Xd, Yd, m, &bad);
                if (bad) {
                    atomicOr(a.err, PIC_EBIT_POSITION);
                    pic_record(a.errinfo, PIC_EBIT_POSITION, p, X, Y, VX, VY, 0, e_old, key, a.species);
                }
                double W[4];
                deposit_weights(Xd, Yd, key, m, W);
                double vxd = (double)VX, vyd = (double)VY, vzd = (double)VZ;
                const double energy = (vxd * vxd + vyd * vyd + vzd * vzd) * a.ek_c;
                long long hv[4];
                int fx = 0;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    v[q] = pic_fix(W[q]);
                    if (!pic_fix_split(W[q] * energy, &v[4 + q], &hv[q])) fx |= 2;
                    if (hv[q] != 0) fx |= 1;
                }
                if (fx) deposit_hi(fx, key, m, hv, a.acc_h, a.err);
            }
            a.cell_next[p] = key;
        }
        wave_deposit(key, m, v, a.acc_w, a.acc_e, none, nullptr, nullptr);
    }
    wave_add(&a.counters[PIC_C_EMIT], ev.emit);
    wave_add(&a.counters[PIC_C_RIGHT], ev.right);
    wave_add(&a.counters[PIC_C_OBJECT], ev.object);
    wave_add(&a.counters[PIC_C_REFLECT], ev.reflect);
    wave_add(&a.counters[PIC_C_DELETED], ev.deleted);
}

// Fused mover + lookup + deposition (DO_PUSH) or lookup + deposition only.
// Register budget per SIMD: the general Cartesian FP32 pass at 4 slots per thread is held to 5
// waves (96 VGPRs; left alone the allocation lands at 98, i.e. 4 waves, and the fifth wave is
// worth more than the two registers spilled around the slot loop); the axisymmetric and FP64
// passes keep 4 (at 5 they spill 9 to 24 registers for no gain). The 8-slot specialisation is
// asked for 4 and allocates 93 VGPRs, so it runs 5 waves without spills (asked for 5 with all
// eight cells held up front, it spilled 36 B per lane).
#define PUSH_WAVES(R, RZ, PPT) ((sizeof(R) == 4 && !(RZ) && (PPT) == 4) ? 5 : 4)
// FAST: the mesh flags of the common layout (power-of-two spacing and map, FP32 lookup and
// node lines exact, every element's spacing exact, fewer than 2^24 elements) are compile-time
// constants, so the general branches and the scalars only they read drop out of the kernel
// (the host launches FAST only where set-up found all of these true; same arithmetic).
// PPT: slots per thread, matching the cell sort's groups of 64 PPT slots (pic_push_ppt). With
// the general branches gone the FP32 Cartesian kernel holds 8 slots at 5 waves per SIMD without
// spills, which halves the per-block costs (window zeroing and flush, the centre element,
// the first loads' latency) per particle.
template <typename R, int DO_PUSH, int RZ, int FAST, int PPT>
__global__ __launch_bounds__(PUSH_BLOCK) __attribute__((amdgpu_waves_per_eu(PUSH_WAVES(R, RZ, PPT), 8)))
void k_push_deposit(PushArgs<R> a_in) {
    __shared__ long long winW[WIN_NODES], winE[WIN_NODES];
    __shared__ int s_center;
    PushArgs<R> a = a_in;
    if (FAST) {
        a.m.pow2_map = 1; a.m.locate_f32 = 1; a.m.nodes_f32 = 1; a.m.unif_r32 = 1; a.m.unif_r64 = 1;
        a.m.ij_exact_div = 0;
    }
    const MeshParams& m = a.m;
    const int64_t base = (int64_t)blockIdx.x * (PUSH_BLOCK * PPT);
    // window around the element of the first live particle at or after the block's middle
    // slot (WIN_NX x WIN_NY nodes)
    if (threadIdx.x < 64) {
        int64_t mid = base + (PUSH_BLOCK * PPT) / 2;
        if (mid >= a.n) mid = base;
        int64_t q = mid + threadIdx.x;
        int c = (q < a.n) ? a.cell[q] : -1;
        unsigned long long b = __ballot(c >= 0);
        int first = b ? (__ffsll((long long)b) - 1) : -1;
        int cc = __shfl(c, first < 0 ? 0 : first, 64);
        if (threadIdx.x == 0) s_center = (first < 0) ? -1 : cc;
    }
    __syncthreads();
    DepWindow w;
    w.nx = 0; w.ny = 0; w.x0 = 0; w.y0 = 0;
    if (s_center >= 0) {
        int ci = s_center / m.ney, cj = s_center - ci * m.ney;
        int tj = cj / SORT_TILE;
        int x0 = ci - PUSH_WIN_MARGIN, y0 = tj * SORT_TILE - PUSH_WIN_MARGIN;
        int x1 = x0 + WIN_NX - 1, y1 = y0 + WIN_NY - 1;
        if (x0 < 0) x0 = 0;
        if (y0 < 0) y0 = 0;
        if (x1 > m.nnx - 1) x1 = m.nnx - 1;
        if (y1 > m.nny - 1) y1 = m.nny - 1;
        w.x0 = x0; w.y0 = y0; w.nx = x1 - x0 + 1; w.ny = y1 - y0 + 1;
    }
    const int wn = w.nx * w.ny;
    for (int k = threadIdx.x; k < wn; k += PUSH_BLOCK) { winW[k] = 0; winE[k] = 0; }

    // ---------------- move, locate, deposit (run accumulation in registers)
    // Wave w of the block owns the 64 PPT-slot group at wbase; iteration k handles
    // slot wbase + 64 k + lane, so every load and store is lane-contiguous. The cell sort lays
    // a full group out transposed (sorted position PPT lane + k at slot 64 k + lane), so
    // a thread's iterations walk consecutive sorted particles, which mostly share an element.
    // Slot arrays are addressed from the block's first slot with 32-bit offsets.
    const int lane = threadIdx.x & 63;
    const unsigned int o0 = (threadIdx.x >> 6) * (64 * PPT) + lane;
    const int64_t nrem = a.n - base;
    const unsigned int nb = nrem < (int64_t)(PUSH_BLOCK * PPT) ? (unsigned int)nrem : (unsigned int)(PUSH_BLOCK * PPT);
    const int* cellb = a.cell + base;
    int* nextb = a.cell_next + base;
    R* xb = a.x + base; R* yb = a.y + base; R* vxb = a.vx + base; R* vyb = a.vy + base;
    R* vzb = a.vz + base;
    R* thb = RZ ? a.th + base : nullptr;
    int rkey = -1;
    long long acc[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) acc[q] = 0;
    // Software pipeline over the PPT slots: the cells of the first two slots are loaded up
    // front (one wait); while slot k is computed, the cell of slot k + 2 and then the state and
    // the element gathers of slot k + 1 are in flight, so a slot waits only for its own loads
    // (the cell load comes first, so waiting for it never waits for the state loads issued
    // after it) and at most three cells are live in registers (8 slots per thread then fit in
    // 93 registers, 5 waves per SIMD). No branch surrounds a load: a slot past the end reads
    // the block's first slot (nb >= 1) and is marked dead, a dead slot's gathers read element
    // 0's; the loop is unrolled so the per-slot registers are indexed statically.
    int cel[PPT];
    auto load_cell = [&](int k) {
        const unsigned int o = o0 + 64 * k;
        const int c = pic_ld(cellb, o < nb ? o : 0u);
        cel[k] = o < nb ? c : -1;
    };
    load_cell(0);
    if (PPT > 1) load_cell(1);
    // vmcnt(0) on every path, so the compiler tracks no cell load as outstanding past here
    __builtin_amdgcn_s_waitcnt(0x0F70);
    R nx, ny, nvx, nvy, nvz, nth;    // state of the slot issued next
    HotIn<R> nh{};                   // its element gathers (DO_PUSH)
    auto issue = [&](int k) {
        const unsigned int o = o0 + 64 * k;
        const unsigned int oc = o < nb ? o : 0u;
        nx = pic_ld((const R*)xb, oc); ny = pic_ld((const R*)yb, oc);
        nvx = pic_ld((const R*)vxb, oc); nvy = pic_ld((const R*)vyb, oc);
        nvz = pic_ld((const R*)vzb, oc);
        nth = (RZ && DO_PUSH) ? pic_ld((const R*)thb, oc) : (R)0;
        if (DO_PUSH) hot_gather<R>(a, cel[k] >= 0 ? cel[k] : 0, nh);
    };
    issue(0);
    unsigned int defer_bits = 0;     // slot k handed to the wall kernel: bit k
    __syncthreads();     // window zeroed before any deposit
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        const unsigned int o = o0 + 64 * k;
        const bool in = o < nb;
        const int e_old = cel[k];
        R X = nx, Y = ny, VX = nvx, VY = nvy, VZ = nvz, TH = nth;
        const HotIn<R> hin = nh;
        if (k + 2 < PPT) load_cell(k + 2);
        if (k + 1 < PPT) issue(k + 1);
        int next = -1, nci = 0, ncj = 0;
        bool dep = false;
        if (DO_PUSH) {
            bool defer = false;
            if (e_old >= 0) {
                defer = RZ ? push_hot_rz<R>(a, e_old, hin, X, Y, VX, VY, VZ, TH, &next, &nci, &ncj)
                           : push_hot<R>(a, e_old, hin, X, Y, VX, VY, &next, &nci, &ncj);
                dep = !defer;
            }
            defer_bits |= (defer ? 1u : 0u) << k;
        } else if (e_old >= 0) {
            int bad = 0;
            next = pic_locate_ij((double)X, (double)Y, m, &bad, &nci, &ncj);
            if (bad) {
                atomicOr(a.err, PIC_EBIT_POSITION);
                pic_record(a.errinfo, PIC_EBIT_POSITION, base + o, X, Y, VX, VY, 0, e_old, next, a.species);
            }
            dep = true;
        }
        if (dep) {
            // GetParChg_2D.f90:788-843: weights and kinetic energy (ParticleOne.f90:339-340, :128)
            double W[4];
            deposit_weights_ij((double)X, (double)Y, nci, ncj, m, W);
            double vxd = (double)VX, vyd = (double)VY, vzd = (double)VZ;
            const double energy = (vxd * vxd + vyd * vyd + vzd * vzd) * a.ek_c;
            long long qv[8];
            pic_fix8(W, energy, qv, next, m, a.acc_h, a.err);
            if (next != rkey) {
                if (rkey >= 0) add_element(rkey, acc, m, a.acc_w, a.acc_e, w, winW, winE);
                rkey = next;
#pragma unroll
                for (int q = 0; q < 8; ++q) acc[q] = qv[q];
            } else {
#pragma unroll
                for (int q = 0; q < 8; ++q) acc[q] += qv[q];
            }
        }
        // write back (a deferred particle keeps its old state for the wall kernel, which
        // also writes its cell_next)
        if (in) {
            pic_st(nextb, o, next);
            if (DO_PUSH && dep) {
                pic_st(xb, o, X); pic_st(yb, o, Y); pic_st(vxb, o, VX); pic_st(vyb, o, VY);
                if (RZ) { pic_st(vzb, o, VZ); pic_st(thb, o, TH); }
            }
        }
    }
    // each thread's last run straight to the window (LDS atomics; lanes of one element add
    // to the same eight words, which the LDS serialises)
    if (rkey >= 0) add_element(rkey, acc, m, a.acc_w, a.acc_e, w, winW, winE);
    // the handed-over slots of all PPT iterations appended to the wall list with one
    // wave-aggregated atomic (its returned base is the only wait; the wall kernel treats the
    // list as an unordered set)
    if (DO_PUSH) {
        unsigned long long wm[PPT];
        unsigned int tot = 0;
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
            wm[k] = __ballot((defer_bits >> k) & 1u);
            tot += (unsigned int)__popcll(wm[k]);
        }
        if (tot) {
            unsigned int wb = 0;
            if (lane == 0) wb = atomicAdd(a.wall_count, tot);
            wb = __shfl(wb, 0, 64);
            const unsigned long long below = (1ull << lane) - 1ull;
#pragma unroll
            for (int k = 0; k < PPT; ++k) {
                if ((defer_bits >> k) & 1u) a.wall_list[wb + __popcll(wm[k] & below)] = (int)(base + o0 + 64 * k);
                wb += (unsigned int)__popcll(wm[k]);
            }
        }
    }

    // ---------------- flush the window
    __syncthreads();
    for (int k = threadIdx.x; k < wn; k += PUSH_BLOCK) {
        long long vw = winW[k], ve = winE[k];
        if (vw == 0 && ve == 0) continue;
        int i = k / w.ny, j = k - i * w.ny;
        int node = (w.x0 + i) * m.nny + (w.y0 + j);
        if (vw != 0) atomicAdd((unsigned long long*)&a.acc_w[node], (unsigned long long)vw);
        if (ve != 0) atomicAdd((unsigned long long*)&a.acc_e[node], (unsigned long long)ve);
    }
}

template <typename R>
static PushArgs<R> make_push_args(int isp) {
    SpeciesHost& S = g.sp[isp];
    PushArgs<R> a;
    a.x = (R*)S.x; a.y = (R*)S.y; a.vx = (R*)S.vx; a.vy = (R*)S.vy; a.vz = (R*)S.vz;
    a.th = g.mesh.delta == 1 ? (R*)S.th : nullptr;
    a.cell = S.cell; a.cell_next = S.cell_next; a.n = S.nslots;
    a.m = g.mesh;
    a.m.xs = g.d_xs; a.m.ys = g.d_ys; a.m.xsf = g.d_xsf; a.m.ysf = g.d_ysf;
    a.m.element_index = g.d_element_index; a.m.node_index = g.d_node_index;
    a.m.elem_clean = g.d_elem_clean;
    a.m.elem_nbr = g.d_elem_nbr;
    a.objs = g.d_objects;
    if (sizeof(R) == 8) {
        a.efx = (const R*)g.d_efx; a.efy = (const R*)g.d_efy; a.coef5 = (const R*)g.d_coef5;
        a.xsR = (const R*)g.d_xs; a.ysR = (const R*)g.d_ys;
    } else {
        a.efx = (const R*)g.d_efxf; a.efy = (const R*)g.d_efyf; a.coef5 = (const R*)g.d_coeff;
        a.xsR = (const R*)g.d_xsf; a.ysR = (const R*)g.d_ysf;
    }
    a.qm = (R)S.qm; a.dt = (R)g.cfg.dt;
    a.xmin = (R)g.mesh.xmin; a.xmax = (R)g.mesh.xmax; a.ymin = (R)g.mesh.ymin; a.ymax = (R)g.mesh.ymax;
    a.emit_eps = (R)(double)10.0e-6f;
    a.hxR = (R)g.mesh.hx; a.hxiR = (R)g.mesh.hxi; a.hyR = (R)g.mesh.hy; a.hyiR = (R)g.mesh.hyi;
    a.x0R = (R)g.mesh.xmin; a.y0R = (R)g.mesh.ymin;
    a.vref_inv = 1.0 / g.v_ref;
    double beta = S.p.mass_kg / (2.0 * PIC_KB * S.p.init_temperature_K);
    a.kappa_c = (1.0 * 2.0 - 1.5) / beta;
    // ParticleOne.f90:333-341 with GetParChg_2D.f90:128 (0.5 Mass |v|^2 VFactor^2 / JtoeV), the
    // constant factors multiplied once here instead of per particle (the product differs from the
    // reference's operation order by at most a few ulps; the oracle uses the same constant)
    a.ek_c = 0.5 * S.p.mass_kg * g.v_ref * g.v_ref / PIC_JTOEV;
    a.seed = g.seed; a.step = (uint32_t)g.step; a.tag_emit = pic_stream_tag(isp, PIC_RNG_EMIT);
    a.acc_w = S.acc_w; a.acc_e = S.acc_e; a.acc_h = S.acc_h;
    a.counters = S.counters; a.err = g.d_err;
    a.errinfo = (PicErrInfo*)g.d_errinfo; a.species = isp;
    a.wall_list = g.d_wall_list; a.wall_count = g.d_wall_count;
    return a;
}

__global__ void k_zero_ll(long long* p, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

int launch_zero_acc(int isp) {
    SpeciesHost& S = g.sp[isp];
    int64_t n = g.mesh.nnodes;
    int blocks = (int)((n + 255) / 256);
    hipLaunchKernelGGL(k_zero_ll, dim3(blocks), dim3(256), 0, g.stream, S.acc_w, n);
    hipLaunchKernelGGL(k_zero_ll, dim3(blocks), dim3(256), 0, g.stream, S.acc_e, n);
    hipLaunchKernelGGL(k_zero_ll, dim3(blocks), dim3(256), 0, g.stream, S.acc_h, n);
    HIPCHK(hipGetLastError());
    S.acc_dirty = false;
    return PIC_OK;
}

// The mesh flags of the common layout hold (the FAST kernels' compile-time constants).
static bool push_fast_layout() {
    const MeshParams& mm = g.mesh;
    return mm.pow2_map && !mm.ij_exact_div && mm.unif_r64 &&
           (g.particle_bits == 64 || (mm.locate_f32 && mm.nodes_f32 && mm.unif_r32));
}

int pic_push_ppt() {
    return (g.particle_bits == 32 && g.mesh.delta == 0 && push_fast_layout()) ? PIC_PPT_WIDE : PIC_PPT_NARROW;
}

int launch_push_deposit(int isp, bool do_push) {
    SpeciesHost& S = g.sp[isp];
    if (S.acc_dirty) { int rc = launch_zero_acc(isp); if (rc) return rc; }
    if (S.nslots == 0) { S.acc_dirty = true; return PIC_OK; }
    const bool rz = g.mesh.delta == 1;
    const bool fast = push_fast_layout() && !rz;
    const int ppt = do_push ? pic_push_ppt() : PIC_PPT_NARROW;
    const int blocks = (int)((S.nslots + PUSH_BLOCK * ppt - 1) / (PUSH_BLOCK * ppt));
    if (do_push) {
        // the wall list holds at most every slot once
        if (g.wall_cap < S.nslots) {
            if (g.d_wall_list) (void)hipFree(g.d_wall_list);
            g.d_wall_list = nullptr;
            HIPCHK(hipMalloc(&g.d_wall_list, sizeof(int) * (size_t)S.capacity));
            g.wall_cap = S.capacity;
        }
        if (!g.d_wall_count) HIPCHK(hipMalloc(&g.d_wall_count, 64 * sizeof(unsigned int)));
        HIPCHK(hipMemsetAsync(g.d_wall_count, 0, sizeof(unsigned int), g.stream));
        HIPCHK(hipEventRecord(g.ev_push[isp][0], g.stream));
    }
    const int wall_blocks = 512;
    if (do_push && rz && !S.th) return pic_fail(PIC_ERR_STATE, "axisymmetric species without azimuth storage");
    // the fused first pass, then the full MoveOne / AdjustParticleBoundary loop for the
    // particles it handed over (axisymmetric: the r-z pass of MoveOne)
    const dim3 grid(blocks), block(PUSH_BLOCK);
    if (g.particle_bits == 64) {
        PushArgs<double> a = make_push_args<double>(isp);
        if (do_push) {
            if (rz) hipLaunchKernelGGL((k_push_deposit<double, 1, 1, 0, 4>), grid, block, 0, g.stream, a);
            else if (fast) hipLaunchKernelGGL((k_push_deposit<double, 1, 0, 1, 4>), grid, block, 0, g.stream, a);
            else hipLaunchKernelGGL((k_push_deposit<double, 1, 0, 0, 4>), grid, block, 0, g.stream, a);
            HIPCHK(hipEventRecord(g.ev_push[isp][1], g.stream));
            hipLaunchKernelGGL((k_push_walls<double>), dim3(wall_blocks), dim3(256), 0, g.stream, a);
        } else {
            hipLaunchKernelGGL((k_push_deposit<double, 0, 0, 0, 4>), grid, block, 0, g.stream, a);
        }
    } else {
        PushArgs<float> a = make_push_args<float>(isp);
        if (do_push) {
            if (rz) hipLaunchKernelGGL((k_push_deposit<float, 1, 1, 0, 4>), grid, block, 0, g.stream, a);
            else if (fast) hipLaunchKernelGGL((k_push_deposit<float, 1, 0, 1, 8>), grid, block, 0, g.stream, a);
            else hipLaunchKernelGGL((k_push_deposit<float, 1, 0, 0, 4>), grid, block, 0, g.stream, a);
            HIPCHK(hipEventRecord(g.ev_push[isp][1], g.stream));
            hipLaunchKernelGGL((k_push_walls<float>), dim3(wall_blocks), dim3(256), 0, g.stream, a);
        } else {
            hipLaunchKernelGGL((k_push_deposit<float, 0, 0, 0, 4>), grid, block, 0, g.stream, a);
        }
    }
    HIPCHK(hipGetLastError());
    if (do_push) { HIPCHK(hipEventRecord(g.ev_push[isp][2], g.stream)); g.ev_push_pending[isp] = true; }
    S.acc_dirty = true;
    return PIC_OK;
}

// ---------------------------------------------------------------- finalisation
// GetParChg_2D.f90:849-911: node sums are already formed (fixed point); convert, check
// Ek != 0 with rho == 0 (the reference STOPs), normalise and total over species; the
// accumulators are cleared for the next deposition.
struct FinalArgs {
    int nnodes, nspecies, delta;
    long long* acc_w[PIC_MAX_SPECIES];
    long long* acc_e[PIC_MAX_SPECIES];
    long long* acc_h[PIC_MAX_SPECIES];
    double weight_qs[PIC_MAX_SPECIES];
    double weight[PIC_MAX_SPECIES];
    double L_ref, n_ref;
    const double* vol;
    double *rho, *rho_s, *rho_p, *ek;
    double* eks;              // Ek_s: mean energy per particle (eV) at the node, before normalisation
    unsigned int* err;
};

__global__ void k_finalize(FinalArgs a) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= a.nnodes) return;
    double vol = a.vol[n];
    double rho = 0.0;
    for (int s = 0; s < a.nspecies; ++s) {
        long long qw = a.acc_w[s][n], qe = a.acc_e[s][n], qh = a.acc_h[s][n];
        a.acc_w[s][n] = 0; a.acc_e[s][n] = 0; a.acc_h[s][n] = 0;
        if (qw < 0 || qe < 0 || qh < 0) atomicOr(a.err, PIC_EBIT_OVERFLOW);
        double w = (double)qw * PIC_FIX_INV;
        double e = (double)qe * PIC_FIX_INV + (double)qh;
        if (e != 0.0 && w == 0.0) atomicOr(a.err, PIC_EBIT_EKRHO);
        double rs, ek;
        if (a.delta == 0) {
            rs = a.weight_qs[s] * w / a.L_ref / a.L_ref / a.n_ref / vol;
            ek = a.weight[s] * e / a.L_ref / a.L_ref / a.n_ref / vol;
        } else {
            rs = a.weight_qs[s] * w / a.L_ref / a.L_ref / a.L_ref / a.n_ref / vol;
            ek = a.weight[s] * e / a.L_ref / a.L_ref / a.L_ref / a.n_ref / vol;
        }
        size_t o = (size_t)s * a.nnodes + n;
        a.rho_p[o] = w; a.rho_s[o] = rs; a.ek[o] = ek;
        a.eks[o] = (e != 0.0 && w != 0.0) ? e / w : 0.0;      // GetParChg_2D.f90:885-886
        rho = rho + rs;
    }
    a.rho[n] = rho;
}

int launch_finalize_deposit() {
    FinalArgs a;
    a.nnodes = g.mesh.nnodes; a.nspecies = g.nspecies; a.delta = g.mesh.delta;
    for (int s = 0; s < g.nspecies; ++s) {
        SpeciesHost& S = g.sp[s];
        if (!S.defined) return pic_fail(PIC_ERR_STATE, "species indices must be contiguous from 0");
        a.acc_w[s] = S.acc_w; a.acc_e[s] = S.acc_e; a.acc_h[s] = S.acc_h;
        a.weight_qs[s] = S.p.weight * S.qs;     // Weight * qs(isp), evaluated left to right
        a.weight[s] = S.p.weight;
        S.acc_dirty = false;
    }
    a.L_ref = g.L_ref; a.n_ref = g.cfg.n_ref;
    a.vol = g.d_cell_volume;
    a.rho = g.d_rho; a.rho_s = g.d_rho_s; a.rho_p = g.d_rho_p; a.ek = g.d_ek; a.eks = g.d_eks;
    a.err = g.d_err;
    hipLaunchKernelGGL(k_finalize, dim3((a.nnodes + 255) / 256), dim3(256), 0, g.stream, a);
    HIPCHK(hipGetLastError());
    return PIC_OK;
}

// ---------------------------------------------------------------- batched drop-ins
__global__ void k_locate_points(int64_t n, const double* x, const double* y, int* out, MeshParams m, int mode, unsigned int* err) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double X = x[i], Y = y[i];
    if (mode == 0) {               // full GetParChg lookup
        int bad = 0;
        int e = pic_locate_elem(X, Y, m, &bad);
        if (bad) atomicOr(err, PIC_EBIT_POSITION);
        out[i] = e + 1;
    } else if (mode == 1) {        // InjectPositioning only (no clamp, like the reference)
        double rx = (X - m.xmin) / m.mapx;
        int ii = (int)rx + 1;
        double ry = (Y - m.ymin) / m.mapy;
        int jj = (int)ry + 1;
        if (ii < 1 || ii > m.nex || jj < 1 || jj > m.ney) { atomicOr(err, PIC_EBIT_POSITION); out[i] = 0; }
        else out[i] = jj + (ii - 1) * m.ney;
    }
}

int launch_locate_points(int64_t n, const double* dx, const double* dy, int* dout, int mode) {
    MeshParams m = g.mesh;
    m.xs = g.d_xs; m.ys = g.d_ys; m.element_index = g.d_element_index; m.node_index = g.d_node_index;
    int blocks = (int)((n + 255) / 256);
    hipLaunchKernelGGL(k_locate_points, dim3(blocks), dim3(256), 0, g.stream, n, dx, dy, dout, m, mode, g.d_err);
    HIPCHK(hipGetLastError());
    return PIC_OK;
}

// Get_ParticleEField_PPR for arbitrary points (FP64 coef_ppr).
__global__ void k_ppr_gather_points(int64_t n, const double* x, const double* y, const int* ele,
                                    double* ex_out, double* ey_out, MeshParams m, const double* coef6) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int e = ele[i] - 1;
    if (e < 0 || e >= m.nelem) { ex_out[i] = 0; ey_out[i] = 0; return; }
    int ci = e / m.ney, cj = e - ci * m.ney;
    int nd[4] = {ci * m.nny + cj, (ci + 1) * m.nny + cj, (ci + 1) * m.nny + cj + 1, ci * m.nny + cj + 1};
    double nx_[4] = {m.xs[ci], m.xs[ci + 1], m.xs[ci + 1], m.xs[ci]};
    double ny_[4] = {m.ys[cj], m.ys[cj], m.ys[cj + 1], m.ys[cj + 1]};
    double X = x[i], Y = y[i], Ex = 0.0, Ey = 0.0;
    bool iface = m.element_index[e] > 0;
    int count = 0;
    for (int k = 0; k < 4; ++k) {
        if (iface && m.node_index[nd[k]] != -1) continue;
        const double* c = coef6 + (size_t)nd[k] * 6;
        count++;
        double xr = X - nx_[k], yr = Y - ny_[k];
        Ex = Ex + c[1] + 2.0 * c[3] * xr + c[4] * yr;
        Ey = Ey + c[2] + 2.0 * c[5] * yr + c[4] * xr;
    }
    if (iface) { Ex = Ex / (double)count; Ey = Ey / (double)count; }
    else { Ex = Ex / 4.0; Ey = Ey / 4.0; }
    ex_out[i] = Ex; ey_out[i] = Ey;
}

int launch_ppr_gather_points(int64_t n, const double* dx, const double* dy, const int* ele, double* ex, double* ey) {
    MeshParams m = g.mesh;
    m.xs = g.d_xs; m.ys = g.d_ys; m.element_index = g.d_element_index; m.node_index = g.d_node_index;
    int blocks = (int)((n + 255) / 256);
    hipLaunchKernelGGL(k_ppr_gather_points, dim3(blocks), dim3(256), 0, g.stream, n, dx, dy, ele, ex, ey, m, g.d_coef);
    HIPCHK(hipGetLastError());
    return PIC_OK;
}

// ---------------------------------------------------------------- wall injection
// InjectBeams_2D (IMPIC/InjectBeams_2D.f90:85-143). Four lanes per injected particle k, each
// with its own Philox stream (slot 4k + c, step it, purpose PIC_RNG_INJECT | species):
// c = 0 Bird's inward flux component (:91-106), c = 1 and 2 Loadv for Vy and Vz
// (Loadv.f90:14-23; the ion drift added after, :121-126), c = 3 PosInit's X then Y
// (ParticleOne.f90:82-89; on an axisymmetric domain r = sqrt(YL^2 + (YU + YL)(YU - YL) R)
// and the azimuth PO%Z = 2 PI R', :95-98, R' from the stream's second block) and
// InjectPositioning (InjectPositioning.f90:13-22). Candidate i
// of a rejection loop takes the two uniforms of Philox block i of its stream (a rejected
// candidate's second uniform goes unused), so a lane evaluates four candidates at a time and
// keeps the first accepted one: the same value as trying them in order. One launch serves
// every registered source.
struct InjectSource {
    InjectHost p;
    int electron;
    int64_t base;                        // first free slot
    int n;                               // particles to add
    uint32_t tag;
    int species;
    void *x, *y, *vx, *vy, *vz;
    void* th;                            // azimuth (axisymmetric domains), else nullptr
    int* cell;
};
struct InjectArgs {
    InjectSource src[PIC_MAX_SPECIES];
    int nsrc;
    int64_t first[PIC_MAX_SPECIES + 1];  // thread offsets of the sources (4 threads per particle)
    uint64_t seed; uint32_t step;
    MeshParams m;
    unsigned int* err;
    PicErrInfo* errinfo;
};

// v if v < hi, else the largest value of v's type below hi
__device__ inline double pic_below(double v, double hi) {
    return v < hi ? v : nextafter(hi, -INFINITY);
}
__device__ inline float pic_below(float v, double hi) {
    if ((double)v < hi) return v;
    float b = (float)hi;
    while ((double)b >= hi) b = nextafterf(b, -INFINITY);
    return b;
}

#define PIC_SAMPLER_MAXIT 1024
#define PIC_SAMPLER_ILP 4

template <typename R>
__global__ void k_inject(InjectArgs a) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.first[a.nsrc]) return;
    int si = 0;
    while (si + 1 < a.nsrc && t >= a.first[si + 1]) ++si;
    const InjectSource& S = a.src[si];
    const InjectHost& p = S.p;
    const int64_t local = t - a.first[si];
    const int64_t k = local >> 2;
    const int c = (int)(local & 3);
    const uint32_t slot = (uint32_t)(4 * k + c);
    const uint32_t k0 = (uint32_t)a.seed, k1 = (uint32_t)(a.seed >> 32);
    const int64_t d = S.base + k;
    if (c < 3) {
        double val = 0.0;
        bool found = false;
        for (int i0 = 0; i0 < PIC_SAMPLER_MAXIT && !found; i0 += PIC_SAMPLER_ILP) {
            double cv[PIC_SAMPLER_ILP];
            bool ok[PIC_SAMPLER_ILP];
#pragma unroll
            for (int j = 0; j < PIC_SAMPLER_ILP; ++j) {
                const philox4 o = pic_philox4x32_10(slot, a.step, S.tag, (uint32_t)(i0 + j), k0, k1);
                const double Ra = pic_u01(o.v[0], o.v[1]), Rb = pic_u01(o.v[2], o.v[3]);
                if (c == 0) {
                    const double UU = -p.qa + 2.0 * p.qa * Ra;
                    const double UN = UU + p.v_r;
                    cv[j] = UN;
                    ok[j] = !(UN < 0.0) && !((2.0 * UN / p.fs1) * exp(p.fs2 - UU * UU) < Rb);
                } else {
                    const double v = (6.0 * Ra - 3.0) * p.vmp;
                    cv[j] = v;
                    ok[j] = !(Rb > exp(-v * v / (p.vmp * p.vmp)));
                }
            }
#pragma unroll
            for (int j = PIC_SAMPLER_ILP - 1; j >= 0; --j)
                if (ok[j]) { val = cv[j]; found = true; }
        }
        if (!found) atomicOr(a.err, PIC_EBIT_SAMPLER);
        if (c == 0) {
            ((R*)S.vx)[d] = (R)(val * p.sq2vt);
        } else {
            if (!S.electron) val = p.vd[c] + val;
            if (c == 1) ((R*)S.vy)[d] = (R)val;
            else ((R*)S.vz)[d] = (R)val;
        }
    } else {
        const philox4 o = pic_philox4x32_10(slot, a.step, S.tag, 0u, k0, k1);
        const double X = p.xl + (p.xu - p.xl) * pic_u01(o.v[0], o.v[1]);
        double Y;
        if (S.th) {
            Y = sqrt(p.y_lo * p.y_lo + (p.y_hi + p.y_lo) * (p.y_hi - p.y_lo) * pic_u01(o.v[2], o.v[3]));
            const philox4 o2 = pic_philox4x32_10(slot, a.step, S.tag, 1u, k0, k1);
            ((R*)S.th)[d] = (R)(2.0 * PIC_PI * pic_u01(o2.v[0], o2.v[1]));
        } else {
            Y = p.y_lo + (p.y_hi - p.y_lo) * pic_u01(o.v[2], o.v[3]);
        }
        // a stored coordinate that rounds onto the slab's upper bound (FP32 slots: Y just below
        // y_hi becomes y_hi) takes the largest value below it, so the particle stays strictly
        // inside the slab, as the sampled value is
        const R Xs = pic_below((R)X, p.xu), Ys = pic_below((R)Y, p.y_hi);
        // InjectPositioning on the stored position (no clamp)
        const MeshParams& m = a.m;
        const double rx = ((double)Xs - m.xmin) / m.mapx;
        const int ii = (int)rx + 1;
        const double ry = ((double)Ys - m.ymin) / m.mapy;
        const int jj = (int)ry + 1;
        int e = -1;
        if (ii < 1 || ii > m.nex || jj < 1 || jj > m.ney) {
            atomicOr(a.err, PIC_EBIT_POSITION);
            pic_record(a.errinfo, PIC_EBIT_POSITION, d, Xs, Ys, 0.0, 0.0, 0, -1, -1, S.species);
        } else {
            e = (jj - 1) + (ii - 1) * m.ney;
        }
        ((R*)S.x)[d] = Xs; ((R*)S.y)[d] = Ys;
        S.cell[d] = e;
    }
}

// add_N of every registered source (InjectBeams_2D.f90:40-64), then one launch;
// the rounding draw is the first word of the species' stream at slot 0xFFFFFFFF.
int run_injection(int it, int64_t* added) {
    int64_t ns[PIC_MAX_SPECIES] = {0};
    for (int s = 0; s < PIC_MAX_SPECIES; ++s) {
        if (added) added[s] = 0;
        const InjectHost& p = g.inj[s];
        if (!p.on || !g.sp[s].defined) continue;
        const SpeciesHost& S = g.sp[s];
        pic_stream rs;
        pic_stream_init(&rs, g.seed, 0xFFFFFFFFu, (uint32_t)it, pic_stream_tag(s, PIC_RNG_INJECT));
        const double ranum = pic_stream_u01(&rs);
        // InjectBeams_2D.f90:56-62: the slab's width (Cartesian) or the disk annulus PI (YU - YL)^2
        // (axisymmetric, as written: the squared radial extent) crossed per step
        const double dy = p.y_hi - p.y_lo;
        const double expect = g.mesh.delta == 1
            ? p.dens * g.cfg.n_ref * std::fabs(p.b_amb) * g.v_ref * g.cfg.dt * g.time_ref * PIC_PI * (dy * dy) *
                  (g.L_ref * g.L_ref) / S.p.weight
            : p.dens * g.cfg.n_ref * std::fabs(p.b_amb) * g.v_ref * g.cfg.dt * g.time_ref * dy * g.L_ref / S.p.weight;
        const int64_t n = (int64_t)(expect + ranum);
        if (n <= 0) continue;
        if (S.nslots + n > S.capacity) return pic_fail(PIC_ERR_CAPACITY, "injection exceeds particle capacity");
        if (n >= (int64_t)1 << 29) return pic_fail(PIC_ERR_CAPACITY, "injection count out of range");
        ns[s] = n;
    }
    InjectArgs a;
    a.nsrc = 0; a.first[0] = 0;
    for (int s = 0; s < PIC_MAX_SPECIES; ++s) {
        if (ns[s] == 0) continue;
        SpeciesHost& S = g.sp[s];
        InjectSource& q = a.src[a.nsrc];
        q.p = g.inj[s]; q.electron = (s == g.e_species) ? 1 : 0;
        q.base = S.nslots; q.n = (int)ns[s]; q.tag = pic_stream_tag(s, PIC_RNG_INJECT);
        q.species = s;
        q.x = S.x; q.y = S.y; q.vx = S.vx; q.vy = S.vy; q.vz = S.vz; q.cell = S.cell;
        q.th = g.mesh.delta == 1 ? S.th : nullptr;
        a.first[a.nsrc + 1] = a.first[a.nsrc] + 4 * ns[s];
        a.nsrc++;
        S.nslots += ns[s]; S.nlive += ns[s]; S.injected += ns[s];
        if (added) added[s] = ns[s];
    }
    if (a.nsrc == 0) return PIC_OK;
    a.seed = g.seed; a.step = (uint32_t)it;
    a.m = g.mesh; a.err = g.d_err; a.errinfo = (PicErrInfo*)g.d_errinfo;
    const int blocks = (int)((a.first[a.nsrc] + 255) / 256);
    if (g.particle_bits == 64) hipLaunchKernelGGL(k_inject<double>, dim3(blocks), dim3(256), 0, g.stream, a);
    else hipLaunchKernelGGL(k_inject<float>, dim3(blocks), dim3(256), 0, g.stream, a);
    HIPCHK(hipGetLastError());
    return PIC_OK;
}
