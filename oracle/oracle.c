/*
 * oracle.c -- CPU restatement of the PIC-IFE_GEC per-timestep hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker the GPU library is compared
 * against; it is loaded by tests/, by __graft_entry__.smoke() and by the cpu_baseline
 * leg of bench.py, and never by the library itself.
 *
 * Each function restates one reference routine (paths relative to the reference
 * checkout) in plain C, in the reference's operation order:
 *   or_philox / or_stream      Philox4x32-10 (Salmon et al. SC'11), pinned by the
 *                              published known-answer vectors
 *   or_move_all_f64 / _f32     MoveAll + MoveOne + Get_ParticleEField_PPR +
 *                              AdjustParticleBoundary (MCCInterface.f90:217-640,
 *                              IMPIC/MoveOne.f90, PPR/Get_ParticleEField_PPR.f90)
 *   or_locate                  GetParChg_2D.f90:137-163 + InjectPositioning.f90 +
 *                              ParticlePositioning.f90 (unrefined mesh)
 *   or_deposit                 GetParChg_2D.f90:122-911, two accumulations: the
 *                              reference's FP64 sequential sums, and the int64 fixed
 *                              point the GPU uses (bit-identical target)
 *   or_ppr                     GetEfield_SIDG_PPR.f90:153-304 + BRINV.f90 (per-node
 *                              least squares redone from the point patches)
 *   or_pcg                     MY_PCG.f90 + MATVEC_1.f90 + MSOLVE_1.f90
 *   or_assemble                FE stiffness / RHS by 2x2 Gauss quadrature
 *                              (Gauss_integration_local_stiffness, FE_RHS_HW_2D) on
 *                              the uniform Q1 mesh
 *   or_mcc_*                   MCCPublic.F90, TypeMCC.f90, MCCElectron.F90, MCCIon.F90,
 *                              MCCChi.f90 with the counter-based selection of the build
 *   or_inject*                 IMPIC/InjectBeams_2D.f90, Loadv.f90, InjectPositioning.f90
 *   or_sincos / or_atan        DCOS / DSIN / DATAN of the axisymmetric mover
 *   or_ife_element             the bilinear IFE interface element (Linear_IFE_Basis_Coeff_2D,
 *                              Gauss_integration_local_stiffness_IFE, IFE_RHS_2D)
 * Every function is pinned by checks under tests/ that do not retype it: published
 * known-answer vectors (Philox), closed forms and invariants (uniform-field pushes, charge
 * conservation, quadratic reproduction of PPR, direct solves for PCG, the defining
 * conditions and the circular-interface convergence of the IFE element), brute force on
 * small inputs (lookup, MCC selection), the C library within 1-2 ulp (trigonometry) and
 * files written by a Fortran compiler (the file formats).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ Philox */
typedef struct { uint32_t v[4]; } or_p4;

or_p4 or_philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1) {
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    or_p4 o = {{c0, c1, c2, c3}};
    return o;
}

void or_philox_block(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    or_p4 o = or_philox(ctr[0], ctr[1], ctr[2], ctr[3], key[0], key[1]);
    memcpy(out, o.v, 16);
}

typedef struct { uint32_t k0, k1, c0, c1, c2, blk; uint32_t buf[4]; int used; } or_stream;

static void or_stream_init(or_stream* s, uint64_t seed, uint32_t slot, uint32_t step, uint32_t tag) {
    s->k0 = (uint32_t)seed; s->k1 = (uint32_t)(seed >> 32);
    s->c0 = slot; s->c1 = step; s->c2 = tag; s->blk = 0; s->used = 4;
}
static uint32_t or_u32(or_stream* s) {
    if (s->used == 4) {
        or_p4 o = or_philox(s->c0, s->c1, s->c2, s->blk, s->k0, s->k1);
        memcpy(s->buf, o.v, 16);
        s->blk++; s->used = 0;
    }
    return s->buf[s->used++];
}
/* 52 random bits centred in their bin: a double in the open interval (0,1) */
static double or_u01(or_stream* s) {
    uint32_t a = or_u32(s), b = or_u32(s);
    uint64_t m = ((uint64_t)a << 20) | (uint64_t)(b >> 12);
    return ((double)m + 0.5) * (1.0 / 4503599627370496.0);
}
static uint32_t or_tag(int species, int purpose) { return ((uint32_t)purpose << 16) | (uint32_t)(species & 0xFFFF); }
enum { OR_RNG_LOAD = 1, OR_RNG_EMIT = 2, OR_RNG_SELECT = 3, OR_RNG_COLLIDE = 4 };

/* ------------------------------------------------------------------ trig */
/* DCOS / DSIN / DATAN of the axisymmetric MoveOne (MoveOne.f90:174-176, :248-262) as the
 * build evaluates them: a fixed sequence of IEEE operations (Cody-Waite reduction by a
 * three-part pi/2, Taylor series through r^15 / r^16 for sin / cos on [-pi/4, pi/4]; atan by
 * the reductions 1/t and (t-1)/(t+1) to |u| <= tan(pi/8) and the series through u^41/41).
 * Pinned against the C library to about 1 ulp by tests/test_oracle_cpu.py. */
void or_sincos(double x, double* s, double* c) {
    const double kd = rint(x * 0x1.45f306dc9c883p-1);
    const double r = ((x - kd * 0x1.921fb54400000p+0) - kd * 0x1.0b4611a600000p-34) - kd * 0x1.3198a2e037073p-69;
    const double z = r * r;
    const double ps = -0x1.5555555555555p-3 + z * (0x1.1111111111111p-7 + z * (-0x1.a01a01a01a01ap-13 +
                      z * (0x1.71de3a556c734p-19 + z * (-0x1.ae64567f544e4p-26 + z * (0x1.6124613a86d09p-33 +
                      z * -0x1.ae7f3e733b81fp-41)))));
    const double sr = r + r * z * ps;
    const double pc = 0x1.5555555555555p-5 + z * (-0x1.6c16c16c16c17p-10 + z * (0x1.a01a01a01a01ap-16 +
                      z * (-0x1.27e4fb7789f5cp-22 + z * (0x1.1eed8eff8d898p-29 + z * (-0x1.93974a8c07c9dp-37 +
                      z * 0x1.ae7f3e733b81fp-45)))));
    const double hz = 0.5 * z, w = 1.0 - hz;
    const double cr = w + (((1.0 - w) - hz) + z * z * pc);
    switch ((int)kd & 3) {
        case 0: *s = sr; *c = cr; break;
        case 1: *s = cr; *c = -sr; break;
        case 2: *s = -sr; *c = -cr; break;
        default: *s = -cr; *c = sr; break;
    }
}

double or_atan(double t) {
    static const double A[20] = {-0x1.5555555555555p-2, 0x1.999999999999ap-3, -0x1.2492492492492p-3,
        0x1.c71c71c71c71cp-4, -0x1.745d1745d1746p-4, 0x1.3b13b13b13b14p-4, -0x1.1111111111111p-4,
        0x1.e1e1e1e1e1e1ep-5, -0x1.af286bca1af28p-5, 0x1.8618618618618p-5, -0x1.642c8590b2164p-5,
        0x1.47ae147ae147bp-5, -0x1.2f684bda12f68p-5, 0x1.1a7b9611a7b96p-5, -0x1.0842108421084p-5,
        0x1.f07c1f07c1f08p-6, -0x1.d41d41d41d41dp-6, 0x1.bacf914c1bad0p-6, -0x1.a41a41a41a41ap-6,
        0x1.8f9c18f9c18fap-6};                                  /* (-1)^n / (2n+1), n = 1..20 */
    if (!(t == t)) return t;
    double a = fabs(t);
    const int inv = a > 1.0;
    if (inv) a = 1.0 / a;
    const int red = a > 0x1.a827999fcef32p-2;                 /* tan(pi/8) */
    const double u = red ? (a - 1.0) / (a + 1.0) : a;
    const double z = u * u;
    double q = A[19];
    for (int n = 18; n >= 0; --n) q = A[n] + z * q;
    double p = u + u * z * q;
    if (red) p = 0x1.921fb54442d18p-1 + (p + 0x1.1a62633145c07p-55);       /* pi/4 hi + lo */
    if (inv) p = 0x1.921fb54442d18p+0 - (p - 0x1.1a62633145c07p-54);       /* pi/2 hi - lo */
    return t < 0.0 ? -p : p;
}

/* ------------------------------------------------------------------ mesh */
typedef struct { int shape; double x0, y0, x1, y1, r2; } or_object;

typedef struct {
    int nnx, nny, nex, ney, nnodes, nelem;
    double xmin, xmax, ymin, ymax, hx, hy;
    const double* xs; const double* ys;   /* node coordinate lines (Cubic_Partition_2D x(i), y(j)) */
    const int* element_index;             /* nelem */
    const int* node_index;                /* nnodes */
    int n_objects;
    or_object obj[8];
} or_mesh;

/* builds an or_mesh from flat arguments (ctypes friendly) */
static void mk_mesh(or_mesh* m, int nnx, int nny, const double* xs, const double* ys,
                    const int* element_index, const int* node_index, int n_objects, const double* objects /* 6 per object */) {
    memset(m, 0, sizeof(*m));
    m->nnx = nnx; m->nny = nny; m->nex = nnx - 1; m->ney = nny - 1;
    m->nnodes = nnx * nny; m->nelem = (nnx - 1) * (nny - 1);
    m->xs = xs; m->ys = ys;
    m->xmin = xs[0]; m->xmax = xs[nnx - 1]; m->ymin = ys[0]; m->ymax = ys[nny - 1];
    m->hx = (m->xmax - m->xmin) / (double)(nnx - 1);
    m->hy = (m->ymax - m->ymin) / (double)(nny - 1);
    m->element_index = element_index; m->node_index = node_index;
    m->n_objects = n_objects;
    for (int k = 0; k < n_objects && k < 8; ++k) {
        const double* o = objects + 6 * k;
        m->obj[k].shape = (int)o[0];
        m->obj[k].x0 = o[1]; m->obj[k].y0 = o[2]; m->obj[k].x1 = o[3]; m->obj[k].y1 = o[4]; m->obj[k].r2 = o[5];
    }
}

/* ------------------------------------------------------------------ threads */
/* OpenMP thread count of the particle, lookup, deposition, RHS, PCG and PPR loops. 1 (the
 * default, used by every parity test) runs each loop serially in the reference's order; > 1
 * exists for the multi-core CPU timing. Integer results are identical for every count;
 * floating-point sums that are split over threads (PCG dot products, RHS scatter) may
 * differ in the last bits. */
static int or_threads = 1;
void or_set_threads(int n) { or_threads = n < 1 ? 1 : n; }

/* Planted mistakes (0 = none), set only by the tests that show their checks catch them:
 * 1 post-collision h vector without its G*Gz*sin(phi) term, 2 SelectProbility reading
 * the table one row too high, 3 VelMaxInit taking the speed from the first uniform of the
 * accepted pair, 4 cell volumes without the hy/4 shift of the first and last radial rows,
 * 5 axisymmetric injection drawing r uniform in r instead of r^2, 6 the kappa load's
 * acceptance density with half its exponent. */
static int or_plant = 0;
void or_set_plant(int code) { or_plant = code; }
int or_get_threads(void) { return or_threads; }

/* ------------------------------------------------------------------ MoveAll */
#define REAL double
#define FMOD fmod
#define FN(name) or_##name##_f64
#include "oracle_push.inc"
#undef REAL
#undef FMOD
#undef FN
#define REAL float
#define FMOD fmodf
#define FN(name) or_##name##_f32
#include "oracle_push.inc"
#undef REAL
#undef FMOD
#undef FN

int or_move_all(int bits, int nnx, int nny, const double* xs, const double* ys, const int* element_index,
                const int* node_index, int n_objects, const double* objects, int delta,
                const void* efx, const void* efy, const void* coef5, double qm, double dt,
                double v_ref, double mass, double init_temp_K, uint64_t seed, uint32_t step, uint32_t tag_emit,
                int64_t n, void* x, void* y, void* vx, void* vy, void* vz, void* th, const int* cell, int* alive,
                int64_t* counters) {
    or_mesh m;
    mk_mesh(&m, nnx, nny, xs, ys, element_index, node_index, n_objects, objects);
    int err = 0;
    if (bits == 64) {
        or_move_all_f64(&m, delta, (const double*)efx, (const double*)efy, (const double*)coef5, xs, ys, qm, dt, v_ref,
                        mass, init_temp_K, seed, step, tag_emit, n, (double*)x, (double*)y, (double*)vx, (double*)vy,
                        (double*)vz, (double*)th, cell, alive, counters, &err);
    } else {
        float* xsf = (float*)malloc(sizeof(float) * nnx);
        float* ysf = (float*)malloc(sizeof(float) * nny);
        for (int i = 0; i < nnx; ++i) xsf[i] = (float)xs[i];
        for (int j = 0; j < nny; ++j) ysf[j] = (float)ys[j];
        or_move_all_f32(&m, delta, (const float*)efx, (const float*)efy, (const float*)coef5, xsf, ysf, (float)qm,
                        (float)dt, v_ref, mass, init_temp_K, seed, step, tag_emit, n, (float*)x, (float*)y, (float*)vx,
                        (float*)vy, (float*)vz, (float*)th, cell, alive, counters, &err);
        free(xsf); free(ysf);
    }
    return err;
}

/* ------------------------------------------------------------------ lookup */
/* GetParChg_2D.f90:137-163 with Vert_o = min - h (Input_2D.f90:76-78), hxi = 1/hx;
 * MOD(rxp, Mapxmin) as the exact remainder; InjectPositioning.f90:13-22 (division);
 * ParticlePositioning on the unrefined mesh returns the estimate. 0-based result. */
static int locate1(const or_mesh* m, double X, double Y, int* bad) {
    double hxi = 1.0 / m->hx, hyi = 1.0 / m->hy;
    double vox = m->xmin - m->hx, voy = m->ymin - m->hy;
    double rxp = (X - vox) * hxi;
    int i = (int)rxp;
    double ryp = (Y - voy) * hyi;
    int j = (int)ryp;
    if (j == m->nny) j = j - 1;
    if (i == m->nnx) i = i - 1;
    double mapx = m->hx / 1.0, mapy = m->hy / 1.0;     /* hx(1) / 2**repeat_refinement */
    if (fmod(rxp, mapx) > 1e-10 && fmod(ryp, mapy) > 1e-10) {
        double rx = (X - m->xmin) / mapx;
        i = (int)rx + 1;
        double ry = (Y - m->ymin) / mapy;
        j = (int)ry + 1;
    }
    if (i < 1 || i > m->nex || j < 1 || j > m->ney) {
        *bad = 1;
        i = i < 1 ? 1 : (i > m->nex ? m->nex : i);
        j = j < 1 ? 1 : (j > m->ney ? m->ney : j);
    }
    return (j - 1) + (i - 1) * m->ney;
}

int or_locate(int nnx, int nny, const double* xs, const double* ys, int64_t n, const double* x, const double* y, int* out) {
    or_mesh m;
    mk_mesh(&m, nnx, nny, xs, ys, NULL, NULL, 0, NULL);
    int bad = 0;
    const int nt = or_threads;
#pragma omp parallel for num_threads(nt) if (nt > 1) schedule(static) reduction(| : bad)
    for (int64_t p = 0; p < n; ++p) out[p] = locate1(&m, x[p], y[p], &bad);
    return bad;
}

/* ------------------------------------------------------------------ deposition */
/* GetParChg_2D.f90:122-911 for one species. Weights of :788-821, moments W and W*Energy
 * (Energy = 0.5*Mass*|v|^2*v_ref^2/JtoeV, ParticleOne.f90:333-341, :128) accumulated to
 * element corners (:823-843), corners summed to nodes in element order (:849-866).
 * fix != 0: every term quantised to round(term*2^40) and summed as int64 (the build's
 * accumulation, order independent); a W*Energy term >= 2^11 eV is held as its whole-eV part
 * (summed apart) plus the quantised fraction, so no single term can wrap the 2^40 sum, and
 * the node value is sum_lo * 2^-40 + sum_hi. fix == 0: the reference's FP64 sums in
 * particle order. Output node sums w_node, e_node (FP64), before normalisation. */
static void or_fix_split(double v, long long* lo, long long* hi) {
    if (fabs(v) < 2048.0) { *hi = 0; *lo = llrint(v * 1099511627776.0); return; }
    double ip = floor(v);
    *hi = (long long)ip;
    *lo = llrint((v - ip) * 1099511627776.0);
}

int or_deposit(int bits, int nnx, int nny, const double* xs, const double* ys, int delta, int64_t n,
               const void* x, const void* y, const void* vx, const void* vy, const void* vz, const int* cell,
               double mass, double v_ref, int fix, double* w_node, double* e_node) {
    const double JtoeV = 1.6022e-19;
    /* 0.5*Mass*|v|^2*v_ref^2/JtoeV with the constant factors multiplied once (the build's
     * order; a few ulps from the reference's left-to-right product) */
    const double Ec = 0.5 * mass * v_ref * v_ref / JtoeV;
    int nex = nnx - 1, ney = nny - 1, nelem = nex * ney, nnodes = nnx * nny;
    double ymin = ys[0];
    double* rs = (double*)calloc((size_t)4 * nelem, sizeof(double));
    double* ek = (double*)calloc((size_t)4 * nelem, sizeof(double));
    long long* qs = (long long*)calloc((size_t)4 * nelem, sizeof(long long));
    long long* qe = (long long*)calloc((size_t)4 * nelem, sizeof(long long));
    long long* qh = (long long*)calloc((size_t)4 * nelem, sizeof(long long));
    /* threads only for the fixed-point sums, which are order independent; the FP64 sums
     * stay sequential in particle order */
    const int nt = fix ? or_threads : 1;
#pragma omp parallel for num_threads(nt) if (nt > 1) schedule(static)
    for (int64_t p = 0; p < n; ++p) {
        int e = cell[p];
        if (e < 0) continue;
        double X, Y, VX, VY, VZ;
        if (bits == 64) {
            X = ((const double*)x)[p]; Y = ((const double*)y)[p];
            VX = ((const double*)vx)[p]; VY = ((const double*)vy)[p]; VZ = ((const double*)vz)[p];
        } else {
            X = ((const float*)x)[p]; Y = ((const float*)y)[p];
            VX = ((const float*)vx)[p]; VY = ((const float*)vy)[p]; VZ = ((const float*)vz)[p];
        }
        double Energy = (VX * VX + VY * VY + VZ * VZ) * Ec;
        int ci = e / ney, cj = e % ney;
        double hxp = xs[ci + 1] - xs[ci], hyp = ys[cj + 1] - ys[cj];
        double dx = (X - xs[ci]) / hxp, dy = (Y - ys[cj]) / hyp;
        double xm = 1.0 - dx, ym = 1.0 - dy;
        double W[4];
        if (delta == 0) { W[0] = xm * ym; W[1] = dx * ym; W[2] = dx * dy; W[3] = xm * dy; }
        else {
            double BETA = (fabs(ys[cj] - ymin) < 1e-12) ? 0.75 : 1.0;
            double R1 = ymin + ys[cj], R2 = ymin + ys[cj + 1], R = Y, den = R2 * R2 - R1 * R1;
            W[0] = xm * (R2 * R2 - R * R) / den * BETA; W[1] = dx * (R2 * R2 - R * R) / den * BETA;
            W[2] = dx * (R * R - R1 * R1) / den; W[3] = xm * (R * R - R1 * R1) / den;
        }
        for (int k = 0; k < 4; ++k) {
            if (fix) {
                long long a = llrint(W[k] * 1099511627776.0), b, h;
                or_fix_split(W[k] * Energy, &b, &h);
                if (nt > 1) {
#pragma omp atomic
                    qs[4 * (size_t)e + k] += a;
#pragma omp atomic
                    qe[4 * (size_t)e + k] += b;
#pragma omp atomic
                    qh[4 * (size_t)e + k] += h;
                } else {
                    qs[4 * (size_t)e + k] += a; qe[4 * (size_t)e + k] += b; qh[4 * (size_t)e + k] += h;
                }
            } else {
                rs[4 * (size_t)e + k] = rs[4 * (size_t)e + k] + W[k];
                ek[4 * (size_t)e + k] = ek[4 * (size_t)e + k] + W[k] * Energy;
            }
        }
    }
    if (fix) {
        long long* nw = (long long*)calloc(nnodes, sizeof(long long));
        long long* ne = (long long*)calloc(nnodes, sizeof(long long));
        long long* nh = (long long*)calloc(nnodes, sizeof(long long));
        if (nt > 1) {
            /* node-side gather of the (up to) four corner sums touching each node */
#pragma omp parallel for num_threads(nt) schedule(static)
            for (int i = 0; i < nnodes; ++i) {
                int ni = i / nny, nj = i % nny;
                long long sw = 0, se = 0, sh = 0;
                for (int k = 0; k < 4; ++k) {
                    int ci = ni - (k == 1 || k == 2), cj = nj - (k >= 2);   /* corner k of element (ci,cj) is node i */
                    if (ci < 0 || ci >= nex || cj < 0 || cj >= ney) continue;
                    size_t e = (size_t)ci * ney + cj;
                    sw += qs[4 * e + k]; se += qe[4 * e + k]; sh += qh[4 * e + k];
                }
                nw[i] = sw; ne[i] = se; nh[i] = sh;
            }
        } else {
            for (int e = 0; e < nelem; ++e) {
                int ci = e / ney, cj = e % ney;
                int ht[4] = {ci * nny + cj, (ci + 1) * nny + cj, (ci + 1) * nny + cj + 1, ci * nny + cj + 1};
                for (int k = 0; k < 4; ++k) {
                    nw[ht[k]] += qs[4 * (size_t)e + k]; ne[ht[k]] += qe[4 * (size_t)e + k]; nh[ht[k]] += qh[4 * (size_t)e + k];
                }
            }
        }
        for (int i = 0; i < nnodes; ++i) {
            w_node[i] = (double)nw[i] * (1.0 / 1099511627776.0);
            e_node[i] = (double)ne[i] * (1.0 / 1099511627776.0) + (double)nh[i];
        }
        free(nw); free(ne); free(nh);
    } else {
        for (int i = 0; i < nnodes; ++i) { w_node[i] = 0.0; e_node[i] = 0.0; }
        for (int e = 0; e < nelem; ++e) {
            int ci = e / ney, cj = e % ney;
            int ht[4] = {ci * nny + cj, (ci + 1) * nny + cj, (ci + 1) * nny + cj + 1, ci * nny + cj + 1};
            for (int k = 0; k < 4; ++k) {
                w_node[ht[k]] = w_node[ht[k]] + rs[4 * (size_t)e + k];
                e_node[ht[k]] = e_node[ht[k]] + ek[4 * (size_t)e + k];
            }
        }
    }
    free(rs); free(ek); free(qs); free(qe); free(qh);
    return 0;
}

/* Normalisation of GetParChg_2D.f90:869-911: rho_s = Weight*qs*sum/L_ref/L_ref/n_ref/CellVol
 * (Cartesian), Ek_tot = Weight*sum_E/..., rho += rho_s over species. */
void or_normalize(int nnodes, int delta, double weight, double qs_norm, double L_ref, double n_ref,
                  const double* vol, const double* w_node, const double* e_node, double* rho_s, double* ek, double* rho) {
    for (int i = 0; i < nnodes; ++i) {
        if (delta == 0) {
            rho_s[i] = weight * qs_norm * w_node[i] / L_ref / L_ref / n_ref / vol[i];
            ek[i] = weight * e_node[i] / L_ref / L_ref / n_ref / vol[i];
        } else {
            rho_s[i] = weight * qs_norm * w_node[i] / L_ref / L_ref / L_ref / n_ref / vol[i];
            ek[i] = weight * e_node[i] / L_ref / L_ref / L_ref / n_ref / vol[i];
        }
        rho[i] = rho[i] + rho_s[i];
    }
}

/* Cell_Volume_zwz(2,:) of SetupGrids_2D_QLL.f90:196-395 for the CG nodes of a mesh without
 * refinement (num == 1, P_flag(3) == 0): Cartesian hx hy / 4 at corners, / 2 on walls, hx hy
 * inside; axisymmetric (y = r) 2 Pi_LY (ymin + hy/4)(hy/2) hx on the first row (/2 at its
 * corners), 2 Pi_LY (ymax - hy/4)(hy/2) hx on the last, 2 Pi_LY y hy hx / 2 on the x walls and
 * 2 Pi_LY y hy hx inside. hx, hy: the spacing of the first element containing the node. */
void or_cell_volume(int nnx, int nny, const double* xs, const double* ys, int delta, double* vol) {
    const double PI_LY = 3.14159265358979, sv = 1e-12;
    const double xmin = xs[0], xmax = xs[nnx - 1], ymin = ys[0], ymax = ys[nny - 1];
    for (int i = 0; i < nnx; ++i)
        for (int j = 0; j < nny; ++j) {
            int ci = i < nnx - 1 ? i : nnx - 2, cj = j < nny - 1 ? j : nny - 2;
            double hx = xs[ci + 1] - xs[ci], hy = ys[cj + 1] - ys[cj], x = xs[i], y = ys[j];
            double q = or_plant == 4 ? 0.0 : hy / 4.0;
            int xl = fabs(x - xmin) < sv, xr = fabs(x - xmax) < sv, yb = fabs(y - ymin) < sv, yt = fabs(y - ymax) < sv;
            double v;
            if ((xl || xr) && (yb || yt)) {
                if (delta == 0) v = hx * hy / 4.0;
                else if (yb) v = 2.0 * PI_LY * (ymin + q) * (hy / 2.0) * hx / 2.0;
                else v = 2.0 * PI_LY * (ymax - q) * (hy / 2.0) * hx / 2.0;
            } else if (xl || xr || yb || yt) {
                if (delta == 0) v = hx * hy / 2.0;
                else if (yb) v = 2.0 * PI_LY * (ymin + q) * (hy / 2.0) * hx;
                else if (yt) v = 2.0 * PI_LY * (ymax - q) * (hy / 2.0) * hx;
                else v = 2.0 * PI_LY * y * hy * hx / 2.0;
            } else {
                if (delta == 0) v = hx * hy / 1.0;
                else v = 2.0 * PI_LY * y * hy * hx / 1.0;
            }
            vol[(size_t)i * nny + j] = v;
        }
}

/* ------------------------------------------------------------------ PPR */
/* BRINV (PPR/BRINV.f90:1-72): full-pivot Gauss-Jordan inverse in place. Returns 0 if
 * singular (the reference prints ERR**NOT INV). */
int or_brinv(double* a, int n) {
    int is[16], js[16];
    for (int k = 0; k < n; ++k) {
        double d = 0.0;
        for (int i = k; i < n; ++i)
            for (int j = k; j < n; ++j) {
                double p = fabs(a[i * n + j]);
                if (p > d) { d = p; is[k] = i; js[k] = j; }
            }
        if (d + 1.0 == 1.0) return 0;
        if (is[k] != k) for (int j = 0; j < n; ++j) { double t = a[k * n + j]; a[k * n + j] = a[is[k] * n + j]; a[is[k] * n + j] = t; }
        if (js[k] != k) for (int i = 0; i < n; ++i) { double t = a[i * n + k]; a[i * n + k] = a[i * n + js[k]]; a[i * n + js[k]] = t; }
        a[k * n + k] = 1.0 / a[k * n + k];
        for (int j = 0; j < n; ++j) if (j != k) a[k * n + j] = a[k * n + j] * a[k * n + k];
        for (int i = 0; i < n; ++i)
            if (i != k)
                for (int j = 0; j < n; ++j)
                    if (j != k) a[i * n + j] = a[i * n + j] - a[i * n + k] * a[k * n + j];
        for (int i = 0; i < n; ++i) if (i != k) a[i * n + k] = -a[i * n + k] * a[k * n + k];
    }
    for (int k = n - 1; k >= 0; --k) {
        if (js[k] != k) for (int j = 0; j < n; ++j) { double t = a[k * n + j]; a[k * n + j] = a[js[k] * n + j]; a[js[k] * n + j] = t; }
        if (is[k] != k) for (int i = 0; i < n; ++i) { double t = a[i * n + k]; a[i * n + k] = a[i * n + is[k]]; a[i * n + is[k]] = t; }
    }
    return 1;
}

/* GetEfield_SIDG_PPR.f90:153-241, 301-304. point_patch is the reference's
 * point_patch(nnodes, 50) in Fortran order (count in column 1, 1-based node ids).
 * coef (6 x nnodes), efx, efy out. Returns the number of singular patches. */
int or_ppr(int nnx, int nny, const double* xs, const double* ys, const int* point_patch, const double* phi,
           double* coef, double* efx, double* efy) {
    int nnodes = nnx * nny, nsing = 0;
    const int nt = or_threads;
#pragma omp parallel for num_threads(nt) if (nt > 1) schedule(static) reduction(+ : nsing)
    for (int i = 0; i < nnodes; ++i) {
        double pnts[2][64], AT[6][64], A_MUL[36], ATn[6][64];
        int cnt = point_patch[i];
        for (int r = 0; r < 6; ++r) coef[6 * (size_t)i + r] = 0.0;
        efx[i] = 0.0; efy[i] = 0.0;
        if (cnt == 0) continue;
        double hxi_ = xs[i / nny], hyi_ = ys[i % nny];
        double h = 0, h_temp;
        double phv[64];
        for (int j = 0; j < cnt; ++j) {
            int k = point_patch[(size_t)(j + 1) * nnodes + i] - 1;
            pnts[0][j] = xs[k / nny]; pnts[1][j] = ys[k % nny];
            phv[j] = phi[k];
            h_temp = fmax(fabs(pnts[0][j] - hxi_), fabs(pnts[1][j] - hyi_));
            if (h_temp > h) h = h_temp;
        }
        for (int j = 0; j < cnt; ++j) { pnts[0][j] = (pnts[0][j] - hxi_) / h; pnts[1][j] = (pnts[1][j] - hyi_) / h; }
        for (int j = 0; j < cnt; ++j) {
            AT[0][j] = 1.0; AT[1][j] = pnts[0][j]; AT[2][j] = pnts[1][j];
            AT[3][j] = pnts[0][j] * pnts[0][j]; AT[4][j] = pnts[0][j] * pnts[1][j]; AT[5][j] = pnts[1][j] * pnts[1][j];
        }
        for (int a = 0; a < 6; ++a)
            for (int b = 0; b < 6; ++b) {
                double s = 0.0;
                for (int j = 0; j < cnt; ++j) s += AT[a][j] * AT[b][j];
                A_MUL[a * 6 + b] = s;
            }
        if (!or_brinv(A_MUL, 6)) { nsing++; continue; }
        for (int a = 0; a < 6; ++a)
            for (int j = 0; j < cnt; ++j) {
                double s = 0.0;
                for (int b = 0; b < 6; ++b) s += A_MUL[a * 6 + b] * AT[b][j];
                ATn[a][j] = s;
            }
        double t[6];
        for (int a = 0; a < 6; ++a) {
            double s = 0.0;
            for (int j = 0; j < cnt; ++j) s += ATn[a][j] * phv[j];
            t[a] = s;
        }
        coef[6 * (size_t)i + 0] = t[0];
        coef[6 * (size_t)i + 1] = t[1] / h; coef[6 * (size_t)i + 2] = t[2] / h;
        coef[6 * (size_t)i + 3] = t[3] / (h * h); coef[6 * (size_t)i + 4] = t[4] / (h * h); coef[6 * (size_t)i + 5] = t[5] / (h * h);
        efx[i] = -coef[6 * (size_t)i + 1];
        efy[i] = -coef[6 * (size_t)i + 2];
    }
    return nsing;
}

/* Get_ParticleEField_PPR.f90:1-50 (returns grad phi; callers negate). e 0-based. */
void or_ppr_gather(int nnx, int nny, const double* xs, const double* ys, const int* element_index, const int* node_index,
                   const double* coef, int64_t n, const double* x, const double* y, const int* ele, double* ex, double* ey) {
    int ney = nny - 1;
    for (int64_t p = 0; p < n; ++p) {
        int e = ele[p];
        int ci = e / ney, cj = e % ney;
        int ht[4] = {ci * nny + cj, (ci + 1) * nny + cj, (ci + 1) * nny + cj + 1, ci * nny + cj + 1};
        double hx_[4] = {xs[ci], xs[ci + 1], xs[ci + 1], xs[ci]}, hy_[4] = {ys[cj], ys[cj], ys[cj + 1], ys[cj + 1]};
        double Ex = 0, Ey = 0;
        int count = 0, iface = element_index[e] > 0;
        for (int i = 0; i < 4; ++i) {
            if (iface && node_index[ht[i]] != -1) continue;
            const double* c = coef + 6 * (size_t)ht[i];
            count = count + 1;
            double x_ref = x[p] - hx_[i], y_ref = y[p] - hy_[i];
            Ex = Ex + c[1] + 2. * c[3] * x_ref + c[4] * y_ref;
            Ey = Ey + c[2] + 2. * c[5] * y_ref + c[4] * x_ref;
        }
        if (iface) { Ex = Ex / count; Ey = Ey / count; } else { Ex = Ex / 4; Ey = Ey / 4; }
        ex[p] = Ex; ey[p] = Ey;
    }
}

/* ------------------------------------------------------------------ MY_PCG */
/* MY_PCG.f90:49-141 with MATVEC_1 (CSR row sums) and MSOLVE_1 (X = Y / diag), 0-based CSR.
 * b: RHS (in) / final residual (out); x: guess (in) / solution (out).
 * Returns IERR (0 ok, 1 max iterations, 2 rho0 == 0, 3 p.Ap == 0); *iter = i-1. */
int or_pcg(int n, const int* rowptr, const int* col, const double* val, const double* diag,
           double* b, double* x, double reltol, int maxiter, int* iter) {
    double *X0 = malloc(sizeof(double) * n), *R0 = malloc(sizeof(double) * n), *W = malloc(sizeof(double) * n);
    double *P = malloc(sizeof(double) * n), *Q = malloc(sizeof(double) * n);
    double alpha, beta, rho = 0, rho0 = 0, normR, normRHS = 0, RELTOL2 = reltol * reltol;
    int i, j, ierr = 0;
    /* with or_threads > 1 every vector loop is split over threads (dot products then sum in
     * per-thread partials); with 1 thread each loop runs in row order */
    const int nt = or_threads;
#pragma omp parallel for num_threads(nt) if (nt > 1) schedule(static) reduction(+ : normRHS)
    for (j = 0; j < n; ++j) normRHS = normRHS + b[j] * b[j];
#pragma omp parallel for num_threads(nt) if (nt > 1) schedule(static)
    for (j = 0; j < n; ++j) {
        X0[j] = x[j];
    }
#pragma omp parallel for num_threads(nt) if (nt > 1) schedule(static)
    for (j = 0; j < n; ++j) {
        double s = 0;
        for (int k = rowptr[j]; k < rowptr[j + 1]; ++k) s = s + val[k] * X0[col[k]];
        R0[j] = b[j] - s;
    }
    for (i = 1; i <= maxiter; ++i) {
        rho = 0;
#pragma omp parallel for num_threads(nt) if (nt > 1) schedule(static) reduction(+ : rho)
        for (j = 0; j < n; ++j) { W[j] = R0[j] / diag[j]; rho = rho + R0[j] * W[j]; }
        if (i == 1) {
#pragma omp parallel for num_threads(nt) if (nt > 1) schedule(static)
            for (j = 0; j < n; ++j) P[j] = W[j];
        } else {
            if (rho0 == 0) { ierr = 2; break; }
            beta = rho / rho0;
#pragma omp parallel for num_threads(nt) if (nt > 1) schedule(static)
            for (j = 0; j < n; ++j) P[j] = W[j] + beta * P[j];
        }
        alpha = 0;
#pragma omp parallel for num_threads(nt) if (nt > 1) schedule(static) reduction(+ : alpha)
        for (j = 0; j < n; ++j) {
            double s = 0;
            for (int k = rowptr[j]; k < rowptr[j + 1]; ++k) s = s + val[k] * P[col[k]];
            Q[j] = s;
            alpha = alpha + P[j] * s;
        }
        if (alpha == 0) { ierr = 3; break; }
        alpha = rho / alpha;
        normR = 0;
#pragma omp parallel for num_threads(nt) if (nt > 1) schedule(static) reduction(+ : normR)
        for (j = 0; j < n; ++j) {
            X0[j] = X0[j] + alpha * P[j];
            R0[j] = R0[j] - alpha * Q[j];
            normR = normR + R0[j] * R0[j];
        }
        rho0 = rho;
        if (normR / normRHS <= RELTOL2) break;
    }
    if (i >= maxiter) ierr = 1;      /* MY_PCG.f90:129-133: ITER left unset */
    else *iter = i - 1;
    for (j = 0; j < n; ++j) { x[j] = X0[j]; b[j] = R0[j]; }
    free(X0); free(R0); free(W); free(P); free(Q);
    return ierr;
}

/* ------------------------------------------------------------------ assembly */
/* Q1 element integrals by 2x2 Gauss (Generate_Gauss_reference with 4 points), the way
 * Gauss_integration_local_stiffness and FE_RHS_HW_2D evaluate them: K_ab = int grad N_a.
 * grad N_b, M_ab = int N_a N_b (times r when axisymmetric). Dense element matrices for an
 * element with lower-left corner (x0,y0), sizes hx, hy; node order LL, LR, UR, UL. */
void or_element_matrices(int delta, double x0, double y0, double hx, double hy, double* K, double* M) {
    const double g = 1.0 / sqrt(3.0);
    const double gp[2] = {-g, g};
    for (int a = 0; a < 16; ++a) { K[a] = 0; M[a] = 0; }
    for (int qi = 0; qi < 2; ++qi)
        for (int qj = 0; qj < 2; ++qj) {
            double xi = gp[qi], et = gp[qj];
            double wq = (hx / 2.0) * (hy / 2.0);
            double yq = y0 + (et + 1.0) * hy / 2.0;
            if (delta == 1) wq *= yq;
            double N[4] = {0.25 * (1 - xi) * (1 - et), 0.25 * (1 + xi) * (1 - et), 0.25 * (1 + xi) * (1 + et), 0.25 * (1 - xi) * (1 + et)};
            double dNxi[4] = {-0.25 * (1 - et), 0.25 * (1 - et), 0.25 * (1 + et), -0.25 * (1 + et)};
            double dNet[4] = {-0.25 * (1 - xi), -0.25 * (1 + xi), 0.25 * (1 + xi), 0.25 * (1 - xi)};
            for (int a = 0; a < 4; ++a)
                for (int b = 0; b < 4; ++b) {
                    double gx = (dNxi[a] * 2.0 / hx) * (dNxi[b] * 2.0 / hx);
                    double gy = (dNet[a] * 2.0 / hy) * (dNet[b] * 2.0 / hy);
                    K[a * 4 + b] += wq * (gx + gy);
                    M[a * 4 + b] += wq * N[a] * N[b];
                }
        }
    (void)x0;
}

/* ------------------------------------------------------------------ IFE interface element
 * Bilinear immersed-finite-element matrices of one square element cut by a circle (the
 * interface of a dielectric object), following the reference's construction:
 *   - classification and vertex order (El_Object_Intersection_info_2D_BL.f90): A1..A4 is the
 *     element's LL, LR, UR, UL rotated so that A1 is the lone vertex (type 1: the circle cuts
 *     two adjacent edges) or, for type 2 (two opposite edges cut), A1 = LL when the bottom and
 *     top edges are cut and A1 = UL when the left and right edges are; E lies on edge A1A2, D
 *     on edge A4A1 (type 1) or A4A3 (type 2); piece 1 holds A1 (and A4 for type 2);
 *   - cut points: the circle's root on each cut edge (Line_Circle_Intersection_BL);
 *   - basis (Linear_IFE_Basis_Coeff_2D.f90): per piece c1 + c2 x + c3 y + c4 xy; nodal values;
 *     continuity at D, E and at the midpoint of DE (c4 equal when DE is axis-parallel); zero
 *     jump of beta * flux integrated along DE;
 *   - stiffness K_ij = sum over pieces beta int grad phi_i . grad phi_j (times r when
 *     axisymmetric, Gauss_integration_local_stiffness_IFE.f90), integrated exactly;
 *   - RHS operator M_ik (IFE_RHS_2D.f90 with Inter_Tetra_Partition_2D and Gauss_Nodes_2D): the
 *     pieces split into the reference's sub-triangles, each integrated by the edge-midpoint
 *     rule (weights area/3), rho_h = sum_k rho_k phi_k, so rhs_i = sum_k M_ik rho_k.
 * Coordinates inside are X = (x - x0)/h, Y = (y - y0)/h. inside[k]: vertex k (LL, LR, UR, UL)
 * inside the circle. Outputs K, M (4x4, row = test node, LL..UL), coef[4][2][4] (basis i,
 * piece, c1..c4 in X, Y), pts[4] = D, E in X, Y, order[4] = local index of A1..A4.
 * Returns the type (1, 2), 0 when the element is not cut, -1 for a saddle (two diagonal
 * vertices inside), -2 when a cut point falls within 1e-12 of a vertex. */
static int or_solve8(double A[8][8], double b[8], double x[8]) {
    for (int c = 0; c < 8; ++c) {
        int p = c;
        for (int r = c + 1; r < 8; ++r) if (fabs(A[r][c]) > fabs(A[p][c])) p = r;
        if (A[p][c] == 0.0) return -1;
        if (p != c) {
            for (int k = 0; k < 8; ++k) { double t = A[c][k]; A[c][k] = A[p][k]; A[p][k] = t; }
            double t = b[c]; b[c] = b[p]; b[p] = t;
        }
        for (int r = c + 1; r < 8; ++r) {
            double f = A[r][c] / A[c][c];
            for (int k = c; k < 8; ++k) A[r][k] -= f * A[c][k];
            b[r] -= f * b[c];
        }
    }
    for (int c = 7; c >= 0; --c) {
        double s = b[c];
        for (int k = c + 1; k < 8; ++k) s -= A[c][k] * x[k];
        x[c] = s / A[c][c];
    }
    return 0;
}

static double or_circle_cut(double xi, double yi, double xo, double yo, double cx, double cy, double r) {
    /* parameter t in (0, 1) of the circle crossing on the segment inside point -> outside point */
    double dx = xo - xi, dy = yo - yi, fx = xi - cx, fy = yi - cy;
    double a = dx * dx + dy * dy, bh = fx * dx + fy * dy, c = fx * fx + fy * fy - r * r;
    return (-bh + sqrt(bh * bh - a * c)) / a;
}

int or_ife_element(int delta, double x0, double y0, double h, const int* inside, double cx, double cy, double rad,
                   double beta_in, double beta_out, double* K, double* M, double* coef, double* pts, int* order) {
    static const double VX[4] = {0, 1, 1, 0}, VY[4] = {0, 0, 1, 1};
    int kin = inside[0] + inside[1] + inside[2] + inside[3];
    if (kin == 0 || kin == 4) return 0;
    int rot, type;
    if (kin == 1 || kin == 3) {
        type = 1; rot = 0;
        for (int k = 0; k < 4; ++k) {
            int same = 0;
            for (int m = 0; m < 4; ++m) same += (inside[m] == inside[k]);
            if (same == 1) rot = k;
        }
    } else {
        if (inside[0] == inside[2]) return -1;
        type = 2;
        rot = (inside[0] == inside[3]) ? 0 : 3;   /* LL, UL alike: bottom and top edges cut */
    }
    int A[4];
    for (int k = 0; k < 4; ++k) A[k] = (rot + k) & 3;
    /* cut points: E on A1A2, D on A4A1 (type 1) or A4A3 (type 2) */
    int eD0 = A[3], eD1 = (type == 1) ? A[0] : A[2], eE0 = A[0], eE1 = A[1];
    double P[2][2];
    int ends[2][2] = {{eD0, eD1}, {eE0, eE1}};
    for (int q = 0; q < 2; ++q) {
        int u = ends[q][0], v = ends[q][1];
        int vi = inside[u] ? u : v, vo = inside[u] ? v : u;
        double t = or_circle_cut(x0 + h * VX[vi], y0 + h * VY[vi], x0 + h * VX[vo], y0 + h * VY[vo], cx, cy, rad);
        if (!(t > 1e-12 && t < 1.0 - 1e-12)) return -2;
        P[q][0] = VX[vi] + t * (VX[vo] - VX[vi]);
        P[q][1] = VY[vi] + t * (VY[vo] - VY[vi]);
    }
    const double Dx = P[0][0], Dy = P[0][1], Ex = P[1][0], Ey = P[1][1];
    int piece[4];
    for (int k = 0; k < 4; ++k) piece[A[k]] = (k == 0 || (type == 2 && k == 3)) ? 0 : 1;
    double beta[2];
    beta[0] = inside[A[0]] ? beta_in : beta_out;
    beta[1] = inside[A[0]] ? beta_out : beta_in;
    double C[4][2][4];
    for (int i = 0; i < 4; ++i) {
        double S[8][8] = {{0}}, b[8] = {0}, x[8];
        for (int k = 0; k < 4; ++k) {
            int o = 4 * piece[k];
            S[k][o] = 1; S[k][o + 1] = VX[k]; S[k][o + 2] = VY[k]; S[k][o + 3] = VX[k] * VY[k];
            b[k] = (k == i) ? 1.0 : 0.0;
        }
        double cp[3][2] = {{Dx, Dy}, {Ex, Ey}, {0.5 * (Dx + Ex), 0.5 * (Dy + Ey)}};
        for (int q = 0; q < 3; ++q) {
            double X = cp[q][0], Y = cp[q][1];
            S[4 + q][0] = 1; S[4 + q][1] = X; S[4 + q][2] = Y; S[4 + q][3] = X * Y;
            S[4 + q][4] = -1; S[4 + q][5] = -X; S[4 + q][6] = -Y; S[4 + q][7] = -X * Y;
        }
        if (Dx == Ex || Dy == Ey) {
            for (int k = 0; k < 8; ++k) S[6][k] = 0;
            S[6][3] = 1; S[6][7] = -1;
        }
        double fq = 0.5 * (Ey * Ey - Dy * Dy - Ex * Ex + Dx * Dx);
        S[7][1] = beta[0] * (Ey - Dy); S[7][2] = -beta[0] * (Ex - Dx); S[7][3] = beta[0] * fq;
        S[7][5] = -beta[1] * (Ey - Dy); S[7][6] = beta[1] * (Ex - Dx); S[7][7] = -beta[1] * fq;
        if (or_solve8(S, b, x)) return -3;
        for (int pc = 0; pc < 2; ++pc)
            for (int k = 0; k < 4; ++k) C[i][pc][k] = x[4 * pc + k];
    }
    /* the reference's sub-triangles, points as indices: 0..3 = A1..A4, 4 = D, 5 = E */
    double Q[6][2];
    for (int k = 0; k < 4; ++k) { Q[k][0] = VX[A[k]]; Q[k][1] = VY[A[k]]; }
    Q[4][0] = Dx; Q[4][1] = Dy; Q[5][0] = Ex; Q[5][1] = Ey;
    static const int T1[4][4] = {{0, 4, 5, 0}, {1, 5, 2, 1}, {5, 2, 4, 1}, {2, 4, 3, 1}};
    static const int T2[4][4] = {{0, 5, 3, 0}, {3, 4, 5, 0}, {5, 2, 4, 1}, {2, 5, 1, 1}};
    const int (*T)[4] = (type == 1) ? T1 : T2;
    for (int a = 0; a < 16; ++a) { K[a] = 0; M[a] = 0; }
    const double gq = sqrt(0.6), gw[3] = {5.0 / 18.0, 8.0 / 18.0, 5.0 / 18.0};
    const double gx[3] = {0.5 * (1 - gq), 0.5, 0.5 * (1 + gq)};
    for (int st = 0; st < 4; ++st) {
        const double *p1 = Q[T[st][0]], *p2 = Q[T[st][1]], *p3 = Q[T[st][2]];
        int pc = T[st][3];
        double area = 0.5 * fabs((p2[0] - p1[0]) * (p3[1] - p1[1]) - (p3[0] - p1[0]) * (p2[1] - p1[1]));
        /* stiffness: collapsed 3x3 Gauss-Legendre, exact to total degree 4 */
        for (int qa = 0; qa < 3; ++qa)
            for (int qb = 0; qb < 3; ++qb) {
                double u = gx[qa], v = gx[qb] * (1.0 - u);
                double w = 2.0 * area * gw[qa] * gw[qb] * (1.0 - u);
                double X = p1[0] + u * (p2[0] - p1[0]) + v * (p3[0] - p1[0]);
                double Y = p1[1] + u * (p2[1] - p1[1]) + v * (p3[1] - p1[1]);
                if (delta == 1) w *= y0 + h * Y;
                double gxv[4], gyv[4];
                for (int i = 0; i < 4; ++i) {
                    gxv[i] = C[i][pc][1] + C[i][pc][3] * Y;
                    gyv[i] = C[i][pc][2] + C[i][pc][3] * X;
                }
                for (int i = 0; i < 4; ++i)
                    for (int j = 0; j < 4; ++j) K[i * 4 + j] += w * beta[pc] * (gxv[i] * gxv[j] + gyv[i] * gyv[j]);
            }
        /* RHS operator: edge midpoints, weight area/3 (physical area h^2 area) */
        const double* tv[3] = {p1, p2, p3};
        for (int g = 0; g < 3; ++g) {
            double X = 0.5 * (tv[g][0] + tv[(g + 1) % 3][0]), Y = 0.5 * (tv[g][1] + tv[(g + 1) % 3][1]);
            double w = h * h * area / 3.0;
            if (delta == 1) w *= y0 + h * Y;
            double phi[4];
            for (int i = 0; i < 4; ++i) phi[i] = C[i][pc][0] + C[i][pc][1] * X + C[i][pc][2] * Y + C[i][pc][3] * X * Y;
            for (int i = 0; i < 4; ++i)
                for (int k = 0; k < 4; ++k) M[i * 4 + k] += w * phi[i] * phi[k];
        }
    }
    if (coef)
        for (int i = 0; i < 4; ++i)
            for (int pc = 0; pc < 2; ++pc)
                for (int k = 0; k < 4; ++k) coef[(i * 2 + pc) * 4 + k] = C[i][pc][k];
    if (pts) { pts[0] = Dx; pts[1] = Dy; pts[2] = Ex; pts[3] = Ey; }
    if (order) for (int k = 0; k < 4; ++k) order[k] = A[k];
    return type;
}

/* rhs_i = sum_e int N_i rho_h - (Dirichlet lift) over unknown rows; node_type2: unknown id
 * (1-based) or <= 0 for Dirichlet nodes with values dphi. Global_RHS_PDE_2D.f90:78-140 +
 * Global_RHS_EBC_2D_bjw.f90:66-78 restated by element loops. */
void or_rhs(int nnx, int nny, const double* xs, const double* ys, int delta, const int* node_type2,
            const double* dphi, const double* rho, int nunk, double* rhs) {
    int nex = nnx - 1, ney = nny - 1;
    for (int u = 0; u < nunk; ++u) rhs[u] = 0.0;
    double* ebc = calloc(nunk, sizeof(double));
    const int nt = or_threads;
    /* with threads, element columns of one parity share no node: even columns, then odd */
    for (int par = 0; par < (nt > 1 ? 2 : 1); ++par) {
        const int c0 = nt > 1 ? par : 0, cs = nt > 1 ? 2 : 1;
#pragma omp parallel for num_threads(nt) if (nt > 1) schedule(static)
        for (int ci = c0; ci < nex; ci += cs)
            for (int cj = 0; cj < ney; ++cj) {
                double K[16], M[16];
                int ht[4] = {ci * nny + cj, (ci + 1) * nny + cj, (ci + 1) * nny + cj + 1, ci * nny + cj + 1};
                or_element_matrices(delta, xs[ci], ys[cj], xs[ci + 1] - xs[ci], ys[cj + 1] - ys[cj], K, M);
                for (int a = 0; a < 4; ++a) {
                    int u = node_type2[ht[a]];
                    if (u <= 0) continue;
                    double s = 0.0;
                    for (int b = 0; b < 4; ++b) s += M[a * 4 + b] * rho[ht[b]];
                    rhs[u - 1] += s;
                    for (int b = 0; b < 4; ++b) if (node_type2[ht[b]] <= 0) ebc[u - 1] += K[a * 4 + b] * dphi[ht[b]];
                }
            }
    }
    for (int u = 0; u < nunk; ++u) rhs[u] = rhs[u] - ebc[u];
    free(ebc);
}

/* y = A x over unknowns with the element-assembled stiffness (matrix free). */
void or_stiffness_apply(int nnx, int nny, const double* xs, const double* ys, int delta, const int* node_type2,
                        int nunk, const double* xu, double* yu) {
    int nex = nnx - 1, ney = nny - 1;
    double K[16], M[16];
    for (int u = 0; u < nunk; ++u) yu[u] = 0.0;
    for (int ci = 0; ci < nex; ++ci)
        for (int cj = 0; cj < ney; ++cj) {
            int ht[4] = {ci * nny + cj, (ci + 1) * nny + cj, (ci + 1) * nny + cj + 1, ci * nny + cj + 1};
            or_element_matrices(delta, xs[ci], ys[cj], xs[ci + 1] - xs[ci], ys[cj + 1] - ys[cj], K, M);
            for (int a = 0; a < 4; ++a) {
                int u = node_type2[ht[a]];
                if (u <= 0) continue;
                for (int b = 0; b < 4; ++b) {
                    int v = node_type2[ht[b]];
                    if (v > 0) yu[u - 1] += K[a * 4 + b] * xu[v - 1];
                }
            }
        }
}

/* ------------------------------------------------------------------ MCC */
#define OR_KB 1.3807e-23
#define OR_EXP 2.71828
#define OR_JTOEV 1.6022e-19
#define OR_PI 3.141592653589793238
#define OR_MINREAL 1.0e-15

/* Selection: slot s of a live particle is selected when word s % 4 of the Philox block
 * with counter (s / 4, step, tag, 0) is below floor(CollisionRatio * 2^32) (a Bernoulli
 * trial per particle standing in for SelectParticle's count + shuffle, MCCPublic.F90:113-171).
 * Writes the selected slots in slot order. */
static int or_selected(int64_t s, const int* cell, uint32_t threshold, uint64_t seed, uint32_t step, uint32_t tag) {
    if (cell[s] < 0) return 0;
    or_p4 o = or_philox((uint32_t)(s >> 2), step, tag, 0u, (uint32_t)seed, (uint32_t)(seed >> 32));
    return o.v[s & 3] < threshold;
}

int64_t or_mcc_select(int64_t n, const int* cell, uint32_t threshold, uint64_t seed, uint32_t step, uint32_t tag, int* out) {
    const int nt = or_threads;
    if (nt <= 1) {
        int64_t k = 0;
        for (int64_t s = 0; s < n; ++s)
            if (or_selected(s, cell, threshold, seed, step, tag)) out[k++] = (int)s;
        return k;
    }
    /* threaded: count per chunk, exclusive prefix, write in slot order (same output) */
    int64_t cnt[65] = {0};
    const int nc = nt < 64 ? nt : 64;
    const int64_t len = (n + nc - 1) / nc;
#pragma omp parallel for num_threads(nc) schedule(static)
    for (int c = 0; c < nc; ++c) {
        int64_t k = 0;
        for (int64_t s = c * len; s < n && s < (c + 1) * len; ++s) k += or_selected(s, cell, threshold, seed, step, tag);
        cnt[c + 1] = k;
    }
    for (int c = 0; c < nc; ++c) cnt[c + 1] += cnt[c];
#pragma omp parallel for num_threads(nc) schedule(static)
    for (int c = 0; c < nc; ++c) {
        int64_t k = cnt[c];
        for (int64_t s = c * len; s < n && s < (c + 1) * len; ++s)
            if (or_selected(s, cell, threshold, seed, step, tag)) out[k++] = (int)s;
    }
    return cnt[nc];
}

typedef struct { double x, y, z; } or_v3;

static or_v3 or_vel_maxwell(or_stream* s, double mass, double temp) {       /* ParticleOne.f90:135-152 */
    double Beta = 1.0 / (OR_KB * temp), FuncA = 1.0, FuncB = 0.0, R = 0.5;
    int it = 0;
    while (FuncA > FuncB) {
        R = or_u01(s); FuncA = R * R;
        double R1 = R;
        R = or_u01(s); FuncB = -OR_EXP * R * log(R);
        if (or_plant == 3 && !(FuncA > FuncB)) R = R1;
        if (++it > 1000) break;
    }
    double V = sqrt(-3.0 * log(R) / Beta / mass);
    R = or_u01(s);                                                               /* :247-261 */
    double CosTheta = 1.0 - 2.0 * R, SinTheta = sqrt(1.0 - CosTheta * CosTheta);
    R = or_u01(s);
    double Fai = 2.0 * OR_PI * R;
    or_v3 v = {V * CosTheta, V * SinTheta * cos(Fai), V * SinTheta * sin(Fai)};
    return v;
}

typedef struct { double Gx, Gy, Gz, Gper, G, Energy, MassRatio; or_v3 v, vt; } or_mcpo;

static void or_post(or_mcpo* M, or_stream* s, int iso) {                     /* TypeMCC.f90:101-133 */
    if (fabs(M->Gper) < OR_MINREAL) {
        double R = or_u01(s), CosTheta = 1.0 - 2.0 * R, SinTheta = sqrt(1.0 - CosTheta * CosTheta);
        R = or_u01(s);
        double Fai = 2.0 * OR_PI * R;
        M->v.x = M->G * CosTheta; M->v.y = M->G * SinTheta * cos(Fai); M->v.z = M->G * SinTheta * sin(Fai);
        return;
    }
    double CosTheta;
    if (!iso) {                                                                  /* MCCChi.f90:13-25 */
        if (M->Energy < OR_MINREAL) CosTheta = 1.0;
        else { double R = or_u01(s); CosTheta = (M->Energy + 2.0 - 2.0 * pow(M->Energy + 1.0, R)) / M->Energy; }
    } else { double R = or_u01(s); CosTheta = 1.0 - 2.0 * R; }
    double FcosTheta = 1.0 - CosTheta, SinTheta = sqrt(1.0 - CosTheta * CosTheta);
    double R = or_u01(s), APhi = 2.0 * OR_PI * R, CosPhi = cos(APhi), SinPhi = sin(APhi);
    double hx = M->Gper * CosPhi;
    double hy = -(M->Gx * M->Gy * CosPhi + (or_plant == 1 ? 0.0 : M->G * M->Gz * SinPhi)) / M->Gper;
    double hz = -(M->Gx * M->Gz * CosPhi - M->G * M->Gy * SinPhi) / M->Gper;
    M->v.x = M->v.x - M->MassRatio * (M->Gx * FcosTheta + hx * SinTheta);
    M->v.y = M->v.y - M->MassRatio * (M->Gy * FcosTheta + hy * SinTheta);
    M->v.z = M->v.z - M->MassRatio * (M->Gz * FcosTheta + hz * SinTheta);
}

static void or_vupd(or_mcpo* M, double f) {                                   /* TypeMCC.f90:84-99 */
    M->Gx *= f; M->Gy *= f; M->Gz *= f; M->Gper *= f; M->G *= f;
    M->v.x *= f; M->v.y *= f; M->v.z *= f;
}

/* SelectProbility (TypeMCC.f90:155-198) before the draw: the reaction probabilities at
 * Energy (eV) from Probility(nreact, nsigma) (column-major, rows on the log10 grid emin +
 * i*eint): N = ceiling((log10 E - emin)/eint); N < 1: the first row; N < nsigma: linear
 * interpolation between rows N and N+1 (1-based) in log10 E; otherwise the last row times
 * E/EnergyMax; zero below each reaction's threshold. */
void or_select_prob(double Energy, int nreact, int nsigma, double emin, double eint, double emax, const double* prob,
                    const double* thr, double* P) {
    double lE = log10(Energy);
    int N = (int)ceil((lE - emin) / eint);
    int sh = or_plant == 2 ? 1 : 0;
    if (N < 1) for (int i = 0; i < nreact; ++i) P[i] = prob[i];
    else if (N < nsigma - sh) {
        double S1 = (double)N - (lE - emin) / eint, S2 = 1.0 - S1;
        for (int i = 0; i < nreact; ++i)
            P[i] = prob[(size_t)(N - 1 + sh) * nreact + i] * S1 + prob[(size_t)(N + sh) * nreact + i] * S2;
    } else for (int i = 0; i < nreact; ++i) P[i] = prob[(size_t)(nsigma - 1) * nreact + i] * Energy / emax;
    for (int i = 0; i < nreact; ++i) if (Energy < thr[i]) P[i] = 0.0;
}

/* Test hooks: PostCollisionVelocity (TypeMCC.f90:101-133) of a projectile at velocity v
 * (m/s) with relative velocity g = v - vt (m/s), energy (eV) and mass ratio M/(m + M),
 * elastic (iso 0: CosKaiVahedi, MCCChi.f90:13-25) or isotropic (iso 1), on the stream
 * (seed; slot, 0, 0); and n target-gas velocities of VelMaxInit (ParticleOne.f90:135-152,
 * :247-261) from the streams (seed; k, 0, 0). */
void or_post_test(const double* g, const double* v, double energy, double mass_ratio, int iso, uint64_t seed,
                  uint32_t slot, double* v_out) {
    or_stream s;
    or_stream_init(&s, seed, slot, 0u, 0u);
    or_mcpo M;
    M.Gx = g[0]; M.Gy = g[1]; M.Gz = g[2];
    M.Gper = sqrt(g[1] * g[1] + g[2] * g[2]);
    M.G = sqrt(M.Gper * M.Gper + g[0] * g[0]);
    M.Energy = energy; M.MassRatio = mass_ratio;
    M.v.x = v[0]; M.v.y = v[1]; M.v.z = v[2];
    or_post(&M, &s, iso);
    v_out[0] = M.v.x; v_out[1] = M.v.y; v_out[2] = M.v.z;
}

void or_vel_maxwell_test(uint64_t seed, int64_t n, double mass, double temp, double* v) {
    for (int64_t k = 0; k < n; ++k) {
        or_stream s;
        or_stream_init(&s, seed, (uint32_t)k, 0u, 0u);
        or_v3 u = or_vel_maxwell(&s, mass, temp);
        v[3 * k] = u.x; v[3 * k + 1] = u.y; v[3 * k + 2] = u.z;
    }
}

/* VelKappaInit (kind 0, ParticleOne.f90:154-204) and VelPolyInit (kind 2, gama = 3, :206-245):
 * n velocities (Vx; Vy = Vz = 0, as the reference stores them) for mass M (kg) at T (K), on the
 * streams (seed; k, 0, 0). */
void or_load_velocity(int kind, uint64_t seed, int64_t n, double T, double M, double* v) {
    const double kB = 1.3807e-23, Te = 11605.0, Me = 9.1095e-31;
    for (int64_t k = 0; k < n; ++k) {
        or_stream s;
        or_stream_init(&s, seed, (uint32_t)k, 0u, 0u);
        double V = 0.0;
        if (kind == 0) {
            const double Kappa = 2.0;
            const double Sigma = ((Kappa - 1.5) / Kappa) * (2.0 * kB * T / M);
            const double Sigma_e = ((Kappa - 1.5) / Kappa) * (2.0 * kB * Te / Me);
            const double ex = or_plant == 6 ? -0.5 * Kappa : -Kappa;
            for (int it = 0; it < 100000; ++it) {
                V = sqrt(Sigma / Sigma_e) * (20.0 * or_u01(&s) - 10.0);
                double P = pow(1.0 + (V * V) / ((2 * Kappa - 3) * Sigma / Sigma_e), ex);
                if (or_u01(&s) <= P) break;
            }
        } else {
            const double gama = 3.0, Sigma = 2.0 * kB * T / M, Sigma_e = 2.0 * kB * Te / Me;
            V = sqrt(Sigma / Sigma_e) * 2.0 * sqrt(gama) * (or_u01(&s) - 0.5);
        }
        v[3 * k] = V; v[3 * k + 1] = 0.0; v[3 * k + 2] = 0.0;
    }
}

/* One collision (MCCPublic.F90:58-90) of selected particle `slot` (normalised velocity
 * v in/out). prob: Probility(nreact, nsigma) column-major. Returns the reaction index
 * (0 = null); for ionisation writes the new electron and ion velocities (normalised)
 * into new_e[3], new_i[3]. */
int or_mcc_collide(int electron, int slot, uint64_t seed, uint32_t step, uint32_t tag,
                   double v_ref, double mass, double gas_mass, double gas_temp,
                   int nreact, int nsigma, double emin, double eint, double emax,
                   const double* prob, const int* rtype, const double* thr,
                   double* v, double* new_e, double* new_i) {
    or_stream s;
    or_stream_init(&s, seed, (uint32_t)slot, step, tag);
    or_mcpo M;
    M.v.x = v[0] * v_ref; M.v.y = v[1] * v_ref; M.v.z = v[2] * v_ref;                  /* VelRes(VFactor) */
    double Miu = mass * gas_mass / (mass + gas_mass);
    M.MassRatio = gas_mass / (mass + gas_mass);
    M.vt = or_vel_maxwell(&s, gas_mass, gas_temp);
    M.Gx = M.v.x - M.vt.x; M.Gy = M.v.y - M.vt.y; M.Gz = M.v.z - M.vt.z;
    double Gtemp = M.Gy * M.Gy + M.Gz * M.Gz;
    M.Gper = sqrt(Gtemp);
    double G2 = Gtemp + M.Gx * M.Gx;
    M.G = sqrt(G2);
    M.Energy = 0.5 * Miu * G2 / OR_JTOEV;
    double P[4] = {0, 0, 0, 0};
    or_select_prob(M.Energy, nreact, nsigma, emin, eint, emax, prob, thr, P);
    double R = or_u01(&s);
    int index = 0;
    for (int i = 0; i < nreact; ++i) if (R < P[i]) { index = i + 1; break; }
    if (index > 0) {
        int type = rtype[index - 1];
        double th = thr[index - 1];
        if (electron) {
            if (type == 11) or_post(&M, &s, 0);
            else if (type == 12) { double f = sqrt(1.0 - th / M.Energy); M.Energy = M.Energy - th; or_vupd(&M, f); or_post(&M, &s, 0); }
            else if (type == 13) {                                                   /* MCCElectron.F90:48-80 */
                or_v3 vsolid = M.v;
                double Eorig = M.Energy, EnergyTemp = M.Energy - th;
                double R1 = or_u01(&s);
                double EnergyOld = 10.0 * tan(R1 * atan(EnergyTemp) / 20.0);
                if (fabs(EnergyOld) < OR_MINREAL) EnergyOld = OR_MINREAL;
                double f = sqrt(EnergyOld / M.Energy);
                or_vupd(&M, f);
                or_post(&M, &s, 0);
                double EnergyNew = EnergyTemp - EnergyOld, f2 = sqrt(EnergyNew / Eorig);
                new_e[0] = vsolid.x * f2 / v_ref; new_e[1] = vsolid.y * f2 / v_ref; new_e[2] = vsolid.z * f2 / v_ref;
                new_i[0] = M.vt.x / v_ref; new_i[1] = M.vt.y / v_ref; new_i[2] = M.vt.z / v_ref;
            }
        } else {
            if (type == 101) or_post(&M, &s, 1);
            else if (type == 103 || type == 102) M.v = M.vt;
        }
    }
    double vf = 1.0 / v_ref;
    v[0] = M.v.x * vf; v[1] = M.v.y * vf; v[2] = M.v.z * vf;
    return index;
}

/* Cumulative probability table of ProbilityNonReactive (MCCInit.F90:115-151) from
 * sigma(nreact, nsigma) (m^2) on the 10^(emin + i*eint) grid. Returns SigmaMax. */
double or_mcc_table(int nreact, int nsigma, double emin, double eint, double mass, double gas_mass, double density,
                    const double* sigma, double* prob) {
    double Miu = mass * gas_mass / (mass + gas_mass), mx = 0.0;
    for (int i = 0; i < nsigma; ++i) {
        double Energy = pow(10.0, emin + (double)i * eint);
        double V = sqrt(2.0 * Energy * OR_JTOEV / Miu);
        for (int j = 0; j < nreact; ++j) prob[(size_t)i * nreact + j] = sigma[(size_t)i * nreact + j] * V * density;
    }
    for (int i = 0; i < nsigma; ++i)
        for (int j = 1; j < nreact; ++j) prob[(size_t)i * nreact + j] = prob[(size_t)i * nreact + j] + prob[(size_t)i * nreact + j - 1];
    for (size_t k = 0; k < (size_t)nsigma * nreact; ++k) if (prob[k] > mx) mx = prob[k];
    for (size_t k = 0; k < (size_t)nsigma * nreact; ++k) prob[k] = prob[k] / mx;
    return mx;
}

/* ------------------------------------------------------------------ wall injection */
/* InjectBeams_2D (IMPIC/InjectBeams_2D.f90:40-143) with the build's per-particle Philox
 * streams. Constants (:42-50, :91-95, Loadv.f90:14; SQRT(2.) is single precision in the
 * reference): consts = {b_amb, v_r, FS1, FS2, QA, sqrt(2) vt, VMP, XL, XU}. */
void or_inject_consts(double vd1, double vt, double tmpj, double xm, double r_start, double dt, double* c) {
    const double s2 = (double)sqrtf(2.0f);
    double b;
    if (vd1 > 0.0)
        b = vt / sqrt(2.0 * OR_PI) * exp(-(vd1 * vd1) / (2.0 * vt * vt)) + vd1 / 2.0 * (1.0 + erf(vd1 / (s2 * vt)));
    else
        b = -vt / sqrt(2.0 * OR_PI) * exp(-(vd1 * vd1) / (2.0 * vt * vt)) + vd1 / 2.0 * (1.0 - erf(vd1 / (s2 * vt)));
    const double v_r = fabs(vd1) / (s2 * vt);
    const double fs1 = v_r + sqrt(v_r * v_r + 2.0);
    c[0] = b; c[1] = v_r; c[2] = fs1; c[3] = 0.5 * (1.0 + v_r * (2.0 * v_r - fs1));
    c[4] = v_r < -3.0 ? fabs(v_r) + 1.0 : 3.0;
    c[5] = s2 * vt;
    c[6] = sqrt(2.0 * tmpj / xm);
    const double x2 = r_start + 0.5 * b * dt;
    c[7] = r_start < x2 ? r_start : x2; c[8] = r_start < x2 ? x2 : r_start;
}

/* add_N = INT(dens0 n_ref |b_amb| v_ref dt t_ref (y_hi - y_lo) L_ref / Weight + R) (:56-57),
 * R = the first draw of the stream at slot 0xFFFFFFFF. */
int64_t or_inject_count(double expect_no_r, uint64_t seed, uint32_t step, uint32_t tag) {
    or_stream s;
    or_stream_init(&s, seed, 0xFFFFFFFFu, step, tag);
    return (int64_t)(expect_no_r + or_u01(&s));
}

/* Candidate i of a rejection loop takes the two uniforms of Philox block i of its stream
 * (a rejected candidate's second uniform goes unused). */
static void or_block_u01(uint64_t seed, uint32_t slot, uint32_t step, uint32_t tag, uint32_t blk, double* ra,
                         double* rb) {
    or_p4 o = or_philox(slot, step, tag, blk, (uint32_t)seed, (uint32_t)(seed >> 32));
    *ra = (((double)((((uint64_t)o.v[0]) << 20) | (uint64_t)(o.v[1] >> 12))) + 0.5) * (1.0 / 4503599627370496.0);
    *rb = (((double)((((uint64_t)o.v[2]) << 20) | (uint64_t)(o.v[3] >> 12))) + 0.5) * (1.0 / 4503599627370496.0);
}

/* v if v < hi, else the largest value below hi of the particle precision (bits 32: float) */
static double or_below(double v, double hi, int bits) {
    if (v < hi) return v;
    if (bits == 64) return nextafter(hi, -INFINITY);
    float b = (float)hi;
    while ((double)b >= hi) b = nextafterf(b, -INFINITY);
    return (double)b;
}

static double or_loadv(uint64_t seed, uint32_t slot, uint32_t step, uint32_t tag, double vmp) {   /* Loadv.f90:14-23 */
    for (uint32_t i = 0;; ++i) {
        double R, R1;
        or_block_u01(seed, slot, step, tag, i, &R, &R1);
        double v = (6.0 * R - 3.0) * vmp;
        double pacc = exp(-v * v / (vmp * vmp));
        if (!(R1 > pacc)) return v;
    }
}

/* n particles of one source, each from four streams (slot 4k + c): c = 0 Bird's inward
 * component (:91-106), c = 1, 2 Loadv for Vy, Vz (+ ion drift, :121-126), c = 3 PosInit's X
 * then Y (ParticleOne.f90:82-89); positions rounded to the particle precision (bits 32:
 * float) before InjectPositioning (InjectPositioning.f90:13-22, no clamp; cell -1 when
 * outside); a stored X or Y that rounds onto the slab's upper bound takes the largest value
 * below it, keeping the particle strictly inside the slab as sampled. vd = drift {vx, vy, vz}; mesh: xmin, ymin, mapx, mapy, nex, ney. */
void or_inject(int bits, int64_t n, int electron, const double* c, const double* vd, double y_lo, double y_hi,
               uint64_t seed, uint32_t step, uint32_t tag, double xmin, double ymin, double mapx, double mapy,
               int nex, int ney, double* x, double* y, double* vx, double* vy, double* vz, int* cell, double* th) {
    /* th != NULL: axisymmetric PosInit (ParticleOne.f90:95-98), r = sqrt(YL^2 + (YU+YL)(YU-YL) R)
     * and the azimuth PO%Z = 2 PI R' (R' the first uniform of block 1 of the position stream) */
    for (int64_t k = 0; k < n; ++k) {
        double UN = 0.0;
        for (uint32_t i = 0;; ++i) {
            double R, R2;
            or_block_u01(seed, (uint32_t)(4 * k), step, tag, i, &R, &R2);
            double UU = -c[4] + 2.0 * c[4] * R;
            UN = UU + c[1];
            if (UN < 0.0) continue;
            double GG = (2.0 * UN / c[2]) * exp(c[3] - UU * UU);
            if (GG < R2) continue;
            break;
        }
        double Vx = UN * c[5];
        double Vy = or_loadv(seed, (uint32_t)(4 * k + 1), step, tag, c[6]);
        double Vz = or_loadv(seed, (uint32_t)(4 * k + 2), step, tag, c[6]);
        if (!electron) { Vy = vd[1] + Vy; Vz = vd[2] + Vz; }
        double RX, RY;
        or_block_u01(seed, (uint32_t)(4 * k + 3), step, tag, 0, &RX, &RY);
        double X = c[7] + (c[8] - c[7]) * RX;
        double Y = (th && or_plant != 5) ? sqrt(y_lo * y_lo + (y_hi + y_lo) * (y_hi - y_lo) * RY) : y_lo + (y_hi - y_lo) * RY;
        if (th) {
            double RZ, RZ2;
            or_block_u01(seed, (uint32_t)(4 * k + 3), step, tag, 1, &RZ, &RZ2);
            th[k] = 2.0 * OR_PI * RZ;
            if (bits == 32) th[k] = (double)(float)th[k];
        }
        if (bits == 32) { X = (double)(float)X; Y = (double)(float)Y; Vx = (double)(float)Vx; Vy = (double)(float)Vy; Vz = (double)(float)Vz; }
        X = or_below(X, c[8], bits); Y = or_below(Y, y_hi, bits);
        int ii = (int)((X - xmin) / mapx) + 1, jj = (int)((Y - ymin) / mapy) + 1;
        x[k] = X; y[k] = Y; vx[k] = Vx; vy[k] = Vy; vz[k] = Vz;
        cell[k] = (ii < 1 || ii > nex || jj < 1 || jj > ney) ? -1 : (jj - 1) + (ii - 1) * ney;
    }
}
