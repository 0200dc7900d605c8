// Null-collision Monte Carlo (MCC_jw/code/mc) and the periodic cell sort.
//
// Selection: every live slot is tested with one Philox word against
// floor(CollisionRatio * 2^32) (a Bernoulli(CollisionRatio) trial per particle, the GPU
// form of SelectParticle's INT(N*P) + Bernoulli(residue) count with a partial
// Fisher-Yates shuffle, MCCPublic.F90:113-171: same expected count, no serial shuffle).
// Slots 4g..4g+3 share one Philox call of counter (g, step, tag, 0); slot s tests word
// v[s & 3]. A mask pass writes one bit per slot and per-block counts, one workgroup
// scans the counts, and a write pass compacts the selected slots in slot order.
// Collision: one thread per selected particle, FP64 kinematics of TypeMCC / MCCElectron /
// MCCIon / MCCChi drawing from the particle's own (slot, step, species) Philox stream.
// Ionisation products are staged per selected particle and appended in selected order.
#include "pic_device.h"

// A wave owns 2048 consecutive slots as 8 chunks of 256; in chunk j lane l tests slots
// wave_base + 256 j + 4 l + (0..3) (one int4 of cell, one Philox call) and keeps the
// four results in bits 4 j .. 4 j + 3 of its mask word.
#define SEL_BLOCK 256
#define SEL_CHUNKS 8
#define SEL_WAVE_SLOTS (SEL_CHUNKS * 256)
#define SEL_BLOCK_SLOTS (SEL_WAVE_SLOTS * (SEL_BLOCK / 64))
#define SCAN_THREADS 1024

__device__ inline uint32_t sel_nibble(int64_t s, int64_t n, const int* cell, uint32_t thr, uint64_t seed,
                                      uint32_t step, uint32_t tag) {
    int c[4];
    if (s + 3 < n) {
        int4 q = *(const int4*)(cell + s);
        c[0] = q.x; c[1] = q.y; c[2] = q.z; c[3] = q.w;
    } else {
        for (int k = 0; k < 4; ++k) c[k] = (s + k < n) ? cell[s + k] : -1;
    }
    if ((c[0] & c[1] & c[2] & c[3]) < 0) return 0u;          // all four dead (or past n)
    philox4 o = pic_philox4x32_10((uint32_t)(s >> 2), step, tag, 0u, (uint32_t)seed, (uint32_t)(seed >> 32));
    uint32_t m = 0;
    for (int k = 0; k < 4; ++k) m |= (c[k] >= 0 && o.v[k] < thr) ? (1u << k) : 0u;
    return m;
}

__device__ inline int wave_sum(int v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// mask word per lane (wave-major), selected count per block
__global__ __launch_bounds__(SEL_BLOCK) void k_sel_mask(int64_t n, const int* cell, uint32_t thr, uint64_t seed,
                                                        uint32_t step, uint32_t tag, uint32_t* mask, int* blockcount) {
    __shared__ int sh[SEL_BLOCK / 64];
    int lane = threadIdx.x & 63;
    int64_t wave = (int64_t)blockIdx.x * (SEL_BLOCK / 64) + (threadIdx.x >> 6);
    int64_t base = wave * SEL_WAVE_SLOTS + 4 * lane;
    uint32_t word = 0;
#pragma unroll
    for (int j = 0; j < SEL_CHUNKS; ++j) word |= sel_nibble(base + 256 * j, n, cell, thr, seed, step, tag) << (4 * j);
    mask[wave * 64 + lane] = word;
    int c = wave_sum(__popc(word));
    if (lane == 0) sh[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < SEL_BLOCK / 64; ++w) t += sh[w];
        blockcount[blockIdx.x] = t;
    }
}

// Exclusive scan of nb counts in one workgroup (each thread a contiguous run of counts);
// the total goes to blockcount[nb] and *total_out, and mul * total is added to *counter.
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_counts(int nb, int* blockcount, int* total_out,
                                                              unsigned long long* counter, int mul) {
    __shared__ int wt[SCAN_THREADS / 64];
    int per = (nb + SCAN_THREADS - 1) / SCAN_THREADS;
    int b0 = threadIdx.x * per, b1 = min(nb, b0 + per);
    int own = 0;
    for (int i = b0; i < b1; ++i) own += blockcount[i];
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int incl = own;
    for (int off = 1; off < 64; off <<= 1) {
        int t = __shfl_up(incl, off, 64);
        if (lane >= off) incl += t;
    }
    if (lane == 63) wt[w] = incl;
    __syncthreads();
    int wbase = 0, total = 0;
    for (int k = 0; k < SCAN_THREADS / 64; ++k) { int v = wt[k]; if (k < w) wbase += v; total += v; }
    int run = wbase + incl - own;
    for (int i = b0; i < b1; ++i) { int v = blockcount[i]; blockcount[i] = run; run += v; }
    if (threadIdx.x == 0) {
        blockcount[nb] = total;
        if (total_out) *total_out = total;
        if (counter) *counter += (unsigned long long)mul * (unsigned long long)total;
    }
}

// compacted slot list in slot order from the mask words and the scanned block offsets
__global__ __launch_bounds__(SEL_BLOCK) void k_sel_write(const uint32_t* mask, const int* blockoff, int* out) {
    __shared__ int sh[SEL_BLOCK / 64];
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int64_t wave = (int64_t)blockIdx.x * (SEL_BLOCK / 64) + w;
    uint32_t word = mask[wave * 64 + lane];
    int c = wave_sum(__popc(word));
    if (lane == 0) sh[w] = c;
    __syncthreads();
    int run = blockoff[blockIdx.x];
    for (int k = 0; k < w; ++k) run += sh[k];
    if (__ballot(word != 0u) == 0ull) return;                  // wave-uniform: nothing selected here
    const uint64_t lt = (1ull << lane) - 1ull;
    int64_t base = wave * SEL_WAVE_SLOTS + 4 * lane;
#pragma unroll
    for (int j = 0; j < SEL_CHUNKS; ++j) {
        uint32_t nib = (word >> (4 * j)) & 15u;
        uint64_t b0 = __ballot(nib & 1u), b1 = __ballot(nib & 2u), b2 = __ballot(nib & 4u), b3 = __ballot(nib & 8u);
        int pos = run + __popcll(b0 & lt) + __popcll(b1 & lt) + __popcll(b2 & lt) + __popcll(b3 & lt);
        for (int q = 0; q < 4; ++q)
            if (nib & (1u << q)) out[pos++] = (int)(base + 256 * j + q);
        run += __popcll(b0) + __popcll(b1) + __popcll(b2) + __popcll(b3);
    }
}

// ---------------------------------------------------------------- collision
#define MCC_PROD_W 7
struct MccArgs {
    int nsel;
    const int* sel;
    int64_t nslots;
    void *x, *y, *vx, *vy, *vz;
    const int* cell;
    int nreact, nsigma;
    double emin, eint, emax;
    const double* prob;     // Probility(nreact, nsigma) column-major
    const int* rtype;
    const double* thr;
    double v_ref, mass, gas_mass, gas_temp;
    int electron;           // Model 1 (electrons) vs 2 (non-reactive ions)
    uint64_t seed; uint32_t step; uint32_t tag;
    unsigned long long* events;   // [32]
    unsigned long long* counters;
    unsigned int* err;
    // staged ionisation products, one pair per selected particle; bit k of newmask marks
    // selected particle k as having created a pair (null when the species creates none)
    unsigned long long* newmask;
    double* new_e;          // MCC_PROD_W per selected: x, y, vx, vy, vz, cell, azimuth
    double* new_i;
    const void* th;         // azimuth of the species (axisymmetric domains), else null
};

struct Vel3 { double x, y, z; };

// VelocityMaxwellianInitializationParticleOne + VelocityRandomInitializationParticleOne
// (ParticleOne.f90:135-152, 247-261): rejection on R^2 > -e R ln R with the 2.71828
// constant, speed from the last R, isotropic direction.
__device__ inline Vel3 vel_maxwell(pic_stream* s, double mass, double temp, unsigned int* err) {
    double Beta = 1.0 / (PIC_KB * temp);
    double FuncA = 1.0, FuncB = 0.0, R = 0.5;
    int it = 0;
    while (FuncA > FuncB) {
        R = pic_stream_u01(s); FuncA = R * R;
        R = pic_stream_u01(s); FuncB = -PIC_EXP_CONST * R * log(R);
        if (++it > 1000) { atomicOr(err, PIC_EBIT_MCCLOOP); break; }
    }
    double V = sqrt(-3.0 * log(R) / Beta / mass);
    R = pic_stream_u01(s);
    double CosTheta = 1.0 - 2.0 * R;
    double SinTheta = sqrt(1.0 - CosTheta * CosTheta);
    R = pic_stream_u01(s);
    double Fai = 2.0 * PIC_PI * R;
    Vel3 v; v.x = V * CosTheta; v.y = V * SinTheta * cos(Fai); v.z = V * SinTheta * sin(Fai);
    return v;
}

struct MccState {
    double Gx, Gy, Gz, Gper, G, Energy, MassRatio, Miu;
    Vel3 v;      // POI velocity (SI)
    Vel3 vt;     // POT velocity (SI)
};

// PostCollisionVelocity, TypeMCC.f90:101-133; kind 0 = CosKaiVahedi, 1 = IsotropicCosKai
__device__ inline void post_collision(MccState& M, pic_stream* s, int kind) {
    if (fabs(M.Gper) < PIC_MINREAL) {
        // VelRanInit(G) on the particle
        double R = pic_stream_u01(s);
        double CosTheta = 1.0 - 2.0 * R;
        double SinTheta = sqrt(1.0 - CosTheta * CosTheta);
        R = pic_stream_u01(s);
        double Fai = 2.0 * PIC_PI * R;
        M.v.x = M.G * CosTheta; M.v.y = M.G * SinTheta * cos(Fai); M.v.z = M.G * SinTheta * sin(Fai);
        return;
    }
    double CosTheta;
    if (kind == 0) {
        if (M.Energy < PIC_MINREAL) CosTheta = 1.0;
        else { double R = pic_stream_u01(s); CosTheta = (M.Energy + 2.0 - 2.0 * pow(M.Energy + 1.0, R)) / M.Energy; }
    } else {
        double R = pic_stream_u01(s); CosTheta = 1.0 - 2.0 * R;
    }
    double FcosTheta = 1.0 - CosTheta;
    double SinTheta = sqrt(1.0 - CosTheta * CosTheta);
    double R = pic_stream_u01(s);
    double APhi = 2.0 * PIC_PI * R;
    double CosPhi = cos(APhi), SinPhi = sin(APhi);
    double hx = M.Gper * CosPhi;
    double hy = -(M.Gx * M.Gy * CosPhi + M.G * M.Gz * SinPhi) / M.Gper;
    double hz = -(M.Gx * M.Gz * CosPhi - M.G * M.Gy * SinPhi) / M.Gper;
    M.v.x = M.v.x - M.MassRatio * (M.Gx * FcosTheta + hx * SinTheta);
    M.v.y = M.v.y - M.MassRatio * (M.Gy * FcosTheta + hy * SinTheta);
    M.v.z = M.v.z - M.MassRatio * (M.Gz * FcosTheta + hz * SinTheta);
}

// UpdateVelocityMCCParticleOne, TypeMCC.f90:84-99
__device__ inline void velocity_updater(MccState& M, double f) {
    M.Gx *= f; M.Gy *= f; M.Gz *= f; M.Gper *= f; M.G *= f;
    M.v.x *= f; M.v.y *= f; M.v.z *= f;
}

// collision of selected particle k; true when it created an electron-ion pair
template <typename R>
__device__ inline bool collide_one(const MccArgs& a, int k) {
    int slot = a.sel[k];
    R* X = (R*)a.x; R* Y = (R*)a.y; R* VX = (R*)a.vx; R* VY = (R*)a.vy; R* VZ = (R*)a.vz;
    pic_stream s; pic_stream_init(&s, a.seed, (uint32_t)slot, a.step, a.tag);
    MccState M;
    // VelRes(VFactor): to SI
    M.v.x = (double)VX[slot] * a.v_ref; M.v.y = (double)VY[slot] * a.v_ref; M.v.z = (double)VZ[slot] * a.v_ref;
    // Updater, TypeMCC.f90:58-82
    M.Miu = a.mass * a.gas_mass / (a.mass + a.gas_mass);
    M.MassRatio = a.gas_mass / (a.mass + a.gas_mass);
    M.vt = vel_maxwell(&s, a.gas_mass, a.gas_temp, a.err);
    M.Gx = M.v.x - M.vt.x; M.Gy = M.v.y - M.vt.y; M.Gz = M.v.z - M.vt.z;
    double Gtemp = M.Gy * M.Gy + M.Gz * M.Gz;
    M.Gper = sqrt(Gtemp);
    double G2 = Gtemp + M.Gx * M.Gx;
    M.G = sqrt(G2);
    M.Energy = 0.5 * M.Miu * G2 / PIC_JTOEV;
    // SelectProbility, TypeMCC.f90:155-198
    int Nr = a.nreact, Ns = a.nsigma;
    double P[4] = {0, 0, 0, 0};
    double lE = log10(M.Energy);
    int N = (int)ceil((lE - a.emin) / a.eint);
    if (N < 1) { for (int i = 0; i < Nr; ++i) P[i] = a.prob[i]; }
    else if (N < Ns) {
        double S1 = (double)N - (lE - a.emin) / a.eint;
        double S2 = 1.0 - S1;
        for (int i = 0; i < Nr; ++i) P[i] = a.prob[(size_t)(N - 1) * Nr + i] * S1 + a.prob[(size_t)N * Nr + i] * S2;
    } else {
        for (int i = 0; i < Nr; ++i) P[i] = a.prob[(size_t)(Ns - 1) * Nr + i] * M.Energy / a.emax;
    }
    for (int i = 0; i < Nr; ++i) if (M.Energy < a.thr[i]) P[i] = 0.0;
    double Rr = pic_stream_u01(&s);
    int index = 0;
    for (int i = 0; i < Nr; ++i) if (Rr < P[i]) { index = i + 1; break; }
    bool created = false;
    if (index > 0) {
        atomicAdd(&a.events[index], 1ull);
        atomicAdd(&a.counters[PIC_C_MCC_REAL], 1ull);
        int type = a.rtype[index - 1];
        double th = a.thr[index - 1];
        if (a.electron) {
            if (type == 11) {
                post_collision(M, &s, 0);
            } else if (type == 12) {
                double f = sqrt(1.0 - th / M.Energy);
                M.Energy = M.Energy - th;
                velocity_updater(M, f);
                post_collision(M, &s, 0);
            } else if (type == 13) {
                // ElectronIonization, MCCElectron.F90:48-80
                Vel3 vsolid = M.v;
                double Eorig = M.Energy;
                double EnergyTemp = M.Energy - th;
                double R1 = pic_stream_u01(&s);
                double EnergyOld = 10.0 * tan(R1 * atan(EnergyTemp) / 20.0);
                if (fabs(EnergyOld) < PIC_MINREAL) EnergyOld = PIC_MINREAL;
                double f = sqrt(EnergyOld / M.Energy);
                velocity_updater(M, f);
                post_collision(M, &s, 0);
                double EnergyNew = EnergyTemp - EnergyOld;
                double f2 = sqrt(EnergyNew / Eorig);
                double* ne = a.new_e + (size_t)k * MCC_PROD_W;
                double* ni = a.new_i + (size_t)k * MCC_PROD_W;
                // both products start at the creator's position, azimuth included
                ne[6] = ni[6] = a.th ? (double)((const R*)a.th)[slot] : 0.0;
                ne[0] = (double)X[slot]; ne[1] = (double)Y[slot];
                ne[2] = vsolid.x * f2 / a.v_ref; ne[3] = vsolid.y * f2 / a.v_ref; ne[4] = vsolid.z * f2 / a.v_ref;
                ne[5] = (double)a.cell[slot];
                ni[0] = (double)X[slot]; ni[1] = (double)Y[slot];
                ni[2] = M.vt.x / a.v_ref; ni[3] = M.vt.y / a.v_ref; ni[4] = M.vt.z / a.v_ref;
                ni[5] = (double)a.cell[slot];
                created = true;
            }
        } else {
            if (type == 101) post_collision(M, &s, 1);
            else if (type == 103 || type == 102) M.v = M.vt;
        }
    }
    // VelRes(1/VFactor)
    double vf = 1.0 / a.v_ref;
    VX[slot] = (R)(M.v.x * vf); VY[slot] = (R)(M.v.y * vf); VZ[slot] = (R)(M.v.z * vf);
    return created;
}

template <typename R>
__global__ __launch_bounds__(128) void k_mcc_collide(MccArgs a) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    bool created = (k < a.nsel) && collide_one<R>(a, k);
    if (a.newmask) {
        unsigned long long b = __ballot(created);
        if ((threadIdx.x & 63) == 0 && k < a.nsel) a.newmask[k >> 6] = b;
    }
}

// Ionisation products: bit k of the collision's mask marks selected particle k as a
// creator; per-block counts, the shared scan, then each word's pairs are written at the
// end of the electron and ion arrays in selected order (MCCAddNewParticle order).
__global__ __launch_bounds__(SEL_BLOCK) void k_mask_count(int nwords, const unsigned long long* mask, int* blockcount) {
    __shared__ int sh[SEL_BLOCK / 64];
    int w = blockIdx.x * SEL_BLOCK + threadIdx.x;
    int c = wave_sum(w < nwords ? __popcll(mask[w]) : 0);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int k = 0; k < SEL_BLOCK / 64; ++k) t += sh[k];
        blockcount[blockIdx.x] = t;
    }
}

struct AppendDst {
    void *x, *y, *vx, *vy, *vz;
    int* cell;
    int64_t base;           // first free slot
    void* th;               // azimuth (axisymmetric domains), else null
};

template <typename R>
__device__ inline void put_product(const AppendDst& d, int64_t i, const double* st) {
    ((R*)d.x)[i] = (R)st[0]; ((R*)d.y)[i] = (R)st[1];
    ((R*)d.vx)[i] = (R)st[2]; ((R*)d.vy)[i] = (R)st[3]; ((R*)d.vz)[i] = (R)st[4];
    d.cell[i] = (int)st[5];
    if (d.th) ((R*)d.th)[i] = (R)st[6];
}

// room = free slots of the fuller species; nothing is written when the pairs do not fit
template <typename R>
__global__ __launch_bounds__(SEL_BLOCK) void k_append_pairs(int nwords, const unsigned long long* mask,
                                                            const int* blockoff, const double* new_e,
                                                            const double* new_i, AppendDst E, AppendDst I,
                                                            int64_t room) {
    __shared__ int sh[SEL_BLOCK / 64];
    if ((int64_t)blockoff[gridDim.x] > room) return;
    int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int w = blockIdx.x * SEL_BLOCK + threadIdx.x;
    unsigned long long word = w < nwords ? mask[w] : 0ull;
    int own = __popcll(word), incl = own;
    for (int off = 1; off < 64; off <<= 1) {
        int t = __shfl_up(incl, off, 64);
        if (lane >= off) incl += t;
    }
    if (lane == 63) sh[wv] = incl;
    __syncthreads();
    int64_t pos = blockoff[blockIdx.x] + incl - own;
    for (int k = 0; k < wv; ++k) pos += sh[k];
    while (word) {
        int b = __ffsll((long long)word) - 1;
        int64_t k = (int64_t)w * 64 + b;
        put_product<R>(E, E.base + pos, new_e + k * MCC_PROD_W);
        put_product<R>(I, I.base + pos, new_i + k * MCC_PROD_W);
        ++pos;
        word &= word - 1ull;
    }
}

__global__ void k_add_counter(unsigned long long* c, unsigned long long v) {
    if (threadIdx.x == 0) *c += v;
}

// Scratch that grows with the particle count: reallocated (which synchronises the device)
// only when outgrown, then with 25% and 1 MiB of headroom so a slowly growing population
// does not reallocate every step.
static int ensure_buf(void** p, int64_t* cap, int64_t need_bytes) {
    if (*cap >= need_bytes) return PIC_OK;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    int64_t bytes = need_bytes + need_bytes / 4 + ((int64_t)1 << 20);
    HIPCHK(hipMalloc(p, (size_t)bytes));
    *cap = bytes;
    return PIC_OK;
}

// mask pass + scan of one species; its selected count lands in *info, its counter
// PIC_C_MCC_SELECTED grows by the same amount on the device
static void launch_select(int isp, int64_t n, int nb, uint32_t* mask, int* bc, int* info) {
    SpeciesHost& S = g.sp[isp];
    uint32_t tag = pic_stream_tag(isp, PIC_RNG_SELECT);
    hipLaunchKernelGGL(k_sel_mask, dim3(nb), dim3(SEL_BLOCK), 0, g.stream, n, S.cell, S.select_threshold, g.seed,
                       (uint32_t)g.step, tag, mask, bc);
    hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(SCAN_THREADS), 0, g.stream, nb, bc, info,
                       S.counters + PIC_C_MCC_SELECTED, 1);
}

template <typename R>
static void launch_collide(int isp, int nsel, const int* sel, unsigned long long* newmask, double* new_e,
                           double* new_i) {
    SpeciesHost& S = g.sp[isp];
    MccArgs a;
    a.nsel = nsel; a.sel = sel; a.nslots = S.nslots;
    a.x = S.x; a.y = S.y; a.vx = S.vx; a.vy = S.vy; a.vz = S.vz; a.cell = S.cell;
    a.nreact = S.nreact; a.nsigma = S.nsigma; a.emin = S.emin; a.eint = S.eint; a.emax = S.emax;
    a.prob = S.prob; a.rtype = S.rtype; a.thr = S.thr;
    a.v_ref = g.v_ref; a.mass = S.p.mass_kg; a.gas_mass = g.gas.mass_kg; a.gas_temp = g.gas.temperature_K;
    a.electron = (isp == g.e_species) ? 1 : 0;
    a.seed = g.seed; a.step = (uint32_t)g.step; a.tag = pic_stream_tag(isp, PIC_RNG_COLLIDE);
    a.events = S.mcc_events; a.counters = S.counters; a.err = g.d_err;
    a.newmask = newmask; a.new_e = new_e; a.new_i = new_i;
    a.th = g.mesh.delta == 1 ? S.th : nullptr;
    hipLaunchKernelGGL((k_mcc_collide<R>), dim3((nsel + 127) / 128), dim3(128), 0, g.stream, a);
}

// MCC order of MCCPublic.F90:50-105 for the one gas: electrons then ions, products of
// the electron collisions appended after both (MCCAddNewParticle). Ion selection only
// looks at ion slots that exist before the products arrive and the two species' random
// streams are disjoint, so selecting and colliding both species before the append gives
// the same result as the reference's interleaving.
// Host synchronisations: one after the two selection counts (sizes of the slot list and
// the staging), one after the product count (new slot counts).
template <typename R>
static int run_mcc_t() {
    int e = g.e_species, ion = g.i_species;
    SpeciesHost& E = g.sp[e];
    SpeciesHost& I = g.sp[ion];
    int64_t n_e = E.has_mcc ? E.nslots : 0, n_i = I.has_mcc ? I.nslots : 0;
    int nb_e = (int)((n_e + SEL_BLOCK_SLOTS - 1) / SEL_BLOCK_SLOTS);
    int nb_i = (int)((n_i + SEL_BLOCK_SLOTS - 1) / SEL_BLOCK_SLOTS);
    if (nb_e + nb_i == 0) return PIC_OK;
    int rc = ensure_buf((void**)&g.d_selmask, &g.selmask_cap, (int64_t)(nb_e + nb_i) * SEL_BLOCK * sizeof(uint32_t));
    if (rc) return rc;
    rc = ensure_buf((void**)&g.d_blockcount, &g.blockcount_cap, (int64_t)(nb_e + nb_i + 2) * sizeof(int));
    if (rc) return rc;
    rc = ensure_buf((void**)&g.d_mcc_info, &g.mcc_info_cap, 4 * sizeof(int));
    if (rc) return rc;
    int* info = g.d_mcc_info;
    HIPCHK(hipMemsetAsync(info, 0, 4 * sizeof(int), g.stream));
    uint32_t* mask_e = g.d_selmask;
    uint32_t* mask_i = mask_e + (size_t)nb_e * SEL_BLOCK;
    int* bc_e = g.d_blockcount;
    int* bc_i = bc_e + nb_e + 1;
    if (nb_e) launch_select(e, n_e, nb_e, mask_e, bc_e, info + 0);
    if (nb_i) launch_select(ion, n_i, nb_i, mask_i, bc_i, info + 1);
    HIPCHK(hipGetLastError());
    int* hinfo = g.h_side_pinned;                   // MCC may run on the side stream during a solve
    HIPCHK(hipMemcpyAsync(hinfo, info, 2 * sizeof(int), hipMemcpyDeviceToHost, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    int nsel_e = hinfo[0], nsel_i = hinfo[1];
    if (nsel_e + nsel_i == 0) return PIC_OK;
    rc = ensure_buf((void**)&g.d_sel, &g.sel_cap, (int64_t)(nsel_e + nsel_i) * sizeof(int));
    if (rc) return rc;
    int nwords = (nsel_e + 63) / 64;
    rc = ensure_buf(&g.d_newbuf, &g.newbuf_cap,
                    (int64_t)nsel_e * 2 * MCC_PROD_W * sizeof(double) + (int64_t)nwords * 8 + 64);
    if (rc) return rc;
    double* new_e = (double*)g.d_newbuf;
    double* new_i = new_e + (size_t)nsel_e * MCC_PROD_W;
    unsigned long long* newmask = (unsigned long long*)(new_e + (size_t)nsel_e * 2 * MCC_PROD_W);
    if (nsel_e) {
        hipLaunchKernelGGL(k_sel_write, dim3(nb_e), dim3(SEL_BLOCK), 0, g.stream, mask_e, bc_e, g.d_sel);
        launch_collide<R>(e, nsel_e, g.d_sel, newmask, new_e, new_i);
    }
    if (nsel_i) {
        hipLaunchKernelGGL(k_sel_write, dim3(nb_i), dim3(SEL_BLOCK), 0, g.stream, mask_i, bc_i, g.d_sel + nsel_e);
        launch_collide<R>(ion, nsel_i, g.d_sel + nsel_e, nullptr, nullptr, nullptr);
    }
    HIPCHK(hipGetLastError());
    if (nsel_e == 0) return PIC_OK;
    // ionisation pairs
    int nbp = (nwords + SEL_BLOCK - 1) / SEL_BLOCK;
    rc = ensure_buf((void**)&g.d_pcount, &g.pcount_cap, (int64_t)(nbp + 1) * sizeof(int));
    if (rc) return rc;
    int64_t room = std::min(E.capacity - E.nslots, I.capacity - I.nslots);
    hipLaunchKernelGGL(k_mask_count, dim3(nbp), dim3(SEL_BLOCK), 0, g.stream, nwords, newmask, g.d_pcount);
    hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(SCAN_THREADS), 0, g.stream, nbp, g.d_pcount, info + 2,
                       (unsigned long long*)nullptr, 0);
    const bool rz = g.mesh.delta == 1;
    AppendDst de{E.x, E.y, E.vx, E.vy, E.vz, E.cell, E.nslots, rz ? E.th : nullptr};
    AppendDst di{I.x, I.y, I.vx, I.vy, I.vz, I.cell, I.nslots, rz ? I.th : nullptr};
    hipLaunchKernelGGL((k_append_pairs<R>), dim3(nbp), dim3(SEL_BLOCK), 0, g.stream, nwords, newmask, g.d_pcount,
                       new_e, new_i, de, di, room);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(hinfo, info + 2, sizeof(int), hipMemcpyDeviceToHost, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    int64_t nnew = hinfo[0];
    if (nnew > room) return pic_fail(PIC_ERR_CAPACITY, "MCC products exceed particle capacity");
    if (nnew == 0) return PIC_OK;
    E.nslots += nnew; E.nlive += nnew;
    I.nslots += nnew; I.nlive += nnew;
    // created counter: the electron species counts both products of each pair
    hipLaunchKernelGGL(k_add_counter, dim3(1), dim3(64), 0, g.stream, E.counters + PIC_C_MCC_CREATED,
                       (unsigned long long)(2 * nnew));
    HIPCHK(hipGetLastError());
    return PIC_OK;
}

int run_mcc() {
    if (g.particle_bits == 64) return run_mcc_t<double>();
    return run_mcc_t<float>();
}

// ---------------------------------------------------------------- cell sort
// Stable LSD radix sort of (tile-major element key, slot) with dead slots keyed past the end,
// 8-bit digits, one pass per digit the largest key needs; then one gather of every SoA array
// into spare arrays that are swapped in (group-transposed layout, pic_sorted_position);
// occupied slots shrink to the live count. A pass is three launches:
//   k_rs_hist    each block of RS_BLOCK_ITEMS consecutive items counts its digits (LDS),
//                written digit-major: hist[d * nblocks + b]
//   scan         exclusive prefix of hist (k_rs_scan_local / k_rs_scan_top / k_rs_scan_add):
//                the first output position of digit d from block b
//   k_rs_scatter the block walks its items in rounds of 256 (one per thread); a wave finds the
//                lanes with its digit by 8 ballots, ranks them by lane, and adds the counts of
//                the waves before it and of the rounds before; so equal digits keep the order
//                of their slots (stable), and the result does not depend on scheduling.
#define RS_THREADS 256
#define RS_ROUNDS 16
#define RS_BLOCK_ITEMS (RS_THREADS * RS_ROUNDS)
#define RS_SCAN_THREADS 1024
#define RS_SCAN_ITEMS (RS_SCAN_THREADS * 4)

__global__ __launch_bounds__(RS_THREADS) void k_rs_hist(int64_t n, const unsigned int* keys, int shift,
                                                       unsigned int* hist, int nblocks) {
    __shared__ unsigned int h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * RS_BLOCK_ITEMS;
    for (int r = 0; r < RS_ROUNDS; ++r) {
        const int64_t i = base + r * RS_THREADS + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// exclusive scan of a[0..len) in place: chunks of RS_SCAN_ITEMS, their totals, their offsets
__device__ inline unsigned int rs_block_exclusive(unsigned int v, unsigned int* total) {
    __shared__ unsigned int wsum[RS_SCAN_THREADS / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    unsigned int incl = v;
    for (int off = 1; off < 64; off <<= 1) {
        unsigned int u = __shfl_up(incl, off, 64);
        if (lane >= off) incl += u;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned int acc = 0;
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) { unsigned int t = wsum[k]; wsum[k] = acc; acc += t; }
        *total = acc;
    }
    __syncthreads();
    const unsigned int r = wsum[w] + incl - v;
    __syncthreads();
    return r;
}
__global__ __launch_bounds__(RS_SCAN_THREADS) void k_rs_scan_local(int64_t len, unsigned int* a, unsigned int* sums) {
    __shared__ unsigned int tot;
    const int64_t b0 = (int64_t)blockIdx.x * RS_SCAN_ITEMS + (int64_t)threadIdx.x * 4;
    unsigned int v[4], s = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) { v[k] = (b0 + k < len) ? a[b0 + k] : 0u; s += v[k]; }
    unsigned int ex = rs_block_exclusive(s, &tot);
#pragma unroll
    for (int k = 0; k < 4; ++k) { if (b0 + k < len) a[b0 + k] = ex; ex += v[k]; }
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}
__global__ __launch_bounds__(RS_SCAN_THREADS) void k_rs_scan_top(int nsums, unsigned int* sums) {
    __shared__ unsigned int tot;
    unsigned int carry = 0;
    for (int c0 = 0; c0 < nsums; c0 += RS_SCAN_THREADS) {
        const int i = c0 + threadIdx.x;
        const unsigned int v = i < nsums ? sums[i] : 0u;
        const unsigned int ex = rs_block_exclusive(v, &tot);
        if (i < nsums) sums[i] = carry + ex;
        carry += tot;
        __syncthreads();
    }
}
__global__ __launch_bounds__(RS_SCAN_THREADS) void k_rs_scan_add(int64_t len, unsigned int* a, const unsigned int* sums) {
    const int64_t b0 = (int64_t)blockIdx.x * RS_SCAN_ITEMS + (int64_t)threadIdx.x * 4;
    const unsigned int off = sums[blockIdx.x];
#pragma unroll
    for (int k = 0; k < 4; ++k) if (b0 + k < len) a[b0 + k] += off;
}

__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(int64_t n, const unsigned int* kin, const unsigned int* vin,
                                                          unsigned int* kout, unsigned int* vout, int shift,
                                                          const unsigned int* hist, int nblocks) {
    __shared__ unsigned int run[256];
    __shared__ unsigned int wcount[RS_THREADS / 64][256];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    run[threadIdx.x] = hist[(size_t)threadIdx.x * nblocks + blockIdx.x];
#pragma unroll
    for (int k = 0; k < RS_THREADS / 64; ++k) wcount[k][threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * RS_BLOCK_ITEMS;
    const unsigned long long below = (1ull << lane) - 1ull;
    for (int r = 0; r < RS_ROUNDS; ++r) {
        const int64_t i = base + r * RS_THREADS + threadIdx.x;
        const bool valid = i < n;
        const unsigned int key = valid ? kin[i] : 0u;
        const unsigned int val = valid ? vin[i] : 0u;
        const unsigned int d = (key >> shift) & 255u;
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const unsigned long long bb = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? bb : ~bb;
        }
        const unsigned int lrank = (unsigned int)__popcll(peers & below);
        if (valid && lrank == 0) wcount[w][d] = (unsigned int)__popcll(peers);
        __syncthreads();
        if (valid) {
            unsigned int pos = run[d] + lrank;
            for (int k = 0; k < w; ++k) pos += wcount[k][d];
            kout[pos] = key; vout[pos] = val;
        }
        __syncthreads();
        unsigned int add = 0;
#pragma unroll
        for (int k = 0; k < RS_THREADS / 64; ++k) { add += wcount[k][threadIdx.x]; wcount[k][threadIdx.x] = 0; }
        run[threadIdx.x] += add;
        __syncthreads();
    }
}
__global__ void k_sort_keys(int64_t n, const int* cell, unsigned int* keys, unsigned int* vals, unsigned int dead_key,
                            int ney, int nty) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int c = cell[i];
    keys[i] = c >= 0 ? pic_sort_key(c, ney, nty) : dead_key;
    vals[i] = (unsigned int)i;
}
// Gather of every particle array in one pass: slot d takes the particle at sorted position
// pic_sorted_position(d) (the group-transposed layout the push kernel walks, ppt slots per thread).
// A block is one group of 64 ppt slots: its waves read the same source lines (sorted positions
// ppt apart per lane), so they share one CU's caches instead of two XCDs' L2s.
template <typename R>
__global__ __launch_bounds__(64 * PIC_PPT_WIDE) void k_sort_gather(int64_t n, int64_t nfull, int ppt, const unsigned int* perm,
                                                     const R* x, const R* y, const R* vx, const R* vy, const R* vz,
                                                     const int* cell, const R* th, R* ox, R* oy, R* ovx, R* ovy,
                                                     R* ovz, int* ocell, R* oth) {
    int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= n) return;
    const unsigned int s = perm[pic_sorted_position(d, nfull, ppt)];
    ox[d] = x[s]; oy[d] = y[s]; ovx[d] = vx[s]; ovy[d] = vy[s]; ovz[d] = vz[s]; ocell[d] = cell[s];
    if (th) oth[d] = th[s];
}
#define TRY_MCC(x) do { int rc_ = (x); if (rc_) return rc_; } while (0)

static int sort_species(int isp) {
    SpeciesHost& S = g.sp[isp];
    int64_t n = S.nslots;
    if (n == 0) return PIC_OK;
    int nty = (g.mesh.ney + SORT_TILE - 1) / SORT_TILE, ntx = (g.mesh.nex + SORT_TILE - 1) / SORT_TILE;
    unsigned int dead_key = (unsigned int)(ntx * nty * SORT_TILE * SORT_TILE);
    int end_bit = 1;
    while ((1u << end_bit) <= dead_key) end_bit++;
    size_t esz = (g.particle_bits == 64) ? 8 : 4;
    // keys and values sized for the species' capacity, so a growing population does not
    // reallocate between sorts
    size_t need = (size_t)std::max<int64_t>(n, S.capacity) * 16 + 256;
    if (g.sort_buf_bytes < need) {
        if (g.d_sort_buf) (void)hipFree(g.d_sort_buf);
        g.d_sort_buf = nullptr;
        HIPCHK(hipMalloc(&g.d_sort_buf, need));
        g.sort_buf_bytes = need;
    }
    unsigned int* k0 = (unsigned int*)g.d_sort_buf;
    unsigned int* k1 = k0 + n;
    unsigned int* v0 = k1 + n;
    unsigned int* v1 = v0 + n;
    int blocks = (int)((n + 255) / 256);
    hipLaunchKernelGGL(k_sort_keys, dim3(blocks), dim3(256), 0, g.stream, n, S.cell, k0, v0, dead_key, g.mesh.ney, nty);
    // digit histograms (256 per block) and the scan's chunk totals, sized for the capacity
    const int64_t cap = std::max<int64_t>(n, S.capacity);
    const int nb_cap = (int)((cap + RS_BLOCK_ITEMS - 1) / RS_BLOCK_ITEMS);
    const int64_t hist_len_cap = (int64_t)256 * nb_cap;
    const size_t tb = (size_t)(hist_len_cap + hist_len_cap / RS_SCAN_ITEMS + 2) * sizeof(unsigned int);
    if (g.sort_tmp_bytes < tb) {
        if (g.d_sort_tmp) (void)hipFree(g.d_sort_tmp);
        g.d_sort_tmp = nullptr;
        const size_t bytes = tb + tb / 4 + ((size_t)1 << 20);
        HIPCHK(hipMalloc(&g.d_sort_tmp, bytes));
        g.sort_tmp_bytes = bytes;
    }
    const int nb = (int)((n + RS_BLOCK_ITEMS - 1) / RS_BLOCK_ITEMS);
    const int64_t hist_len = (int64_t)256 * nb;
    const int nchunks = (int)((hist_len + RS_SCAN_ITEMS - 1) / RS_SCAN_ITEMS);
    unsigned int* hist = (unsigned int*)g.d_sort_tmp;
    unsigned int* sums = hist + hist_len;
    unsigned int *kin = k0, *vin = v0, *kout = k1, *vout = v1;
    for (int shift = 0; shift < end_bit; shift += 8) {
        hipLaunchKernelGGL(k_rs_hist, dim3(nb), dim3(RS_THREADS), 0, g.stream, n, kin, shift, hist, nb);
        hipLaunchKernelGGL(k_rs_scan_local, dim3(nchunks), dim3(RS_SCAN_THREADS), 0, g.stream, hist_len, hist, sums);
        hipLaunchKernelGGL(k_rs_scan_top, dim3(1), dim3(RS_SCAN_THREADS), 0, g.stream, nchunks, sums);
        hipLaunchKernelGGL(k_rs_scan_add, dim3(nchunks), dim3(RS_SCAN_THREADS), 0, g.stream, hist_len, hist, sums);
        hipLaunchKernelGGL(k_rs_scatter, dim3(nb), dim3(RS_THREADS), 0, g.stream, n, kin, vin, kout, vout, shift, hist, nb);
        std::swap(kin, kout); std::swap(vin, vout);
    }
    HIPCHK(hipGetLastError());
    unsigned int* perm = vin;             // the sorted slot order after the last pass
    // live count; the live particles occupy the first `live` sorted positions
    int64_t live = 0;
    TRY_MCC(count_live_slots(n, S.cell, &live));
    // gather into the spare arrays (and cell_next), then swap them in
    const bool rz = g.mesh.delta == 1 && S.th;
    for (int a = 0; a < (rz ? 6 : 5); ++a)
        if (!S.spare[a]) HIPCHK(hipMalloc(&S.spare[a], (size_t)S.capacity * esz));
    const int ppt = pic_push_ppt();
    const int64_t nfull = live / (64 * ppt) * (64 * ppt);
    const int gthreads = 64 * ppt;
    const int gblocks = (int)((live + gthreads - 1) / gthreads);
    if (live > 0) {
        if (esz == 8)
            hipLaunchKernelGGL(k_sort_gather<double>, dim3(gblocks), dim3(gthreads), 0, g.stream, live, nfull, ppt, perm,
                               (const double*)S.x, (const double*)S.y, (const double*)S.vx, (const double*)S.vy,
                               (const double*)S.vz, S.cell, rz ? (const double*)S.th : nullptr,
                               (double*)S.spare[0], (double*)S.spare[1], (double*)S.spare[2], (double*)S.spare[3],
                               (double*)S.spare[4], S.cell_next, rz ? (double*)S.spare[5] : nullptr);
        else
            hipLaunchKernelGGL(k_sort_gather<float>, dim3(gblocks), dim3(gthreads), 0, g.stream, live, nfull, ppt, perm,
                               (const float*)S.x, (const float*)S.y, (const float*)S.vx, (const float*)S.vy,
                               (const float*)S.vz, S.cell, rz ? (const float*)S.th : nullptr,
                               (float*)S.spare[0], (float*)S.spare[1], (float*)S.spare[2], (float*)S.spare[3],
                               (float*)S.spare[4], S.cell_next, rz ? (float*)S.spare[5] : nullptr);
        HIPCHK(hipGetLastError());
    }
    std::swap(S.x, S.spare[0]); std::swap(S.y, S.spare[1]); std::swap(S.vx, S.spare[2]);
    std::swap(S.vy, S.spare[3]); std::swap(S.vz, S.spare[4]); std::swap(S.cell, S.cell_next);
    if (rz) std::swap(S.th, S.spare[5]);
    HIPCHK(hipStreamSynchronize(g.stream));
    S.nslots = (int64_t)live;
    S.nlive = (int64_t)live;
    return PIC_OK;
}

int run_sort_all() {
    for (int s = 0; s < g.nspecies; ++s) {
        if (!g.sp[s].defined) continue;
        int rc = sort_species(s);
        if (rc) return rc;
    }
    return PIC_OK;
}
