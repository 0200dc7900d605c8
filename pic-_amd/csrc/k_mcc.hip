// Null-collision Monte Carlo (MCC_jw/code/mc) and the periodic cell sort.
//
// Selection: every live slot is tested with one Philox word against
// floor(CollisionRatio * 2^32) (a Bernoulli(CollisionRatio) trial per particle, the GPU
// form of SelectParticle's INT(N*P) + Bernoulli(residue) count with a partial
// Fisher-Yates shuffle, MCCPublic.F90:113-171: same expected count, no serial shuffle).
// The selected slots are compacted in slot order (block counts, one-block scan, write).
// Collision: one thread per selected particle, FP64 kinematics of TypeMCC / MCCElectron /
// MCCIon / MCCChi drawing from the particle's own (slot, step, species) Philox stream.
// Ionisation products are staged per selected particle and appended in selected order.
#include "pic_device.h"
#include <hipcub/hipcub.hpp>

#define SEL_PER_THREAD 4
#define SEL_BLOCK 256
#define SEL_TILE (SEL_PER_THREAD * SEL_BLOCK)

__device__ inline bool sel_test(int64_t slot, const int* cell, uint32_t thr, uint64_t seed, uint32_t step, uint32_t tag) {
    if (cell[slot] < 0) return false;
    philox4 o = pic_philox4x32_10((uint32_t)slot, step, tag, 0u, (uint32_t)seed, (uint32_t)(seed >> 32));
    return o.v[0] < thr;
}

__global__ __launch_bounds__(SEL_BLOCK) void k_sel_count(int64_t n, const int* cell, uint32_t thr, uint64_t seed,
                                                         uint32_t step, uint32_t tag, int* blockcount) {
    __shared__ int sh[SEL_BLOCK / 64];
    int64_t base = (int64_t)blockIdx.x * SEL_TILE + (int64_t)threadIdx.x * SEL_PER_THREAD;
    int c = 0;
    for (int k = 0; k < SEL_PER_THREAD; ++k) { int64_t s = base + k; if (s < n && sel_test(s, cell, thr, seed, step, tag)) c++; }
    for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) { int t = 0; for (int w = 0; w < SEL_BLOCK / 64; ++w) t += sh[w]; blockcount[blockIdx.x] = t; }
}

// exclusive scan of nb block counts in one block; total at blockcount[nb]
__global__ __launch_bounds__(1024) void k_scan_counts(int nb, int* blockcount) {
    __shared__ int carry;
    __shared__ int sh[1024];
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < nb; base += 1024) {
        int i = base + threadIdx.x;
        int v = i < nb ? blockcount[i] : 0;
        sh[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            int t = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
            __syncthreads();
            sh[threadIdx.x] += t;
            __syncthreads();
        }
        int incl = sh[threadIdx.x];
        if (i < nb) blockcount[i] = carry + incl - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry += incl;
        __syncthreads();
    }
    if (threadIdx.x == 0) blockcount[nb] = carry;
}

__global__ __launch_bounds__(SEL_BLOCK) void k_sel_write(int64_t n, const int* cell, uint32_t thr, uint64_t seed,
                                                         uint32_t step, uint32_t tag, const int* blockoff, int* out) {
    __shared__ int sh[SEL_BLOCK];
    int64_t base = (int64_t)blockIdx.x * SEL_TILE + (int64_t)threadIdx.x * SEL_PER_THREAD;
    bool f[SEL_PER_THREAD];
    int c = 0;
    for (int k = 0; k < SEL_PER_THREAD; ++k) { int64_t s = base + k; f[k] = s < n && sel_test(s, cell, thr, seed, step, tag); c += f[k]; }
    sh[threadIdx.x] = c;
    __syncthreads();
    for (int off = 1; off < SEL_BLOCK; off <<= 1) {
        int t = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
        __syncthreads();
        sh[threadIdx.x] += t;
        __syncthreads();
    }
    int pos = blockoff[blockIdx.x] + sh[threadIdx.x] - c;
    for (int k = 0; k < SEL_PER_THREAD; ++k) if (f[k]) out[pos++] = (int)(base + k);
}

// ---------------------------------------------------------------- collision
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
    // staged ionisation products, one pair per selected particle
    unsigned char* newflag;
    double* new_e;          // 6 per selected: x, y, vx, vy, vz, cell
    double* new_i;
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

template <typename R>
__global__ __launch_bounds__(128) void k_mcc_collide(MccArgs a) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.nsel) return;
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
                double* ne = a.new_e + (size_t)k * 6;
                double* ni = a.new_i + (size_t)k * 6;
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
    a.newflag[k] = created ? 1 : 0;
    // VelRes(1/VFactor)
    double vf = 1.0 / a.v_ref;
    VX[slot] = (R)(M.v.x * vf); VY[slot] = (R)(M.v.y * vf); VZ[slot] = (R)(M.v.z * vf);
}

// append staged products: flags -> offsets (single block scan is enough: nsel is small)
template <typename R>
__global__ __launch_bounds__(1024) void k_append(int nsel, const unsigned char* flag, const double* stage, int64_t base,
                                                 R* x, R* y, R* vx, R* vy, R* vz, int* cell, int* count_out) {
    __shared__ int sh[1024];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int b0 = 0; b0 < nsel; b0 += 1024) {
        int k = b0 + threadIdx.x;
        int f = (k < nsel) ? flag[k] : 0;
        sh[threadIdx.x] = f;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            int t = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
            __syncthreads();
            sh[threadIdx.x] += t;
            __syncthreads();
        }
        int incl = sh[threadIdx.x];
        if (f) {
            int64_t d = base + carry + incl - 1;
            const double* st = stage + (size_t)k * 6;
            x[d] = (R)st[0]; y[d] = (R)st[1]; vx[d] = (R)st[2]; vy[d] = (R)st[3]; vz[d] = (R)st[4]; cell[d] = (int)st[5];
        }
        __syncthreads();
        if (threadIdx.x == 1023) carry += incl;
        __syncthreads();
    }
    if (threadIdx.x == 0) *count_out = carry;
}

static int ensure_buf(void** p, int64_t* cap, int64_t need_bytes) {
    if (*cap >= need_bytes) return PIC_OK;
    if (*p) hipFree(*p);
    *p = nullptr;
    HIPCHK(hipMalloc(p, (size_t)need_bytes));
    *cap = need_bytes;
    return PIC_OK;
}

static int select_species(int isp, int* nsel_out) {
    SpeciesHost& S = g.sp[isp];
    int64_t n = S.nslots;
    if (n == 0) { *nsel_out = 0; return PIC_OK; }
    int nb = (int)((n + SEL_TILE - 1) / SEL_TILE);
    int rc = ensure_buf((void**)&g.d_blockcount, &g.blockcount_cap, (int64_t)(nb + 1) * sizeof(int));
    if (rc) return rc;
    uint32_t tag = pic_stream_tag(isp, PIC_RNG_SELECT);
    hipLaunchKernelGGL(k_sel_count, dim3(nb), dim3(SEL_BLOCK), 0, g.stream, n, S.cell, S.select_threshold, g.seed,
                       (uint32_t)g.step, tag, g.d_blockcount);
    hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, g.stream, nb, g.d_blockcount);
    HIPCHK(hipGetLastError());
    int total = 0;
    HIPCHK(hipMemcpyAsync(&total, g.d_blockcount + nb, sizeof(int), hipMemcpyDeviceToHost, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    rc = ensure_buf((void**)&g.d_sel, &g.sel_cap, (int64_t)std::max(total, 1) * sizeof(int));
    if (rc) return rc;
    hipLaunchKernelGGL(k_sel_write, dim3(nb), dim3(SEL_BLOCK), 0, g.stream, n, S.cell, S.select_threshold, g.seed,
                       (uint32_t)g.step, tag, g.d_blockcount, g.d_sel);
    HIPCHK(hipGetLastError());
    *nsel_out = total;
    return PIC_OK;
}

template <typename R>
static int collide_species(int isp, int nsel, int* ncreated) {
    SpeciesHost& S = g.sp[isp];
    *ncreated = 0;
    if (nsel == 0) return PIC_OK;
    int64_t need = (int64_t)nsel * (1 + 12 * sizeof(double)) + 64;
    int rc = ensure_buf(&g.d_newbuf, &g.newbuf_cap, need);
    if (rc) return rc;
    MccArgs a;
    a.nsel = nsel; a.sel = g.d_sel; a.nslots = S.nslots;
    a.x = S.x; a.y = S.y; a.vx = S.vx; a.vy = S.vy; a.vz = S.vz; a.cell = S.cell;
    a.nreact = S.nreact; a.nsigma = S.nsigma; a.emin = S.emin; a.eint = S.eint; a.emax = S.emax;
    a.prob = S.prob; a.rtype = S.rtype; a.thr = S.thr;
    a.v_ref = g.v_ref; a.mass = S.p.mass_kg; a.gas_mass = g.gas.mass_kg; a.gas_temp = g.gas.temperature_K;
    a.electron = (isp == g.e_species) ? 1 : 0;
    a.seed = g.seed; a.step = (uint32_t)g.step; a.tag = pic_stream_tag(isp, PIC_RNG_COLLIDE);
    a.events = S.mcc_events; a.counters = S.counters; a.err = g.d_err;
    double* dbl = (double*)g.d_newbuf;
    a.new_e = dbl; a.new_i = dbl + (size_t)nsel * 6;
    a.newflag = (unsigned char*)(dbl + (size_t)nsel * 12);
    hipLaunchKernelGGL((k_mcc_collide<R>), dim3((nsel + 127) / 128), dim3(128), 0, g.stream, a);
    HIPCHK(hipGetLastError());
    return PIC_OK;
}

template <typename R>
static int append_products(int nsel, int esp, int isp_ion) {
    if (nsel == 0) return PIC_OK;
    double* dbl = (double*)g.d_newbuf;
    const double* new_e = dbl; const double* new_i = dbl + (size_t)nsel * 6;
    const unsigned char* flag = (const unsigned char*)(dbl + (size_t)nsel * 12);
    // count first (host) to check capacity before writing
    std::vector<unsigned char> hf(nsel);
    HIPCHK(hipMemcpyAsync(hf.data(), flag, nsel, hipMemcpyDeviceToHost, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    int nnew = 0;
    for (int k = 0; k < nsel; ++k) nnew += hf[k];
    if (nnew == 0) return PIC_OK;
    SpeciesHost& E = g.sp[esp];
    SpeciesHost& I = g.sp[isp_ion];
    if (E.nslots + nnew > E.capacity || I.nslots + nnew > I.capacity)
        return pic_fail(PIC_ERR_CAPACITY, "MCC products exceed particle capacity");
    int* dcount = (int*)g.d_blockcount;
    hipLaunchKernelGGL((k_append<R>), dim3(1), dim3(1024), 0, g.stream, nsel, flag, new_e, E.nslots, (R*)E.x, (R*)E.y,
                       (R*)E.vx, (R*)E.vy, (R*)E.vz, E.cell, dcount);
    hipLaunchKernelGGL((k_append<R>), dim3(1), dim3(1024), 0, g.stream, nsel, flag, new_i, I.nslots, (R*)I.x, (R*)I.y,
                       (R*)I.vx, (R*)I.vy, (R*)I.vz, I.cell, dcount);
    HIPCHK(hipGetLastError());
    E.nslots += nnew; E.nlive += nnew;
    I.nslots += nnew; I.nlive += nnew;
    unsigned long long add = (unsigned long long)nnew;
    // created counters (electron species counts both products it created)
    HIPCHK(hipMemcpyAsync(g.h_scal_pinned, E.counters + PIC_C_MCC_CREATED, 8, hipMemcpyDeviceToHost, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    unsigned long long cur; memcpy(&cur, g.h_scal_pinned, 8); cur += 2 * add;
    HIPCHK(hipMemcpy(E.counters + PIC_C_MCC_CREATED, &cur, 8, hipMemcpyHostToDevice));
    return PIC_OK;
}

// MCC order of MCCPublic.F90:50-105: for each gas (one), species 0..Ns with a table;
// products of all species are appended after the loop (MCCAddNewParticle).
template <typename R>
static int run_mcc_t() {
    int e = g.e_species, ion = g.i_species;
    int nsel_e = 0, nsel_i = 0, nc = 0;
    int rc;
    // electrons: select, collide (staging kept in d_newbuf), then ions would overwrite the
    // shared selection buffers, so the electron products are appended before the ions
    // are selected. Appending after selection of ions or before is equivalent because ion
    // selection only looks at ion slots that existed before this call.
    if (g.sp[e].has_mcc) {
        int64_t ion_slots_before = g.sp[ion].nslots;
        rc = select_species(e, &nsel_e); if (rc) return rc;
        rc = collide_species<R>(e, nsel_e, &nc); if (rc) return rc;
        unsigned long long c = nsel_e;
        (void)c;
        // ions selected against the slots present before the products arrive
        int64_t save = g.sp[ion].nslots;
        rc = append_products<R>(nsel_e, e, ion); if (rc) return rc;
        int64_t added = g.sp[ion].nslots - save;
        if (g.sp[ion].has_mcc) {
            g.sp[ion].nslots = ion_slots_before;
            rc = select_species(ion, &nsel_i);
            g.sp[ion].nslots = ion_slots_before + added;
            if (rc) return rc;
            rc = collide_species<R>(ion, nsel_i, &nc); if (rc) return rc;
        }
    } else if (g.sp[ion].has_mcc) {
        rc = select_species(ion, &nsel_i); if (rc) return rc;
        rc = collide_species<R>(ion, nsel_i, &nc); if (rc) return rc;
    }
    // selected counters
    for (int s = 0; s < g.nspecies; ++s) {
        int ns = (s == e) ? nsel_e : (s == ion ? nsel_i : 0);
        if (!ns) continue;
        unsigned long long cur;
        HIPCHK(hipMemcpy(&cur, g.sp[s].counters + PIC_C_MCC_SELECTED, 8, hipMemcpyDeviceToHost));
        cur += (unsigned long long)ns;
        HIPCHK(hipMemcpy(g.sp[s].counters + PIC_C_MCC_SELECTED, &cur, 8, hipMemcpyHostToDevice));
    }
    return PIC_OK;
}

int run_mcc() {
    if (g.particle_bits == 64) return run_mcc_t<double>();
    return run_mcc_t<float>();
}

// ---------------------------------------------------------------- cell sort
// Stable radix sort of (tile-major element key, slot) with dead slots keyed past the end,
// then a gather of every SoA array; occupied slots shrink to the live count.
__global__ void k_sort_keys(int64_t n, const int* cell, unsigned int* keys, unsigned int* vals, unsigned int dead_key,
                            int ney, int nty) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int c = cell[i];
    keys[i] = c >= 0 ? pic_sort_key(c, ney, nty) : dead_key;
    vals[i] = (unsigned int)i;
}
template <typename T>
__global__ void k_gather(int64_t n, const unsigned int* perm, const T* src, T* dst) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[perm[i]];
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
    size_t need = (size_t)n * 16 + (size_t)n * esz + 256;
    if (g.sort_buf_bytes < need) {
        if (g.d_sort_buf) hipFree(g.d_sort_buf);
        g.d_sort_buf = nullptr;
        HIPCHK(hipMalloc(&g.d_sort_buf, need));
        g.sort_buf_bytes = need;
    }
    unsigned int* k0 = (unsigned int*)g.d_sort_buf;
    unsigned int* k1 = k0 + n;
    unsigned int* v0 = k1 + n;
    unsigned int* v1 = v0 + n;
    void* tmpdata = (void*)(v1 + n);
    int blocks = (int)((n + 255) / 256);
    hipLaunchKernelGGL(k_sort_keys, dim3(blocks), dim3(256), 0, g.stream, n, S.cell, k0, v0, dead_key, g.mesh.ney, nty);
    size_t tb = 0;
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, k0, k1, v0, v1, (int)n, 0, end_bit, g.stream));
    if (g.sort_tmp_bytes < tb) {
        if (g.d_sort_tmp) hipFree(g.d_sort_tmp);
        g.d_sort_tmp = nullptr;
        HIPCHK(hipMalloc(&g.d_sort_tmp, tb));
        g.sort_tmp_bytes = tb;
    }
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(g.d_sort_tmp, tb, k0, k1, v0, v1, (int)n, 0, end_bit, g.stream));
    // live count
    int64_t live = 0;
    TRY_MCC(count_live_slots(n, S.cell, &live));
    // gather arrays through the temporary
    void* arrays[5] = {S.x, S.y, S.vx, S.vy, S.vz};
    for (int a = 0; a < 5; ++a) {
        if (esz == 8) hipLaunchKernelGGL(k_gather<double>, dim3(blocks), dim3(256), 0, g.stream, n, v1, (const double*)arrays[a], (double*)tmpdata);
        else hipLaunchKernelGGL(k_gather<float>, dim3(blocks), dim3(256), 0, g.stream, n, v1, (const float*)arrays[a], (float*)tmpdata);
        HIPCHK(hipMemcpyAsync(arrays[a], tmpdata, n * esz, hipMemcpyDeviceToDevice, g.stream));
    }
    hipLaunchKernelGGL(k_gather<int>, dim3(blocks), dim3(256), 0, g.stream, n, v1, (const int*)S.cell, (int*)tmpdata);
    HIPCHK(hipMemcpyAsync(S.cell, tmpdata, n * 4, hipMemcpyDeviceToDevice, g.stream));
    HIPCHK(hipGetLastError());
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
