// On-disk formats of the reference, for restart and interchange with its own tools:
//
//   particle bundle   MCC_jw DumpParticleBundle / LoadParticleBundle (ParticleBundle.f90:286-375):
//                     a Fortran unformatted sequential file of records [NPar, NParNormal]
//                     (INTEGER(4) x 2), [Charge, Mass, Weight] (REAL(8) x 3) and one record per
//                     particle [X, Y, Z, Vx, Vy, Vz, Ax, Ay, Az, Wq] (REAL(8) x 10) + Location
//                     (INTEGER), each framed by 4-byte little-endian length markers. The dump
//                     rescales the bundle before writing (PosRes / VelRes toggled, MCCInterface
//                     sets XFactor = L_ref, VFactor = v_ref): X * L_ref (PositionRescale-
//                     ParticleOne rescales X only, so Y and Z stay normalised) and Vx, Vy, Vz *
//                     v_ref; the load multiplies X by 1/L_ref and the velocities by 1/v_ref
//                     (MCCInterface.f90:136-141).
//   restart dump      Dump_2D (code/PIC/Dump_2D.f90:35-74) / Restart_2D: var<lap>dump, list-
//                     directed (lap, ntot, ns(1:nspecies), nx ny, Period_res, iStep_previous,
//                     nPeriod), and phi<lap>dump, one unformatted record of Phi(1:Field_Size).
//   Tecplot field     Output_To_Tecplot_IJK_2D (Output_To_Tecplot_IJK_2D.f90:144-161, Output_
//                     Option 1): field_IJ_<it>.dat, FEPOINT quadrilateral zone, node records in
//                     FORMAT 50 (E15.6 columns) and element records in FORMAT 44 (4 x I15).
#include "pic_device.h"
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#define TRY(x) do { int rc_ = (x); if (rc_) return rc_; } while (0)

namespace {

bool domain_ready() { return g.inited && g.have_domain; }

struct File {
    FILE* f = nullptr;
    explicit File(const char* path, const char* mode) { f = std::fopen(path, mode); }
    ~File() { if (f) std::fclose(f); }
};

// one Fortran unformatted sequential record: length, payload, length
bool put_record(FILE* f, const void* data, uint32_t bytes) {
    return std::fwrite(&bytes, 4, 1, f) == 1 && (bytes == 0 || std::fwrite(data, bytes, 1, f) == 1) &&
           std::fwrite(&bytes, 4, 1, f) == 1;
}

// reads one record of exactly `bytes` payload bytes; false on a short file or a length mismatch
bool get_record(FILE* f, void* data, uint32_t bytes) {
    uint32_t head = 0, tail = 0;
    if (std::fread(&head, 4, 1, f) != 1 || head != bytes) return false;
    if (bytes && std::fread(data, bytes, 1, f) != 1) return false;
    return std::fread(&tail, 4, 1, f) == 1 && tail == bytes;
}

// Fortran Ew.d edit descriptor (w = 15, d = 6 here): [-]0.dddddd E+ee, right-justified; the
// exponent drops its E above 99 (Fortran 2018, 13.7.2.3.3). Rounding from printf("%.*e").
std::string fortran_e(double v, int w, int d) {
    char buf[64];
    std::string s;
    if (v == 0.0 || !std::isfinite(v)) {
        if (!std::isfinite(v)) s = std::isnan(v) ? "NaN" : (v > 0 ? "Infinity" : "-Infinity");
        else s = std::string(std::signbit(v) ? "-" : "") + "0." + std::string(d, '0') + "E+00";
    } else {
        std::snprintf(buf, sizeof buf, "%.*e", d - 1, v);      // [-]D.DDDDDe[+-]XX
        const char* p = buf;
        const bool neg = *p == '-';
        if (neg) ++p;
        std::string digits;
        digits += p[0];
        const char* q = p + 2;
        while (*q && *q != 'e') digits += *q++;
        const int ex = std::atoi(q + 1) + 1;
        char e[16];
        if (std::abs(ex) <= 99) std::snprintf(e, sizeof e, "E%c%02d", ex < 0 ? '-' : '+', std::abs(ex));
        else std::snprintf(e, sizeof e, "%c%03d", ex < 0 ? '-' : '+', std::abs(ex));
        s = std::string(neg ? "-" : "") + "0." + digits + e;
    }
    if ((int)s.size() > w) return std::string(w, '*');
    return std::string(w - s.size(), ' ') + s;
}

std::string fortran_i(long long v, int w) {
    std::string s = std::to_string(v);
    if ((int)s.size() > w) return std::string(w, '*');
    return std::string(w - s.size(), ' ') + s;
}

int species_ok(int isp) {
    if (isp < 0 || isp >= PIC_MAX_SPECIES || !g.sp[isp].defined) return pic_fail(PIC_ERR_ARG, "species not defined");
    return PIC_OK;
}

}  // namespace

extern "C" {

int pic_format_fortran_e(double v, int w, int d, char* out) {
    if (!out || w < 1 || w > 60 || d < 1 || d > 30) return pic_fail(PIC_ERR_ARG, "bad Ew.d");
    const std::string s = fortran_e(v, w, d);
    std::memcpy(out, s.c_str(), s.size() + 1);
    return PIC_OK;
}

int pic_DumpParticleBundle(int isp, const char* path) {
    if (!domain_ready()) return pic_fail(PIC_ERR_STATE, "pic_setup_domain not called");
    TRY(species_ok(isp));
    if (!path) return pic_fail(PIC_ERR_ARG, "no file name");
    const SpeciesHost& S = g.sp[isp];
    const int64_t cap = std::max<int64_t>(S.nslots, 1);
    std::vector<double> x(cap), y(cap), vx(cap), vy(cap), vz(cap), th;
    std::vector<int32_t> loc(cap);
    int64_t n = 0;
    TRY(pic_get_particles(isp, x.data(), y.data(), vx.data(), vy.data(), vz.data(), loc.data(), cap, &n));
    if (g.mesh.delta == 1) {
        th.resize(cap);
        int64_t nz = 0;
        TRY(pic_get_azimuth(isp, th.data(), cap, &nz));
    }
    if (n > (int64_t)INT32_MAX) return pic_fail(PIC_ERR_CAPACITY, "bundle larger than an INTEGER(4) count");
    File F(path, "wb");
    if (!F.f) return pic_fail(PIC_ERR_ARG, std::string("cannot open ") + path);
    const int32_t head[2] = {(int32_t)n, (int32_t)std::min<int64_t>(S.capacity, INT32_MAX)};
    const double phys[3] = {S.p.charge_C, S.p.mass_kg, S.p.weight};
    bool ok = put_record(F.f, head, 8) && put_record(F.f, phys, 24);
    unsigned char rec[84];
    for (int64_t i = 0; ok && i < n; ++i) {
        const double v[10] = {x[i] * g.L_ref, y[i], th.empty() ? 0.0 : th[i], vx[i] * g.v_ref, vy[i] * g.v_ref,
                              vz[i] * g.v_ref, 0.0, 0.0, 0.0, 0.0};
        std::memcpy(rec, v, 80);
        std::memcpy(rec + 80, &loc[i], 4);
        ok = put_record(F.f, rec, 84);
    }
    if (!ok) return pic_fail(PIC_ERR_ARG, std::string("write failed: ") + path);
    return PIC_OK;
}

int pic_LoadParticleBundle(int isp, const char* path, int64_t* npar) {
    if (!domain_ready()) return pic_fail(PIC_ERR_STATE, "pic_setup_domain not called");
    TRY(species_ok(isp));
    if (!path) return pic_fail(PIC_ERR_ARG, "no file name");
    const SpeciesHost& S = g.sp[isp];
    File F(path, "rb");
    if (!F.f) return pic_fail(PIC_ERR_ARG, std::string("cannot open ") + path);
    int32_t head[2];
    double phys[3];
    if (!get_record(F.f, head, 8) || !get_record(F.f, phys, 24) || head[0] < 0)
        return pic_fail(PIC_ERR_ARG, std::string("not a particle-bundle file: ") + path);
    // the bundle belongs to this species only if charge, mass and weight agree
    for (int k = 0; k < 3; ++k) {
        const double want = k == 0 ? S.p.charge_C : (k == 1 ? S.p.mass_kg : S.p.weight);
        if (!(std::fabs(phys[k] - want) <= 1e-12 * std::fabs(want)))
            return pic_fail(PIC_ERR_ARG, "bundle charge, mass or weight differs from the species");
    }
    const int64_t n = head[0];
    if (n > S.capacity) return pic_fail(PIC_ERR_CAPACITY, "bundle exceeds the species capacity");
    std::vector<double> x(n), y(n), z(n), vx(n), vy(n), vz(n);
    std::vector<int32_t> loc(n);
    const double xf = 1.0 / g.L_ref, vf = 1.0 / g.v_ref;
    unsigned char rec[84];
    for (int64_t i = 0; i < n; ++i) {
        if (!get_record(F.f, rec, 84)) return pic_fail(PIC_ERR_ARG, std::string("truncated particle bundle: ") + path);
        double v[10];
        std::memcpy(v, rec, 80);
        std::memcpy(&loc[i], rec + 80, 4);
        x[i] = v[0] * xf; y[i] = v[1]; z[i] = v[2];
        vx[i] = v[3] * vf; vy[i] = v[4] * vf; vz[i] = v[5] * vf;
    }
    TRY(pic_load_particles(isp, n, x.data(), y.data(), vx.data(), vy.data(), vz.data(), loc.data()));
    if (g.mesh.delta == 1) TRY(pic_set_azimuth(isp, n, z.data()));
    if (npar) *npar = n;
    return PIC_OK;
}

int pic_Dump_2D(int lap, const char* dir) {
    if (!domain_ready()) return pic_fail(PIC_ERR_STATE, "pic_setup_domain not called");
    if (lap < 0 || lap > 9999999) return pic_fail(PIC_ERR_ARG, "lap out of the I7.7 range");
    const std::string d = dir ? dir : ".";
    char name[32];
    int64_t ns[PIC_MAX_SPECIES] = {0}, ntot = 0;
    for (int s = 0; s < g.nspecies; ++s) {
        TRY(pic_get_counts(s, &ns[s], nullptr, nullptr));
        ntot += ns[s];
    }
    std::snprintf(name, sizeof name, "/var%07ddump", lap);
    {
        File F((d + name).c_str(), "w");
        if (!F.f) return pic_fail(PIC_ERR_ARG, "cannot open " + d + name);
        // list-directed: one item group per WRITE; Period_res, iStep_previous and nPeriod (the
        // RF-cycle bookkeeping of the reference's driver) are not on this path and are written 0
        std::fprintf(F.f, " %d\n %lld\n", lap, (long long)ntot);
        for (int s = 0; s < g.nspecies; ++s) std::fprintf(F.f, " %lld", (long long)ns[s]);
        std::fprintf(F.f, "\n %d %d\n 0.0\n 0\n 0\n", g.mesh.nnx, g.mesh.nny);
    }
    std::vector<double> phi(g.mesh.nnodes);
    TRY(pic_get_field(PIC_F_PHI, phi.data(), (int64_t)phi.size()));
    std::snprintf(name, sizeof name, "/phi%07ddump", lap);
    File F((d + name).c_str(), "wb");
    if (!F.f) return pic_fail(PIC_ERR_ARG, "cannot open " + d + name);
    if (phi.size() * 8 > 0x7FFFFFFFu) return pic_fail(PIC_ERR_CAPACITY, "Phi exceeds one Fortran record");
    if (!put_record(F.f, phi.data(), (uint32_t)(phi.size() * 8))) return pic_fail(PIC_ERR_ARG, "write failed: " + d + name);
    return PIC_OK;
}

int pic_Restart_2D(int lap, const char* dir) {
    if (!domain_ready()) return pic_fail(PIC_ERR_STATE, "pic_setup_domain not called");
    if (lap < 0 || lap > 9999999) return pic_fail(PIC_ERR_ARG, "lap out of the I7.7 range");
    const std::string d = dir ? dir : ".";
    char name[32];
    std::snprintf(name, sizeof name, "/var%07ddump", lap);
    {
        File F((d + name).c_str(), "r");
        if (!F.f) return pic_fail(PIC_ERR_ARG, "cannot open " + d + name);
        long long lapd = 0, ntot = 0;
        if (std::fscanf(F.f, "%lld %lld", &lapd, &ntot) != 2) return pic_fail(PIC_ERR_ARG, "bad " + d + name);
        long long ns = 0;
        for (int s = 0; s < g.nspecies; ++s)
            if (std::fscanf(F.f, "%lld", &ns) != 1) return pic_fail(PIC_ERR_ARG, "bad " + d + name);
        int nx = 0, ny = 0;
        if (std::fscanf(F.f, "%d %d", &nx, &ny) != 2) return pic_fail(PIC_ERR_ARG, "bad " + d + name);
        if (nx != g.mesh.nnx || ny != g.mesh.nny) return pic_fail(PIC_ERR_ARG, "restart mesh differs from the domain");
    }
    std::vector<double> phi(g.mesh.nnodes);
    std::snprintf(name, sizeof name, "/phi%07ddump", lap);
    File F((d + name).c_str(), "rb");
    if (!F.f) return pic_fail(PIC_ERR_ARG, "cannot open " + d + name);
    if (!get_record(F.f, phi.data(), (uint32_t)(phi.size() * 8))) return pic_fail(PIC_ERR_ARG, "bad " + d + name);
    return pic_set_field(PIC_F_PHI, phi.data(), (int64_t)phi.size());
}

int pic_Output_To_Tecplot_IJ(int it, const char* dir) {
    if (!domain_ready()) return pic_fail(PIC_ERR_STATE, "pic_setup_domain not called");
    if (it < 0 || it > 999999) return pic_fail(PIC_ERR_ARG, "it out of the I6.6 range");
    const int64_t nn = g.mesh.nnodes, ne = g.mesh.nelem;
    std::vector<double> phi(nn), rho(nn), efx(nn), efy(nn), rs(4 * nn, 0.0), eks(4 * nn, 0.0), ekt(4 * nn, 0.0);
    TRY(pic_get_field(PIC_F_PHI, phi.data(), nn));
    TRY(pic_get_field(PIC_F_RHO, rho.data(), nn));
    TRY(pic_get_field(PIC_F_EFX, efx.data(), nn));
    TRY(pic_get_field(PIC_F_EFY, efy.data(), nn));
    if (g.nspecies > 0) {
        TRY(pic_get_field(PIC_F_RHO_S, rs.data(), nn * g.nspecies));
        TRY(pic_get_field(PIC_F_EK_S, eks.data(), nn * g.nspecies));
        TRY(pic_get_field(PIC_F_EK_TOT, ekt.data(), nn * g.nspecies));
    }
    auto sp = [&](const std::vector<double>& a, int s, int64_t i) { return s < g.nspecies ? a[s * nn + i] : 0.0; };
    char name[48];
    std::snprintf(name, sizeof name, "/field_IJ_%06d.dat", it);
    const std::string path = std::string(dir ? dir : ".") + name;
    File F(path.c_str(), "w");
    if (!F.f) return pic_fail(PIC_ERR_ARG, "cannot open " + path);
    std::string var = "VARIABLES = \"x\" \"y\" \"R\" \"Phi\" \"Rho\" \"Rho_ele\" \"Rho_H\" \"Rho_C\" \"efx\" \"efy\" "
                      "\"Ek_ele\" \"Ek_ion\" \"Ek_tot_ele\" \"Ek_tot_ion\" \"node_type1\"  \"node_type2\"";
    var.resize(150, ' ');                                         // (A150)
    auto pad50 = [](long long v) { std::string s = fortran_i(v, 10); s.resize(50, ' '); return s; };   // I10 into CHARACTER*50
    std::fprintf(F.f, " TITLE = \"Field Plot\"\n%s\n ZONE N=%s, E=%s,F=FEPOINT,ET=QUADRILATERAL\n", var.c_str(),
                 pad50(nn).c_str(), pad50(ne).c_str());
    const MeshParams& m = g.mesh;
    std::string line;
    for (int64_t i = 0; i < nn; ++i) {
        const int ix = (int)(i / m.nny), jy = (int)(i % m.nny);
        const double x = g.h_xs[ix], y = g.h_ys[jy];
        const double vals[14] = {x, y, std::sqrt(x * x + y * y), phi[i], rho[i], sp(rs, 0, i), sp(rs, 1, i), sp(rs, 2, i),
                                 efx[i], efy[i], sp(eks, 0, i), sp(eks, 1, i), sp(ekt, 0, i), sp(ekt, 1, i)};
        // FORMAT 50: blanks between the E15.6 fields except after the 7th ('') and the 10th
        line.clear();
        for (int k = 0; k < 14; ++k) {
            line += fortran_e(vals[k], 15, 6);
            if (k < 13 && k != 6 && k != 9) line += ' ';
        }
        // node_type(1, i): 0 for an unknown, 1 for a Dirichlet node; node_type(2, i): unknown id or <= 0
        const int nt2 = g.h_node_type2[i];
        line += ' ' + fortran_i(nt2 > 0 ? 0 : 1, 15) + ' ' + fortran_i(nt2, 15);
        std::fprintf(F.f, "%s\n", line.c_str());
    }
    for (int64_t e = 0; e < ne; ++e) {
        const int ci = (int)(e / m.ney), cj = (int)(e % m.ney);
        const long long n1 = (long long)ci * m.nny + cj + 1, n2 = n1 + m.nny, n3 = n2 + 1, n4 = n1 + 1;   // HT 1..4
        std::fprintf(F.f, "%s%s%s%s\n", fortran_i(n1, 15).c_str(), fortran_i(n2, 15).c_str(), fortran_i(n3, 15).c_str(),
                     fortran_i(n4, 15).c_str());
    }
    if (std::ferror(F.f)) return pic_fail(PIC_ERR_ARG, "write failed: " + path);
    return PIC_OK;
}

}  // extern "C"
