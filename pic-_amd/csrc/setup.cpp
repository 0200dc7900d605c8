// Host-side FP64 set-up of the IFE-PIC hot path: the geometry-only work the reference
// redoes inside its setup routines (and, for PPR, every step) is done once here and
// handed to the device as compact operators.
//
//   mesh / nodes / elements  : Cubic_Partition_2D.f90:30-87 (node (i,j) -> (i-1)ny+j,
//                              element -> row + (col-1)(ny-1), HT = LL, LR, UR, UL)
//   node / element classes   : Update_Locate_Node_2D, Update_Mesh_Objects_Intersection_2D
//                              (element_index >0 interface, -1 vacuum, -2 object;
//                              node_index -1 vacuum, -2 object), classified by nodes
//   Dirichlet / unknowns     : Classify_Nodes_2D.f90:33-80 (unknowns numbered in node order)
//   stiffness + diagonal     : IFE_Start_2D.f90:313-483, Diagonal_Preconditioner.f90
//   rhs operator             : Global_RHS_PDE_2D.f90:52-143 (rhs_i = int N_i rho_h, linear
//                              in rho -> matrix B) and Global_RHS_EBC_2D_bjw.f90:45-78
//                              (Dirichlet lift -> constant c)
//   cell volumes             : SetupGrids_2D_QLL.f90:196-395
//   PPR point patches        : Get_PointPatch.f90:83-590, least squares of
//                              GetEfield_SIDG_PPR.f90:153-241 (geometry only -> rows)
//   MCC tables               : MCCInit.F90:14-152 ProbilityNonReactive, MCCSigma.f90:9-120
#include "pic_device.h"
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <set>

static inline int node_id0(int i0, int j0, int nny) { return i0 * nny + j0; }   // 0-based

// ------------------------------------------------------------------ geometry helpers
static bool inside_object(const pic_object& o, double x, double y) {
    if (o.shape == 1) {
        double dx = x - o.locations[0][0], dy = y - o.locations[0][1];
        return dx * dx + dy * dy < o.dimensions[0] * o.dimensions[0];
    }
    if (o.shape == 3) {
        return x > o.locations[0][0] && x < o.locations[1][0] &&
               y > o.locations[0][1] && y < o.locations[1][1];
    }
    return false;
}

// Bilinear Q1 element matrices on an hx x hy rectangle, node order LL, LR, UR, UL.
static void q1_stiffness_cart(double hx, double hy, double K[4][4]) {
    static const double Mx[4][4] = {{2, -2, -1, 1}, {-2, 2, 1, -1}, {-1, 1, 2, -2}, {1, -1, -2, 2}};
    static const double My[4][4] = {{2, 1, -1, -2}, {1, 2, -2, -1}, {-1, -2, 2, 1}, {-2, -1, 1, 2}};
    double ax = hy / (6.0 * hx), ay = hx / (6.0 * hy);
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b) K[a][b] = ax * Mx[a][b] + ay * My[a][b];
}
static void q1_mass_cart(double hx, double hy, double M[4][4]) {
    static const double P[4][4] = {{4, 2, 1, 2}, {2, 4, 2, 1}, {1, 2, 4, 2}, {2, 1, 2, 4}};
    double s = hx * hy / 36.0;
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b) M[a][b] = s * P[a][b];
}
// Axisymmetric (r = y): 2x2 Gauss with weight r, exact for these polynomial integrands.
static void q1_axi(double x0, double y0, double hx, double hy, double K[4][4], double M[4][4]) {
    const double g = 1.0 / std::sqrt(3.0);
    const double gp[2] = {0.5 * (1.0 - g), 0.5 * (1.0 + g)};
    for (int a = 0; a < 4; ++a) for (int b = 0; b < 4; ++b) { K[a][b] = 0; M[a][b] = 0; }
    for (int qi = 0; qi < 2; ++qi)
        for (int qj = 0; qj < 2; ++qj) {
            double s = gp[qi], t = gp[qj];
            double r = y0 + t * hy;
            double w = 0.25 * hx * hy * r;
            double N[4] = {(1 - s) * (1 - t), s * (1 - t), s * t, (1 - s) * t};
            double Nx[4] = {-(1 - t) / hx, (1 - t) / hx, t / hx, -t / hx};
            double Ny[4] = {-(1 - s) / hy, -s / hy, s / hy, (1 - s) / hy};
            for (int a = 0; a < 4; ++a)
                for (int b = 0; b < 4; ++b) {
                    K[a][b] += w * (Nx[a] * Nx[b] + Ny[a] * Ny[b]);
                    M[a][b] += w * N[a] * N[b];
                }
        }
    (void)x0;
}

// ------------------------------------------------------------------ IFE interface element
// A square element (side h, lower-left (x0, y0)) cut by the circle of a dielectric object:
// bilinear immersed-finite-element basis on the reference square of the reference's element
// orientation (El_Object_Intersection_info_2D_BL.f90: A1 = the lone vertex for a corner cut;
// for a cut through two opposite edges A1 = LL when bottom and top are cut, UL when left and
// right are; E on A1A2, D on A4A1 or A4A3; piece 1 = the side of A1), defined by
// Retangular_reference_basis_IFE's conditions: nodal values, continuity at D and E, one
// cross coefficient on both pieces, beta * normal flux continuous at the midpoint of DE.
// K = sum over pieces beta int grad phi_i . grad phi_j (times r when axisymmetric) by a
// degree-4 rule, exact here; M = the RHS operator of IFE_RHS_2D (rho_h = sum rho_k phi_k on
// the reference's sub-triangles, edge-midpoint rule). K and M rows/columns are LL, LR, UR, UL.
// Returns 1 or 2 (the cut type), -1 for two diagonal vertices inside (no bilinear IFE
// element), -2 if a cut point lies within 1e-12 h of a vertex.
static int ife_element(int delta, double x0, double y0, double h, const bool inside[4], double cx, double cy,
                       double rad, double beta_in, double beta_out, double K[4][4], double M[4][4]) {
    static const double VX[4] = {0, 1, 1, 0}, VY[4] = {0, 0, 1, 1};
    int kin = (int)inside[0] + inside[1] + inside[2] + inside[3];
    int type, rot = 0;
    if (kin == 1 || kin == 3) {
        type = 1;
        for (int k = 0; k < 4; ++k)
            if (inside[k] != inside[(k + 1) & 3] && inside[k] != inside[(k + 3) & 3]) rot = k;
    } else {
        if (inside[0] == inside[2]) return -1;
        type = 2;
        rot = (inside[0] == inside[3]) ? 0 : 3;
    }
    int A[4];
    for (int k = 0; k < 4; ++k) A[k] = (rot + k) & 3;
    // reference frame: local (X, Y) = V[A1] + xi e1 + eta e2
    const double e1x = VX[A[1]] - VX[A[0]], e1y = VY[A[1]] - VY[A[0]];
    const double e2x = VX[A[3]] - VX[A[0]], e2y = VY[A[3]] - VY[A[0]];
    auto toY = [&](double xi, double eta) { return VY[A[0]] + xi * e1y + eta * e2y; };
    // edge crossing: parameter from vertex u towards vertex v
    auto cross = [&](int u, int v) {
        const int vi = inside[u] ? u : v, vo = inside[u] ? v : u;
        const double px = x0 + h * VX[vi], py = y0 + h * VY[vi];
        const double dx = h * (VX[vo] - VX[vi]), dy = h * (VY[vo] - VY[vi]);
        const double fx = px - cx, fy = py - cy;
        const double a2 = dx * dx + dy * dy, b1 = fx * dx + fy * dy, c0 = fx * fx + fy * fy - rad * rad;
        const double t = (-b1 + std::sqrt(b1 * b1 - a2 * c0)) / a2;      // from the inside vertex
        return (vi == u) ? t : 1.0 - t;
    };
    const double a = cross(A[0], A[1]);                                      // E = (a, 0)
    const double b = (type == 1) ? 1.0 - cross(A[3], A[0]) : cross(A[3], A[2]);  // D = (0, b) or (b, 1)
    if (!(a > 1e-12 && a < 1 - 1e-12 && b > 1e-12 && b < 1 - 1e-12)) return -2;
    const double Dxi = (type == 1) ? 0.0 : b, Deta = (type == 1) ? b : 1.0, Exi = a, Eeta = 0.0;
    const double beta1 = inside[A[0]] ? beta_in : beta_out, beta2 = inside[A[0]] ? beta_out : beta_in;
    const double bp[2] = {beta1, beta2};
    static const double RX[4] = {0, 1, 1, 0}, RY[4] = {0, 0, 1, 1};
    auto pieceof = [&](int k) { return (k == 0 || (type == 2 && k == 3)) ? 0 : 1; };
    // basis coefficients c[k][piece][0..3] of c0 + c1 xi + c2 eta + c3 xi eta for reference vertex k
    double c[4][2][4];
    const double mx = 0.5 * (Dxi + Exi), my = 0.5 * (Deta + Eeta);
    const double nx = Eeta - Deta, ny = -(Exi - Dxi);
    for (int k = 0; k < 4; ++k) {
        double S[8][9] = {{0}};
        for (int m = 0; m < 4; ++m) {
            const int o = 4 * pieceof(m);
            S[m][o] = 1; S[m][o + 1] = RX[m]; S[m][o + 2] = RY[m]; S[m][o + 3] = RX[m] * RY[m];
            S[m][8] = (m == k) ? 1.0 : 0.0;
        }
        const double cp[2][2] = {{Dxi, Deta}, {Exi, Eeta}};
        for (int q = 0; q < 2; ++q) {
            const double X = cp[q][0], Y = cp[q][1], f[4] = {1, X, Y, X * Y};
            for (int j = 0; j < 4; ++j) { S[4 + q][j] = f[j]; S[4 + q][4 + j] = -f[j]; }
        }
        S[6][3] = 1; S[6][7] = -1;
        // beta1 grad p1 . n - beta2 grad p2 . n at the midpoint: grad = (c1 + c3 eta, c2 + c3 xi)
        for (int pc = 0; pc < 2; ++pc) {
            const double sg = pc == 0 ? bp[0] : -bp[1];
            S[7][4 * pc + 1] = sg * nx; S[7][4 * pc + 2] = sg * ny; S[7][4 * pc + 3] = sg * (nx * my + ny * mx);
        }
        for (int col = 0; col < 8; ++col) {                 // Gauss-Jordan with partial pivoting
            int pv = col;
            for (int r = col + 1; r < 8; ++r) if (std::fabs(S[r][col]) > std::fabs(S[pv][col])) pv = r;
            if (S[pv][col] == 0.0) return -3;
            if (pv != col) for (int j = 0; j < 9; ++j) std::swap(S[pv][j], S[col][j]);
            for (int r = 0; r < 8; ++r) {
                if (r == col || S[r][col] == 0.0) continue;
                const double f = S[r][col] / S[col][col];
                for (int j = col; j < 9; ++j) S[r][j] -= f * S[col][j];
            }
        }
        for (int pc = 0; pc < 2; ++pc)
            for (int j = 0; j < 4; ++j) c[k][pc][j] = S[4 * pc + j][8] / S[4 * pc + j][4 * pc + j];
    }
    // sub-triangles of Inter_Tetra_Partition_2D: points 0..3 = A1..A4, 4 = D, 5 = E; last = piece
    const double P[6][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}, {Dxi, Deta}, {Exi, Eeta}};
    static const int T1[4][4] = {{0, 4, 5, 0}, {1, 5, 2, 1}, {5, 2, 4, 1}, {2, 4, 3, 1}};
    static const int T2[4][4] = {{0, 5, 3, 0}, {3, 4, 5, 0}, {5, 2, 4, 1}, {2, 5, 1, 1}};
    const int (*T)[4] = (type == 1) ? T1 : T2;
    // symmetric 6-point rule on the triangle, degree 4 (Dunavant 1985)
    static const double DW[2] = {0.223381589678011, 0.109951743655322};
    static const double DA[2] = {0.445948490915965, 0.091576213509771};
    double Kr[4][4] = {{0}}, Mr[4][4] = {{0}};
    for (int st = 0; st < 4; ++st) {
        const double* p1 = P[T[st][0]]; const double* p2 = P[T[st][1]]; const double* p3 = P[T[st][2]];
        const int pc = T[st][3];
        const double area = 0.5 * std::fabs((p2[0] - p1[0]) * (p3[1] - p1[1]) - (p3[0] - p1[0]) * (p2[1] - p1[1]));
        for (int g2 = 0; g2 < 2; ++g2)
            for (int perm = 0; perm < 3; ++perm) {
                double l[3] = {DA[g2], DA[g2], DA[g2]};
                l[perm] = 1.0 - 2.0 * DA[g2];
                const double xi = l[0] * p1[0] + l[1] * p2[0] + l[2] * p3[0];
                const double eta = l[0] * p1[1] + l[1] * p2[1] + l[2] * p3[1];
                double w = DW[g2] * area * bp[pc];
                if (delta == 1) w *= y0 + h * toY(xi, eta);
                double gx[4], gy[4];
                for (int k = 0; k < 4; ++k) { gx[k] = c[k][pc][1] + c[k][pc][3] * eta; gy[k] = c[k][pc][2] + c[k][pc][3] * xi; }
                for (int i = 0; i < 4; ++i)
                    for (int j = 0; j < 4; ++j) Kr[i][j] += w * (gx[i] * gx[j] + gy[i] * gy[j]);
            }
        const double* tv[3] = {p1, p2, p3};
        for (int e = 0; e < 3; ++e) {
            const double xi = 0.5 * (tv[e][0] + tv[(e + 1) % 3][0]), eta = 0.5 * (tv[e][1] + tv[(e + 1) % 3][1]);
            double w = h * h * area / 3.0;
            if (delta == 1) w *= y0 + h * toY(xi, eta);
            double ph[4];
            for (int k = 0; k < 4; ++k) ph[k] = c[k][pc][0] + c[k][pc][1] * xi + c[k][pc][2] * eta + c[k][pc][3] * xi * eta;
            for (int i = 0; i < 4; ++i)
                for (int j = 0; j < 4; ++j) Mr[i][j] += w * ph[i] * ph[j];
        }
    }
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) { K[A[i]][A[j]] = Kr[i][j]; M[A[i]][A[j]] = Mr[i][j]; }
    return type;
}

// ------------------------------------------------------------------ domain build
int setup_build_domain(const pic_domain_config* c) {
    if (c->nnx < 3 || c->nny < 3) return pic_fail(PIC_ERR_ARG, "nnx, nny must be >= 3");
    // node-indexed device arrays are addressed with 32-bit byte offsets (8-byte entries)
    if ((int64_t)c->nnx * (int64_t)c->nny >= ((int64_t)1 << 29))
        return pic_fail(PIC_ERR_ARG, "nnx * nny must be below 2^29");
    if (!(c->xmax > c->xmin) || !(c->ymax > c->ymin)) return pic_fail(PIC_ERR_ARG, "empty domain");
    if (c->n_objects < 0 || c->n_objects > PIC_MAX_OBJECTS) return pic_fail(PIC_ERR_ARG, "n_objects out of range");
    if (c->delta != 0 && c->delta != 1) return pic_fail(PIC_ERR_ARG, "delta must be 0 or 1");
    if (!(c->dt > 0) || !(c->n_ref > 0) || !(c->T_ref_eV > 0)) return pic_fail(PIC_ERR_ARG, "dt, n_ref, T_ref must be > 0");
    for (int k = 0; k < c->n_objects; ++k) {
        const pic_object& o = c->objects[k];
        if (o.shape != 1 && o.shape != 3) return pic_fail(PIC_ERR_UNSUPPORTED, "object shape must be 1 (circle) or 3 (box)");
        if (!o.conductor && !(o.eps > 0.0)) return pic_fail(PIC_ERR_ARG, "a dielectric needs eps > 0");
        if (o.eps != 1.0 && o.shape != 1)
            return pic_fail(PIC_ERR_UNSUPPORTED, "objects with eps != 1 must be circles (shape 1)");
    }
    g.cfg = *c;
    g.n_ife_elements = 0;
    const int nnx = c->nnx, nny = c->nny, nex = nnx - 1, ney = nny - 1;
    const int nnodes = nnx * nny, nelem = nex * ney;

    // Normalisation, Input_2D.f90:356-398 with Constant_Variable_2D.f90 constants.
    const double T_ref = c->T_ref_eV * PIC_EVTOK;
    g.L_ref = std::sqrt(PIC_EPSILON0 * PIC_KB * T_ref / c->n_ref / PIC_ELECTRON_CHARGE / PIC_ELECTRON_CHARGE);
    g.time_ref = std::sqrt(PIC_EPSILON0 * PIC_ELECTRON_MASS / c->n_ref / PIC_ELECTRON_CHARGE / PIC_ELECTRON_CHARGE);
    g.v_ref = std::sqrt(PIC_KB * T_ref / PIC_ELECTRON_MASS);
    g.phi_ref = PIC_KB * T_ref / PIC_ELECTRON_CHARGE;

    const double hx = (c->xmax - c->xmin) / (double)nex;
    const double hy = (c->ymax - c->ymin) / (double)ney;
    for (int k = 0; k < c->n_objects; ++k)
        if (c->objects[k].eps != 1.0 && std::fabs(hx - hy) > 1e-12 * hx)
            return pic_fail(PIC_ERR_UNSUPPORTED, "objects with eps != 1 need square elements (hx == hy)");
    g.h_xs.resize(nnx); g.h_ys.resize(nny);
    for (int i = 0; i < nnx; ++i) g.h_xs[i] = c->xmin + hx * (double)i;     // x_s + h_x*(count-1)
    for (int j = 0; j < nny; ++j) g.h_ys[j] = c->ymin + hy * (double)j;

    MeshParams& m = g.mesh;
    m.nnx = nnx; m.nny = nny; m.nex = nex; m.ney = ney; m.nnodes = nnodes; m.nelem = nelem;
    m.delta = c->delta;
    m.xmin = c->xmin; m.xmax = c->xmax; m.ymin = c->ymin; m.ymax = c->ymax;
    m.hx = hx; m.hy = hy; m.hxi = 1.0 / hx; m.hyi = 1.0 / hy;
    m.vox = c->xmin - hx; m.voy = c->ymin - hy;                 // Vert_o = min - h
    m.mapx = hx; m.mapy = hy;                                    // hx / 2^0
    m.imapx = 1.0 / hx; m.imapy = 1.0 / hy;
    int ex_;
    m.pow2_map = (std::frexp(hx, &ex_) == 0.5 && std::frexp(hy, &ex_) == 0.5) ? 1 : 0;
    {
        const double kx = c->xmin / hx, ky = c->ymin / hy;
        m.locate_f32 = (m.pow2_map && m.mapx == hx && m.mapy == hy && std::floor(kx) == kx && std::floor(ky) == ky &&
                        std::fabs(kx) < 1048576.0 && std::fabs(ky) < 1048576.0) ? 1 : 0;
    }
    m.nodes_f32 = 1;
    {
        const float x0 = (float)c->xmin, y0 = (float)c->ymin, hxf = (float)hx, hyf = (float)hy;
        for (int i = 0; i < nnx; ++i) if ((float)g.h_xs[i] != x0 + hxf * (float)i) m.nodes_f32 = 0;
        for (int j = 0; j < nny; ++j) if ((float)g.h_ys[j] != y0 + hyf * (float)j) m.nodes_f32 = 0;
    }
    // the spacing of every element from the node lines as the push recomputes them (FP32:
    // x0 + hx i in FP32 when nodes_f32, else the FP64 line rounded; FP64: xmin + hx i)
    m.unif_r32 = m.pow2_map; m.unif_r64 = m.pow2_map;
    {
        const float x0 = (float)c->xmin, y0 = (float)c->ymin, hxf = (float)hx, hyf = (float)hy;
        auto nx32 = [&](int i) { return m.nodes_f32 ? x0 + hxf * (float)i : (float)(c->xmin + hx * (double)i); };
        auto ny32 = [&](int j) { return m.nodes_f32 ? y0 + hyf * (float)j : (float)(c->ymin + hy * (double)j); };
        for (int i = 0; i + 1 < nnx; ++i) {
            if (nx32(i + 1) - nx32(i) != hxf) m.unif_r32 = 0;
            if ((c->xmin + hx * (double)(i + 1)) - (c->xmin + hx * (double)i) != hx) m.unif_r64 = 0;
        }
        for (int j = 0; j + 1 < nny; ++j) {
            if (ny32(j + 1) - ny32(j) != hyf) m.unif_r32 = 0;
            if ((c->ymin + hy * (double)(j + 1)) - (c->ymin + hy * (double)j) != hy) m.unif_r64 = 0;
        }
    }
    m.ney_rcp = 1.0f / (float)ney;
    m.ij_exact_div = ((int64_t)nelem >= ((int64_t)1 << 24)) ? 1 : 0;
    m.n_objects = c->n_objects;
    for (int k = 0; k < c->n_objects; ++k) {
        const pic_object& o = c->objects[k];
        DevObject& d = m.obj[k];
        d.shape = o.shape;
        d.x0 = o.locations[0][0]; d.y0 = o.locations[0][1];
        d.x1 = o.locations[1][0]; d.y1 = o.locations[1][1];
        d.r2 = o.dimensions[0] * o.dimensions[0];
    }

    // node_index / element_index
    g.h_node_index.assign(nnodes, -1);
    std::vector<int> node_obj(nnodes, -1);
    for (int i = 0; i < nnx; ++i)
        for (int j = 0; j < nny; ++j) {
            int n = node_id0(i, j, nny);
            for (int k = 0; k < c->n_objects; ++k)
                if (inside_object(c->objects[k], g.h_xs[i], g.h_ys[j])) { g.h_node_index[n] = -2; node_obj[n] = k; break; }
        }
    g.h_element_index.assign(nelem, -1);
    int n_interface = 0;
    for (int ci = 0; ci < nex; ++ci)
        for (int cj = 0; cj < ney; ++cj) {
            int e = cj + ci * ney;
            int nodes[4] = {node_id0(ci, cj, nny), node_id0(ci + 1, cj, nny), node_id0(ci + 1, cj + 1, nny), node_id0(ci, cj + 1, nny)};
            int kin = 0;
            for (int a = 0; a < 4; ++a) kin += (g.h_node_index[nodes[a]] == -2);
            if (kin == 4) g.h_element_index[e] = -2;
            else if (kin == 0) g.h_element_index[e] = -1;
            else g.h_element_index[e] = ++n_interface;
        }

    // Dirichlet nodes: conductor interiors at the object potential, then Dirichlet walls.
    g.h_dirichlet_phi.assign(nnodes, 0.0);
    std::vector<char> is_dir(nnodes, 0);
    for (int n = 0; n < nnodes; ++n)
        if (node_obj[n] >= 0 && c->objects[node_obj[n]].conductor) { is_dir[n] = 1; g.h_dirichlet_phi[n] = c->objects[node_obj[n]].phi; }
    for (int i = 0; i < nnx; ++i)
        for (int j = 0; j < nny; ++j) {
            int n = node_id0(i, j, nny);
            if (is_dir[n]) continue;
            bool wall[4] = {i == 0, i == nnx - 1, j == 0, j == nny - 1};
            for (int w = 0; w < 4; ++w)
                if (wall[w] && c->wall_dirichlet[w]) { is_dir[n] = 1; g.h_dirichlet_phi[n] = c->wall_phi[w]; break; }
        }
    int nd = 0;
    for (int n = 0; n < nnodes; ++n) nd += is_dir[n];
    if (nd == 0) return pic_fail(PIC_ERR_ARG, "no Dirichlet node: the Poisson problem is singular");
    g.h_node_type2.assign(nnodes, 0);
    g.h_unknown_node.clear();
    int nu = 0;
    for (int n = 0; n < nnodes; ++n) {
        if (is_dir[n]) g.h_node_type2[n] = (node_obj[n] >= 0) ? -2 : -1;
        else { g.h_node_type2[n] = ++nu; g.h_unknown_node.push_back(n + 1); }
    }
    g.nunknowns = nu;

    // Cell volumes (Cell_Volume_zwz row 2), SetupGrids_2D_QLL.f90:196-395 for CG nodes.
    g.h_cell_volume.assign(2 * (size_t)nnodes, 0.0);
    const double sv = 1e-12;
    for (int i = 0; i < nnx; ++i)
        for (int j = 0; j < nny; ++j) {
            int n = node_id0(i, j, nny);
            double x = g.h_xs[i], y = g.h_ys[j];
            // element used for the local spacing: the first element containing the node
            int ci = std::min(i, nex - 1), cj = std::min(j, ney - 1);
            double hxp = g.h_xs[ci + 1] - g.h_xs[ci];
            double hyp = g.h_ys[cj + 1] - g.h_ys[cj];
            bool xl = std::fabs(x - c->xmin) < sv, xr = std::fabs(x - c->xmax) < sv;
            bool yb = std::fabs(y - c->ymin) < sv, yt = std::fabs(y - c->ymax) < sv;
            double v;
            if ((xl || xr) && (yb || yt)) {
                if (c->delta == 0) v = hxp * hyp / 4.0;
                else if (yb) v = 2.0 * PIC_PI_LY * (c->ymin + hyp / 4.0) * (hyp / 2.0) * hxp / 2.0;
                else v = 2.0 * PIC_PI_LY * (c->ymax - hyp / 4.0) * (hyp / 2.0) * hxp / 2.0;
            } else if (xl || xr || yb || yt) {
                if (c->delta == 0) v = hxp * hyp / 2.0;
                else if (yb) v = 2.0 * PIC_PI_LY * (c->ymin + hyp / 4.0) * (hyp / 2.0) * hxp;
                else if (yt) v = 2.0 * PIC_PI_LY * (c->ymax - hyp / 4.0) * (hyp / 2.0) * hxp;
                else v = 2.0 * PIC_PI_LY * y * hyp * hxp / 2.0;
            } else {
                if (c->delta == 0) v = hxp * hyp / 1.0;
                else v = 2.0 * PIC_PI_LY * y * hyp * hxp / 1.0;
            }
            g.h_cell_volume[2 * (size_t)n] = (double)(n + 1);
            g.h_cell_volume[2 * (size_t)n + 1] = v;
        }

    // Stiffness on unknowns, lift constant, RHS operator B.
    std::vector<std::map<int, double>> Arows(nu), Brows(nu);
    std::vector<double> cvec(nu, 0.0);
    for (int ci = 0; ci < nex; ++ci)
        for (int cj = 0; cj < ney; ++cj) {
            int nodes[4] = {node_id0(ci, cj, nny), node_id0(ci + 1, cj, nny), node_id0(ci + 1, cj + 1, nny), node_id0(ci, cj + 1, nny)};
            double hxp = g.h_xs[ci + 1] - g.h_xs[ci], hyp = g.h_ys[cj + 1] - g.h_ys[cj];
            double K[4][4], M[4][4];
            if (c->delta == 0) { q1_stiffness_cart(hxp, hyp, K); q1_mass_cart(hxp, hyp, M); }
            else q1_axi(g.h_xs[ci], g.h_ys[cj], hxp, hyp, K, M);
            // elements inside or cut by an object of eps != 1: beta = eps inside, the IFE element on
            // the cut (Generate_stiffness_matrix_local_IFE.f90:79-118, IFE_RHS_2D.f90). A conductor
            // keeps its inside nodes Dirichlet; with eps = 1 its cut elements are plain Q1
            bool ins[4];
            int kin = 0, ko = -1;
            for (int a = 0; a < 4; ++a) {
                const int k = node_obj[nodes[a]];
                ins[a] = k >= 0;
                if (k < 0) continue;
                ++kin;
                if (ko < 0) ko = k;
                else if (ko != k) return pic_fail(PIC_ERR_UNSUPPORTED, "an element touches two objects");
            }
            if (ko >= 0 && c->objects[ko].eps != 1.0) {
                const pic_object& o = c->objects[ko];
                if (kin == 4) {
                    for (int a = 0; a < 4; ++a)
                        for (int b = 0; b < 4; ++b) K[a][b] *= o.eps;
                } else {
                    double Ki[4][4], Mi[4][4];
                    const int t = ife_element(c->delta, g.h_xs[ci], g.h_ys[cj], hxp, ins, o.locations[0][0],
                                              o.locations[0][1], o.dimensions[0], o.eps, 1.0, Ki, Mi);
                    if (t == -1) return pic_fail(PIC_ERR_UNSUPPORTED, "a dielectric cuts an element through two diagonal vertices");
                    if (t == -3) return pic_fail(PIC_ERR_ARG, "singular IFE basis system");
                    if (t > 0) {                 // t == -2 (a cut within 1e-12 h of a vertex): Q1 kept
                        std::memcpy(K, Ki, sizeof(K)); std::memcpy(M, Mi, sizeof(M));
                        ++g.n_ife_elements;
                    }
                }
            }
            for (int a = 0; a < 4; ++a) {
                int u = g.h_node_type2[nodes[a]];
                if (u <= 0) continue;
                for (int b = 0; b < 4; ++b) {
                    int nb = nodes[b];
                    int ub = g.h_node_type2[nb];
                    if (ub > 0) Arows[u - 1][ub - 1] += K[a][b];
                    else cvec[u - 1] += K[a][b] * g.h_dirichlet_phi[nb];
                    Brows[u - 1][nb] += M[a][b];
                }
            }
        }
    g.h_A_rowptr.assign(nu + 1, 0); g.h_A_col.clear(); g.h_A_val.clear(); g.h_A_diag.assign(nu, 0.0);
    g.h_B_rowptr.assign(nu + 1, 0); g.h_B_col.clear(); g.h_B_val.clear();
    for (int u = 0; u < nu; ++u) {
        for (auto& kv : Arows[u]) {
            g.h_A_col.push_back(kv.first); g.h_A_val.push_back(kv.second);
            if (kv.first == u) g.h_A_diag[u] = kv.second;
        }
        g.h_A_rowptr[u + 1] = (int)g.h_A_col.size();
        for (auto& kv : Brows[u]) { g.h_B_col.push_back(kv.first); g.h_B_val.push_back(kv.second); }
        g.h_B_rowptr[u + 1] = (int)g.h_B_col.size();
    }
    g.h_c = cvec;
    return PIC_OK;
}

// ------------------------------------------------------------------ PPR patches
// Sorted unique union (Bubble_Sort_new + the duplicate-removal loop of Get_PointPatch).
static void sorted_union(std::vector<int>& acc, const std::vector<int>& add) {
    std::vector<int> t = acc;
    t.insert(t.end(), add.begin(), add.end());
    std::sort(t.begin(), t.end());
    t.erase(std::unique(t.begin(), t.end()), t.end());
    acc.swap(t);
}

// Gauss-Jordan inverse with full pivoting of a 6x6 matrix (in place). Returns false if
// singular (the reference's BRINV prints "ERR**NOT INV" and returns, PPR/BRINV.f90:22-27).
static bool inv6(double a[6][6]) {
    int is[6], js[6];
    for (int k = 0; k < 6; ++k) {
        double d = 0.0;
        for (int i = k; i < 6; ++i)
            for (int j = k; j < 6; ++j)
                if (std::fabs(a[i][j]) > d) { d = std::fabs(a[i][j]); is[k] = i; js[k] = j; }
        if (d + 1.0 == 1.0) return false;
        if (is[k] != k) for (int j = 0; j < 6; ++j) std::swap(a[k][j], a[is[k]][j]);
        if (js[k] != k) for (int i = 0; i < 6; ++i) std::swap(a[i][k], a[i][js[k]]);
        a[k][k] = 1.0 / a[k][k];
        for (int j = 0; j < 6; ++j) if (j != k) a[k][j] *= a[k][k];
        for (int i = 0; i < 6; ++i)
            if (i != k)
                for (int j = 0; j < 6; ++j)
                    if (j != k) a[i][j] -= a[i][k] * a[k][j];
        for (int i = 0; i < 6; ++i) if (i != k) a[i][k] = -a[i][k] * a[k][k];
    }
    for (int k = 5; k >= 0; --k) {
        if (js[k] != k) for (int j = 0; j < 6; ++j) std::swap(a[k][j], a[js[k]][j]);
        if (is[k] != k) for (int i = 0; i < 6; ++i) std::swap(a[i][k], a[i][is[k]]);
    }
    return true;
}

int setup_build_ppr_classes() {
    const MeshParams& m = g.mesh;
    const int nnx = m.nnx, nny = m.nny, nex = m.nex, ney = m.ney, nnodes = m.nnodes;
    const double sv = 1e-12;
    auto elem1 = [&](int ci, int cj) { return cj + ci * ney + 1; };   // 1-based element id
    // point_patch_element: elements (1-based) and type, Get_PointPatch.f90:101-231
    std::vector<std::vector<int>> pelem(nnodes);
    std::vector<int> ptype(nnodes, 0);
    for (int i = 0; i < nnx; ++i)
        for (int j = 0; j < nny; ++j) {
            int n = node_id0(i, j, nny);
            double x = g.h_xs[i], y = g.h_ys[j];
            bool xl = std::fabs(x - m.xmin) < sv, xr = std::fabs(x - m.xmax) < sv;
            bool yb = std::fabs(y - m.ymin) < sv, yt = std::fabs(y - m.ymax) < sv;
            // ParPosition_ppr_elem: ex_1 (x-b,y-b), ex_2 (x+b,y-b), ex_3 (x+b,y+b), ex_4 (x-b,y+b)
            auto ex = [&](int k) {
                int ci = (k == 1 || k == 4) ? i - 1 : i;
                int cj = (k == 1 || k == 2) ? j - 1 : j;
                ci = std::max(0, std::min(ci, nex - 1));
                cj = std::max(0, std::min(cj, ney - 1));
                return elem1(ci, cj);
            };
            std::vector<int> el;
            if (xl && yb) { el = {1}; ptype[n] = 2; }
            else if (xl && yt) { el = {elem1(0, ney - 1)}; ptype[n] = 2; }
            else if (xr && yb) { el = {elem1(nex - 1, 0)}; ptype[n] = 2; }
            else if (xr && yt) { el = {elem1(nex - 1, ney - 1)}; ptype[n] = 2; }
            else if (xl) { el = {ex(2), ex(3)}; ptype[n] = 2; }
            else if (yt) { el = {ex(1), ex(2)}; ptype[n] = 2; }
            else if (xr) { el = {ex(1), ex(4)}; ptype[n] = 2; }
            else if (yb) { el = {ex(4), ex(3)}; ptype[n] = 2; }
            else {
                int e4[4] = {ex(1), ex(2), ex(3), ex(4)};
                bool allin = true, allout = true;
                for (int k = 0; k < 4; ++k) {
                    int ei = g.h_element_index[e4[k] - 1];
                    allin = allin && (ei == -2);
                    allout = allout && (ei == -1);
                }
                if (allin || allout) { el.assign(e4, e4 + 4); ptype[n] = 1; }
                else {
                    ptype[n] = 3;
                    for (int k = 0; k < 4; ++k) if (g.h_element_index[e4[k] - 1] < 0) el.push_back(e4[k]);
                }
            }
            std::sort(el.begin(), el.end());
            el.erase(std::unique(el.begin(), el.end()), el.end());
            pelem[n] = el;
        }
    // initial patches: union of the element nodes (1-based node ids)
    std::vector<std::vector<int>> patch(nnodes), big;
    for (int n = 0; n < nnodes; ++n)
        for (int e1 : pelem[n]) {
            int e = e1 - 1, ci = e / ney, cj = e % ney;
            std::vector<int> ht = {node_id0(ci, cj, nny) + 1, node_id0(ci + 1, cj, nny) + 1,
                                   node_id0(ci + 1, cj + 1, nny) + 1, node_id0(ci, cj + 1, nny) + 1};
            sorted_union(patch[n], ht);
        }
    big = patch;   // tmp_big_patch
    // nodes with fewer than 6 samples: union of the initial patches of their patch nodes
    for (int n = 0; n < nnodes; ++n) {
        if (patch[n].size() >= 6 || patch[n].empty()) continue;
        std::vector<int> t;
        for (int k : big[n]) sorted_union(t, big[k - 1]);
        patch[n] = t;
    }
    // boundary / interface nodes with fewer than 12 samples: same union (Get_PointPatch.f90:447-500)
    for (int n = 0; n < nnodes; ++n) {
        if (!(ptype[n] == 2 || ptype[n] == 3)) continue;
        if (patch[n].size() >= 12) continue;
        std::vector<int> t;
        for (int k : big[n]) sorted_union(t, big[k - 1]);
        if (!t.empty()) patch[n] = t;
    }
    g.h_point_patch.assign((size_t)nnodes * PIC_PATCH_MAX, 0);
    for (int n = 0; n < nnodes; ++n) {
        int cnt = (int)std::min<size_t>(patch[n].size(), PIC_PATCH_MAX - 1);
        g.h_point_patch[(size_t)n] = cnt;                                  // point_patch(n,1)
        for (int k = 0; k < cnt; ++k) g.h_point_patch[(size_t)(k + 1) * nnodes + n] = patch[n][k];
    }

    // recovery rows per node, grouped into classes of identical (offsets, h)
    std::map<std::vector<long long>, int> key2class;
    g.h_ppr_classes.clear();
    g.h_ppr_class.assign(nnodes, -1);
    for (int n = 0; n < nnodes; ++n) {
        int cnt = (int)patch[n].size();
        if (cnt == 0 || cnt >= PIC_PATCH_MAX) continue;
        int i = n / nny, j = n % nny;
        double xi = g.h_xs[i], yi = g.h_ys[j];
        double h = 0.0;
        std::vector<long long> key;
        key.push_back(cnt);
        for (int k = 0; k < cnt; ++k) {
            int p = patch[n][k] - 1, pi = p / nny, pj = p % nny;
            double ht = std::max(std::fabs(g.h_xs[pi] - xi), std::fabs(g.h_ys[pj] - yi));
            if (ht > h) h = ht;
            key.push_back(pi - i); key.push_back(pj - j);
        }
        long long hb; std::memcpy(&hb, &h, 8); key.push_back(hb);
        auto it = key2class.find(key);
        if (it != key2class.end()) { g.h_ppr_class[n] = it->second; continue; }
        PprClass pc; std::memset(&pc, 0, sizeof(pc));
        pc.n = cnt; pc.h = h;
        double A[PIC_PATCH_MAX][6];
        for (int k = 0; k < cnt; ++k) {
            int p = patch[n][k] - 1, pi = p / nny, pj = p % nny;
            pc.di[k] = pi - i; pc.dj[k] = pj - j;
            double z = (g.h_xs[pi] - xi) / h, e = (g.h_ys[pj] - yi) / h;
            A[k][0] = 1.0; A[k][1] = z; A[k][2] = e; A[k][3] = z * z; A[k][4] = z * e; A[k][5] = e * e;
        }
        double AtA[6][6];
        for (int a = 0; a < 6; ++a)
            for (int b = 0; b < 6; ++b) {
                double s = 0.0;
                for (int k = 0; k < cnt; ++k) s += A[k][a] * A[k][b];
                AtA[a][b] = s;
            }
        int cls = (int)g.h_ppr_classes.size();
        if (!inv6(AtA)) {
            // singular patch: recovery left at zero like an un-invertible BRINV result would be unusable
            key2class[key] = -1;
            continue;
        }
        for (int a = 0; a < 6; ++a)
            for (int k = 0; k < cnt; ++k) {
                double s = 0.0;
                for (int b = 0; b < 6; ++b) s += AtA[a][b] * A[k][b];
                pc.R[a][k] = s;
            }
        g.h_ppr_classes.push_back(pc);
        key2class[key] = cls;
        g.h_ppr_class[n] = cls;
    }
    // nodes whose coefficients feed the interface-element gather
    g.h_coef_needed.assign(nnodes, 0);
    for (int e = 0; e < m.nelem; ++e) {
        if (g.h_element_index[e] <= 0) continue;
        int ci = e / ney, cj = e % ney;
        int nodes[4] = {node_id0(ci, cj, nny), node_id0(ci + 1, cj, nny), node_id0(ci + 1, cj + 1, nny), node_id0(ci, cj + 1, nny)};
        for (int a = 0; a < 4; ++a) g.h_coef_needed[nodes[a]] = 1;
    }
    return PIC_OK;
}

// ------------------------------------------------------------------ MCC tables
// Analytic argon cross sections (m^2, E in eV): smooth forms chosen for the synthetic
// workloads, tabulated exactly like SigmaNormalizedUpdate + ProbilityNonReactive.
static double sig_e_elastic(double E) { return 5.0e-21 + 1.5e-19 * std::sqrt(E / (E + 1.0)) * 12.0 / (E + 12.0); }
static double sig_e_excite(double E) { const double t = 11.55; return E > t ? 1.2e-17 * std::log(E / t) / (E * t) : 0.0; }
static double sig_e_ionize(double E) { const double t = 15.76; return E > t ? 2.7e-17 * std::log(E / t) / (E * t) : 0.0; }
static double sig_i_isotropic(double E) { return 1.5e-19 + 3.5e-19 / std::sqrt(1.0 + E); }
static double sig_i_cx(double E) { double s = 7.5e-10 - 0.4e-10 * std::log(E); return s * s; }

void setup_analytic_tables(int isp, bool electron, const pic_gas_params& gas, double mass_kg, double dt_SI) {
    SpeciesHost& S = g.sp[isp];
    const int Ns = 10001;
    double emin = electron ? -3.0 : -1.0, emax = 3.0;
    double eint = (emax - emin) / (double)(Ns - 1);
    int nr = electron ? 3 : 2;
    S.nreact = nr; S.nsigma = Ns; S.emin = emin; S.eint = eint; S.emax = std::pow(10.0, emax);
    S.rtype_host = electron ? std::vector<int>{11, 12, 13} : std::vector<int>{101, 103};
    S.thr_host = electron ? std::vector<double>{0.0, 11.55, 15.76} : std::vector<double>{0.0, 0.0};
    S.prob_host.assign((size_t)nr * Ns, 0.0);
    double miu = mass_kg * gas.mass_kg / (mass_kg + gas.mass_kg);
    for (int i = 0; i < Ns; ++i) {
        double E = std::pow(10.0, emin + (double)i * eint);
        double V = std::sqrt(2.0 * E * PIC_JTOEV / miu);
        double sig[3];
        if (electron) { sig[0] = sig_e_elastic(E); sig[1] = sig_e_excite(E); sig[2] = sig_e_ionize(E); }
        else { sig[0] = sig_i_isotropic(E); sig[1] = sig_i_cx(E); }
        for (int r = 0; r < nr; ++r) S.prob_host[(size_t)i * nr + r] = sig[r] * V * gas.density_m3;
    }
    for (int i = 0; i < Ns; ++i)
        for (int r = 1; r < nr; ++r) S.prob_host[(size_t)i * nr + r] += S.prob_host[(size_t)i * nr + r - 1];
    double mx = 0.0;
    for (double v : S.prob_host) mx = std::max(mx, v);
    for (double& v : S.prob_host) v = v / mx;
    S.sigma_max = mx;
    S.collision_ratio = 1.0 - std::exp(-mx * dt_SI);
}

// ------------------------------------------------------------------ particle loader
// The other initial velocity laws of AllInitializationParticleBundleIFE (MaxKappa,
// MCCInterface.f90:168-176), on the stream (seed; slot, 0, PIC_RNG_LOADV | species).
// kind 0, VelocityKappaInitializationParticleOne (ParticleOne.f90:154-204): kappa = 2, Vx by
// rejection on s (20 R - 10) against (1 + Vx^2 / ((2 kappa - 3) s^2))^-kappa, s^2 =
// Sigma / Sigma_e with Sigma = ((kappa - 1.5)/kappa) 2 kB T / M and Sigma_e the same for an
// electron at 11605 K; Vy = Vz = 0. kind 2, VelocityPolytropicInitializationParticleOne
// (:206-245, gama = 3): Vx = s 2 sqrt(3) (R - 0.5), s^2 = (2 kB T / M) / (2 kB 11605 / Me);
// Vy = Vz = 0. Both store Vx as drawn (the reference's rescale to v_ref is commented out).
static void load_velocity_law(int kind, uint64_t seed, uint32_t slot, uint32_t tag, double T, double M,
                              double* vx, double* vy, double* vz) {
    const double kB = 1.3807e-23, Te = 11605.0, Me = 9.1095e-31;
    pic_stream s; pic_stream_init(&s, seed, slot, 0u, tag);
    double V = 0.0;
    if (kind == 0) {
        const double Kappa = 2.0;
        const double Sigma = ((Kappa - 1.5) / Kappa) * (2.0 * kB * T / M);
        const double Sigma_e = ((Kappa - 1.5) / Kappa) * (2.0 * kB * Te / Me);
        for (int it = 0; it < 100000; ++it) {
            V = std::sqrt(Sigma / Sigma_e) * (20.0 * pic_stream_u01(&s) - 10.0);
            const double P = std::pow(1.0 + (V * V) / ((2 * Kappa - 3) * Sigma / Sigma_e), -Kappa);
            if (pic_stream_u01(&s) <= P) break;
        }
    } else {
        const double gama = 3.0;
        const double Sigma = 2.0 * kB * T / M, Sigma_e = 2.0 * kB * Te / Me;
        const double V_max = std::sqrt(gama);
        V = std::sqrt(Sigma / Sigma_e) * 2.0 * V_max * (pic_stream_u01(&s) - 0.5);
    }
    *vx = V; *vy = 0.0; *vz = 0.0;
}

// ppc particles per vacuum element (r-z: ppc on average, in proportion to the element's
// volume), elements in sort order; uniform position inside the element,
// Maxwellian velocity (three Box-Muller normals) at temp_eV (kind 1), or the kappa (kind 0)
// or polytropic (kind 2) law above. Particles falling inside an object are dropped (the
// reference's AdjustBoundary_2D pre-filter, Main_IFE_Test_2.f90:218).
void setup_maxwellian(int isp, int ppc, double temp_eV, uint64_t seed,
                      std::vector<double>& x, std::vector<double>& y,
                      std::vector<double>& vx, std::vector<double>& vy, std::vector<double>& vz,
                      std::vector<int>& cell, std::vector<double>* th, int kind) {
    const MeshParams& m = g.mesh;
    const SpeciesHost& S = g.sp[isp];
    double vth = std::sqrt(temp_eV / g.cfg.T_ref_eV / S.xm);   // normalised thermal speed
    x.clear(); y.clear(); vx.clear(); vy.clear(); vz.clear(); cell.clear();
    if (th) th->clear();
    size_t cap = (size_t)m.nelem * ppc;
    x.reserve(cap); y.reserve(cap); vx.reserve(cap); vy.reserve(cap); vz.reserve(cap); cell.reserve(cap);
    const uint32_t tag = pic_stream_tag(isp, PIC_RNG_LOAD);
    // elements in the tile-major order of the cell sort, so the loaded slots start sorted
    int nty = (m.ney + SORT_TILE - 1) / SORT_TILE;
    std::vector<int> order(m.nelem);
    for (int e = 0; e < m.nelem; ++e) order[e] = e;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return pic_sort_key(a, m.ney, nty) < pic_sort_key(b, m.ney, nty);
    });
    // axisymmetric: an element's count follows its volume share (ppc on average), so the
    // density is uniform as with the reference's volume-uniform draw over the domain
    const double vol_all = m.ymax * m.ymax - m.ymin * m.ymin;
    const int stride = m.delta == 1 ? 2 * ppc + 1 : ppc;
    for (int e : order) {
        if (g.h_element_index[e] == -2) continue;
        int ci = e / m.ney, cj = e % m.ney;
        double x0 = g.h_xs[ci], x1 = g.h_xs[ci + 1], y0 = g.h_ys[cj], y1 = g.h_ys[cj + 1];
        int count = ppc;
        if (m.delta == 1) {
            const double q = (double)ppc * m.ney * (y1 * y1 - y0 * y0) / vol_all;
            count = std::min(stride, (int)(std::floor((ci + 1) * q) - std::floor(ci * q)));
        }
        for (int k = 0; k < count; ++k) {
            uint32_t slot = (uint32_t)((uint64_t)e * stride + k);
            pic_stream s; pic_stream_init(&s, seed, slot, 0u, tag);
            double px = x0 + (x1 - x0) * pic_stream_u01(&s);
            double uy = pic_stream_u01(&s);
            // axisymmetric: uniform in r^2 over the cell (ParticleOne.f90:97-101)
            double py = m.delta == 1 ? std::sqrt(y0 * y0 + (y1 + y0) * (y1 - y0) * uy) : y0 + (y1 - y0) * uy;
            double u1 = pic_stream_u01(&s), u2 = pic_stream_u01(&s);
            double u3 = pic_stream_u01(&s), u4 = pic_stream_u01(&s);
            double r1 = std::sqrt(-2.0 * std::log(u1)), r2 = std::sqrt(-2.0 * std::log(u3));
            double a = 2.0 * PIC_PI * u2, b = 2.0 * PIC_PI * u4;
            bool inside = false;
            for (int o = 0; o < g.cfg.n_objects; ++o) inside = inside || inside_object(g.cfg.objects[o], px, py);
            if (inside) continue;
            x.push_back(px); y.push_back(py);
            if (th) th->push_back(2.0 * PIC_PI * pic_stream_u01(&s));
            if (kind == 1) {
                vx.push_back(vth * r1 * std::cos(a)); vy.push_back(vth * r1 * std::sin(a)); vz.push_back(vth * r2 * std::cos(b));
            } else {
                double u, v, w;
                load_velocity_law(kind, seed, slot, pic_stream_tag(isp, PIC_RNG_LOADV), temp_eV * PIC_EVTOK, S.p.mass_kg,
                                  &u, &v, &w);
                vx.push_back(u); vy.push_back(v); vz.push_back(w);
            }
            cell.push_back(e);
        }
    }
}
