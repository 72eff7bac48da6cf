// ORACLE — TEST INFRASTRUCTURE ONLY.
//
// CPU FP64 restatement of the reference RKDG algorithm (AsterJoe/RKDG_heter_calculate,
// multi_gpuRkdg/).  Only the tests, the smoke check and the CPU-baseline leg of the
// benchmark may load this library; the product path never links or calls it.
//
// What is restated, and from where (paths relative to multi_gpuRkdg/):
//   * per-element geometry, 7-point volume / 3-point edge quadrature,
//     per-element Gram-Schmidt basis, diagonal mass, basis tables
//       src/triangleinfo.cpp:77-121 (evaluation), 153-278 (initInformation),
//       281-430 (13-point moments), 433-602 (Gram-Schmidt systems), 605-624 (mass)
//   * partial-pivot Gauss elimination               src/gauss.cpp:3-107
//   * the per-stage kernel chain                    src/kernelfunctions.cu:84-748
//       conserved values at points 84-156, boundary ghosts 158-275,
//       volume integral 278-351, LF coefficient 354-390, edge F/G 392-433,
//       LF flux 510-582, edge integral 584-637, SSP-RK3 stages 640-748
//   * CFL time step with the reduction's dropped subset   src/kernelfunctions.cu:3-81
//   * residual (max change of the cell means)              src/kernelfunctions.cu:751-867
//   * free-stream initial state                            src/cudarkdgsolver.cu:326-368
//   * far-field state used by the boundary kernel          src/rkdgadvance.cpp:31-34
//
// Extensions that have no reference counterpart (and are pinned by the tests through
// invariants and closed forms, not by the reference):
//   * P1 = first three modes of the P2 construction; P3 = the same Gram-Schmidt continued
//     to cubic monomials, 12-point Dunavant volume rule, 4-point Gauss edge rule;
//   * Roe flux (Roe 1981) with the Harten-Hyman entropy fix on the acoustic waves;
//   * TVB minmod limiter (Cockburn & Shu 1998, triangles), component-wise, followed by the
//     positivity-preserving scaling of Zhang & Shu (2010);
//   * boundary handling per edge (the reference creates one ghost per element);
//   * L2 projection of the isentropic vortex and of the four-quadrant Riemann data.
//
// Build: oracle/Makefile (-O3 -fopenmp -ffp-contract=off so the operation order of the
// restated expressions is the one written).

#include <cmath>
#include <cstdint>
#include <cstring>
#include <cstdio>
#include <vector>
#include <algorithm>
#include <omp.h>

namespace {

constexpr int MAXNP = 10;
constexpr int MAXQV = 12;
constexpr int MAXQE = 4;
constexpr int NVAR = 4;

enum { BC_INTERIOR = 0, BC_FARFIELD = 1, BC_OUTFLOW = 2, BC_REFLECTION = 3 };
enum { FLUX_LF = 0, FLUX_ROE = 1 };
enum { DT_REFERENCE = 0, DT_MAX = 1, DT_FIXED = 2 };

inline int np_of(int P) { return (P + 1) * (P + 2) / 2; }
inline int nqv_of(int P) { return P <= 2 ? 7 : 12; }
inline int nqe_of(int P) { return P <= 2 ? 3 : 4; }

// ---------------------------------------------------------------------------
// Gauss elimination with partial (column) pivoting; src/gauss.cpp:3-107.
// The first row holding the largest |a[i][k]| is chosen (strict '>').
void gauss_solve(int n, const double* A, double* B) {
    double a[MAXNP][MAXNP], b[MAXNP], x[MAXNP], m[MAXNP];
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < n; ++j) a[i][j] = A[i * n + j];
        b[i] = B[i];
    }
    for (int k = 0; k < n - 1; ++k) {
        int piv = k;
        double major = a[k][k];
        for (int i = k + 1; i < n; ++i)
            if (std::fabs(a[i][k]) > std::fabs(major)) { piv = i; major = a[i][k]; }
        for (int j = k; j < n; ++j) std::swap(a[k][j], a[piv][j]);
        std::swap(b[k], b[piv]);
        for (int i = k + 1; i < n; ++i) {
            m[i] = a[i][k] / a[k][k];
            for (int j = k + 1; j < n; ++j) a[i][j] = a[i][j] - m[i] * a[k][j];
            b[i] = b[i] - m[i] * b[k];
        }
    }
    x[n - 1] = b[n - 1] / a[n - 1][n - 1];
    for (int r = n - 2; r >= 0; --r) {
        double sum = 0.0;
        for (int j = r + 1; j < n; ++j) sum = sum + a[r][j] * x[j];
        x[r] = (b[r] - sum) / a[r][r];
    }
    for (int i = 0; i < n; ++i) B[i] = x[i];
}

// ---------------------------------------------------------------------------
// Monomials.  Gram-Schmidt order m_k: 1, X, Y, X^2, XY, Y^2, X^3, X^2Y, XY^2, Y^3.
// "POLY" storage order of a basis function's coefficients (highest degree first,
// then rising power of Y; the reference's POLYA..POLYF are X^2, XY, Y^2, X, Y, 1):
//   X^3, X^2Y, XY^2, Y^3, X^2, XY, Y^2, X, Y, 1
const int GS_PX[10] = {0, 1, 0, 2, 1, 0, 3, 2, 1, 0};
const int GS_PY[10] = {0, 0, 1, 0, 1, 2, 0, 1, 2, 3};
const int POLY_PX[10] = {3, 2, 1, 0, 2, 1, 0, 1, 0, 0};
const int POLY_PY[10] = {0, 1, 2, 3, 0, 1, 2, 0, 1, 0};

inline double mono(double x, double y, int a, int b) {
    // evaluated as x*x*...*y*y from the left (the reference's Hanshu1..14)
    double r = 1.0;
    bool first = true;
    for (int i = 0; i < a; ++i) { r = first ? x : r * x; first = false; }
    for (int i = 0; i < b; ++i) { r = first ? y : r * y; first = false; }
    return r;
}

struct ElemInfo {
    int P, Np, Nqv, Nqe;
    double area, perimeter, bx, by;
    double normal[6];
    double len[3];
    double coeff[MAXNP][MAXNP];  // physical monomial coefficients about the barycenter
    double mass[MAXNP];
    double vol_w[MAXQV];
    double edge_w[3 * MAXQE];
    double vol_x[MAXQV], vol_y[MAXQV];
    double edge_x[3 * MAXQE], edge_y[3 * MAXQE];
    double vol_bf[MAXNP][MAXQV];
    double vol_bdf[2 * MAXNP][MAXQV];
    double edge_bf[MAXNP][3 * MAXQE];
};

// Physical monomial layout of coeff[i][*]: 1, dx, dy, dx^2, dxdy, dy^2, dx^3, dx^2dy, dxdy^2, dy^3
double basis_value(const ElemInfo& I, int i, double x, double y) {
    const double* c = I.coeff[i];
    double dx = x - I.bx, dy = y - I.by;
    double r = c[0] + c[1] * dx + c[2] * dy + c[3] * dx * dx + c[4] * dx * dy + c[5] * dy * dy;
    if (I.P >= 3)
        r = r + c[6] * dx * dx * dx + c[7] * dx * dx * dy + c[8] * dx * dy * dy + c[9] * dy * dy * dy;
    return r;
}

double basis_deriv(const ElemInfo& I, int i, int d, double x, double y) {
    const double* c = I.coeff[i];
    double dx = x - I.bx, dy = y - I.by;
    double r;
    if (d == 0) {
        r = c[1] + 2 * c[3] * dx + c[4] * dy;
        if (I.P >= 3) r = r + 3 * c[6] * dx * dx + 2 * c[7] * dx * dy + c[8] * dy * dy;
    } else {
        r = c[2] + c[4] * dx + 2 * c[5] * dy;
        if (I.P >= 3) r = r + c[7] * dx * dx + 2 * c[8] * dx * dy + 3 * c[9] * dy * dy;
    }
    return r;
}

// 13-point moments of X^a Y^b (scaled coordinates about the barycenter);
// src/triangleinfo.cpp:281-430.  The moment of the constant is taken as exactly 1.
struct Moments {
    double s[7][7];  // s[a][b], a+b <= 6
    double get(int a, int b) const { return (a == 0 && b == 0) ? 1.0 : s[a][b]; }
};

void compute_moments(const double* x, const double* y, double area, double bx, double by, Moments& M) {
    const double A = 0.065130102902216, B = 0.869739794195568, C = 0.312865496004875,
                 D = 0.638444188569809, Ec = 0.048690315425316, F = 0.260345966079038,
                 G = 0.479308067841923;
    const double w1 = -0.149570044467670, w2 = 0.053347235608839, w3 = 0.175615257433204,
                 w4 = 0.077113760890257;
    double ss = 1.0 / (std::sqrt(area));
    double n[13][2];
    n[0][0] = 0; n[0][1] = 0;
    n[1][0] = (B * x[0] + A * (x[1] + x[2]) - bx) * ss;  n[1][1] = (B * y[0] + A * (y[1] + y[2]) - by) * ss;
    n[2][0] = (B * x[1] + A * (x[0] + x[2]) - bx) * ss;  n[2][1] = (B * y[1] + A * (y[0] + y[2]) - by) * ss;
    n[3][0] = (B * x[2] + A * (x[0] + x[1]) - bx) * ss;  n[3][1] = (B * y[2] + A * (y[0] + y[1]) - by) * ss;
    n[4][0] = (G * x[0] + F * (x[1] + x[2]) - bx) * ss;  n[4][1] = (G * y[0] + F * (y[1] + y[2]) - by) * ss;
    n[5][0] = (G * x[1] + F * (x[0] + x[2]) - bx) * ss;  n[5][1] = (G * y[1] + F * (y[0] + y[2]) - by) * ss;
    n[6][0] = (G * x[2] + F * (x[0] + x[1]) - bx) * ss;  n[6][1] = (G * y[2] + F * (y[0] + y[1]) - by) * ss;
    n[7][0] = (C * x[0] + D * x[1] + Ec * x[2] - bx) * ss;  n[7][1] = (C * y[0] + D * y[1] + Ec * y[2] - by) * ss;
    n[8][0] = (D * x[0] + C * x[1] + Ec * x[2] - bx) * ss;  n[8][1] = (D * y[0] + C * y[1] + Ec * y[2] - by) * ss;
    n[9][0] = (D * x[0] + Ec * x[1] + C * x[2] - bx) * ss;  n[9][1] = (D * y[0] + Ec * y[1] + C * y[2] - by) * ss;
    n[10][0] = (Ec * x[0] + D * x[1] + C * x[2] - bx) * ss; n[10][1] = (Ec * y[0] + D * y[1] + C * y[2] - by) * ss;
    n[11][0] = (Ec * x[0] + C * x[1] + D * x[2] - bx) * ss; n[11][1] = (Ec * y[0] + C * y[1] + D * y[2] - by) * ss;
    n[12][0] = (C * x[0] + Ec * x[1] + D * x[2] - bx) * ss; n[12][1] = (C * y[0] + Ec * y[1] + D * y[2] - by) * ss;
    for (int a = 0; a <= 6; ++a)
        for (int b = 0; a + b <= 6; ++b) {
            auto H = [&](int q) { return mono(n[q][0], n[q][1], a, b); };
            M.s[a][b] = w1 * H(0) + w2 * (H(1) + H(2) + H(3)) + w3 * (H(4) + H(5) + H(6)) +
                        w4 * (H(7) + H(8) + H(9) + H(10) + H(11) + H(12));
        }
}

// Inner product of basis function phi (POLY-ordered coefficients) with monomial X^a Y^b.
double inner_poly_mono(const double* poly, const Moments& M, int a, int b) {
    double r = 0.0;
    bool first = true;
    for (int t = 0; t < 10; ++t) {
        if (poly[t] == 0.0) continue;  // absent terms contribute exactly nothing
        double term = poly[t] == 1.0 ? M.get(POLY_PX[t] + a, POLY_PY[t] + b)
                                     : poly[t] * M.get(POLY_PX[t] + a, POLY_PY[t] + b);
        r = first ? term : r + term;
        first = false;
    }
    return r;
}

int poly_slot(int k) {  // POLY index of Gram-Schmidt monomial k
    for (int t = 0; t < 10; ++t)
        if (POLY_PX[t] == GS_PX[k] && POLY_PY[t] == GS_PY[k]) return t;
    return -1;
}

// Gram-Schmidt basis by successive pivoted solves; src/triangleinfo.cpp:433-602.
// phi_k = m_k + sum_{j<k} c_j m_j, with <phi_l, phi_k> = 0 for l < k.
// Unknowns are ordered as in POLY; equations follow phi_0, phi_1, ...
void solve_basis(int Np, const Moments& M, double poly[MAXNP][10]) {
    for (int k = 0; k < Np; ++k)
        for (int t = 0; t < 10; ++t) poly[k][t] = 0.0;
    poly[0][poly_slot(0)] = 1.0;
    if (Np > 1) poly[1][poly_slot(1)] = 1.0;
    for (int k = 2; k < Np; ++k) {
        int unk[MAXNP], nu = 0;
        for (int t = 0; t < 10; ++t)
            for (int j = 0; j < k; ++j)
                if (poly_slot(j) == t) unk[nu++] = t;
        double A[MAXNP * MAXNP], B[MAXNP];
        for (int l = 0; l < k; ++l) {
            for (int u = 0; u < nu; ++u) {
                int a = POLY_PX[unk[u]], b = POLY_PY[unk[u]];
                double v = inner_poly_mono(poly[l], M, a, b);
                // src/triangleinfo.cpp:482 pairs the constant part of phi_2 with the
                // X moment (not the Y moment) in the phi_3 system, row phi_2, column Y.
                if (k == 3 && l == 2 && a == 0 && b == 1)
                    v = poly[2][7] * M.get(1, 1) + poly[2][8] * M.get(0, 2) + poly[2][9] * M.get(1, 0);
                A[l * nu + u] = v;
            }
            B[l] = -inner_poly_mono(poly[l], M, GS_PX[k], GS_PY[k]);
        }
        gauss_solve(nu, A, B);
        poly[k][poly_slot(k)] = 1.0;
        for (int u = 0; u < nu; ++u) poly[k][unk[u]] = B[u];
    }
}

// P3 (no reference counterpart): the same Gram-Schmidt, with exact moments and the
// solves in long double; the monomial Gram matrix up to degree 3 is too ill-conditioned
// for the 15-digit 13-point rule in double (orthogonality would only hold to ~1e-9).
void solve_basis_exact(int Np, const double* x, const double* y, double area, double bx, double by,
                       double poly[MAXNP][10]) {
    // exact mean of X^a Y^b over the triangle: collapsed 10x10 Gauss rule in long double
    const int n = 10;
    long double t[n], w[n];
    for (int i = 0; i < n; ++i) {
        long double z = std::cos(M_PI * (i + 0.75L) / (n + 0.5L)), pp = 0;
        for (int it = 0; it < 100; ++it) {
            long double p1 = 1, p2 = 0;
            for (int j = 1; j <= n; ++j) { long double p3 = p2; p2 = p1; p1 = ((2 * j - 1) * z * p2 - (j - 1) * p3) / j; }
            pp = n * (z * p1 - p2) / (z * z - 1);
            long double dz = p1 / pp;
            z -= dz;
            if (std::fabs((double)dz) < 1e-19) break;
        }
        t[i] = 0.5L * (1 - z);
        w[i] = 1.0L / ((1 - z * z) * pp * pp);
    }
    long double S[7][7] = {};
    long double isq = 1.0L / std::sqrt((long double)area);
    for (int a = 0; a < n; ++a)
        for (int b = 0; b < n; ++b) {
            long double xi = t[a], eta = t[b] * (1 - t[a]), ww = w[a] * w[b] * (1 - t[a]) * 2;
            long double px = x[0] + ((long double)x[1] - x[0]) * xi + ((long double)x[2] - x[0]) * eta;
            long double py = y[0] + ((long double)y[1] - y[0]) * xi + ((long double)y[2] - y[0]) * eta;
            long double X = (px - bx) * isq, Y = (py - by) * isq;
            long double xp[7], yp[7];
            xp[0] = yp[0] = 1;
            for (int k = 1; k < 7; ++k) { xp[k] = xp[k - 1] * X; yp[k] = yp[k - 1] * Y; }
            for (int i = 0; i <= 6; ++i)
                for (int j = 0; i + j <= 6; ++j) S[i][j] += ww * xp[i] * yp[j];
        }
    long double pl[MAXNP][10] = {};
    for (int k = 0; k < Np; ++k) {
        // phi_k = m_k - sum_l <m_k, phi_l>/<phi_l, phi_l> phi_l, coefficients over POLY slots
        long double c[10] = {};
        c[poly_slot(k)] = 1;
        for (int l = 0; l < k; ++l) {
            long double ip = 0, nn = 0;
            for (int t1 = 0; t1 < 10; ++t1) {
                if (pl[l][t1] == 0) continue;
                ip += pl[l][t1] * S[POLY_PX[t1] + GS_PX[k]][POLY_PY[t1] + GS_PY[k]];
                for (int t2 = 0; t2 < 10; ++t2)
                    if (pl[l][t2] != 0) nn += pl[l][t1] * pl[l][t2] * S[POLY_PX[t1] + POLY_PX[t2]][POLY_PY[t1] + POLY_PY[t2]];
            }
            long double f = ip / nn;
            for (int t1 = 0; t1 < 10; ++t1) c[t1] -= f * pl[l][t1];
        }
        for (int t1 = 0; t1 < 10; ++t1) pl[k][t1] = c[t1];
    }
    for (int k = 0; k < Np; ++k)
        for (int t1 = 0; t1 < 10; ++t1) poly[k][t1] = (double)pl[k][t1];
}

void element_info(int P, const double* x, const double* y, ElemInfo& I) {
    I.P = P; I.Np = np_of(P); I.Nqv = nqv_of(P); I.Nqe = nqe_of(P);
    const int Np = I.Np, Nqv = I.Nqv, Nqe = I.Nqe;
    double l[3];
    I.area = 0.5 * std::fabs((x[2] - x[0]) * (y[1] - y[0]) - (y[2] - y[0]) * (x[1] - x[0]));
    l[0] = std::sqrt((x[1] - x[2]) * (x[1] - x[2]) + (y[1] - y[2]) * (y[1] - y[2]));
    l[1] = std::sqrt((x[0] - x[2]) * (x[0] - x[2]) + (y[0] - y[2]) * (y[0] - y[2]));
    l[2] = std::sqrt((x[0] - x[1]) * (x[0] - x[1]) + (y[0] - y[1]) * (y[0] - y[1]));
    for (int k = 0; k < 3; ++k) I.len[k] = l[k];
    I.perimeter = l[0] + l[1] + l[2];
    I.bx = (x[0] + x[1] + x[2]) / 3.0;
    I.by = (y[0] + y[1] + y[2]) / 3.0;
    I.normal[0] = (y[2] - y[1]) / l[0]; I.normal[1] = (x[1] - x[2]) / l[0];
    I.normal[2] = (y[0] - y[2]) / l[1]; I.normal[3] = (x[2] - x[0]) / l[1];
    I.normal[4] = (y[1] - y[0]) / l[2]; I.normal[5] = (x[0] - x[1]) / l[2];

    // volume quadrature
    if (P <= 2) {
        const double w1 = 0.797426985353087, w2 = 0.101286507323456, w3 = 0.059715871789770,
                     w4 = 0.470142064105115;
        I.vol_x[0] = (x[0] + x[1] + x[2]) / 3.0; I.vol_y[0] = (y[0] + y[1] + y[2]) / 3.0;
        I.vol_x[1] = w1 * x[0] + w2 * (x[1] + x[2]); I.vol_y[1] = w1 * y[0] + w2 * (y[1] + y[2]);
        I.vol_x[2] = w1 * x[1] + w2 * (x[0] + x[2]); I.vol_y[2] = w1 * y[1] + w2 * (y[0] + y[2]);
        I.vol_x[3] = w1 * x[2] + w2 * (x[0] + x[1]); I.vol_y[3] = w1 * y[2] + w2 * (y[0] + y[1]);
        I.vol_x[4] = w3 * x[0] + w4 * (x[1] + x[2]); I.vol_y[4] = w3 * y[0] + w4 * (y[1] + y[2]);
        I.vol_x[5] = w3 * x[1] + w4 * (x[0] + x[2]); I.vol_y[5] = w3 * y[1] + w4 * (y[0] + y[2]);
        I.vol_x[6] = w3 * x[2] + w4 * (x[0] + x[1]); I.vol_y[6] = w3 * y[2] + w4 * (y[0] + y[1]);
        const double wq[7] = {0.225000000000000, 0.125939180544827, 0.125939180544827, 0.125939180544827,
                              0.132394152788506, 0.132394152788506, 0.132394152788506};
        for (int q = 0; q < 7; ++q) I.vol_w[q] = I.area * wq[q];
    } else {
        // Dunavant 12-point rule, exact to degree 6 (no reference counterpart)
        const double a1 = 0.063089014491502228340331602870819, b1 = 0.873821971016995543319336794258362;
        const double a2 = 0.249286745170910421291638553107019, b2 = 0.501426509658179157416722893785962;
        const double c1 = 0.053145049844816947353249671631398, c2 = 0.310352451033784405416607733956552,
                     c3 = 0.636502499121398647230142594411050;
        const double wa = 0.050844906370206816920936809106869, wb = 0.116786275726379366030690538687593,
                     wc = 0.082851075618373575193553456420442;
        double bc[12][3] = {{b1, a1, a1}, {a1, b1, a1}, {a1, a1, b1}, {b2, a2, a2}, {a2, b2, a2}, {a2, a2, b2},
                            {c1, c2, c3}, {c1, c3, c2}, {c2, c1, c3}, {c2, c3, c1}, {c3, c1, c2}, {c3, c2, c1}};
        double wq[12] = {wa, wa, wa, wb, wb, wb, wc, wc, wc, wc, wc, wc};
        for (int q = 0; q < 12; ++q) {
            I.vol_x[q] = bc[q][0] * x[0] + bc[q][1] * x[1] + bc[q][2] * x[2];
            I.vol_y[q] = bc[q][0] * y[0] + bc[q][1] * y[1] + bc[q][2] * y[2];
            I.vol_w[q] = I.area * wq[q];
        }
    }
    // edge quadrature: edge k runs from vertex k+1 to vertex k+2
    double sa[MAXQE], sb[MAXQE], we[MAXQE];
    if (P <= 2) {
        double cp = 0.5 + std::sqrt(15.0) / 10.0;
        sa[0] = cp; sb[0] = 1 - cp; sa[1] = 0.5; sb[1] = 0.5; sa[2] = 1 - cp; sb[2] = cp;
        we[0] = 0.55555555555556; we[1] = 0.88888888888889; we[2] = 0.55555555555556;
    } else {
        const double t0 = 0.861136311594052575223946488892809, t1 = 0.339981043584856264802665759103245;
        const double v0 = 0.347854845137453857373063949221999, v1 = 0.652145154862546142626936050778001;
        double tt[4] = {t0, t1, -t1, -t0};
        double ww[4] = {v0, v1, v1, v0};
        for (int j = 0; j < 4; ++j) { sa[j] = 0.5 * (1.0 + tt[j]); sb[j] = 0.5 * (1.0 - tt[j]); we[j] = ww[j]; }
    }
    for (int k = 0; k < 3; ++k) {
        int a = (k + 1) % 3, b = (k + 2) % 3;
        for (int j = 0; j < Nqe; ++j) {
            I.edge_x[k * Nqe + j] = sa[j] * x[a] + sb[j] * x[b];
            I.edge_y[k * Nqe + j] = sa[j] * y[a] + sb[j] * y[b];
            I.edge_w[k * Nqe + j] = we[j] * 0.5 * l[k];
        }
    }
    // basis: P1 and P2 from the P2 construction (P1 keeps its first three modes)
    Moments M;
    compute_moments(x, y, I.area, I.bx, I.by, M);
    double poly[MAXNP][10];
    if (P <= 2) solve_basis(6, M, poly);
    else solve_basis_exact(10, x, y, I.area, I.bx, I.by, poly);
    double sq = std::sqrt(I.area);
    for (int i = 0; i < MAXNP; ++i)
        for (int c = 0; c < MAXNP; ++c) I.coeff[i][c] = 0.0;
    // POLY slot -> physical slot, scale 1/sqrt(A) per degree (degree-2 terms use 1/A)
    const int phys_of_poly[10] = {6, 7, 8, 9, 3, 4, 5, 1, 2, 0};
    for (int i = 0; i < Np; ++i)
        for (int t = 0; t < 10; ++t) {
            if (poly[i][t] == 0.0) continue;
            int deg = POLY_PX[t] + POLY_PY[t];
            double v;
            if (deg == 0) v = poly[i][t];
            else if (deg == 1) v = poly[i][t] == 1.0 ? 1 / sq : poly[i][t] / sq;
            else if (deg == 2) v = poly[i][t] == 1.0 ? 1 / I.area : poly[i][t] / I.area;
            else v = poly[i][t] / (I.area * sq);
            I.coeff[i][phys_of_poly[t]] = v;
        }
    // mass (src/triangleinfo.cpp:605-624)
    for (int i = 0; i < Np; ++i) {
        double m = 0;
        for (int q = 0; q < Nqv; ++q) {
            double bf = basis_value(I, i, I.vol_x[q], I.vol_y[q]);
            m += I.vol_w[q] * bf * bf;
        }
        I.mass[i] = m;
    }
    for (int i = 0; i < Np; ++i) {
        for (int p = 0; p < 3 * Nqe; ++p) I.edge_bf[i][p] = basis_value(I, i, I.edge_x[p], I.edge_y[p]);
        for (int q = 0; q < Nqv; ++q) {
            I.vol_bf[i][q] = basis_value(I, i, I.vol_x[q], I.vol_y[q]);
            I.vol_bdf[2 * i][q] = basis_deriv(I, i, 0, I.vol_x[q], I.vol_y[q]);
            I.vol_bdf[2 * i + 1][q] = basis_deriv(I, i, 1, I.vol_x[q], I.vol_y[q]);
        }
    }
}

// ---------------------------------------------------------------------------
// Gauss-Legendre nodes on [0,1] (Newton on P_n) and the collapsed (Duffy) rule on
// the reference triangle used for the L2 projection of initial data.
void gauss_legendre01(int n, double* t, double* w) {
    for (int i = 0; i < n; ++i) {
        long double z = std::cos(M_PI * (i + 0.75L) / (n + 0.5L)), pp = 0;
        for (int it = 0; it < 100; ++it) {
            long double p1 = 1, p2 = 0;
            for (int j = 1; j <= n; ++j) { long double p3 = p2; p2 = p1; p1 = ((2 * j - 1) * z * p2 - (j - 1) * p3) / j; }
            pp = n * (z * p1 - p2) / (z * z - 1);
            long double dz = p1 / pp;
            z -= dz;
            if (std::fabs((double)dz) < 1e-19) break;
        }
        t[i] = (double)(0.5L * (1 - z));
        w[i] = (double)(1.0L / ((1 - z * z) * pp * pp));
    }
}

constexpr int PROJ_N = 6;  // 36-point collapsed rule

struct VortexParams { double gamma, beta, uinf, vinf, xc, yc, Lx, Ly, t; };

void vortex_state(const VortexParams& V, double x, double y, double* U) {
    double g = V.gamma;
    double du = 0, dv = 0, dT = 0;
    int nimg = (V.Lx > 0 && V.Ly > 0) ? 1 : 0;
    double cx = V.xc + V.uinf * V.t, cy = V.yc + V.vinf * V.t;
    if (nimg) {
        cx = cx - V.Lx * std::floor((cx - 0.0) / V.Lx);
        cy = cy - V.Ly * std::floor((cy - 0.0) / V.Ly);
    }
    for (int ix = -nimg; ix <= nimg; ++ix)
        for (int iy = -nimg; iy <= nimg; ++iy) {
            double rx = x - (cx + ix * V.Lx), ry = y - (cy + iy * V.Ly);
            double r2 = rx * rx + ry * ry;
            double e = std::exp(0.5 * (1.0 - r2));
            du += -V.beta / (2 * M_PI) * e * ry;
            dv += V.beta / (2 * M_PI) * e * rx;
            dT += -(g - 1) * V.beta * V.beta / (8 * g * M_PI * M_PI) * e * e;
        }
    double rho = std::pow(1.0 + dT, 1.0 / (g - 1));
    double u = V.uinf + du, v = V.vinf + dv;
    double p = std::pow(rho, g);
    U[0] = rho; U[1] = rho * u; U[2] = rho * v; U[3] = p / (g - 1) + 0.5 * rho * (u * u + v * v);
}

// Area of triangle (x, y) inside the box [x0, x1] x [y0, y1]: the triangle clipped by the
// four box sides in turn (each a half-plane a*x + b*y <= c), then the shoelace formula.
double triangle_box_area(const double* x, const double* y, double x0, double x1, double y0, double y1) {
    std::vector<double> px(x, x + 3), py(y, y + 3);
    const double hp[4][3] = {{-1, 0, -x0}, {1, 0, x1}, {0, -1, -y0}, {0, 1, y1}};
    for (const auto& h : hp) {
        std::vector<double> qx, qy;
        const size_t n = px.size();
        for (size_t i = 0; i < n; ++i) {
            const size_t j = (i + 1) % n;
            const double fi = h[2] - (h[0] * px[i] + h[1] * py[i]);
            const double fj = h[2] - (h[0] * px[j] + h[1] * py[j]);
            if (fi >= 0) { qx.push_back(px[i]); qy.push_back(py[i]); }
            if ((fi >= 0) != (fj >= 0)) {
                const double t = fi / (fi - fj);
                qx.push_back(px[i] + t * (px[j] - px[i]));
                qy.push_back(py[i] + t * (py[j] - py[i]));
            }
        }
        px.swap(qx); py.swap(qy);
        if (px.size() < 3) return 0.0;
    }
    double a = 0;
    for (size_t i = 0; i < px.size(); ++i) {
        const size_t j = (i + 1) % px.size();
        a += px[i] * py[j] - px[j] * py[i];
    }
    return 0.5 * std::fabs(a);
}

// Exact cell average of the four-quadrant data over a triangle; prm: xs, ys, then the four
// quadrants' primitive states (rho, u, v, p) in the order UR, UL, LL, LR.
void riemann4_average(const double* prm, double gamma, const double* x, const double* y, double* U) {
    const double big = 1e300;
    const double xs = prm[0], ys = prm[1];
    // quadrants in the state order UR, UL, LL, LR
    const double box[4][4] = {{xs, big, ys, big}, {-big, xs, ys, big}, {-big, xs, -big, ys}, {xs, big, -big, ys}};
    double a[4], tot = 0;
    for (int q = 0; q < 4; ++q) tot += (a[q] = triangle_box_area(x, y, box[q][0], box[q][1], box[q][2], box[q][3]));
    for (int v = 0; v < 4; ++v) U[v] = 0;
    for (int q = 0; q < 4; ++q) {
        const double* s = prm + 2 + 4 * q;
        const double Uq[4] = {s[0], s[0] * s[1], s[0] * s[2],
                              s[3] / (gamma - 1) + 0.5 * s[0] * (s[1] * s[1] + s[2] * s[2])};
        for (int v = 0; v < 4; ++v) U[v] += a[q] / tot * Uq[v];
    }
}

// ---------------------------------------------------------------------------
// Physics helpers following src/kernelfunctions.cu's expressions.
inline void flux_fg_edge(double g, const double* U, double* F, double* G) {
    // kernel_calculateEdgeFG (392-433)
    double rho = U[0], u = U[1] / rho, v = U[2] / rho, rhoE = U[3];
    double p = (g - 1) * (rhoE - 0.5 * rho * (u * u + v * v));
    F[0] = rho * u; F[1] = rho * u * u + p; F[2] = rho * u * v; F[3] = u * (rhoE + p);
    G[0] = rho * v; G[1] = rho * v * u; G[2] = rho * v * v + p; G[3] = v * (rhoE + p);
}

// Roe (1981) flux in the normal direction with the Harten-Hyman (1983) entropy fix on the
// two acoustic waves.  No reference counterpart.
void roe_flux(double g, const double* UL, const double* UR, double nx, double ny, double* Fh) {
    double FL[4], GL[4], FR[4], GR[4];
    flux_fg_edge(g, UL, FL, GL);
    flux_fg_edge(g, UR, FR, GR);
    double rl = UL[0], ul = UL[1] / rl, vl = UL[2] / rl;
    double pl = (g - 1) * (UL[3] - 0.5 * rl * (ul * ul + vl * vl));
    double rr = UR[0], ur = UR[1] / rr, vr = UR[2] / rr;
    double pr = (g - 1) * (UR[3] - 0.5 * rr * (ur * ur + vr * vr));
    double hl = (UL[3] + pl) / rl, hr = (UR[3] + pr) / rr;
    double sl = std::sqrt(rl), sr = std::sqrt(rr), isum = 1.0 / (sl + sr);
    double rt = sl * sr;
    double ut = (sl * ul + sr * ur) * isum, vt = (sl * vl + sr * vr) * isum, ht = (sl * hl + sr * hr) * isum;
    double q2 = ut * ut + vt * vt;
    double c2 = (g - 1) * (ht - 0.5 * q2);
    double ct = std::sqrt(c2);
    double unt = ut * nx + vt * ny;
    double utt = -ut * ny + vt * nx;
    double dp = pr - pl, dr = rr - rl;
    double dun = (ur * nx + vr * ny) - (ul * nx + vl * ny);
    double dut = (-ur * ny + vr * nx) - (-ul * ny + vl * nx);
    double a1 = (dp - rt * ct * dun) / (2 * c2);
    double a2 = dr - dp / c2;
    double a3 = rt * dut;
    double a4 = (dp + rt * ct * dun) / (2 * c2);
    // Harten-Hyman entropy fix on the acoustic waves: delta = max(0, lam - lam_L, lam_R - lam),
    // which vanishes across shocks (lam_L > lam > lam_R) and keeps them exact
    double cl = std::sqrt(g * pl / rl), cr = std::sqrt(g * pr / rr);
    double unl = ul * nx + vl * ny, unr = ur * nx + vr * ny;
    double l1 = std::fabs(unt - ct), l2 = std::fabs(unt), l4 = std::fabs(unt + ct);
    double d1 = std::max(0.0, std::max((unt - ct) - (unl - cl), (unr - cr) - (unt - ct)));
    double d4 = std::max(0.0, std::max((unt + ct) - (unl + cl), (unr + cr) - (unt + ct)));
    if (l1 < d1) l1 = (l1 * l1 + d1 * d1) / (2 * d1);
    if (l4 < d4) l4 = (l4 * l4 + d4 * d4) / (2 * d4);
    double d[4];
    d[0] = l1 * a1 + l2 * a2 + l4 * a4;
    d[1] = l1 * a1 * (ut - ct * nx) + l2 * a2 * ut + l2 * a3 * (-ny) + l4 * a4 * (ut + ct * nx);
    d[2] = l1 * a1 * (vt - ct * ny) + l2 * a2 * vt + l2 * a3 * nx + l4 * a4 * (vt + ct * ny);
    d[3] = l1 * a1 * (ht - unt * ct) + l2 * a2 * 0.5 * q2 + l2 * a3 * utt + l4 * a4 * (ht + unt * ct);
    for (int v = 0; v < 4; ++v) Fh[v] = 0.5 * ((FL[v] + FR[v]) * nx + (GL[v] + GR[v]) * ny - d[v]);
}

// ---------------------------------------------------------------------------
struct Params {
    int P, flux, limiter, dt_policy, reduction_threads;
    double gamma, cfl, tvb_M, dt_fixed;
    double fs[4];  // far-field conserved state
};

struct Solver {
    int E;
    Params prm;
    int Np, Nqv, Nqe;
    std::vector<ElemInfo> info;
    std::vector<int> nbr, se, bc;   // [3][E]
    std::vector<unsigned char> dtflag;
    std::vector<double> U, U0, R;   // [4][Np][E]
    std::vector<int> flags;         // limiter troubled-cell flags
    // limiter geometry
    std::vector<double> lim_a;      // [E][3][2]
    std::vector<int> lim_pair;      // [E][3][2]
    std::vector<double> lim_h2;
    std::vector<double> mid_bf;     // [E][3 edges][3 modes]

    double& u(std::vector<double>& A, int v, int i, int e) { return A[((size_t)v * Np + i) * E + e]; }
};

void build_limiter_geometry(Solver& S, const double* ex, const double* ey) {
    int E = S.E;
    S.lim_a.assign((size_t)E * 6, 0.0);
    S.lim_pair.assign((size_t)E * 6, -1);
    S.lim_h2.assign(E, 0.0);
    S.mid_bf.assign((size_t)E * 9, 0.0);
    #pragma omp parallel for schedule(static)
    for (int e = 0; e < E; ++e) {
        const ElemInfo& I = S.info[e];
        double lmax = std::max(I.len[0], std::max(I.len[1], I.len[2]));
        S.lim_h2[e] = lmax * lmax;
        double b0x = I.bx, b0y = I.by;
        double bnx[3], bny[3];
        for (int k = 0; k < 3; ++k) {
            int n = S.nbr[k * E + e];
            if (n >= 0) {
                // neighbour barycenter seen from this element (periodic images are
                // handled by mirroring the own barycenter across the shared edge
                // midpoint when the neighbour is far away)
                double mx = 0.5 * (ex[3 * e + (k + 1) % 3] + ex[3 * e + (k + 2) % 3]);
                double my = 0.5 * (ey[3 * e + (k + 1) % 3] + ey[3 * e + (k + 2) % 3]);
                double nbx = S.info[n].bx, nby = S.info[n].by;
                // periodic shift: bring the neighbour barycenter next to the shared midpoint
                const ElemInfo& In = S.info[n];
                int sn = S.se[k * E + e];
                double nmx = 0.5 * (ex[3 * n + (sn + 1) % 3] + ex[3 * n + (sn + 2) % 3]);
                double nmy = 0.5 * (ey[3 * n + (sn + 1) % 3] + ey[3 * n + (sn + 2) % 3]);
                (void)In;
                bnx[k] = nbx + (mx - nmx);
                bny[k] = nby + (my - nmy);
            } else {
                double mx = 0.5 * (ex[3 * e + (k + 1) % 3] + ex[3 * e + (k + 2) % 3]);
                double my = 0.5 * (ey[3 * e + (k + 1) % 3] + ey[3 * e + (k + 2) % 3]);
                bnx[k] = 2 * mx - b0x;
                bny[k] = 2 * my - b0y;
            }
        }
        for (int k = 0; k < 3; ++k) {
            double mx = 0.5 * (ex[3 * e + (k + 1) % 3] + ex[3 * e + (k + 2) % 3]);
            double my = 0.5 * (ey[3 * e + (k + 1) % 3] + ey[3 * e + (k + 2) % 3]);
            double rx = mx - b0x, ry = my - b0y;
            bool found = false;
            for (int off = 1; off <= 2 && !found; ++off) {
                int k1 = k, k2 = (k + off) % 3;
                double ax = bnx[k1] - b0x, ay = bny[k1] - b0y, cx = bnx[k2] - b0x, cy = bny[k2] - b0y;
                double det = ax * cy - ay * cx;
                if (std::fabs(det) < 1e-300) continue;
                double al1 = (rx * cy - ry * cx) / det, al2 = (ax * ry - ay * rx) / det;
                if (al1 >= -1e-12 && al2 >= -1e-12) {
                    S.lim_a[e * 6 + 2 * k] = std::max(al1, 0.0);
                    S.lim_a[e * 6 + 2 * k + 1] = std::max(al2, 0.0);
                    S.lim_pair[e * 6 + 2 * k] = k1;
                    S.lim_pair[e * 6 + 2 * k + 1] = k2;
                    found = true;
                }
            }
            if (!found) {  // fall back to the neighbour across the edge alone
                S.lim_a[e * 6 + 2 * k] = 0.5;
                S.lim_pair[e * 6 + 2 * k] = k;
                S.lim_pair[e * 6 + 2 * k + 1] = k;
                S.lim_a[e * 6 + 2 * k + 1] = 0.0;
            }
            for (int i = 0; i < 3; ++i) S.mid_bf[e * 9 + 3 * k + i] = basis_value(I, i, mx, my);
        }
    }
}

// R = residual of state A (before division by the mass); restates the kernel chain.
void residual(Solver& S, const std::vector<double>& A, std::vector<double>& Rout) {
    const int E = S.E, Np = S.Np, Nqv = S.Nqv, Nqe = S.Nqe, NE = 3 * Nqe;
    const double g = S.prm.gamma;
    auto at = [&](int v, int i, int e) { return A[((size_t)v * Np + i) * E + e]; };
    #pragma omp parallel for schedule(static)
    for (int e = 0; e < E; ++e) {
        const ElemInfo& I = S.info[e];
        // kernel_calculateConVars
        double Uv[MAXQV][4], Ue[3 * MAXQE][4];
        for (int p = 0; p < NE; ++p) {
            double vc[4] = {0, 0, 0, 0}, ec[4] = {0, 0, 0, 0};
            for (int j = 0; j < Np; ++j) {
                double f[4] = {at(0, j, e), at(1, j, e), at(2, j, e), at(3, j, e)};
                if (p < Nqv) {
                    double bfv = I.vol_bf[j][p];
                    for (int v = 0; v < 4; ++v) vc[v] += bfv * f[v];
                }
                double bfv = I.edge_bf[j][p];
                for (int v = 0; v < 4; ++v) ec[v] += bfv * f[v];
            }
            if (p < Nqv) for (int v = 0; v < 4; ++v) Uv[p][v] = vc[v];
            for (int v = 0; v < 4; ++v) Ue[p][v] = ec[v];
        }
        if (Nqv > NE) {  // P3: 12 volume points, 12 edge points
            for (int p = NE; p < Nqv; ++p) {
                double vc[4] = {0, 0, 0, 0};
                for (int j = 0; j < Np; ++j) {
                    double bfv = I.vol_bf[j][p];
                    for (int v = 0; v < 4; ++v) vc[v] += bfv * at(v, j, e);
                }
                for (int v = 0; v < 4; ++v) Uv[p][v] = vc[v];
            }
        }
        // kernel_calculateVolumeRHS
        double Rv[MAXNP][4];
        for (int i = 0; i < Np; ++i) {
            double rhs[4] = {0, 0, 0, 0}, f[4];
            for (int q = 0; q < Nqv; ++q) {
                double rho = Uv[q][0], u = Uv[q][1] / rho, v = Uv[q][2] / rho, rhoE = Uv[q][3];
                double w = I.vol_w[q], dvx = I.vol_bdf[2 * i][q], dvy = I.vol_bdf[2 * i + 1][q];
                double p = (g - 1) * (rhoE - 0.5 * rho * (u * u + v * v));
                f[0] = rho * u; f[1] = rho * u * u + p; f[2] = rho * u * v; f[3] = u * (rhoE + p);
                for (int c = 0; c < 4; ++c) rhs[c] += w * f[c] * dvx;
                f[0] = rho * v; f[1] = f[2]; f[2] = rho * v * v + p; f[3] = v * (rhoE + p);
                for (int c = 0; c < 4; ++c) rhs[c] += w * f[c] * dvy;
            }
            for (int c = 0; c < 4; ++c) Rv[i][c] = rhs[c];
        }
        // fluxes on the three edges
        double Fh[3 * MAXQE][4];
        for (int k = 0; k < 3; ++k) {
            int n = S.nbr[k * E + e];
            double nx = I.normal[2 * k], ny = I.normal[2 * k + 1];
            double mean_n[4];
            double UR[MAXQE][4];
            if (n >= 0) {
                int sn = S.se[k * E + e];
                const ElemInfo& In = S.info[n];
                for (int j = 0; j < Nqe; ++j) {
                    int pr = sn * Nqe + (Nqe - 1 - j);
                    for (int v = 0; v < 4; ++v) {
                        double ec = 0;
                        for (int i = 0; i < Np; ++i) ec += In.edge_bf[i][pr] * at(v, i, n);
                        UR[j][v] = ec;
                    }
                }
                for (int v = 0; v < 4; ++v) mean_n[v] = at(v, 0, n);
            } else {
                int t = S.bc[k * E + e];
                // kernel_boundaryCondition (158-275), applied per boundary edge: the ghost
                // edge 0 point (Nqe-1-j) holds the interior trace at point j
                if (t == BC_FARFIELD) {
                    for (int j = 0; j < Nqe; ++j) for (int v = 0; v < 4; ++v) UR[j][v] = S.prm.fs[v];
                    for (int v = 0; v < 4; ++v) mean_n[v] = S.prm.fs[v];
                } else if (t == BC_REFLECTION) {
                    for (int j = 0; j < Nqe; ++j) {
                        const double* L = Ue[k * Nqe + j];
                        UR[j][0] = L[0]; UR[j][3] = L[3];
                        UR[j][1] = L[1] * (ny * ny - nx * nx) - 2.0 * nx * ny * L[2];
                        UR[j][2] = L[2] * (nx * nx - ny * ny) - 2.0 * nx * ny * L[1];
                    }
                    double m[4] = {at(0, 0, e), at(1, 0, e), at(2, 0, e), at(3, 0, e)};
                    mean_n[0] = m[0]; mean_n[3] = m[3];
                    mean_n[1] = m[1] * (ny * ny - nx * nx) - 2.0 * nx * ny * m[2];
                    mean_n[2] = m[2] * (nx * nx - ny * ny) - 2.0 * nx * ny * m[1];
                } else {  // BC_OUTFLOW (transmissive)
                    for (int j = 0; j < Nqe; ++j) for (int v = 0; v < 4; ++v) UR[j][v] = Ue[k * Nqe + j][v];
                    for (int v = 0; v < 4; ++v) mean_n[v] = at(v, 0, e);
                }
            }
            if (S.prm.flux == FLUX_LF) {
                // kernel_calculateLFCoeff (354-390)
                double rho = 0.5 * (at(0, 0, e) + mean_n[0]);
                double u = 0.5 * (at(1, 0, e) + mean_n[1]) / rho;
                double v = 0.5 * (at(2, 0, e) + mean_n[2]) / rho;
                double rhoE = 0.5 * (at(3, 0, e) + mean_n[3]);
                double p = (g - 1) * (rhoE - 0.5 * rho * (u * u + v * v));
                double ma = std::sqrt(g * p / rho);
                double alpha = std::fabs(u * nx + v * ny) + ma;
                // kernel_calculateEdgeFG + kernel_calculateFlux (392-433, 510-582)
                for (int j = 0; j < Nqe; ++j) {
                    double FL[4], GL[4], FR[4], GR[4];
                    flux_fg_edge(g, Ue[k * Nqe + j], FL, GL);
                    flux_fg_edge(g, UR[j], FR, GR);
                    for (int c = 0; c < 4; ++c)
                        Fh[k * Nqe + j][c] = 0.5 * ((FL[c] + FR[c]) * nx + (GL[c] + GR[c]) * ny -
                                                    alpha * (UR[j][c] - Ue[k * Nqe + j][c]));
                }
            } else {
                for (int j = 0; j < Nqe; ++j) roe_flux(g, Ue[k * Nqe + j], UR[j], nx, ny, Fh[k * Nqe + j]);
            }
        }
        // kernel_calculateEdgeRHS (584-637)
        for (int i = 0; i < Np; ++i) {
            double rhs[4] = {0, 0, 0, 0};
            for (int p = 0; p < NE; ++p) {
                double w = I.edge_w[p], bfv = I.edge_bf[i][p];
                for (int c = 0; c < 4; ++c) rhs[c] += w * Fh[p][c] * bfv;
            }
            for (int c = 0; c < 4; ++c) Rout[((size_t)c * Np + i) * E + e] = Rv[i][c] - rhs[c];
        }
    }
}

// Positivity-preserving scaling (Zhang & Shu 2010), applied after the TVB limiter:
// the modes >= 1 are scaled toward the cell mean so that density, then pressure, stay
// >= eps at every volume and edge quadrature point.  Pressure is concave in the
// conserved variables, so theta = (p_mean - eps) / (p_mean - p_q) is a valid bound.
constexpr double POS_EPS = 1e-13;

void positivity(Solver& S, std::vector<double>& A) {
    const int E = S.E, Np = S.Np, Nqv = S.Nqv, NE = 3 * S.Nqe;
    const double g = S.prm.gamma;
    #pragma omp parallel for schedule(static)
    for (int e = 0; e < E; ++e) {
        const ElemInfo& I = S.info[e];
        double c[4][MAXNP];
        for (int v = 0; v < 4; ++v)
            for (int i = 0; i < Np; ++i) c[v][i] = A[((size_t)v * Np + i) * E + e];
        auto point = [&](int q, double* U) {
            const double* bf;
            for (int v = 0; v < 4; ++v) {
                double s = c[v][0];
                for (int i = 1; i < Np; ++i) {
                    bf = q < Nqv ? &I.vol_bf[i][q] : &I.edge_bf[i][q - Nqv];
                    s += *bf * c[v][i];
                }
                U[v] = s;
            }
        };
        const double rm = c[0][0];
        double rmin = rm;
        for (int q = 0; q < Nqv + NE; ++q) { double U[4]; point(q, U); rmin = std::min(rmin, U[0]); }
        double th1 = 1.0;
        if (rmin < POS_EPS) th1 = std::min(1.0, (rm - POS_EPS) / (rm - rmin));
        if (th1 < 1.0)
            for (int i = 1; i < Np; ++i) c[0][i] *= th1;
        const double pm = (g - 1) * (c[3][0] - 0.5 * (c[1][0] * c[1][0] + c[2][0] * c[2][0]) / rm);
        double th2 = 1.0;
        for (int q = 0; q < Nqv + NE; ++q) {
            double U[4];
            point(q, U);
            double pq = (g - 1) * (U[3] - 0.5 * (U[1] * U[1] + U[2] * U[2]) / U[0]);
            if (pq < POS_EPS) th2 = std::min(th2, (pm - POS_EPS) / (pm - pq));
        }
        if (th1 < 1.0 || th2 < 1.0) {
            for (int v = 0; v < 4; ++v)
                for (int i = 1; i < Np; ++i)
                    A[((size_t)v * Np + i) * E + e] = (v == 0 ? c[0][i] : c[v][i]) * (th2 < 1.0 ? th2 : 1.0);
            S.flags[e] |= 2;
        }
    }
}

// TVB minmod limiter on the P1 part (Cockburn & Shu 1998), component-wise, in place.
void limit(Solver& S, std::vector<double>& A) {
    const int E = S.E, Np = S.Np;
    std::vector<double> mean((size_t)4 * E);
    for (int v = 0; v < 4; ++v)
        for (int e = 0; e < E; ++e) mean[(size_t)v * E + e] = A[((size_t)v * Np) * E + e];
    const double nu = 1.5, M = S.prm.tvb_M;
    #pragma omp parallel for schedule(static)
    for (int e = 0; e < E; ++e) {
        double nm[3][4];
        for (int k = 0; k < 3; ++k) {
            int n = S.nbr[k * E + e];
            for (int v = 0; v < 4; ++v) nm[k][v] = n >= 0 ? mean[(size_t)v * E + n] : mean[(size_t)v * E + e];
        }
        bool troubled = false;
        for (int v = 0; v < 4; ++v) {
            double m0 = mean[(size_t)v * E + e];
            double d[3];
            bool changed = false;
            for (int k = 0; k < 3; ++k) {
                double ut = 0;
                for (int i = 1; i < 3; ++i) ut += S.mid_bf[e * 9 + 3 * k + i] * A[((size_t)v * Np + i) * E + e];
                int k1 = S.lim_pair[e * 6 + 2 * k], k2 = S.lim_pair[e * 6 + 2 * k + 1];
                double du = S.lim_a[e * 6 + 2 * k] * (nm[k1][v] - m0) + S.lim_a[e * 6 + 2 * k + 1] * (nm[k2][v] - m0);
                double b = nu * du;
                double r;
                if (std::fabs(ut) <= M * S.lim_h2[e]) r = ut;          // TVB modification
                else if (ut > 0 && b > 0) r = std::min(ut, b);
                else if (ut < 0 && b < 0) r = std::max(ut, b);
                else r = 0;
                if (r != ut) changed = true;
                d[k] = r;
            }
            if (!changed) continue;
            troubled = true;
            double s = d[0] + d[1] + d[2];
            if (s != 0.0) {
                double pos = 0, neg = 0;
                for (int k = 0; k < 3; ++k) { pos += std::max(0.0, d[k]); neg += std::max(0.0, -d[k]); }
                double tp = pos > 0 ? std::min(1.0, neg / pos) : 0.0;
                double tn = neg > 0 ? std::min(1.0, pos / neg) : 0.0;
                for (int k = 0; k < 3; ++k) d[k] = tp * std::max(0.0, d[k]) - tn * std::max(0.0, -d[k]);
            }
            // P1 slopes from the midpoint deviations (consistent since they sum to zero);
            // modes >= 3 of a limited component are dropped
            double a11 = 0, a12 = 0, a22 = 0, r1 = 0, r2 = 0;
            for (int k = 0; k < 3; ++k) {
                double p1 = S.mid_bf[e * 9 + 3 * k + 1], p2 = S.mid_bf[e * 9 + 3 * k + 2];
                a11 += p1 * p1; a12 += p1 * p2; a22 += p2 * p2; r1 += p1 * d[k]; r2 += p2 * d[k];
            }
            double det = a11 * a22 - a12 * a12;
            A[((size_t)v * Np + 1) * E + e] = (r1 * a22 - r2 * a12) / det;
            A[((size_t)v * Np + 2) * E + e] = (a11 * r2 - a12 * r1) / det;
            for (int i = 3; i < Np; ++i) A[((size_t)v * Np + i) * E + e] = 0.0;
        }
        S.flags[e] = troubled ? 1 : 0;
    }
    positivity(S, A);
}

double compute_dt(Solver& S, const std::vector<double>& A) {
    // kernel_getTimeStep (3-81): dt = cfl / max over the included cells of
    // (|u| + c) * perimeter / area, on the cell means
    const int E = S.E, Np = S.Np;
    const double g = S.prm.gamma;
    double mx = 0.0;
    #pragma omp parallel for reduction(max : mx) schedule(static)
    for (int e = 0; e < E; ++e) {
        if (!S.dtflag[e]) continue;
        double rho = A[(0 * (size_t)Np) * E + e];
        double u = A[(1 * (size_t)Np) * E + e] / rho;
        double v = A[(2 * (size_t)Np) * E + e] / rho;
        double p = (g - 1) * (A[(3 * (size_t)Np) * E + e] - 0.5 * rho * (u * u + v * v));
        double a = std::sqrt(g * p / rho);
        double dx = (std::sqrt(u * u + v * v) + a) * S.info[e].perimeter / S.info[e].area;
        if (dx > mx) mx = dx;
    }
    return S.prm.cfl / mx;
}

void rk_stage(Solver& S, int s, double dt) {
    // kernel_rkdgStepOne/Two/Three (640-748)
    residual(S, S.U, S.R);
    const int E = S.E, Np = S.Np;
    #pragma omp parallel for schedule(static)
    for (int e = 0; e < E; ++e)
        for (int i = 0; i < Np; ++i) {
            double mass = S.info[e].mass[i];
            for (int v = 0; v < 4; ++v) {
                size_t idx = ((size_t)v * Np + i) * E + e;
                if (s == 0) S.U[idx] += dt * (S.R[idx]) / mass;
                else if (s == 1) S.U[idx] = 0.75 * S.U0[idx] + 0.25 * S.U[idx] + 0.25 * dt * (S.R[idx]) / mass;
                else S.U[idx] = 1.0 / 3.0 * S.U0[idx] + 2.0 / 3.0 * S.U[idx] + 2.0 / 3.0 * dt * (S.R[idx]) / mass;
            }
        }
    if (S.prm.limiter) limit(S, S.U);
}

}  // namespace

extern "C" {

int orc_info_size(int P) {
    int Np = np_of(P), Nqv = nqv_of(P), Nqe = nqe_of(P);
    return 10 + Np * Np + Np + Nqv + 3 * Nqe + Np * Nqv + 2 * Np * Nqv + Np * 3 * Nqe;
}

// Per-element precompute; out layout:
//  area, perimeter, n0x, n0y, n1x, n1y, n2x, n2y, bx, by,
//  coeff[Np][Np] (physical monomials about the barycenter), mass[Np], vol_w[Nqv],
//  edge_w[3Nqe], vol_bf[Np][Nqv], vol_bdf[2Np][Nqv], edge_bf[Np][3Nqe]
int orc_triangle_info(int P, const double* x, const double* y, double* out) {
    if (P < 1 || P > 3) return -1;
    ElemInfo I;
    element_info(P, x, y, I);
    int Np = I.Np, Nqv = I.Nqv, Nqe = I.Nqe;
    double* o = out;
    *o++ = I.area; *o++ = I.perimeter;
    for (int k = 0; k < 6; ++k) *o++ = I.normal[k];
    *o++ = I.bx; *o++ = I.by;
    int ncoef = P <= 2 ? 6 : 10;
    for (int i = 0; i < Np; ++i) for (int c = 0; c < Np; ++c) *o++ = c < ncoef ? I.coeff[i][c] : 0.0;
    for (int i = 0; i < Np; ++i) *o++ = I.mass[i];
    for (int q = 0; q < Nqv; ++q) *o++ = I.vol_w[q];
    for (int p = 0; p < 3 * Nqe; ++p) *o++ = I.edge_w[p];
    for (int i = 0; i < Np; ++i) for (int q = 0; q < Nqv; ++q) *o++ = I.vol_bf[i][q];
    for (int i = 0; i < 2 * Np; ++i) for (int q = 0; q < Nqv; ++q) *o++ = I.vol_bdf[i][q];
    for (int i = 0; i < Np; ++i) for (int p = 0; p < 3 * Nqe; ++p) *o++ = I.edge_bf[i][p];
    return 0;
}

// Basis value of mode i of an element at a physical point (used for L2 norms in tests).
double orc_basis_value(int P, const double* x, const double* y, int i, double px, double py) {
    ElemInfo I;
    element_info(P, x, y, I);
    return basis_value(I, i, px, py);
}

void orc_set_threads(int n) { if (n > 0) omp_set_num_threads(n); }
int orc_get_threads(void) { return omp_get_max_threads(); }

// ex, ey: [E][3] vertex coordinates per element (CCW, edge k opposite vertex k)
// nbr, se, bc: [3][E]; nbr = -1 marks a boundary edge whose type is bc
// dt_id: [E] id used for the reduction-subset rule (NULL: element index)
// iparams: P, flux, limiter, dt_policy, reduction_threads
// dparams: gamma, cfl, tvb_M, dt_fixed, fs[4]
void* orc_create(int E, const double* ex, const double* ey, const int* nbr, const int* se, const int* bc,
                 const int* dt_id, const int* iparams, const double* dparams) {
    Solver* S = new Solver();
    S->E = E;
    S->prm.P = iparams[0]; S->prm.flux = iparams[1]; S->prm.limiter = iparams[2];
    S->prm.dt_policy = iparams[3]; S->prm.reduction_threads = iparams[4];
    S->prm.gamma = dparams[0]; S->prm.cfl = dparams[1]; S->prm.tvb_M = dparams[2]; S->prm.dt_fixed = dparams[3];
    for (int v = 0; v < 4; ++v) S->prm.fs[v] = dparams[4 + v];
    S->Np = np_of(S->prm.P); S->Nqv = nqv_of(S->prm.P); S->Nqe = nqe_of(S->prm.P);
    S->info.resize(E);
    #pragma omp parallel for schedule(static)
    for (int e = 0; e < E; ++e) element_info(S->prm.P, ex + 3 * e, ey + 3 * e, S->info[e]);
    S->nbr.assign(nbr, nbr + 3 * (size_t)E);
    S->se.assign(se, se + 3 * (size_t)E);
    S->bc.assign(bc, bc + 3 * (size_t)E);
    S->dtflag.assign(E, 1);
    if (S->prm.dt_policy == DT_REFERENCE) {
        // the reduction tree of kernel_getTimeStep keeps 64 of every 128 threads when
        // its block has >= 128 threads (power of two)
        int R = S->prm.reduction_threads;
        if (R >= 128)
            for (int e = 0; e < E; ++e) {
                int id = dt_id ? dt_id[e] : e;
                S->dtflag[e] = (id % 128) < 64;
            }
    }
    size_t n = (size_t)4 * S->Np * E;
    S->U.assign(n, 0.0); S->U0.assign(n, 0.0); S->R.assign(n, 0.0);
    S->flags.assign(E, 0);
    if (S->prm.limiter) build_limiter_geometry(*S, ex, ey);
    return S;
}

void orc_destroy(void* h) { delete (Solver*)h; }

int orc_set_state(void* h, const double* U) {
    Solver* S = (Solver*)h;
    std::memcpy(S->U.data(), U, S->U.size() * sizeof(double));
    return 0;
}

int orc_get_state(void* h, double* U) {
    Solver* S = (Solver*)h;
    std::memcpy(U, S->U.data(), S->U.size() * sizeof(double));
    return 0;
}

int orc_get_mass(void* h, double* M) {  // [Np][E]
    Solver* S = (Solver*)h;
    for (int i = 0; i < S->Np; ++i)
        for (int e = 0; e < S->E; ++e) M[(size_t)i * S->E + e] = S->info[e].mass[i];
    return 0;
}

int orc_get_flags(void* h, int* f) {
    Solver* S = (Solver*)h;
    std::memcpy(f, S->flags.data(), S->E * sizeof(int));
    return 0;
}

// Residual R(U) for the current state (before division by the mass)
int orc_residual(void* h, double* Rout) {
    Solver* S = (Solver*)h;
    residual(*S, S->U, S->R);
    std::memcpy(Rout, S->R.data(), S->R.size() * sizeof(double));
    return 0;
}

double orc_compute_dt(void* h) {
    Solver* S = (Solver*)h;
    if (S->prm.dt_policy == DT_FIXED) return S->prm.dt_fixed;
    return compute_dt(*S, S->U);
}

// One stage with a host dt.  Stage 0 saves the step-start state first (the
// cudaMemcpy2DAsync U -> U_old of src/rkdgadvance.cpp:104-107).
int orc_stage(void* h, int s, double dt) {
    Solver* S = (Solver*)h;
    if (s == 0) S->U0 = S->U;
    rk_stage(*S, s, dt);
    return 0;
}

// n SSP-RK3 steps (t_end > 0 clips the last step); returns the number of steps taken
int orc_advance(void* h, int nsteps, double t_end, double* t_io, double* dt_log) {
    Solver* S = (Solver*)h;
    double t = t_io ? *t_io : 0.0;
    int taken = 0;
    for (int n = 0; n < nsteps; ++n) {
        double dt = S->prm.dt_policy == DT_FIXED ? S->prm.dt_fixed : compute_dt(*S, S->U);
        if (t_end > 0) {
            if (t >= t_end) break;
            if (t_end - t < dt) dt = t_end - t;
        }
        S->U0 = S->U;
        for (int s = 0; s < 3; ++s) rk_stage(*S, s, dt);
        if (dt_log) dt_log[n] = dt;
        t += dt;
        ++taken;
    }
    if (t_io) *t_io = t;
    return taken;
}

// kernel_calculateResidual (751-867): max |change| of the rho and rhoE means over the
// included cells, against the state saved at the start of the last step.
int orc_residual_change(void* h, double* res) {
    Solver* S = (Solver*)h;
    const int E = S->E, Np = S->Np;
    double r0 = 0, r1 = 0;
    for (int e = 0; e < E; ++e) {
        if (!S->dtflag[e]) continue;
        double a = std::fabs(S->U[e] - S->U0[e]);
        double b = std::fabs(S->U[(3 * (size_t)Np) * E + e] - S->U0[(3 * (size_t)Np) * E + e]);
        if (r0 < a) r0 = a;
        if (r1 < b) r1 = b;
    }
    res[0] = r0; res[1] = r1;
    return 0;
}

// L2 projection of initial data onto each element's basis (36-point collapsed rule).
// kind 0: free stream of src/cudarkdgsolver.cu:326-368 (prm: gamma, alpha_rad, mach, rhoref, pref)
// kind 1: isentropic vortex (prm: gamma, beta, uinf, vinf, xc, yc, Lx, Ly, t)
// kind 2: four-quadrant Riemann data (prm: gamma, xs, ys, 4 x (rho,u,v,p) for UR, UL, LL, LR),
//         exact cell averages (the triangle clipped to each quadrant)
int orc_project(void* h, const double* ex, const double* ey, int kind, const double* prm) {
    Solver* S = (Solver*)h;
    const int E = S->E, Np = S->Np;
    if (kind == 0) {
        double g = prm[0], alpha = prm[1], mach = prm[2], rhoref = prm[3], pref = prm[4];
        double ut = std::sqrt(g * pref / rhoref) * mach;
        double u = ut * std::cos(alpha), v = ut * std::sin(alpha);
        std::fill(S->U.begin(), S->U.end(), 0.0);
        for (int e = 0; e < E; ++e) {
            S->U[(0 * (size_t)Np) * E + e] = rhoref;
            S->U[(1 * (size_t)Np) * E + e] = rhoref * u;
            S->U[(2 * (size_t)Np) * E + e] = rhoref * v;
            S->U[(3 * (size_t)Np) * E + e] = rhoref * (ut * ut) / 2 + pref / (rhoref * (g - 1));
        }
        return 0;
    }
    if (kind == 2) {
        // discontinuous four-quadrant data start from their exact cell averages (no slopes;
        // phi_0 = 1, so mode 0 is the mean)
        std::fill(S->U.begin(), S->U.end(), 0.0);
        for (int e = 0; e < E; ++e) {
            double U[4];
            riemann4_average(prm + 1, prm[0], ex + 3 * e, ey + 3 * e, U);
            for (int v = 0; v < 4; ++v) S->U[((size_t)v * Np) * E + e] = U[v];
        }
        return 0;
    }
    double gt[PROJ_N], gw[PROJ_N];
    gauss_legendre01(PROJ_N, gt, gw);
    #pragma omp parallel for schedule(static)
    for (int e = 0; e < E; ++e) {
        const ElemInfo& I = S->info[e];
        const double* x = ex + 3 * e;
        const double* y = ey + 3 * e;
        double num[MAXNP][4] = {}, den[MAXNP] = {};
        for (int a = 0; a < PROJ_N; ++a)
            for (int b = 0; b < PROJ_N; ++b) {
                double xi = gt[a], eta = gt[b] * (1 - gt[a]);
                double w = gw[a] * gw[b] * (1 - gt[a]) * 2.0 * I.area;  // weights sum to area
                double px = x[0] + (x[1] - x[0]) * xi + (x[2] - x[0]) * eta;
                double py = y[0] + (y[1] - y[0]) * xi + (y[2] - y[0]) * eta;
                double U[4];
                VortexParams V{prm[0], prm[1], prm[2], prm[3], prm[4], prm[5], prm[6], prm[7], prm[8]};
                vortex_state(V, px, py, U);
                for (int i = 0; i < Np; ++i) {
                    double bf = basis_value(I, i, px, py);
                    for (int v = 0; v < 4; ++v) num[i][v] += w * U[v] * bf;
                    den[i] += w * bf * bf;
                }
            }
        for (int i = 0; i < Np; ++i)
            for (int v = 0; v < 4; ++v) S->U[((size_t)v * Np + i) * E + e] = num[i][v] / den[i];
    }
    return 0;
}

// Exact isentropic vortex at a list of points (prm as kind 1 above).
int orc_vortex_exact(int n, const double* px, const double* py, const double* prm, double* U) {
    VortexParams V{prm[0], prm[1], prm[2], prm[3], prm[4], prm[5], prm[6], prm[7], prm[8]};
    for (int i = 0; i < n; ++i) vortex_state(V, px[i], py[i], U + 4 * i);
    return 0;
}

// Roe flux at one point (exposed for the consistency / contact / shock tests).
void orc_roe_flux(double gamma, const double* UL, const double* UR, double nx, double ny, double* F) {
    roe_flux(gamma, UL, UR, nx, ny, F);
}

}  // extern "C"
