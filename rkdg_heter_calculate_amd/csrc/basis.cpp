// Reference-element basis, quadrature tables, element geometry, the reference's
// per-element basis (for state exchange in the reference's coefficients) and the
// L2 projection of initial data.
//
// The device evolves modal coefficients in one orthonormal basis psi on the reference
// triangle; every element is its affine image.  The reference instead builds a
// Gram-Schmidt basis per element (src/triangleinfo.cpp:433-602).  Both span P_k on
// each element, the quadrature points are affine images of the same barycentric
// points, and the mass is diagonal in both, so the two schemes evolve the same
// polynomial up to round-off; states are exchanged through an exact L2 projection.
#include <cmath>
#include <cstring>
#include <mutex>
#include <algorithm>
#include "internal.h"

namespace rkdg {

namespace {

const int MX[10] = {0, 1, 0, 2, 1, 0, 3, 2, 1, 0};  // monomial exponents (xi / X)
const int MY[10] = {0, 0, 1, 0, 1, 2, 0, 1, 2, 3};  // (eta / Y)

long double factl(int n) { long double r = 1; for (int i = 2; i <= n; ++i) r *= i; return r; }

// Volume rule of src/triangleinfo.cpp:159-163,199-205,219-225 (P1, P2): barycentric
// points with the reference's literals; Dunavant's 12-point degree-6 rule for P3.
void volume_rule(int P, int& n, double* xi, double* eta, double* w) {
    if (P <= 2) {
        const double w1 = 0.797426985353087, w2 = 0.101286507323456, w3 = 0.059715871789770,
                     w4 = 0.470142064105115;
        const double bc[7][3] = {{1.0 / 3, 1.0 / 3, 1.0 / 3}, {w1, w2, w2}, {w2, w1, w2}, {w2, w2, w1},
                                 {w3, w4, w4}, {w4, w3, w4}, {w4, w4, w3}};
        const double wq[7] = {0.225000000000000, 0.125939180544827, 0.125939180544827, 0.125939180544827,
                              0.132394152788506, 0.132394152788506, 0.132394152788506};
        n = 7;
        for (int q = 0; q < 7; ++q) { xi[q] = bc[q][1]; eta[q] = bc[q][2]; w[q] = wq[q]; }
    } else {
        const double a1 = 0.063089014491502228340331602870819, b1 = 0.873821971016995543319336794258362;
        const double a2 = 0.249286745170910421291638553107019, b2 = 0.501426509658179157416722893785962;
        const double c1 = 0.053145049844816947353249671631398, c2 = 0.310352451033784405416607733956552,
                     c3 = 0.636502499121398647230142594411050;
        const double wa = 0.050844906370206816920936809106869, wb = 0.116786275726379366030690538687593,
                     wc = 0.082851075618373575193553456420442;
        const double bc[12][3] = {{b1, a1, a1}, {a1, b1, a1}, {a1, a1, b1}, {b2, a2, a2}, {a2, b2, a2},
                                  {a2, a2, b2}, {c1, c2, c3}, {c1, c3, c2}, {c2, c1, c3}, {c2, c3, c1},
                                  {c3, c1, c2}, {c3, c2, c1}};
        const double wq[12] = {wa, wa, wa, wb, wb, wb, wc, wc, wc, wc, wc, wc};
        n = 12;
        for (int q = 0; q < 12; ++q) { xi[q] = bc[q][1]; eta[q] = bc[q][2]; w[q] = wq[q]; }
    }
}

// Edge rule: 3-point Gauss with the reference's literals (src/triangleinfo.cpp:165,
// 209-217,227-232) for P1/P2; 4-point Gauss-Legendre for P3.  Point j of an edge is
// sa[j]*v_a + sb[j]*v_b with v_a the edge's first vertex.
void edge_rule(int P, int& n, double* sa, double* sb, double* w) {
    if (P <= 2) {
        double cp = 0.5 + std::sqrt(15.0) / 10.0;
        n = 3;
        sa[0] = cp; sb[0] = 1 - cp; sa[1] = 0.5; sb[1] = 0.5; sa[2] = 1 - cp; sb[2] = cp;
        w[0] = 0.55555555555556; w[1] = 0.88888888888889; w[2] = 0.55555555555556;
    } else {
        const double t0 = 0.861136311594052575223946488892809, t1 = 0.339981043584856264802665759103245;
        const double v0 = 0.347854845137453857373063949221999, v1 = 0.652145154862546142626936050778001;
        const double tt[4] = {t0, t1, -t1, -t0}, ww[4] = {v0, v1, v1, v0};
        n = 4;
        for (int j = 0; j < 4; ++j) { sa[j] = 0.5 * (1 + tt[j]); sb[j] = 0.5 * (1 - tt[j]); w[j] = ww[j]; }
    }
}

const double RV[3][2] = {{0, 0}, {1, 0}, {0, 1}};  // reference vertices

long double mono_ld(long double x, long double y, int m) {
    long double r = 1;
    for (int i = 0; i < MX[m]; ++i) r *= x;
    for (int i = 0; i < MY[m]; ++i) r *= y;
    return r;
}

void build_ref_basis(int P, RefBasis& B) {
    B.P = P; B.Np = np_of(P); B.Nqv = nqv_of(P); B.Nqe = nqe_of(P);
    const int Np = B.Np;
    // Gram matrix of the monomials in the mean inner product, exact
    long double G[MAXNP][MAXNP];
    for (int a = 0; a < Np; ++a)
        for (int b = 0; b < Np; ++b) {
            int px = MX[a] + MX[b], py = MY[a] + MY[b];
            G[a][b] = 2.0L * factl(px) * factl(py) / factl(px + py + 2);
        }
    // Cholesky G = L L^T; psi = L^{-1} m
    long double L[MAXNP][MAXNP] = {};
    for (int i = 0; i < Np; ++i)
        for (int j = 0; j <= i; ++j) {
            long double s = G[i][j];
            for (int k = 0; k < j; ++k) s -= L[i][k] * L[j][k];
            L[i][j] = (i == j) ? std::sqrt(s) : s / L[j][j];
        }
    long double Li[MAXNP][MAXNP] = {};
    for (int i = 0; i < Np; ++i) {
        Li[i][i] = 1.0L / L[i][i];
        for (int j = 0; j < i; ++j) {
            long double s = 0;
            for (int k = j; k < i; ++k) s += L[i][k] * Li[k][j];
            Li[i][j] = -s / L[i][i];
        }
    }
    for (int i = 0; i < MAXNP; ++i)
        for (int m = 0; m < MAXNP; ++m) B.C[i][m] = (i < Np && m < Np) ? Li[i][m] : 0.0L;

    DevTables& T = B.tab;
    std::memset(&T, 0, sizeof T);
    int nq;
    double xi[MAXQV], eta[MAXQV], wv[MAXQV];
    volume_rule(P, nq, xi, eta, wv);
    for (int q = 0; q < nq; ++q) {
        T.wv[q] = wv[q];
        for (int i = 0; i < Np; ++i) {
            T.Vv[q][i] = psi_value(B, i, xi[q], eta[q]);
            psi_grad(B, i, xi[q], eta[q], &T.Dxi[q][i], &T.Deta[q][i]);
        }
    }
    for (int i = 0; i < Np; ++i) {
        long double m = 0;
        for (int q = 0; q < nq; ++q) m += (long double)wv[q] * T.Vv[q][i] * T.Vv[q][i];
        T.minv[i] = (double)(1.0L / m);
    }
    int ne;
    double sa[MAXQE], sb[MAXQE], we[MAXQE];
    edge_rule(P, ne, sa, sb, we);
    for (int j = 0; j < ne; ++j) T.we[j] = we[j];
    for (int k = 0; k < 3; ++k) {
        int a = (k + 1) % 3, b = (k + 2) % 3;
        for (int j = 0; j < ne; ++j) {
            double px = sa[j] * RV[a][0] + sb[j] * RV[b][0];
            double py = sa[j] * RV[a][1] + sb[j] * RV[b][1];
            for (int i = 0; i < Np; ++i) T.Ve[k][j][i] = psi_value(B, i, px, py);
        }
    }
    // psi_0 = 1 exactly (G_00 = 1 in the mean inner product): the kernels take mode 0 as the
    // constant term of a point value without multiplying it by the table
    for (int q = 0; q < nq; ++q) T.Vv[q][0] = 1.0;
    for (int k = 0; k < 3; ++k)
        for (int j = 0; j < ne; ++j) T.Ve[k][j][0] = 1.0;
    // limiter: P1 modes at the edge midpoints and the least-squares slope recovery
    double A[3][2];
    for (int k = 0; k < 3; ++k) {
        int a = (k + 1) % 3, b = (k + 2) % 3;
        double mx = 0.5 * (RV[a][0] + RV[b][0]), my = 0.5 * (RV[a][1] + RV[b][1]);
        for (int i = 0; i < 3; ++i) T.Vmid[k][i] = psi_value(B, i, mx, my);
        A[k][0] = T.Vmid[k][1]; A[k][1] = T.Vmid[k][2];
    }
    double a11 = 0, a12 = 0, a22 = 0;
    for (int k = 0; k < 3; ++k) { a11 += A[k][0] * A[k][0]; a12 += A[k][0] * A[k][1]; a22 += A[k][1] * A[k][1]; }
    double det = a11 * a22 - a12 * a12;
    for (int k = 0; k < 3; ++k) {
        T.recon[0][k] = (a22 * A[k][0] - a12 * A[k][1]) / det;
        T.recon[1][k] = (a11 * A[k][1] - a12 * A[k][0]) / det;
    }
    // the edge rule applied to a constant flux (same operation order as a rolled accumulation
    // over the points)
    for (int i = 0; i < Np; ++i) {
        for (int k = 0; k < 3; ++k) {
            double se = 0.0;
            for (int j = 0; j < ne; ++j) se += T.we[j] * T.Ve[k][j][i];
            T.Se[k][i] = se;
        }
    }
}

}  // namespace

double psi_value(const RefBasis& B, int i, double xi, double eta) {
    long double r = 0;
    for (int m = 0; m < B.Np; ++m) r += B.C[i][m] * mono_ld(xi, eta, m);
    return (double)r;
}

void psi_grad(const RefBasis& B, int i, double xi, double eta, double* gx, double* ge) {
    long double rx = 0, ry = 0;
    for (int m = 0; m < B.Np; ++m) {
        if (MX[m] > 0) {
            long double t = MX[m];
            for (int a = 0; a < MX[m] - 1; ++a) t *= xi;
            for (int b = 0; b < MY[m]; ++b) t *= eta;
            rx += B.C[i][m] * t;
        }
        if (MY[m] > 0) {
            long double t = MY[m];
            for (int a = 0; a < MX[m]; ++a) t *= xi;
            for (int b = 0; b < MY[m] - 1; ++b) t *= eta;
            ry += B.C[i][m] * t;
        }
    }
    *gx = (double)rx;
    *ge = (double)ry;
}

const RefBasis& ref_basis(int P) {
    static RefBasis bases[3];
    static std::once_flag once[3];
    std::call_once(once[P - 1], [P] { build_ref_basis(P, bases[P - 1]); });
    return bases[P - 1];
}

const TriRule& projection_rule() {
    static TriRule R;
    static std::once_flag once;
    std::call_once(once, [] {
        const int n = 6;
        long double t[n], w[n];
        for (int i = 0; i < n; ++i) {  // Gauss-Legendre on [0,1] by Newton on P_n
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
        R.n = n * n;
        for (int a = 0; a < n; ++a)
            for (int b = 0; b < n; ++b) {
                R.xi.push_back((double)t[a]);
                R.eta.push_back((double)(t[b] * (1 - t[a])));
                R.w.push_back((double)(w[a] * w[b] * (1 - t[a]) * 2.0L));  // sums to 1
            }
    });
    return R;
}

// psi_j at the points of the projection rule (element-independent), [q][j]
const std::vector<double>& projection_psi(int P) {
    static std::vector<double> tab[3];
    static std::once_flag once[3];
    std::call_once(once[P - 1], [P] {
        const RefBasis& B = ref_basis(P);
        const TriRule& R = projection_rule();
        tab[P - 1].resize((size_t)R.n * B.Np);
        for (int q = 0; q < R.n; ++q)
            for (int j = 0; j < B.Np; ++j) tab[P - 1][(size_t)q * B.Np + j] = psi_value(B, j, R.xi[q], R.eta[q]);
    });
    return tab[P - 1];
}

void element_geometry(const double* x, const double* y, double* g) {
    double detJ = (x[1] - x[0]) * (y[2] - y[0]) - (x[2] - x[0]) * (y[1] - y[0]);
    g[0] = (y[2] - y[0]) / detJ;    // xi_x
    g[1] = -(x[2] - x[0]) / detJ;   // xi_y
    g[2] = -(y[1] - y[0]) / detJ;   // eta_x
    g[3] = (x[1] - x[0]) / detJ;    // eta_y
    // area, lengths and outward normals as src/triangleinfo.cpp:171-189
    g[4] = 0.5 * std::fabs((x[2] - x[0]) * (y[1] - y[0]) - (y[2] - y[0]) * (x[1] - x[0]));
    double l0 = std::sqrt((x[1] - x[2]) * (x[1] - x[2]) + (y[1] - y[2]) * (y[1] - y[2]));
    double l1 = std::sqrt((x[0] - x[2]) * (x[0] - x[2]) + (y[0] - y[2]) * (y[0] - y[2]));
    double l2 = std::sqrt((x[0] - x[1]) * (x[0] - x[1]) + (y[0] - y[1]) * (y[0] - y[1]));
    g[5] = l0; g[6] = l1; g[7] = l2;
    g[8] = (y[2] - y[1]) / l0;  g[9] = (x[1] - x[2]) / l0;
    g[10] = (y[0] - y[2]) / l1; g[11] = (x[2] - x[0]) / l1;
    g[12] = (y[1] - y[0]) / l2; g[13] = (x[0] - x[1]) / l2;
}

double ElemBasis::value(int i, double x, double y) const {
    double X = (x - xc) * isq, Y = (y - yc) * isq;
    double r = 0;
    for (int m = 0; m < Np; ++m) {
        double t = c[i][m];
        if (t == 0.0) continue;
        for (int a = 0; a < MX[m]; ++a) t *= X;
        for (int b = 0; b < MY[m]; ++b) t *= Y;
        r += t;
    }
    return r;
}

// Classical Gram-Schmidt of the scaled monomials in the element's L2 product
// (evaluated exactly with the collapsed rule).  Produces the same functions as the
// reference's successive pivoted solves: phi_k = m_k + (lower monomials), orthogonal
// to phi_0..phi_{k-1}.
void element_basis(int P, const double* x, const double* y, ElemBasis& B) {
    const int Np = np_of(P);
    B.Np = Np;
    double area = 0.5 * std::fabs((x[2] - x[0]) * (y[1] - y[0]) - (y[2] - y[0]) * (x[1] - x[0]));
    B.xc = (x[0] + x[1] + x[2]) / 3.0;
    B.yc = (y[0] + y[1] + y[2]) / 3.0;
    B.isq = 1.0 / std::sqrt(area);
    const TriRule& R = projection_rule();
    std::vector<double> mv((size_t)Np * R.n), pv((size_t)Np * R.n);
    for (int q = 0; q < R.n; ++q) {
        double px = x[0] + (x[1] - x[0]) * R.xi[q] + (x[2] - x[0]) * R.eta[q];
        double py = y[0] + (y[1] - y[0]) * R.xi[q] + (y[2] - y[0]) * R.eta[q];
        double X = (px - B.xc) * B.isq, Y = (py - B.yc) * B.isq;
        for (int m = 0; m < Np; ++m) {
            double t = 1;
            for (int a = 0; a < MX[m]; ++a) t *= X;
            for (int b = 0; b < MY[m]; ++b) t *= Y;
            mv[(size_t)m * R.n + q] = t;
        }
    }
    double nrm[MAXNP];
    for (int k = 0; k < Np; ++k) {
        for (int m = 0; m < MAXNP; ++m) B.c[k][m] = (m == k) ? 1.0 : 0.0;
        for (int q = 0; q < R.n; ++q) pv[(size_t)k * R.n + q] = mv[(size_t)k * R.n + q];
        for (int l = 0; l < k; ++l) {
            double ip = 0;
            for (int q = 0; q < R.n; ++q) ip += R.w[q] * mv[(size_t)k * R.n + q] * pv[(size_t)l * R.n + q];
            double f = ip / nrm[l];
            for (int m = 0; m < Np; ++m) B.c[k][m] -= f * B.c[l][m];
            for (int q = 0; q < R.n; ++q) pv[(size_t)k * R.n + q] -= f * pv[(size_t)l * R.n + q];
        }
        double s = 0;
        for (int q = 0; q < R.n; ++q) s += R.w[q] * pv[(size_t)k * R.n + q] * pv[(size_t)k * R.n + q];
        nrm[k] = s;
    }
}

void element_to_native(int P, const double* x, const double* y, const double* c, double* d) {
    const int Np = np_of(P);
    ElemBasis EB;
    element_basis(P, x, y, EB);
    const TriRule& R = projection_rule();
    const std::vector<double>& ps = projection_psi(P);
    double num[NVAR][MAXNP] = {}, den[MAXNP] = {};
    for (int q = 0; q < R.n; ++q) {
        double px = x[0] + (x[1] - x[0]) * R.xi[q] + (x[2] - x[0]) * R.eta[q];
        double py = y[0] + (y[1] - y[0]) * R.xi[q] + (y[2] - y[0]) * R.eta[q];
        double u[NVAR] = {0, 0, 0, 0};
        for (int i = 0; i < Np; ++i) {
            double f = EB.value(i, px, py);
            for (int v = 0; v < NVAR; ++v) u[v] += c[v * Np + i] * f;
        }
        for (int j = 0; j < Np; ++j) {
            double s = ps[(size_t)q * Np + j];
            for (int v = 0; v < NVAR; ++v) num[v][j] += R.w[q] * u[v] * s;
            den[j] += R.w[q] * s * s;
        }
    }
    for (int v = 0; v < NVAR; ++v)
        for (int j = 0; j < Np; ++j) d[v * Np + j] = num[v][j] / den[j];
}

void native_to_element(int P, const double* x, const double* y, const double* d, double* c) {
    const int Np = np_of(P);
    ElemBasis EB;
    element_basis(P, x, y, EB);
    const TriRule& R = projection_rule();
    const std::vector<double>& ps = projection_psi(P);
    double num[NVAR][MAXNP] = {}, den[MAXNP] = {};
    for (int q = 0; q < R.n; ++q) {
        double px = x[0] + (x[1] - x[0]) * R.xi[q] + (x[2] - x[0]) * R.eta[q];
        double py = y[0] + (y[1] - y[0]) * R.xi[q] + (y[2] - y[0]) * R.eta[q];
        double u[NVAR] = {0, 0, 0, 0};
        for (int j = 0; j < Np; ++j) {
            double s = ps[(size_t)q * Np + j];
            for (int v = 0; v < NVAR; ++v) u[v] += d[v * Np + j] * s;
        }
        for (int i = 0; i < Np; ++i) {
            double f = EB.value(i, px, py);
            for (int v = 0; v < NVAR; ++v) num[v][i] += R.w[q] * u[v] * f;
            den[i] += R.w[q] * f * f;
        }
    }
    for (int v = 0; v < NVAR; ++v)
        for (int i = 0; i < Np; ++i) c[v * Np + i] = num[v][i] / den[i];
}

void ic_state(int kind, const double* p, int np, double x, double y, double* U) {
    if (kind == RKDG_IC_VORTEX) {
        // isentropic vortex, free stream rho = p = 1 (temperature perturbation form)
        double g = p[0], beta = p[1], uinf = p[2], vinf = p[3];
        double Lx = np > 6 ? p[6] : 0, Ly = np > 7 ? p[7] : 0, t = np > 8 ? p[8] : 0;
        double cx = p[4] + uinf * t, cy = p[5] + vinf * t;
        int img = (Lx > 0 && Ly > 0) ? 1 : 0;
        if (img) { cx -= Lx * std::floor(cx / Lx); cy -= Ly * std::floor(cy / Ly); }
        double du = 0, dv = 0, dT = 0;
        for (int ix = -img; ix <= img; ++ix)
            for (int iy = -img; iy <= img; ++iy) {
                double rx = x - (cx + ix * Lx), ry = y - (cy + iy * Ly);
                double e = std::exp(0.5 * (1.0 - (rx * rx + ry * ry)));
                du -= beta / (2 * M_PI) * e * ry;
                dv += beta / (2 * M_PI) * e * rx;
                dT -= (g - 1) * beta * beta / (8 * g * M_PI * M_PI) * e * e;
            }
        double rho = std::pow(1.0 + dT, 1.0 / (g - 1));
        double u = uinf + du, v = vinf + dv, pr = std::pow(rho, g);
        U[0] = rho; U[1] = rho * u; U[2] = rho * v; U[3] = pr / (g - 1) + 0.5 * rho * (u * u + v * v);
    } else if (kind == RKDG_IC_RIEMANN4) {
        double g = p[0];
        int q = (x >= p[1]) ? (y >= p[2] ? 0 : 3) : (y >= p[2] ? 1 : 2);
        const double* s = p + 3 + 4 * q;
        U[0] = s[0]; U[1] = s[0] * s[1]; U[2] = s[0] * s[2];
        U[3] = s[3] / (g - 1) + 0.5 * s[0] * (s[1] * s[1] + s[2] * s[2]);
    } else {
        // free stream as CCUDARkdgSolver::initRKDG (src/cudarkdgsolver.cu:330-355)
        double g = p[0], alpha = p[1], mach = p[2], rhoref = p[3], pref = p[4];
        double ut = std::sqrt(g * pref / rhoref) * mach;
        U[0] = rhoref;
        U[1] = rhoref * (ut * std::cos(alpha));
        U[2] = rhoref * (ut * std::sin(alpha));
        U[3] = rhoref * (ut * ut) / 2 + pref / (rhoref * (g - 1));
    }
}

// Area of a polygon (shoelace) and its clip to the half-plane s * (c - c0) >= 0 of one
// coordinate (axis 0: x, 1: y; Sutherland-Hodgman, at most n + 1 vertices out).
static double poly_area(const double* px, const double* py, int n) {
    double a = 0.0;
    for (int i = 0; i < n; ++i) {
        const int j = (i + 1) % n;
        a += px[i] * py[j] - px[j] * py[i];
    }
    return 0.5 * std::fabs(a);
}

static int clip_half(const double* px, const double* py, int n, int axis, double c0, double s, double* ox,
                     double* oy) {
    int m = 0;
    for (int i = 0; i < n; ++i) {
        const int j = (i + 1) % n;
        const double ci = s * ((axis ? py[i] : px[i]) - c0), cj = s * ((axis ? py[j] : px[j]) - c0);
        if (ci >= 0) { ox[m] = px[i]; oy[m] = py[i]; ++m; }
        if ((ci >= 0) != (cj >= 0)) {
            const double t = ci / (ci - cj);
            ox[m] = px[i] + t * (px[j] - px[i]);
            oy[m] = py[i] + t * (py[j] - py[i]);
            ++m;
        }
    }
    return m;
}

// Exact cell average of the four-quadrant data over one triangle: the areas of the
// triangle's parts in the four quadrants about (xs, ys) (p as ic_state's RKDG_IC_RIEMANN4).
void riemann4_average(const double* x, const double* y, const double* p, double* U) {
    double ax[8], ay[8], bx[8], by[8];
    double area[4] = {0, 0, 0, 0};
    const double xs = p[1], ys = p[2];
    for (int sx = 0; sx < 2; ++sx) {  // sx 0: x >= xs, 1: x < xs
        const int na = clip_half(x, y, 3, 0, xs, sx ? -1.0 : 1.0, ax, ay);
        for (int sy = 0; sy < 2; ++sy) {  // sy 0: y >= ys, 1: y < ys
            const int nb = na >= 3 ? clip_half(ax, ay, na, 1, ys, sy ? -1.0 : 1.0, bx, by) : 0;
            const int q = sx == 0 ? (sy == 0 ? 0 : 3) : (sy == 0 ? 1 : 2);
            area[q] = nb >= 3 ? poly_area(bx, by, nb) : 0.0;
        }
    }
    const double tot = area[0] + area[1] + area[2] + area[3];
    for (int v = 0; v < NVAR; ++v) U[v] = 0.0;
    for (int q = 0; q < 4; ++q) {
        double Uq[NVAR];
        const double* s = p + 3 + 4 * q;
        const double g = p[0];
        Uq[0] = s[0]; Uq[1] = s[0] * s[1]; Uq[2] = s[0] * s[2];
        Uq[3] = s[3] / (g - 1) + 0.5 * s[0] * (s[1] * s[1] + s[2] * s[2]);
        for (int v = 0; v < NVAR; ++v) U[v] += area[q] / tot * Uq[v];
    }
}

void project_native(int P, const double* x, const double* y, int kind, const double* p, int np, double* d) {
    const int Np = np_of(P);
    if (kind == RKDG_IC_RIEMANN4) {
        // discontinuous four-quadrant data start from their exact cell averages: the L2
        // projection of a jump onto P_k over- and undershoots and can make the pressure
        // negative at flux points (psi_0 = 1, so mode 0 is the mean)
        double U[NVAR];
        riemann4_average(x, y, p, U);
        for (int v = 0; v < NVAR; ++v)
            for (int j = 0; j < Np; ++j) d[v * Np + j] = j == 0 ? U[v] : 0.0;
        return;
    }
    const TriRule& R = projection_rule();
    const std::vector<double>& ps = projection_psi(P);
    double num[NVAR][MAXNP] = {}, den[MAXNP] = {};
    for (int q = 0; q < R.n; ++q) {
        double px = x[0] + (x[1] - x[0]) * R.xi[q] + (x[2] - x[0]) * R.eta[q];
        double py = y[0] + (y[1] - y[0]) * R.xi[q] + (y[2] - y[0]) * R.eta[q];
        double U[NVAR];
        ic_state(kind, p, np, px, py, U);
        for (int j = 0; j < Np; ++j) {
            double s = ps[(size_t)q * Np + j];
            for (int v = 0; v < NVAR; ++v) num[v][j] += R.w[q] * U[v] * s;
            den[j] += R.w[q] * s * s;
        }
    }
    for (int v = 0; v < NVAR; ++v)
        for (int j = 0; j < Np; ++j) d[v * Np + j] = num[v][j] / den[j];
}

// Cockburn-Shu choice of two neighbour barycentres per edge midpoint: m_k - b0 =
// a1 (b_k1 - b0) + a2 (b_k2 - b0) with a1, a2 >= 0 and k1 = k.  Pairs are packed as
// 2 bits each, edge k at bits 4k..4k+3.  nbx/nby are the neighbour barycentres in this
// element's frame (a boundary side uses the mirror image of b0).
void limiter_geometry(const double* x, const double* y, const double nbx[3], const double nby[3],
                      double* lim, int* pairs) {
    double b0x = (x[0] + x[1] + x[2]) / 3.0, b0y = (y[0] + y[1] + y[2]) / 3.0;
    int pk = 0;
    double lmax = 0;
    for (int k = 0; k < 3; ++k) {
        int a = (k + 1) % 3, b = (k + 2) % 3;
        double l = std::hypot(x[a] - x[b], y[a] - y[b]);
        lmax = std::max(lmax, l);
        double rx = 0.5 * (x[a] + x[b]) - b0x, ry = 0.5 * (y[a] + y[b]) - b0y;
        bool found = false;
        for (int off = 1; off <= 2 && !found; ++off) {
            int k1 = k, k2 = (k + off) % 3;
            double ax = nbx[k1] - b0x, ay = nby[k1] - b0y, cx = nbx[k2] - b0x, cy = nby[k2] - b0y;
            double det = ax * cy - ay * cx;
            if (std::fabs(det) < 1e-300) continue;
            double a1 = (rx * cy - ry * cx) / det, a2 = (ax * ry - ay * rx) / det;
            if (a1 >= -1e-12 && a2 >= -1e-12) {
                lim[2 * k] = std::max(a1, 0.0);
                lim[2 * k + 1] = std::max(a2, 0.0);
                pk |= (k1 | (k2 << 2)) << (4 * k);
                found = true;
            }
        }
        if (!found) {
            lim[2 * k] = 0.5;
            lim[2 * k + 1] = 0.0;
            pk |= (k | (k << 2)) << (4 * k);
        }
    }
    lim[6] = lmax * lmax;
    *pairs = pk;
}

}  // namespace rkdg
