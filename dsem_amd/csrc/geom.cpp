// Host-side set-up of operators and geometry for the DSEM library.
//
// Restated from the reference (FP64, same operation order, no fused
// multiply-add so that the values equal the reference's bit for bit):
//   Chebyshev points              src/Grid_Routines/stac3m_geomLib.f90:7-47
//   Lagrange basis and derivative src/Independentroutines/baselib.f90:18-80
//   dermat / intrpmat             src/Spectral/stac3m/matrices.f90:144-207
//   hex8 map, gradient, inverse   src/Grid_Routines/GeomLib3D.f90:60-140,685-734
//   set_geometry (hex8 branch)    src/Grid_Routines/set_domain_geometry.f90:77-345
#pragma clang fp contract(off)

#include "geom.h"

#include <cfloat>
#include <cmath>

namespace dsem {

static double lagrange(int k, double x, int n, const double *z)
{
    double r;
    if (k == 1) {
        r = (x - z[1]) / (z[0] - z[1]);
        for (int j = 3; j <= n; ++j) r = r * (x - z[j - 1]) / (z[0] - z[j - 1]);
    } else {
        r = (x - z[0]) / (z[k - 1] - z[0]);
        for (int j = 2; j <= k - 1; ++j) r = r * (x - z[j - 1]) / (z[k - 1] - z[j - 1]);
        for (int j = k + 1; j <= n; ++j) r = r * (x - z[j - 1]) / (z[k - 1] - z[j - 1]);
    }
    return r;
}

static double lagrange_der(int k, double x, int n, const double *z)
{
    double hp = 0.0;
    for (int l = 1; l <= n; ++l) {
        if (l == k) continue;
        double poly = 1.0;
        for (int m = 1; m <= n; ++m) {
            if (m == l || m == k) continue;
            poly = poly * (x - z[m - 1]) / (z[k - 1] - z[m - 1]);
        }
        hp = hp + poly / (z[k - 1] - z[l - 1]);
    }
    return hp;
}

bool make_ops(int p, Ops &o)
{
    if (p < 1 || p > 8) return false;
    const double pi = 3.141592653589793238462643;
    o.p = p;
    o.ncg = p + 1;
    o.nc = p + 2;
    const int ncg = o.ncg, nc = o.nc;
    for (int n = 1; n <= nc; ++n) o.cx[n - 1] = 0.5 * (1.0 - std::cos((n - 1) * pi / (nc - 1)));
    for (int n = 1; n <= ncg; ++n)
        o.cxg[n - 1] = 0.5 * (1.0 - std::cos((2.0 * n - 1.0) * pi / (2.0 * (ncg - 1.0) + 2.0)));
    // derivative of the Lobatto basis at the Gauss points; entry (k,k) of the
    // row is minus the running sum of the others, entries below eps are zero
    for (int k = 1; k <= ncg; ++k) {
        double sum = 0.0;
        for (int j = 1; j <= nc; ++j) {
            if (j == k) continue;
            double b = lagrange_der(j, o.cxg[k - 1], nc, o.cx);
            if (std::fabs(b) < DBL_EPSILON) b = 0.0;
            o.D[(k - 1) * nc + (j - 1)] = b;
            sum = sum - b;
            o.D[(k - 1) * nc + (k - 1)] = sum;
        }
    }
    for (int j = 1; j <= nc; ++j)
        for (int k = 1; k <= ncg; ++k) o.B[(j - 1) * ncg + (k - 1)] = lagrange(k, o.cx[j - 1], ncg, o.cxg);
    for (int k = 0; k < ncg; ++k)
        for (int k2 = 0; k2 < ncg; ++k2) {
            long double s = 0.0L;
            for (int j = 1; j < nc - 1; ++j) s += (long double)o.D[k * nc + j] * (long double)o.B[j * ncg + k2];
            o.M[k * ncg + k2] = (double)s;
        }
    return true;
}

static void grad_hex8(const double *c, const double *X, double g[3][3])
{
    for (int i = 0; i < 3; ++i) {
        g[i][0] = -(c[0 * 3 + i] * (1.0 - X[1]) * (1.0 - X[2])) + c[1 * 3 + i] * (1.0 - X[1]) * (1.0 - X[2])
                + c[2 * 3 + i] * X[1] * (1.0 - X[2]) - c[3 * 3 + i] * X[1] * (1.0 - X[2])
                - c[4 * 3 + i] * (1.0 - X[1]) * X[2] + c[5 * 3 + i] * (1.0 - X[1]) * X[2]
                + c[6 * 3 + i] * X[1] * X[2] - c[7 * 3 + i] * X[1] * X[2];
        g[i][1] = -(c[0 * 3 + i] * (1.0 - X[0]) * (1.0 - X[2])) - c[1 * 3 + i] * X[0] * (1.0 - X[2])
                + c[2 * 3 + i] * X[0] * (1.0 - X[2]) + c[3 * 3 + i] * (1.0 - X[0]) * (1.0 - X[2])
                - c[4 * 3 + i] * (1.0 - X[0]) * X[2] - c[5 * 3 + i] * X[0] * X[2]
                + c[6 * 3 + i] * X[0] * X[2] + c[7 * 3 + i] * (1.0 - X[0]) * X[2];
        g[i][2] = -(c[0 * 3 + i] * (1.0 - X[0]) * (1.0 - X[1])) - c[1 * 3 + i] * X[0] * (1.0 - X[1])
                - c[2 * 3 + i] * X[0] * X[1] - c[3 * 3 + i] * (1.0 - X[0]) * X[1]
                + c[4 * 3 + i] * (1.0 - X[0]) * (1.0 - X[1]) + c[5 * 3 + i] * X[0] * (1.0 - X[1])
                + c[6 * 3 + i] * X[0] * X[1] + c[7 * 3 + i] * (1.0 - X[0]) * X[1];
    }
}

static void map_hex8(const double *c, const double *X, double *x)
{
    for (int j = 0; j < 3; ++j)
        x[j] = c[0 * 3 + j] * (1.0 - X[0]) * (1.0 - X[1]) * (1.0 - X[2]) + c[1 * 3 + j] * X[0] * (1.0 - X[1]) * (1.0 - X[2])
             + c[2 * 3 + j] * X[0] * X[1] * (1.0 - X[2]) + c[3 * 3 + j] * (1.0 - X[0]) * X[1] * (1.0 - X[2])
             + c[4 * 3 + j] * (1.0 - X[0]) * (1.0 - X[1]) * X[2] + c[5 * 3 + j] * X[0] * (1.0 - X[1]) * X[2]
             + c[6 * 3 + j] * X[0] * X[1] * X[2] + c[7 * 3 + j] * (1.0 - X[0]) * X[1] * X[2];
}

// inverse metric gi[i][j] = d X_i / d x_j; returns the Jacobian determinant
static double invert(const double g[3][3], double gi[3][3])
{
    const double small = 100.0 * DBL_EPSILON;
    double J = g[0][0] * g[1][1] * g[2][2] - g[0][1] * g[1][0] * g[2][2] - g[0][0] * g[1][2] * g[2][1]
             + g[0][2] * g[1][0] * g[2][1] + g[0][1] * g[1][2] * g[2][0] - g[0][2] * g[1][1] * g[2][0];
    gi[0][0] = (g[1][1] * g[2][2] - g[1][2] * g[2][1]) / J;
    gi[1][0] = (-(g[1][0] * g[2][2]) + g[1][2] * g[2][0]) / J;
    gi[2][0] = (g[1][0] * g[2][1] - g[2][0] * g[1][1]) / J;
    gi[0][1] = (-(g[0][1] * g[2][2]) + g[0][2] * g[2][1]) / J;
    gi[1][1] = (g[0][0] * g[2][2] - g[0][2] * g[2][0]) / J;
    gi[2][1] = (-(g[0][0] * g[2][1]) + g[0][1] * g[2][0]) / J;
    gi[0][2] = (g[0][1] * g[1][2] - g[1][1] * g[0][2]) / J;
    gi[1][2] = (-(g[0][0] * g[1][2]) + g[0][2] * g[1][0]) / J;
    gi[2][2] = (g[0][0] * g[1][1] - g[0][1] * g[1][0]) / J;
    for (int m = 0; m < 3; ++m)
        for (int n = 0; n < 3; ++n)
            if (std::fabs(gi[n][m]) <= small) gi[n][m] = 0.0;
    return J;
}

// A hex8 element is a parallelepiped when its bilinear and trilinear map
// coefficients vanish; then every metric term is constant in the element.
static bool is_affine(const double *c)
{
    double scale = 0.0;
    for (int a = 0; a < 8; ++a)
        for (int i = 0; i < 3; ++i) scale = std::fmax(scale, std::fabs(c[a * 3 + i] - c[i]));
    const double tol = 1e-12 * (scale > 0 ? scale : 1.0);
    for (int i = 0; i < 3; ++i) {
        const double c1 = c[0 * 3 + i], c2 = c[1 * 3 + i], c3 = c[2 * 3 + i], c4 = c[3 * 3 + i];
        const double c5 = c[4 * 3 + i], c6 = c[5 * 3 + i], c7 = c[6 * 3 + i], c8 = c[7 * 3 + i];
        const double xy = c1 - c2 + c3 - c4, xz = c1 - c2 - c5 + c6, yz = c1 - c4 - c5 + c8;
        const double xyz = -c1 + c2 - c3 + c4 + c5 - c6 + c7 - c8;
        if (std::fabs(xy) > tol || std::fabs(xz) > tol || std::fabs(yz) > tol || std::fabs(xyz) > tol) return false;
    }
    return true;
}

bool element_geometry(const Ops &o, int K, const double *corners, Geometry &g)
{
    const int ncg = o.ncg, nc = o.nc, np = ncg * ncg * ncg, ns = nc * ncg * ncg;
    g.K = K;
    g.p = o.p;
    g.xg.assign((size_t)K * np * 3, 0.0);
    g.jacob.assign((size_t)K * np, 0.0);
    g.gmetg.assign((size_t)K * 9 * np, 0.0);
    for (int d = 0; d < 3; ++d) g.gmet[d].assign((size_t)K * 3 * ns, 0.0);
    bool affine = true;
#pragma omp parallel for schedule(static) reduction(&& : affine)
    for (int e = 0; e < K; ++e) {
        const double *c = corners + (size_t)e * 24;
        affine = affine && is_affine(c);
        double X[3], gr[3][3], gi[3][3];
        for (int l = 0; l < ncg; ++l)
            for (int m = 0; m < ncg; ++m)
                for (int n = 0; n < ncg; ++n) {
                    const int pt = n + ncg * (m + ncg * l);
                    X[0] = o.cxg[n]; X[1] = o.cxg[m]; X[2] = o.cxg[l];
                    map_hex8(c, X, &g.xg[((size_t)e * np + pt) * 3]);
                    grad_hex8(c, X, gr);
                    const double J = invert(gr, gi);
                    g.jacob[(size_t)e * np + pt] = 1.0 / J;
                    for (int a = 0; a < 3; ++a)
                        for (int b = 0; b < 3; ++b) g.gmetg[((size_t)e * 9 + a * 3 + b) * np + pt] = gi[a][b];
                }
        for (int d = 0; d < 3; ++d) {
            for (int b = 0; b < ncg; ++b)
                for (int a = 0; a < ncg; ++a)
                    for (int j = 0; j < nc; ++j) {
                        int s;
                        if (d == 0) { X[0] = o.cx[j]; X[1] = o.cxg[a]; X[2] = o.cxg[b]; s = j + nc * (a + ncg * b); }
                        else if (d == 1) { X[0] = o.cxg[a]; X[1] = o.cx[j]; X[2] = o.cxg[b]; s = a + ncg * (j + nc * b); }
                        else { X[0] = o.cxg[a]; X[1] = o.cxg[b]; X[2] = o.cx[j]; s = a + ncg * (b + ncg * j); }
                        grad_hex8(c, X, gr);
                        const double J = invert(gr, gi);
                        for (int q = 0; q < 3; ++q) g.gmet[d][((size_t)e * 3 + q) * ns + s] = J * gi[d][q];
                    }
        }
    }
    g.affine = affine;
    return true;
}

void mortar_geometry(const Ops &o, int M, const double *corners, const int32_t *m_elem,
                     const int32_t *m_face, std::vector<double> &nhat, std::vector<double> &norm)
{
    static const double end_vals[6] = {0.0, 1.0, 0.0, 1.0, 1.0, 0.0};
    const int ncg = o.ncg, nf = ncg * ncg;
    nhat.assign((size_t)M * 3 * nf, 0.0);
    norm.assign((size_t)M * nf, 0.0);
#pragma omp parallel for schedule(static)
    for (int mm = 0; mm < M; ++mm) {
        const int e = m_elem[2 * mm], f = m_face[2 * mm];
        const double *c = corners + (size_t)e * 24;
        int fixed, a0, a1, row;
        if (f == 1 || f == 2) { fixed = 1; a0 = 0; a1 = 2; row = 1; }
        else if (f == 3 || f == 5) { fixed = 2; a0 = 0; a1 = 1; row = 2; }
        else { fixed = 0; a0 = 1; a1 = 2; row = 0; }
        double X[3], gr[3][3], gi[3][3];
        for (int j = 0; j < ncg; ++j)
            for (int i = 0; i < ncg; ++i) {
                X[fixed] = end_vals[f - 1];
                X[a0] = o.cxg[i];
                X[a1] = o.cxg[j];
                grad_hex8(c, X, gr);
                const double J = invert(gr, gi);
                const double sqr = std::sqrt(gi[row][0] * gi[row][0] + gi[row][1] * gi[row][1] + gi[row][2] * gi[row][2]);
                const int q = i + ncg * j;
                for (int k = 0; k < 3; ++k) nhat[((size_t)mm * 3 + k) * nf + q] = gi[row][k] / sqr;
                norm[(size_t)mm * nf + q] = J * sqr;
            }
    }
}

// Gauss-Legendre rule on [-1,1] (ascending nodes) by Newton's method on the
// Legendre recurrence; the reference takes the same rule from gaussq
// (src/Independentroutines/legendre_routines.f90:98-150).
static void gauss_legendre(int n, double *t, double *w)
{
    const double pi = 3.141592653589793238462643;
    auto legendre = [n](double x, double &pn, double &dpn) {
        double p0 = 1.0, p1 = x;
        for (int k = 2; k <= n; ++k) {
            const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        pn = p1;
        dpn = (n == 1) ? 1.0 : n * (x * p1 - p0) / (x * x - 1.0);
    };
    for (int i = 0; i < n; ++i) {
        double x = std::cos(pi * (i + 0.75) / (n + 0.5)), pn, dpn;
        for (int it = 0; it < 100; ++it) {
            legendre(x, pn, dpn);
            const double dx = pn / dpn;
            x -= dx;
            if (std::fabs(dx) < 1e-16) break;
        }
        legendre(x, pn, dpn);
        t[n - 1 - i] = x;
        w[n - 1 - i] = 2.0 / ((1.0 - x * x) * dpn * dpn);
    }
}

// bmatu1b (caseFiles/tgv/BackstepProblem-stoch.f90:345-361): the triple
// Gauss-Legendre sum of products of the Chebyshev-Gauss Lagrange bases at
// x = (t+1)/2, in the reference's summation order.
std::vector<double> ke_weights(const Ops &o)
{
    const int n = o.ncg;
    double t[16], w[16];
    gauss_legendre(n, t, w);
    std::vector<double> b((size_t)n * n * n);
    for (int i = 1; i <= n; ++i)
        for (int j = 1; j <= n; ++j)
            for (int l = 1; l <= n; ++l) {
                double sum = 0.0;
                for (int k = 1; k <= n; ++k)
                    for (int m = 1; m <= n; ++m)
                        for (int q = 1; q <= n; ++q) {
                            const double hl = lagrange(l, 0.5 * (t[q - 1] + 1.0), n, o.cxg);
                            const double hj = lagrange(j, 0.5 * (t[m - 1] + 1.0), n, o.cxg);
                            const double hi = lagrange(i, 0.5 * (t[k - 1] + 1.0), n, o.cxg);
                            sum = sum + 0.1250 * w[k - 1] * w[m - 1] * w[q - 1] * hi * hj * hl;
                        }
                b[(i - 1) + n * ((j - 1) + n * (l - 1))] = sum;
            }
    return b;
}

}  // namespace dsem
