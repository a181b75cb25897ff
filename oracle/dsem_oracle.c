/*
 * dsem_oracle.c -- CPU oracle (plain C, FP64) for the DSEM staggered-grid
 * spectral-element RHS and the Carpenter-Kennedy 5-stage RK4 2N step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for tests/, smoke() and the
 * cpu_baseline leg of bench.py.  Not part of the product library.
 *
 * Every routine restates the reference routine cited next to it, in the
 * reference's floating-point operation order (left-to-right sums, the same
 * association of products), so that it can be compared bit-for-bit with the
 * reference routines that compile here (see oracle/ref_harness.f90).  Build
 * with -ffp-contract=off so that no multiply-add is fused.
 *
 * Parity pins (tests/test_oracle_*.py):
 *   or_ops, or_geometry, or_mortar_geometry, gauss_deriv, contractions,
 *   or_dtcalc, or_rkcoefs  -> bit-for-bit against the compiled reference
 *                             leaf routines (oracle/_ref) + closed forms;
 *   or_euler_flux          -> tensor-form Euler flux (closed form);
 *   or_riemann             -> textbook eigen-decomposed Roe flux where the
 *                             entropy fix is inactive, consistency, symmetry;
 *   or_vis_flux            -> tensor-form Navier-Stokes stress/heat flux;
 *   or_rhs / or_rk_step    -> free-stream preservation, discrete conservation,
 *                             spectral convergence to an independent FFT
 *                             evaluation of the compressible NS RHS, RK4 order.
 */
#include "dsem_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NEQ 5
static const double PI_REF = 3.141592653589793238462643;

/* ------------------------------------------------------------------ */
/* Lagrange basis (src/Independentroutines/baselib.f90:18-48, 53-80)   */
/* ------------------------------------------------------------------ */
static double polyn(int k, double x, int n, const double *z)
{
    /* k is 1-based like the reference */
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

static double polynder(int k, double x, int n, const double *z)
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

/* ------------------------------------------------------------------ */
/* Points and operators                                                */
/* set_points: src/Grid_Routines/stac3m_geomLib.f90:7-47               */
/* dermat/intrpmat: src/Spectral/stac3m/matrices.f90:144-207           */
/* dmx is stored [ncg][nc] (row = Gauss point), bx [nc][ncg].          */
/* ------------------------------------------------------------------ */
int or_ops(int p, double *cx, double *cxg, double *dmx, double *bx)
{
    if (p < 1 || p > 8) return -1;
    const int ncg = p + 1, nc = p + 2;
    const int na = nc - 1;
    for (int n = 1; n <= nc; ++n) {
        double xx = cos((n - 1) * PI_REF / na);
        cx[n - 1] = 0.5 * (1.0 - xx);
    }
    for (int n = 1; n <= ncg; ++n) {
        double xx = cos((2.0 * n - 1.0) * PI_REF / (2.0 * (ncg - 1.0) + 2.0));
        cxg[n - 1] = 0.5 * (1.0 - xx);
    }
    const double eps = DBL_EPSILON;
    /* dermat(ncol, no=nc, xo=cx, nnew=ncg, xnew=cxg, b): b(k,j), k Gauss, j Lobatto.
     * The diagonal entry b(k,k) (column k of the Lobatto basis) is replaced by
     * minus the running sum of the other entries of the row. */
    for (int k = 1; k <= ncg; ++k) {
        double sum = 0.0;
        for (int j = 1; j <= nc; ++j) {
            if (j == k) continue;
            double b = polynder(j, cxg[k - 1], nc, cx);
            if (fabs(b) < eps) b = 0.0;
            dmx[(k - 1) * nc + (j - 1)] = b;
            sum = sum - b;
            dmx[(k - 1) * nc + (k - 1)] = sum;
        }
    }
    /* intrpmat(ncol, no=ncg, xo=cxg, nnew=nc, xnew=cx, b): b(k,j) = h_j(cx_k) */
    for (int k = 1; k <= nc; ++k)
        for (int j = 1; j <= ncg; ++j)
            bx[(k - 1) * ncg + (j - 1)] = polyn(j, cx[k - 1], ncg, cxg);
    return 0;
}

/* Carpenter-Kennedy RK4(5) 2N coefficients, baselib.f90:135-152 (iord=4). */
void or_rkcoefs(double *ark, double *brk, double *crk)
{
    ark[0] = 0.0; ark[1] = -0.4178904745; ark[2] = -1.192151694643;
    ark[3] = -1.697784692471; ark[4] = -1.514183444257;
    brk[0] = 0.0; brk[1] = 0.1496590219993; brk[2] = 0.3704009573644;
    brk[3] = 0.6222557631345; brk[4] = 0.9582821306748;
    crk[0] = 0.1496590219993; crk[1] = 0.3792103129999; crk[2] = 0.8229550293869;
    crk[3] = 0.6994504559488; crk[4] = 0.1530572479681;
}

/* ------------------------------------------------------------------ */
/* hex8 geometry: src/Grid_Routines/GeomLib3D.f90:60-140,685-734       */
/* corners[8][3] in standard FE order; cx in [0,1]^3.                  */
/* ------------------------------------------------------------------ */
static void hex8_map(const double *c, const double *X, double *x)
{
    for (int j = 0; j < 3; ++j) {
        x[j] = c[0 * 3 + j] * (1.0 - X[0]) * (1.0 - X[1]) * (1.0 - X[2])
             + c[1 * 3 + j] * X[0] * (1.0 - X[1]) * (1.0 - X[2])
             + c[2 * 3 + j] * X[0] * X[1] * (1.0 - X[2])
             + c[3 * 3 + j] * (1.0 - X[0]) * X[1] * (1.0 - X[2])
             + c[4 * 3 + j] * (1.0 - X[0]) * (1.0 - X[1]) * X[2]
             + c[5 * 3 + j] * X[0] * (1.0 - X[1]) * X[2]
             + c[6 * 3 + j] * X[0] * X[1] * X[2]
             + c[7 * 3 + j] * (1.0 - X[0]) * X[1] * X[2];
    }
}

/* g[i][j] = d x_i / d X_j */
static void grad_hex8_map(const double *c, const double *X, double g[3][3])
{
    for (int i = 0; i < 3; ++i) {
        g[i][0] = -(c[0 * 3 + i] * (1.0 - X[1]) * (1.0 - X[2]))
                + c[1 * 3 + i] * (1.0 - X[1]) * (1.0 - X[2])
                + c[2 * 3 + i] * X[1] * (1.0 - X[2])
                - c[3 * 3 + i] * X[1] * (1.0 - X[2])
                - c[4 * 3 + i] * (1.0 - X[1]) * X[2]
                + c[5 * 3 + i] * (1.0 - X[1]) * X[2]
                + c[6 * 3 + i] * X[1] * X[2]
                - c[7 * 3 + i] * X[1] * X[2];
        g[i][1] = -(c[0 * 3 + i] * (1.0 - X[0]) * (1.0 - X[2]))
                - c[1 * 3 + i] * X[0] * (1.0 - X[2])
                + c[2 * 3 + i] * X[0] * (1.0 - X[2])
                + c[3 * 3 + i] * (1.0 - X[0]) * (1.0 - X[2])
                - c[4 * 3 + i] * (1.0 - X[0]) * X[2]
                - c[5 * 3 + i] * X[0] * X[2]
                + c[6 * 3 + i] * X[0] * X[2]
                + c[7 * 3 + i] * (1.0 - X[0]) * X[2];
        g[i][2] = -(c[0 * 3 + i] * (1.0 - X[0]) * (1.0 - X[1]))
                - c[1 * 3 + i] * X[0] * (1.0 - X[1])
                - c[2 * 3 + i] * X[0] * X[1]
                - c[3 * 3 + i] * (1.0 - X[0]) * X[1]
                + c[4 * 3 + i] * (1.0 - X[0]) * (1.0 - X[1])
                + c[5 * 3 + i] * X[0] * (1.0 - X[1])
                + c[6 * 3 + i] * X[0] * X[1]
                + c[7 * 3 + i] * (1.0 - X[0]) * X[1];
    }
}

/* gi[i][j] = d X_i / d x_j, returns J = det(g) */
static double invert_metric(const double g[3][3], double gi[3][3])
{
    const double small = 100.0 * DBL_EPSILON;
    double J = g[0][0] * g[1][1] * g[2][2]
             - g[0][1] * g[1][0] * g[2][2]
             - g[0][0] * g[1][2] * g[2][1]
             + g[0][2] * g[1][0] * g[2][1]
             + g[0][1] * g[1][2] * g[2][0]
             - g[0][2] * g[1][1] * g[2][0];
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
            if (fabs(gi[n][m]) <= small) gi[n][m] = 0.0;
    return J;
}

/* set_geometry, all-straight hex8 branch: set_domain_geometry.f90:77-145 */
int or_geometry(int K, int p, const double *corners, double *xg, double *jacob,
                double *gmetg, double *gmetx, double *gmety, double *gmetz)
{
    double cx[10], cxg[9], dmx[90], bx[90];
    if (or_ops(p, cx, cxg, dmx, bx)) return -1;
    const int ncg = p + 1, nc = p + 2, np = ncg * ncg * ncg, ns = nc * ncg * ncg;
#pragma omp parallel for schedule(static)
    for (int e = 0; e < K; ++e) {
        const double *c = corners + (size_t)e * 24;
        double X[3], g[3][3], gi[3][3];
        for (int l = 0; l < ncg; ++l)
            for (int m = 0; m < ncg; ++m)
                for (int n = 0; n < ncg; ++n) {
                    const int pt = n + ncg * (m + ncg * l);
                    X[0] = cxg[n]; X[1] = cxg[m]; X[2] = cxg[l];
                    if (xg) hex8_map(c, X, xg + ((size_t)e * np + pt) * 3);
                    grad_hex8_map(c, X, g);
                    double J = invert_metric(g, gi);
                    jacob[(size_t)e * np + pt] = 1.0 / J;
                    for (int a = 0; a < 3; ++a)
                        for (int b = 0; b < 3; ++b)
                            gmetg[((size_t)e * np + pt) * 9 + a * 3 + b] = gi[a][b];
                }
        for (int l = 0; l < ncg; ++l)
            for (int m = 0; m < ncg; ++m)
                for (int j = 0; j < nc; ++j) {
                    X[0] = cx[j]; X[1] = cxg[m]; X[2] = cxg[l];
                    grad_hex8_map(c, X, g);
                    double J = invert_metric(g, gi);
                    double *o = gmetx + ((size_t)e * ns + j + nc * (m + ncg * l)) * 3;
                    for (int b = 0; b < 3; ++b) o[b] = J * gi[0][b];
                }
        for (int l = 0; l < ncg; ++l)
            for (int j = 0; j < nc; ++j)
                for (int n = 0; n < ncg; ++n) {
                    X[0] = cxg[n]; X[1] = cx[j]; X[2] = cxg[l];
                    grad_hex8_map(c, X, g);
                    double J = invert_metric(g, gi);
                    double *o = gmety + ((size_t)e * ns + n + ncg * (j + nc * l)) * 3;
                    for (int b = 0; b < 3; ++b) o[b] = J * gi[1][b];
                }
        for (int j = 0; j < nc; ++j)
            for (int m = 0; m < ncg; ++m)
                for (int n = 0; n < ncg; ++n) {
                    X[0] = cxg[n]; X[1] = cxg[m]; X[2] = cx[j];
                    grad_hex8_map(c, X, g);
                    double J = invert_metric(g, gi);
                    double *o = gmetz + ((size_t)e * ns + n + ncg * (m + ncg * j)) * 3;
                    for (int b = 0; b < 3; ++b) o[b] = J * gi[2][b];
                }
    }
    return 0;
}

/* Mortar normals from the master face: set_domain_geometry.f90:254-345.
 * end_vals = (0,1,0,1,1,0) for faces 1..6. */
int or_mortar_geometry(int M, int p, const double *corners, const int32_t *m_elem,
                       const int32_t *m_face, double *nhat, double *norm)
{
    double cx[10], cxg[9], dmx[90], bx[90];
    if (or_ops(p, cx, cxg, dmx, bx)) return -1;
    static const double end_vals[6] = {0.0, 1.0, 0.0, 1.0, 1.0, 0.0};
    const int ncg = p + 1, nf = ncg * ncg;
#pragma omp parallel for schedule(static)
    for (int mm = 0; mm < M; ++mm) {
        const int e = m_elem[2 * mm], f = m_face[2 * mm];
        const double *c = corners + (size_t)e * 24;
        double X[3], g[3][3], gi[3][3];
        int row, a0, a1, fixed;
        if (f == 1 || f == 2) { fixed = 1; a0 = 0; a1 = 2; row = 1; }
        else if (f == 3 || f == 5) { fixed = 2; a0 = 0; a1 = 1; row = 2; }
        else { fixed = 0; a0 = 1; a1 = 2; row = 0; }
        for (int j = 0; j < ncg; ++j)
            for (int i = 0; i < ncg; ++i) {
                X[fixed] = end_vals[f - 1];
                X[a0] = cxg[i];
                X[a1] = cxg[j];
                grad_hex8_map(c, X, g);
                double J = invert_metric(g, gi);
                double sqr = sqrt(gi[row][0] * gi[row][0] + gi[row][1] * gi[row][1]
                                  + gi[row][2] * gi[row][2]);
                double *nh = nhat + ((size_t)mm * nf + i + ncg * j) * 3;
                nh[0] = gi[row][0] / sqr;
                nh[1] = gi[row][1] / sqr;
                nh[2] = gi[row][2] / sqr;
                norm[(size_t)mm * nf + i + ncg * j] = J * sqr;
            }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Pointwise physics: src/Parabolic3D/NS/NS/PhysLib.f90                */
/* ------------------------------------------------------------------ */
/* fflux/gflux/hflux: PhysLib.f90:17-124 */
void or_euler_flux(const or_phys *ph, const double *Q, double *f, double *g, double *h)
{
    const double gamma = ph->gamma;
    const double rho = Q[0], rhou = Q[1], rhov = Q[2], rhow = Q[3], rhoe = Q[4];
    const double u = rhou / rho, v = rhov / rho, w = rhow / rho;
    const double p = (gamma - 1.0) * (rhoe - 0.5 * rho * (u * u + v * v + w * w));
    if (f) {
        f[0] = rhou; f[1] = p + rhou * u; f[2] = rhou * v; f[3] = rhou * w; f[4] = u * (rhoe + p);
    }
    if (g) {
        g[0] = rhov; g[1] = rhou * v; g[2] = p + rhov * v; g[3] = rhow * v; g[4] = v * (rhoe + p);
    }
    if (h) {
        h[0] = rhow; h[1] = rhou * w; h[2] = rhov * w; h[3] = p + rhow * w; h[4] = w * (rhoe + p);
    }
}

/* Planted mistakes for the tests' controls (0 = none, the reference):
 * OR_MUT_ENTROPY_FIX_SIGN flips the sign of the entropy-fix term hd1m/hd4,
 * OR_MUT_SWAP_SIDES swaps the left/right states of interface_flux. */
static int g_mutation = 0;

void or_set_mutation(int m) { g_mutation = m; }

/* Roe flux with entropy fix: PhysLib.f90:231-320.  eta (nullable) receives
 * the entropy-fix eta1 (qtd >= 0, branch +1) or eta4 (qtd < 0, branch -1). */
static void riemann_core(const or_phys *ph, const double *nhat, double ds, const double *Ql,
                         const double *Qr, double *flux, double *eta, double *branch)
{
    const double fsgn = (g_mutation & OR_MUT_ENTROPY_FIX_SIGN) ? -1.0 : 1.0;
    const double gamma = ph->gamma;
    const double rho = Ql[0], rhou = Ql[1], rhov = Ql[2], rhow = Ql[3], rhoe = Ql[4];
    const double rhon = Qr[0], rhoun = Qr[1], rhovn = Qr[2], rhown = Qr[3], rhoen = Qr[4];
    const double ul = rhou / rho, vl = rhov / rho, wl = rhow / rho;
    const double pleft = (gamma - 1.0) * (rhoe - 0.5 / rho * (rhou * rhou + rhov * rhov + rhow * rhow));
    const double ur = rhoun / rhon, vr = rhovn / rhon, wr = rhown / rhon;
    const double pright = (gamma - 1.0) * (rhoen - 0.5 / rhon * (rhoun * rhoun + rhovn * rhovn + rhown * rhown));
    const double ql = nhat[0] * ul + nhat[1] * vl + nhat[2] * wl;
    const double qr = nhat[0] * ur + nhat[1] * vr + nhat[2] * wr;
    const double hl = 0.5 * (ul * ul + vl * vl + wl * wl) + gamma / (gamma - 1.0) * pleft / rho;
    const double hr = 0.5 * (ur * ur + vr * vr + wr * wr) + gamma / (gamma - 1.0) * pright / rhon;
    const double rtd = sqrt(rho * rhon);
    const double betal = rho / (rho + rtd);
    const double betar = 1.0 - betal;
    const double utd = betal * ul + betar * ur;
    const double vtd = betal * vl + betar * vr;
    const double wtd = betal * wl + betar * wr;
    const double htd = betal * hl + betar * hr;
    const double atd2 = (gamma - 1.0) * (htd - 0.5 * (utd * utd + vtd * vtd + wtd * wtd));
    const double atd = sqrt(atd2);
    const double qtd = utd * nhat[0] + vtd * nhat[1] + wtd * nhat[2];
    if (qtd >= 0.0) {
        const double dw1 = 0.5 * ((pright - pleft) / atd2 - (qr - ql) * rtd / atd);
        const double sp1 = qtd - atd;
        const double sp1m = fmin(sp1, 0.0);
        const double hd1m = fsgn * (((gamma + 1.0) / 4.0 * atd / rtd) * dw1);
        const double eta1 = fmax(-fabs(sp1) - hd1m, 0.0);
        if (eta) { *eta = eta1; *branch = 1.0; }
        const double udw1 = dw1 * (sp1m - 0.5 * eta1);
        const double rql = rho * ql;
        flux[0] = ds * (rql + udw1);
        flux[1] = ds * (rql * ul + pleft * nhat[0] + udw1 * (utd - atd * nhat[0]));
        flux[2] = ds * (rql * vl + pleft * nhat[1] + udw1 * (vtd - atd * nhat[1]));
        flux[3] = ds * (rql * wl + pleft * nhat[2] + udw1 * (wtd - atd * nhat[2]));
        flux[4] = ds * (rql * hl + udw1 * (htd - qtd * atd));
    } else {
        const double dw4 = 0.5 * ((pright - pleft) / atd2 + (qr - ql) * rtd / atd);
        const double sp4 = qtd + atd;
        const double sp4p = fmax(sp4, 0.0);
        const double hd4 = fsgn * (((gamma + 1.0) / 4.0 * atd / rtd) * dw4);
        const double eta4 = fmax(-fabs(sp4) + hd4, 0.0);
        if (eta) { *eta = eta4; *branch = -1.0; }
        const double udw4 = dw4 * (sp4p + 0.5 * eta4);
        const double rqr = rhon * qr;
        flux[0] = ds * (rqr - udw4);
        flux[1] = ds * (rqr * ur + pright * nhat[0] - udw4 * (utd + atd * nhat[0]));
        flux[2] = ds * (rqr * vr + pright * nhat[1] - udw4 * (vtd + atd * nhat[1]));
        flux[3] = ds * (rqr * wr + pright * nhat[2] - udw4 * (wtd + atd * nhat[2]));
        flux[4] = ds * (rqr * hr - udw4 * (htd + qtd * atd));
    }
}

void or_riemann(const or_phys *ph, const double *nhat, double ds, const double *Ql,
                const double *Qr, double *flux)
{
    riemann_core(ph, nhat, ds, Ql, Qr, flux, NULL, NULL);
}

/* x/y/z_vis_flux: PhysLib.f90:419-583 (each recomputes the same primitives). */
void or_vis_flux(const or_phys *ph, const double *Q, const double *Qx, const double *Qy,
                 const double *Qz, double amu, double *f, double *g, double *h)
{
    const double gamma = ph->gamma, mach = ph->mach, pr = ph->pr;
    const double rho = Q[0], rhou = Q[1], rhov = Q[2], rhow = Q[3], rhoe = Q[4];
    const double u = rhou / rho, v = rhov / rho, w = rhow / rho;
    const double p = (gamma - 1.0) * (rhoe - 0.5 * rho * (u * u + v * v + w * w));
    const double ux = (Qx[1] - u * Qx[0]) / rho, uy = (Qy[1] - u * Qy[0]) / rho, uz = (Qz[1] - u * Qz[0]) / rho;
    const double vx = (Qx[2] - v * Qx[0]) / rho, vy = (Qy[2] - v * Qy[0]) / rho, vz = (Qz[2] - v * Qz[0]) / rho;
    const double wx = (Qx[3] - w * Qx[0]) / rho, wy = (Qy[3] - w * Qy[0]) / rho, wz = (Qz[3] - w * Qz[0]) / rho;
    const double akappa = 1.0 + (amu - 1.0) * 2.0e-1 / (gamma - 1.0);
    const double div = ux + vy + wz;
    const double cden = pr * mach * mach * (gamma - 1.0);
    if (f) {
        const double px = (gamma - 1.0) * (Qx[4] - 0.5 * Qx[0] * (u * u + v * v + w * w) - rho * (u * ux + v * vx + w * wx));
        const double tx = (px - p * Qx[0] / rho) * mach * mach * gamma / rho;
        f[0] = 0.0;
        f[1] = 2.0 * amu * (ux - div / 3.0);
        f[2] = amu * (uy + vx);
        f[3] = amu * (wx + uz);
        f[4] = u * f[1] + v * f[2] + w * f[3] + akappa * tx / cden;
    }
    if (g) {
        const double py = (gamma - 1.0) * (Qy[4] - 0.5 * Qy[0] * (u * u + v * v + w * w) - rho * (u * uy + v * vy + w * wy));
        const double ty = (py - p * Qy[0] / rho) * mach * mach * gamma / rho;
        g[0] = 0.0;
        g[1] = amu * (uy + vx);
        g[2] = 2.0 * amu * (vy - div / 3.0);
        g[3] = amu * (wy + vz);
        g[4] = v * g[2] + u * g[1] + w * g[3] + akappa * ty / cden;
    }
    if (h) {
        const double pz = (gamma - 1.0) * (Qz[4] - 0.5 * Qz[0] * (u * u + v * v + w * w) - rho * (u * uz + v * vz + w * wz));
        const double tz = (pz - p * Qz[0] / rho) * mach * mach * gamma / rho;
        h[0] = 0.0;
        h[1] = amu * (wx + uz);
        h[2] = amu * (wy + vz);
        h[3] = 2.0 * amu * (wz - div / 3.0);
        h[4] = v * h[2] + u * h[1] + w * h[3] + akappa * tz / cden;
    }
}

/* ------------------------------------------------------------------ */
/* Face slots: which staggered-grid entry holds face point (i,j).       */
/* send_q.f90:21-280, send_flux.f90:22-236                             */
/* ------------------------------------------------------------------ */
static void face_slot(int face, int i, int j, int ncg, int nc, int *grid, int *idx)
{
    switch (face) {
    case 1: *grid = 1; *idx = i + ncg * (0 + nc * j); break;
    case 2: *grid = 1; *idx = i + ncg * ((nc - 1) + nc * j); break;
    case 3: *grid = 2; *idx = i + ncg * (j + ncg * 0); break;
    case 5: *grid = 2; *idx = i + ncg * (j + ncg * (nc - 1)); break;
    case 4: *grid = 0; *idx = (nc - 1) + nc * (i + ncg * j); break;
    default: *grid = 0; *idx = 0 + nc * (i + ncg * j); break; /* face 6 */
    }
}

/* Slave face point (i,j) that mortar point (I,J) takes under orientation o:
 * rotate_to_mortar (mortar_operations.f90:461-605); its inverse is
 * rotate_flux_to_face (:616-758) and rotate_fv_to_mortar (:851-995) is the
 * same map for the viscous traces.  o = DFLT, B1F2, B1B2, F1B2, F2F1, B2F1,
 * B2B1, F2B1 = 0..7 (element_library.f90:99-101). */
static void slave_point(int o, int I, int J, int n, int *i, int *j)
{
    const int m = n - 1;
    if (g_mutation & OR_MUT_ORIENT_INVERSE) o = (o == 5) ? 7 : (o == 7 ? 5 : o);
    switch (o) {
    case 1: *i = m - I; *j = J; break;
    case 2: *i = m - I; *j = m - J; break;
    case 3: *i = I; *j = m - J; break;
    case 4: *i = J; *j = I; break;
    case 5: *i = J; *j = m - I; break;
    case 6: *i = m - J; *j = m - I; break;
    case 7: *i = m - J; *j = I; break;
    default: *i = I; *j = J; break;
    }
}

/* Staggered index of Lobatto position j on the line (a,b) of grid d. */
static inline int sidx(int d, int j, int a, int b, int ncg, int nc)
{
    if (d == 0) return j + nc * (a + ncg * b);         /* (j, m=a, l=b) */
    if (d == 1) return a + ncg * (j + nc * b);         /* (n=a, j, l=b) */
    return a + ncg * (b + ncg * j);                    /* (n=a, m=b, j) */
}
/* Gauss index of position k on line (a,b) in direction d. */
static inline int gidx(int d, int k, int a, int b, int ncg)
{
    if (d == 0) return k + ncg * (a + ncg * b);
    if (d == 1) return a + ncg * (k + ncg * b);
    return a + ncg * (b + ncg * k);
}

typedef struct {
    int p, ncg, nc, np, ns, nf;
    double cx[10], cxg[9], dmx[90], bx[90];
} ops_t;

static int make_ops(int p, ops_t *o)
{
    o->p = p; o->ncg = p + 1; o->nc = p + 2;
    o->np = o->ncg * o->ncg * o->ncg; o->ns = o->nc * o->ncg * o->ncg; o->nf = o->ncg * o->ncg;
    return or_ops(p, o->cx, o->cxg, o->dmx, o->bx);
}

/*
 * The RHS.  Order of operations follows Fluxes (fluxesH_newa.f90:17-199):
 *   Int_fluxes per element (fluxesH_newa.f90:221-355)
 *   send_q_to_mortars -> per mortar: interface_flux/riemann -> send_flux_to_faces,
 *   V_Avg_Soln -> send_q_to_faces  (mortar_operations.f90:87-238,781-837)
 *   Vis_Fluxes (Vis_Fluxes_new.f90:16-382) with Gauss_Deriv (SolnDerivsc.f90:14-83)
 *   send_flux_to_mortars, mortar_vis_flux (ViscousBC_slide.f90:231-271),
 *   send_flux_to_faces_add (send_flux.f90:699-921)
 *   update_interior divergence (updatea.f90:100-113).
 * muh: optional entropy viscosity [K][3][ns] on the three staggered grids.
 */
int or_rhs(const or_mesh *mesh, const or_phys *ph, const double *JQ, const double *muh,
           double *Qt, or_debug *dbg, int nthreads)
{
    return or_rhs_les(mesh, ph, JQ, muh, NULL, Qt, dbg, nthreads);
}

/* or_rhs with a static-Smagorinsky eddy viscosity nut [K][3][ns] on the
 * staggered grids: amu = 1 + nut * rho + mu_h (Vis_Fluxes_new.f90:123-125,
 * 149-151, 177-179; rho is the face-averaged staggered density). */
int or_rhs_les(const or_mesh *mesh, const or_phys *ph, const double *JQ, const double *muh,
               const double *nut, double *Qt, or_debug *dbg, int nthreads)
{
    ops_t o;
    if (make_ops(mesh->p, &o)) return -1;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    const int K = mesh->K, ncg = o.ncg, nc = o.nc, np = o.np, ns = o.ns, nf = o.nf;
    const double re = ph->re;
    const size_t sz_stag = (size_t)K * 3 * NEQ * ns;
    double *S = dbg && dbg->stag ? dbg->stag : (double *)malloc(sizeof(double) * sz_stag);
    double *F = dbg && dbg->flux ? dbg->flux : (double *)malloc(sizeof(double) * sz_stag);
    double *R = dbg && dbg->visc ? dbg->visc : (double *)malloc(sizeof(double) * sz_stag);
    double *Gd = dbg && dbg->grad ? dbg->grad : (double *)malloc(sizeof(double) * (size_t)K * 3 * NEQ * np);
    double *mQ = (double *)malloc(sizeof(double) * (size_t)mesh->nmort * 2 * NEQ * nf);
    double *mFV = (double *)malloc(sizeof(double) * (size_t)mesh->nmort * 2 * NEQ * nf);
    if (!S || !F || !R || !Gd || !mQ || !mFV) return -2;
    memset(F, 0, sizeof(double) * sz_stag);
    const double *gm[3] = {mesh->gmetx, mesh->gmety, mesh->gmetz};

    /* ---------------- Int_fluxes ---------------- */
#pragma omp parallel for schedule(static)
    for (int e = 0; e < K; ++e) {
        double Qp[NEQ * 729];
        const double *jq = JQ + (size_t)e * NEQ * np;
        const double *jac = mesh->jacob + (size_t)e * np;
        for (int v = 0; v < NEQ; ++v)
            for (int pt = 0; pt < np; ++pt) Qp[v * np + pt] = jq[v * np + pt] * jac[pt];
        for (int d = 0; d < 3; ++d) {
            double *Sd = S + ((size_t)e * 3 + d) * NEQ * ns;
            double *Fd = F + ((size_t)e * 3 + d) * NEQ * ns;
            /* interpxa/ya/za: FInterpLiba.f90:174-275; mxm left-to-right sums */
            for (int v = 0; v < NEQ; ++v)
                for (int b = 0; b < ncg; ++b)
                    for (int a = 0; a < ncg; ++a)
                        for (int j = 0; j < nc; ++j) {
                            double s = 0.0;
                            for (int k = 0; k < ncg; ++k) {
                                double t = o.bx[j * ncg + k] * Qp[v * np + gidx(d, k, a, b, ncg)];
                                s = (k == 0) ? t : s + t;
                            }
                            Sd[v * ns + sidx(d, j, a, b, ncg, nc)] = s;
                        }
            /* interior Lobatto points: contravariant Euler flux */
            for (int b = 0; b < ncg; ++b)
                for (int a = 0; a < ncg; ++a)
                    for (int j = 1; j < nc - 1; ++j) {
                        const int si = sidx(d, j, a, b, ncg, nc);
                        double Qv[NEQ], ff[NEQ], gg[NEQ], hh[NEQ];
                        for (int v = 0; v < NEQ; ++v) Qv[v] = Sd[v * ns + si];
                        or_euler_flux(ph, Qv, ff, gg, hh);
                        const double *gmt = gm[d] + ((size_t)e * ns + si) * 3;
                        for (int v = 0; v < NEQ; ++v)
                            Fd[v * ns + si] = gmt[0] * ff[v] + gmt[1] * gg[v] + gmt[2] * hh[v];
                    }
        }
    }

    /* ---------------- send_q_to_mortars (+ rotate_to_mortar) ---------------- */
#pragma omp parallel for schedule(static)
    for (int mm = 0; mm < mesh->nmort; ++mm) {
        for (int side = 0; side < 2; ++side) {
            const int e = mesh->m_elem[2 * mm + side], f = mesh->m_face[2 * mm + side];
            const int o = (side == 1 && mesh->m_orient) ? mesh->m_orient[mm] : 0;
            for (int J = 0; J < ncg; ++J)
                for (int I = 0; I < ncg; ++I) {
                    int grid, idx, i, j;
                    slave_point(o, I, J, ncg, &i, &j);
                    face_slot(f, i, j, ncg, nc, &grid, &idx);
                    for (int v = 0; v < NEQ; ++v)
                        mQ[(((size_t)mm * 2 + side) * NEQ + v) * nf + I + ncg * J] =
                            S[(((size_t)e * 3 + grid) * NEQ + v) * ns + idx];
                }
        }
    }

    /* ---------------- mortar fluxes + solution average ---------------- */
#pragma omp parallel for schedule(static)
    for (int mm = 0; mm < mesh->nmort; ++mm) {
        const int em = mesh->m_elem[2 * mm], fm = mesh->m_face[2 * mm];
        const int es = mesh->m_elem[2 * mm + 1], fs = mesh->m_face[2 * mm + 1];
        const double sgn = (double)mesh->m_sign[mm];
        for (int j = 0; j < ncg; ++j)
            for (int i = 0; i < ncg; ++i) {
                const int q = i + ncg * j;
                double Qm[NEQ], Qs[NEQ], fl[NEQ], avg[NEQ];
                for (int v = 0; v < NEQ; ++v) {
                    Qm[v] = mQ[(((size_t)mm * 2 + 0) * NEQ + v) * nf + q];
                    Qs[v] = mQ[(((size_t)mm * 2 + 1) * NEQ + v) * nf + q];
                }
                const double *nh = mesh->nhat + ((size_t)mm * nf + q) * 3;
                const double ds = mesh->norm[(size_t)mm * nf + q];
                /* left state = the minus side: the slave on master faces 1, 3, 6
                 * (mortar_operations.f90:190-236) */
                int slave_left = (fm == 1 || fm == 3 || fm == 6);
                if (g_mutation & OR_MUT_SWAP_SIDES) slave_left = !slave_left;
                double *eta = (dbg && dbg->eta) ? dbg->eta + ((size_t)mm * nf + q) * 2 : NULL;
                if (slave_left) riemann_core(ph, nh, ds, Qs, Qm, fl, eta, eta ? eta + 1 : NULL);
                else riemann_core(ph, nh, ds, Qm, Qs, fl, eta, eta ? eta + 1 : NULL);
                for (int v = 0; v < NEQ; ++v) avg[v] = (Qm[v] + Qs[v]) * 0.5;
                int grid, idx;
                face_slot(fm, i, j, ncg, nc, &grid, &idx);
                for (int v = 0; v < NEQ; ++v) {
                    F[(((size_t)em * 3 + grid) * NEQ + v) * ns + idx] = fl[v];
                    S[(((size_t)em * 3 + grid) * NEQ + v) * ns + idx] = avg[v];
                }
                /* rotate_flux_to_face: the slave's face point of this mortar point */
                int is, js;
                slave_point(mesh->m_orient ? mesh->m_orient[mm] : 0, i, j, ncg, &is, &js);
                face_slot(fs, is, js, ncg, nc, &grid, &idx);
                for (int v = 0; v < NEQ; ++v) {
                    F[(((size_t)es * 3 + grid) * NEQ + v) * ns + idx] = sgn * fl[v];
                    S[(((size_t)es * 3 + grid) * NEQ + v) * ns + idx] = avg[v];
                }
            }
    }

    /* ---------------- Vis_Fluxes: per element ---------------- */
#pragma omp parallel for schedule(static)
    for (int e = 0; e < K; ++e) {
        double W[3][NEQ * 729];           /* w1x, w2y, w3z per variable */
        double *Ge = Gd + (size_t)e * 3 * NEQ * np;
        const double *gmg = mesh->gmetg + (size_t)e * np * 9;
        /* Gauss_Deriv: dx3dga/dy3dga/dz3dga (FDerivLibb.f90:153-241) */
        for (int d = 0; d < 3; ++d) {
            const double *Sd = S + ((size_t)e * 3 + d) * NEQ * ns;
            for (int v = 0; v < NEQ; ++v)
                for (int b = 0; b < ncg; ++b)
                    for (int a = 0; a < ncg; ++a)
                        for (int k = 0; k < ncg; ++k) {
                            double s = 0.0;
                            for (int j = 0; j < nc; ++j) {
                                double t = o.dmx[k * nc + j] * Sd[v * ns + sidx(d, j, a, b, ncg, nc)];
                                s = (j == 0) ? t : s + t;
                            }
                            W[d][v * np + gidx(d, k, a, b, ncg)] = s;
                        }
        }
        for (int v = 0; v < NEQ; ++v)
            for (int pt = 0; pt < np; ++pt) {
                const double *G9 = gmg + (size_t)pt * 9;   /* G9[i*3+j] = gmetg(i+1,j+1) */
                const double w1 = W[0][v * np + pt], w2 = W[1][v * np + pt], w3 = W[2][v * np + pt];
                Ge[(0 * NEQ + v) * np + pt] = G9[0 * 3 + 0] * w1 + G9[1 * 3 + 0] * w2 + G9[2 * 3 + 0] * w3;
                Ge[(1 * NEQ + v) * np + pt] = G9[0 * 3 + 1] * w1 + G9[1 * 3 + 1] * w2 + G9[2 * 3 + 1] * w3;
                Ge[(2 * NEQ + v) * np + pt] = G9[0 * 3 + 2] * w1 + G9[1 * 3 + 2] * w2 + G9[2 * 3 + 2] * w3;
            }
        /* prolong the 15 gradient fields (interp -> mxv) and viscous fluxes */
        for (int d = 0; d < 3; ++d) {
            const double *Sd = S + ((size_t)e * 3 + d) * NEQ * ns;
            double *Rd = R + ((size_t)e * 3 + d) * NEQ * ns;
            double *Fd = F + ((size_t)e * 3 + d) * NEQ * ns;
            const double *gmt = gm[d] + (size_t)e * ns * 3;
            for (int b = 0; b < ncg; ++b)
                for (int a = 0; a < ncg; ++a)
                    for (int j = 0; j < nc; ++j) {
                        const int si = sidx(d, j, a, b, ncg, nc);
                        double Qv[NEQ], Gx[3][NEQ], ff[NEQ], gg[NEQ], hh[NEQ];
                        for (int v = 0; v < NEQ; ++v) Qv[v] = Sd[v * ns + si];
                        for (int c = 0; c < 3; ++c)
                            for (int v = 0; v < NEQ; ++v) {
                                double s = 0.0;
                                for (int k = 0; k < ncg; ++k) {
                                    double t = o.bx[j * ncg + k] * Ge[(c * NEQ + v) * np + gidx(d, k, a, b, ncg)];
                                    s = (k == 0) ? t : s + t;
                                }
                                Gx[c][v] = s;
                            }
                        double amu = 1.0;
                        if (nut) amu = amu + nut[((size_t)e * 3 + d) * ns + si] * Qv[0];
                        if (muh) amu = amu + muh[((size_t)e * 3 + d) * ns + si];
                        or_vis_flux(ph, Qv, Gx[0], Gx[1], Gx[2], amu, ff, gg, hh);
                        const double *gq = gmt + (size_t)si * 3;
                        for (int v = 0; v < NEQ; ++v)
                            Rd[v * ns + si] = gq[0] * ff[v] + gq[1] * gg[v] + gq[2] * hh[v];
                    }
            /* interior: f = f - r/re (Vis_Fluxes_new.f90:195-228) */
            for (int v = 0; v < NEQ; ++v)
                for (int b = 0; b < ncg; ++b)
                    for (int a = 0; a < ncg; ++a)
                        for (int j = 1; j < nc - 1; ++j) {
                            const int si = sidx(d, j, a, b, ncg, nc);
                            Fd[v * ns + si] = Fd[v * ns + si] - Rd[v * ns + si] / re;
                        }
        }
    }

    /* ---------------- send_flux_to_mortars: fv = r/norm (+ rotate_fv_to_mortar) ---------------- */
#pragma omp parallel for schedule(static)
    for (int mm = 0; mm < mesh->nmort; ++mm) {
        for (int side = 0; side < 2; ++side) {
            const int e = mesh->m_elem[2 * mm + side], f = mesh->m_face[2 * mm + side];
            const int o = (side == 1 && mesh->m_orient) ? mesh->m_orient[mm] : 0;
            for (int J = 0; J < ncg; ++J)
                for (int I = 0; I < ncg; ++I) {
                    int grid, idx, i, j;
                    slave_point(o, I, J, ncg, &i, &j);
                    face_slot(f, i, j, ncg, nc, &grid, &idx);
                    const double nrm = mesh->norm[(size_t)mm * nf + I + ncg * J];
                    for (int v = 0; v < NEQ; ++v)
                        mFV[(((size_t)mm * 2 + side) * NEQ + v) * nf + I + ncg * J] =
                            R[(((size_t)e * 3 + grid) * NEQ + v) * ns + idx] / nrm;
                }
        }
    }

    /* ---------------- mortar_vis_flux + send_flux_to_faces_add ---------------- */
#pragma omp parallel for schedule(static)
    for (int mm = 0; mm < mesh->nmort; ++mm) {
        const int em = mesh->m_elem[2 * mm], fm = mesh->m_face[2 * mm];
        const int es = mesh->m_elem[2 * mm + 1], fs = mesh->m_face[2 * mm + 1];
        for (int j = 0; j < ncg; ++j)
            for (int i = 0; i < ncg; ++i) {
                const int q = i + ncg * j;
                const double nrm = mesh->norm[(size_t)mm * nf + q];
                double Qb[NEQ];
                for (int v = 0; v < NEQ; ++v) {
                    double Qs = mFV[(((size_t)mm * 2 + 1) * NEQ + v) * nf + q];
                    if (mesh->m_nsign[mm] == -1) Qs = -Qs;
                    const double Qm = mFV[(((size_t)mm * 2 + 0) * NEQ + v) * nf + q];
                    Qb[v] = 0.5 * (Qm + Qs) * nrm;
                }
                int grid, idx;
                face_slot(fm, i, j, ncg, nc, &grid, &idx);
                for (int v = 0; v < NEQ; ++v) {
                    double *t = &F[(((size_t)em * 3 + grid) * NEQ + v) * ns + idx];
                    *t = *t - Qb[v] / re;
                }
                int is, js;
                slave_point(mesh->m_orient ? mesh->m_orient[mm] : 0, i, j, ncg, &is, &js);
                face_slot(fs, is, js, ncg, nc, &grid, &idx);
                for (int v = 0; v < NEQ; ++v) {
                    const double fl = (mesh->m_sign[mm] == -1) ? -Qb[v] : Qb[v];
                    double *t = &F[(((size_t)es * 3 + grid) * NEQ + v) * ns + idx];
                    *t = *t - fl / re;
                }
            }
    }

    /* ---------------- update_interior: divergence ---------------- */
#pragma omp parallel for schedule(static)
    for (int e = 0; e < K; ++e) {
        double Dv[3][729];
        for (int v = 0; v < NEQ; ++v) {
            for (int d = 0; d < 3; ++d) {
                const double *Fd = F + (((size_t)e * 3 + d) * NEQ + v) * ns;
                for (int b = 0; b < ncg; ++b)
                    for (int a = 0; a < ncg; ++a)
                        for (int k = 0; k < ncg; ++k) {
                            double s = 0.0;
                            for (int j = 0; j < nc; ++j) {
                                double t = o.dmx[k * nc + j] * Fd[sidx(d, j, a, b, ncg, nc)];
                                s = (j == 0) ? t : s + t;
                            }
                            Dv[d][gidx(d, k, a, b, ncg)] = s;
                        }
            }
            double *qt = Qt + ((size_t)e * NEQ + v) * np;
            for (int pt = 0; pt < np; ++pt) qt[pt] = -(Dv[0][pt] + Dv[1][pt] + Dv[2][pt]);
        }
    }

    if (!(dbg && dbg->stag)) free(S);
    if (!(dbg && dbg->flux)) free(F);
    if (!(dbg && dbg->visc)) free(R);
    if (!(dbg && dbg->grad)) free(Gd);
    free(mQ);
    free(mFV);
    return 0;
}

/*
 * One time step: tstep (src/tstep.f90:12-120) with update_interior's RK
 * register update (updatea.f90:213-227).  G is zeroed at the start of the
 * step.  resid follows the reference: update_interior resets it per element,
 * so the value left is max|Qt| over the last element of the last stage.
 * stag_out (nullable, [K][3][5][ns]) receives the last stage's face-averaged
 * staggered states, which the reference leaves in d%Qlgg/Qglg/Qggl.
 */
int or_rk_step(const or_mesh *mesh, const or_phys *ph, double *JQ, double *G, double dt,
               const double *muh, double *stag_out, double *resid, int nthreads)
{
    double ark[5], brk[5], crk[5];
    or_rkcoefs(ark, brk, crk);
    const int np = (mesh->p + 1) * (mesh->p + 1) * (mesh->p + 1);
    const size_t n = (size_t)mesh->K * NEQ * np;
    double *Qt = (double *)malloc(sizeof(double) * n);
    if (!Qt) return -2;
    memset(G, 0, sizeof(double) * n);
    double last = 0.0;
    for (int kk = 0; kk < 5; ++kk) {
        or_debug dbg = {stag_out, NULL, NULL, NULL, NULL};
        int rc = or_rhs(mesh, ph, JQ, muh, Qt, (kk == 4 && stag_out) ? &dbg : NULL, nthreads);
        if (rc) { free(Qt); return rc; }
        const double a = ark[kk], c = crk[kk];
#pragma omp parallel for schedule(static)
        for (int e = 0; e < mesh->K; ++e) {
            for (int v = 0; v < NEQ; ++v)
                for (int pt = 0; pt < np; ++pt) {
                    const size_t i = ((size_t)e * NEQ + v) * np + pt;
                    G[i] = a * G[i] + Qt[i];
                    JQ[i] = JQ[i] + c * dt * G[i];
                }
        }
        last = 0.0;
        const double *q = Qt + (size_t)(mesh->K - 1) * NEQ * np;
        for (int i = 0; i < NEQ * np; ++i) last = fmax(fabs(q[i]), last);
    }
    if (resid) *resid = last;
    free(Qt);
    return 0;
}

/* dtcalc (src/Parabolic3D/NS/NS/stac3m/dtcalc.f90:9-58): minimum over all
 * elements of 1/dti, before the CFL factor.  It reads gmet(d,:) at the
 * Lobatto index equal to the Gauss index, as the reference does. */
double or_dtcalc(const or_mesh *mesh, const or_phys *ph, const double *JQ)
{
    ops_t o;
    if (make_ops(mesh->p, &o)) return -1.0;
    const int ncg = o.ncg, nc = o.nc, np = o.np, ns = o.ns;
    const double gamma = ph->gamma;
    double dt = DBL_MAX;
    for (int e = 0; e < mesh->K; ++e) {
        const double *q = JQ + (size_t)e * NEQ * np;
        for (int l = 0; l < ncg; ++l) {
            const double dcz = o.cx[l + 1] - o.cx[l];
            for (int m = 0; m < ncg; ++m) {
                const double dcy = o.cx[m + 1] - o.cx[m];
                for (int n = 0; n < ncg; ++n) {
                    const double dcx = o.cx[n + 1] - o.cx[n];
                    const int pt = n + ncg * (m + ncg * l);
                    const double u0 = fabs(q[1 * np + pt] / q[0 * np + pt]);
                    const double v0 = fabs(q[2 * np + pt] / q[0 * np + pt]);
                    const double w0 = fabs(q[3 * np + pt] / q[0 * np + pt]);
                    const double p = (gamma - 1.0) * (q[4 * np + pt] - 0.5 * q[0 * np + pt] * (u0 * u0 + v0 * v0 + w0 * w0));
                    const double a = sqrt(gamma * p / q[0 * np + pt]);
                    const double *g1 = mesh->gmetx + ((size_t)e * ns + n + nc * (m + ncg * l)) * 3;
                    const double *g2 = mesh->gmety + ((size_t)e * ns + n + ncg * (m + nc * l)) * 3;
                    const double *g3 = mesh->gmetz + ((size_t)e * ns + n + ncg * (m + ncg * l)) * 3;
                    const double uu = (u0 + a) * g1[0] + (v0 + a) * g1[1] + (w0 + a) * g1[2];
                    const double vv = (u0 + a) * g2[0] + (v0 + a) * g2[1] + (w0 + a) * g2[2];
                    const double ww = (u0 + a) * g3[0] + (v0 + a) * g3[1] + (w0 + a) * g3[2];
                    const double jac = mesh->jacob[(size_t)e * np + pt];
                    const double xlam = fabs(uu * jac), ylam = fabs(vv * jac), zlam = fabs(ww * jac);
                    const double dti = xlam / dcx + ylam / dcy + zlam / dcz + 1.0e-14;
                    dt = fmin(dt, 1.0 / dti);
                }
            }
        }
    }
    return dt;
}

/* ------------------------------------------------------------------ */
/* Entropy viscosity                                                   */
/* ------------------------------------------------------------------ */
/* Gauss-point gradient of one field given on the three staggered grids:
 * Gauss_Deriv (SolnDerivsc.f90:14-83). w[d][np] scratch, out[3][np]. */
static void gauss_deriv_field(const ops_t *o, const double *gmg, const double *s0, const double *s1,
                              const double *s2, double *out_x, double *out_y, double *out_z)
{
    const int ncg = o->ncg, nc = o->nc, np = o->np;
    double W[3][729];
    const double *Sd[3] = {s0, s1, s2};
    for (int d = 0; d < 3; ++d)
        for (int b = 0; b < ncg; ++b)
            for (int a = 0; a < ncg; ++a)
                for (int k = 0; k < ncg; ++k) {
                    double s = 0.0;
                    for (int j = 0; j < nc; ++j) {
                        double t = o->dmx[k * nc + j] * Sd[d][sidx(d, j, a, b, ncg, nc)];
                        s = (j == 0) ? t : s + t;
                    }
                    W[d][gidx(d, k, a, b, ncg)] = s;
                }
    for (int pt = 0; pt < np; ++pt) {
        const double *G9 = gmg + (size_t)pt * 9;
        const double w1 = W[0][pt], w2 = W[1][pt], w3 = W[2][pt];
        out_x[pt] = G9[0] * w1 + G9[3] * w2 + G9[6] * w3;
        out_y[pt] = G9[1] * w1 + G9[4] * w2 + G9[7] * w3;
        out_z[pt] = G9[2] * w1 + G9[5] * w2 + G9[8] * w3;
    }
}

/* calculate_terms (EV_routines/Main/RHS.f90:207-278): residual and Ducros
 * sensor.  The second-derivative term (lam) and the density sensor are
 * computed there but reach no output of the EV path, so they are omitted. */
static void ev_terms(const or_phys *ph, const double *Q, const double *Qx, const double *Qy, const double *Qz,
                     const double *T1, double *rhs, double *ducros)
{
    const double gamma = ph->gamma;
    const double rho = Q[0];
    const double u = Q[1] / rho, v = Q[2] / rho, w = Q[3] / rho;
    const double p = (gamma - 1.0) * (Q[4] - 0.5 * rho * (u * u + v * v + w * w));
    const double T = p * gamma / rho;
    const double ux = (Qx[1] - u * Qx[0]) / rho, uy = (Qy[1] - u * Qy[0]) / rho, uz = (Qz[1] - u * Qz[0]) / rho;
    const double vx = (Qx[2] - v * Qx[0]) / rho, vy = (Qy[2] - v * Qy[0]) / rho, vz = (Qz[2] - v * Qz[0]) / rho;
    const double wx = (Qx[3] - w * Qx[0]) / rho, wy = (Qy[3] - w * Qy[0]) / rho, wz = (Qz[3] - w * Qz[0]) / rho;
    const double tx = T1[0], ty = T1[1], tz = T1[2];
    const double s1 = ux + vy + wz;
    const double phi = (2.0 * (ux * ux + vy * vy + wz * wz) + ((vx + uy) * (vx + uy) + (uz + wx) * (uz + wx)
                        + (vz + wy) * (vz + wy)) - 2.0 * (s1 * s1) / 3.0);
    const double gam = (tx * tx + ty * ty + tz * tz) / (T * (gamma - 1.0) * ph->pr);
    const double div = ux + vy + wz;
    const double cur = sqrt((wy - vz) * (wy - vz) + (uz - wx) * (uz - wx) + (vx - uy) * (vx - uy));
    *rhs = (((phi + gam) / T) / ph->re);
    *ducros = fmax(0.0, fmin(1.0, (div * div / (div * div + cur * cur + 1.0e-9))));
}

/* entropy_viscosity_smooting (EV_routines/Filter/viscosity_smoothing.f90:4-195):
 * in-place (Gauss-Seidel order) smoothing of one element's mu at the Gauss
 * points: corners, the six faces, then the interior. */
static double jq_of(const double *JQ, int e, int v, int pt, int np)
{
    return JQ[((size_t)e * NEQ + v) * np + pt];
}

static void ev_smooth(double *mu, int ncg)
{
#define MU(n, m, l) mu[((n) - 1) + ncg * (((m) - 1) + ncg * ((l) - 1))]
    const double SIXTH = 1.0 / 6.0, TENTH = 0.10, TWELVETH = 1.0 / 12.0;
    const int N = ncg;
    int n, m, l;
    n = 1; m = 1; l = 1;
    MU(1, 1, 1) = SIXTH * (3.0 * MU(n, m, l) + MU(n + 1, m, l) + MU(n, m + 1, l) + MU(n, m, l + 1));
    n = N; m = 1; l = 1;
    MU(n, m, l) = SIXTH * (3.0 * MU(n, m, l) + MU(n - 1, m, l) + MU(n, m + 1, l) + MU(n, m, l + 1));
    n = 1; m = N; l = 1;
    MU(n, m, l) = SIXTH * (3.0 * MU(n, m, l) + MU(n + 1, m, l) + MU(n, m - 1, l) + MU(n, m, l + 1));
    n = 1; m = 1; l = N;
    MU(n, m, l) = SIXTH * (3.0 * MU(n, m, l) + MU(n + 1, m, l) + MU(n, m + 1, l) + MU(n, m, l - 1));
    n = 1; m = N; l = N;
    MU(n, m, l) = SIXTH * (3.0 * MU(n, m, l) + MU(n + 1, m, l) + MU(n, m - 1, l) + MU(n, m, l - 1));
    n = N; m = 1; l = N;
    MU(n, m, l) = SIXTH * (3.0 * MU(n, m, l) + MU(n - 1, m, l) + MU(n, m + 1, l) + MU(n, m, l - 1));
    n = N; m = N; l = 1;
    MU(n, m, l) = SIXTH * (3.0 * MU(n, m, l) + MU(n - 1, m, l) + MU(n, m - 1, l) + MU(n, m, l + 1));
    n = N; m = N; l = N;
    MU(n, m, l) = SIXTH * (3.0 * MU(n, m, l) + MU(n - 1, m, l) + MU(n, m - 1, l) + MU(n, m, l - 1));
    n = 1;
    for (m = 2; m <= N - 1; ++m)
        for (l = 2; l <= N - 1; ++l)
            MU(n, m, l) = TENTH * (5.0 * MU(n, m, l) + MU(n + 1, m, l) + MU(n, m - 1, l) + MU(n, m + 1, l)
                                   + MU(n, m, l + 1) + MU(n, m, l - 1));
    m = 1;
    for (n = 2; n <= N - 1; ++n)
        for (l = 2; l <= N - 1; ++l)
            MU(n, m, l) = TENTH * (5.0 * MU(n, m, l) + MU(n, m + 1, l) + MU(n - 1, m, l) + MU(n + 1, m, l)
                                   + MU(n, m, l + 1) + MU(n, m, l - 1));
    l = 1;
    for (n = 2; n <= N - 1; ++n)
        for (m = 2; m <= N - 1; ++m)
            MU(n, m, l) = TENTH * (5.0 * MU(n, m, l) + MU(n, m, l + 1) + MU(n, m - 1, l) + MU(n, m + 1, l)
                                   + MU(n - 1, m, l) + MU(n + 1, m, l));
    n = N;
    for (m = 2; m <= N - 1; ++m)
        for (l = 2; l <= N - 1; ++l)
            MU(n, m, l) = TENTH * (5.0 * MU(n, m, l) + MU(n - 1, m, l) + MU(n, m - 1, l) + MU(n, m + 1, l)
                                   + MU(n, m, l + 1) + MU(n, m, l - 1));
    m = N;
    for (n = 2; n <= N - 1; ++n)
        for (l = 2; l <= N - 1; ++l)
            MU(n, m, l) = TENTH * (5.0 * MU(n, m, l) + MU(n, m - 1, l) + MU(n - 1, m, l) + MU(n + 1, m, l)
                                   + MU(n, m, l + 1) + MU(n, m, l - 1));
    l = N;
    for (n = 2; n <= N - 1; ++n)
        for (m = 2; m <= N - 1; ++m)
            MU(n, m, l) = TENTH * (5.0 * MU(n, m, l) + MU(n, m, l - 1) + MU(n, m - 1, l) + MU(n, m + 1, l)
                                   + MU(n - 1, m, l) + MU(n + 1, m, l));
    for (n = 2; n <= N - 1; ++n)
        for (m = 2; m <= N - 1; ++m)
            for (l = 2; l <= N - 1; ++l)
                MU(n, m, l) = TWELVETH * (6.0 * MU(n, m, l) + MU(n + 1, m, l) + MU(n - 1, m, l) + MU(n, m + 1, l)
                                          + MU(n, m - 1, l) + MU(n, m, l + 1) + MU(n, m, l - 1));
#undef MU
}

/* RHS_ent_trans (RHS.f90:21-190) + Entropy_Viscosity (EViscosity_GL.f90:19-101)
 * + smoothing + EV_Prolong (Prolong_gtol.f90:7-51).  delta_g: staticSmag.f90:191-201.
 * The Strain routine called in between (EViscosity_GL.f90:314-348) stores a
 * field that nothing on this path reads, so it is not restated. */
/* The density-based sensor of calculate_terms (RHS.f90:263-275), from the
 * Gauss-point density gradient: sqrt of the sum over the three direction
 * pairs of 1/(1/(r_a^2+eps) + 1/(r_b^2+eps) + 1), eps = 1e-9.  It feeds the
 * Smagorinsky model's rhoSensor switch (staticSmag.f90:69). */
static double rho_sensor(double rx, double ry, double rz)
{
    const double eps = 1.0e-9;
    const double rhoX2 = rx * rx, rhoY2 = ry * ry, rhoZ2 = rz * rz;
    return sqrt((1.0 / ((1.0 / (rhoX2 + eps)) + (1.0 / (rhoY2 + eps)) + 1.0)) +
                (1.0 / ((1.0 / (rhoX2 + eps)) + (1.0 / (rhoZ2 + eps)) + 1.0)) +
                (1.0 / ((1.0 / (rhoY2 + eps)) + (1.0 / (rhoZ2 + eps)) + 1.0)));
}

int or_ev_viscosity(const or_mesh *mesh, const or_phys *ph, const or_ev *ev, const double *JQ,
                    const double *stag_lag, double *muhg, double *muh, double *resid, double *ducros,
                    double *rhosens)
{
    ops_t o;
    if (make_ops(mesh->p, &o)) return -1;
    const int K = mesh->K, ncg = o.ncg, nc = o.nc, np = o.np, ns = o.ns;
    const double gamma = ph->gamma;
#pragma omp parallel for schedule(static)
    for (int e = 0; e < K; ++e) {
        double Tst[3][810], G[6][3][729], mu[729];
        const double *S = stag_lag + (size_t)e * 3 * NEQ * ns;
        const double *gmg = mesh->gmetg + (size_t)e * np * 9;
        /* temperature on the staggered grids: P*gamma/rho */
        for (int d = 0; d < 3; ++d)
            for (int s = 0; s < ns; ++s) {
                const double *q = S + (size_t)d * NEQ * ns;
                const double q1 = q[0 * ns + s];
                const double P = (gamma - 1.0) * (q[4 * ns + s] - 0.5 * q1 * ((q[1 * ns + s] / q1) * (q[1 * ns + s] / q1)
                                  + (q[2 * ns + s] / q1) * (q[2 * ns + s] / q1) + (q[3 * ns + s] / q1) * (q[3 * ns + s] / q1)));
                Tst[d][s] = P * gamma / q1;
            }
        for (int v = 0; v < NEQ; ++v)
            gauss_deriv_field(&o, gmg, S + (size_t)(0 * NEQ + v) * ns, S + (size_t)(1 * NEQ + v) * ns,
                              S + (size_t)(2 * NEQ + v) * ns, G[v][0], G[v][1], G[v][2]);
        gauss_deriv_field(&o, gmg, Tst[0], Tst[1], Tst[2], G[5][0], G[5][1], G[5][2]);
        const double *jq = JQ + (size_t)e * NEQ * np;
        const double *jac = mesh->jacob + (size_t)e * np;
        const double c3 = pow((double)(ncg * ncg * ncg), 1.0 / 3.0);
        /* cmax_cap: the cap cmax * max(rho) * h * max(|u| + c) over the element
         * (the commented-out line EViscosity_GL.f90:82 with its maxima from
         * :44-67), with the physical pressure and sound speed c = sqrt(gamma p / rho) */
        double rhomax = 0.0, utmax = 0.0;
        if (ev->cmax_cap)
            for (int pt = 0; pt < np; ++pt) {
                double Q[NEQ];
                for (int v = 0; v < NEQ; ++v) Q[v] = jq_of(JQ, e, v, pt, np) * mesh->jacob[(size_t)e * np + pt];
                const double rho = Q[0], u = Q[1] / rho, v = Q[2] / rho, w = Q[3] / rho;
                const double p = (gamma - 1.0) * (Q[4] - 0.5 * rho * (u * u + v * v + w * w));
                rhomax = fmax(rhomax, rho);
                utmax = fmax(utmax, sqrt(u * u + v * v + w * w) + sqrt(gamma * p / rho));
            }
        for (int pt = 0; pt < np; ++pt) {
            double Q[NEQ], Qx[NEQ], Qy[NEQ], Qz[NEQ], T1[3], rhs, duc;
            for (int v = 0; v < NEQ; ++v) {
                Q[v] = jq[v * np + pt] * jac[pt];
                Qx[v] = G[v][0][pt];
                Qy[v] = G[v][1][pt];
                Qz[v] = G[v][2][pt];
            }
            T1[0] = G[5][0][pt]; T1[1] = G[5][1][pt]; T1[2] = G[5][2][pt];
            ev_terms(ph, Q, Qx, Qy, Qz, T1, &rhs, &duc);
            if (resid) resid[(size_t)e * np + pt] = rhs;
            if (ducros) ducros[(size_t)e * np + pt] = duc;
            if (rhosens) rhosens[(size_t)e * np + pt] = rho_sensor(Qx[0], Qy[0], Qz[0]);
            const double rho = jq[0 * np + pt] * jac[pt];
            const double h = pow(fabs(1.0 / jac[pt]), 1.0 / 3.0) / c3;
            double mur = ev->cmu * ph->re * rho * (h * h) * fabs(rhs);
            if (ev->ducros) mur = mur * duc;
            const double cap = ev->cmax_cap ? ev->cmax * rhomax * h * utmax : ev->mumax;
            mu[pt] = fmin(fmax(mur, 0.0), cap);
        }
        if (ev->smoothing) ev_smooth(mu, ncg);
        if (muhg)
            for (int pt = 0; pt < np; ++pt) muhg[(size_t)e * np + pt] = mu[pt];
        /* prolong to the three staggered grids: interp -> mxv */
        for (int d = 0; d < 3; ++d)
            for (int b = 0; b < ncg; ++b)
                for (int a = 0; a < ncg; ++a)
                    for (int j = 0; j < nc; ++j) {
                        double s = 0.0;
                        for (int k = 0; k < ncg; ++k) {
                            double t = o.bx[j * ncg + k] * mu[gidx(d, k, a, b, ncg)];
                            s = (k == 0) ? t : s + t;
                        }
                        muh[((size_t)e * 3 + d) * ns + sidx(d, j, a, b, ncg, nc)] = s;
                    }
    }
    return 0;
}

int or_rk_step_ev(const or_mesh *mesh, const or_phys *ph, const or_ev *ev, double *JQ, double *G, double dt,
                  double *stag_lag, int have_lag, double *muh, double *resid, int nthreads)
{
    double ark[5], brk[5], crk[5];
    or_rkcoefs(ark, brk, crk);
    const int np = (mesh->p + 1) * (mesh->p + 1) * (mesh->p + 1);
    const int ns = (mesh->p + 2) * (mesh->p + 1) * (mesh->p + 1);
    const size_t n = (size_t)mesh->K * NEQ * np;
    if (have_lag) {
        int rc = or_ev_viscosity(mesh, ph, ev, JQ, stag_lag, NULL, muh, NULL, NULL, NULL);
        if (rc) return rc;
    } else {
        memset(muh, 0, sizeof(double) * (size_t)mesh->K * 3 * ns);
    }
    double *Qt = (double *)malloc(sizeof(double) * n);
    if (!Qt) return -2;
    memset(G, 0, sizeof(double) * n);
    double last = 0.0;
    for (int kk = 0; kk < 5; ++kk) {
        or_debug dbg = {kk == 4 ? stag_lag : NULL, NULL, NULL, NULL, NULL};
        int rc = or_rhs(mesh, ph, JQ, muh, Qt, kk == 4 ? &dbg : NULL, nthreads);
        if (rc) { free(Qt); return rc; }
        const double a = ark[kk], c = crk[kk];
#pragma omp parallel for schedule(static)
        for (int e = 0; e < mesh->K; ++e)
            for (int i = 0; i < NEQ * np; ++i) {
                const size_t g = (size_t)e * NEQ * np + i;
                G[g] = a * G[g] + Qt[g];
                JQ[g] = JQ[g] + c * dt * G[g];
            }
        last = 0.0;
        const double *q = Qt + (size_t)(mesh->K - 1) * NEQ * np;
        for (int i = 0; i < NEQ * np; ++i) last = fmax(fabs(q[i]), last);
    }
    if (resid) *resid = last;
    free(Qt);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Kinetic energy and enstrophy-dissipation diagnostics               */
/* ------------------------------------------------------------------ */
/* Gauss-Legendre rule on [-1,1], nodes ascending: Newton iteration on the
 * three-term recurrence of P_n (textbook).  The reference obtains the same
 * rule from gaussq (src/Independentroutines/legendre_routines.f90:98-150,
 * Golub-Welsch); both agree to rounding. */
void or_gauss_legendre(int n, double *t, double *w)
{
    for (int i = 0; i < n; ++i) {
        double x = cos(PI_REF * (i + 0.75) / (n + 0.5)), p0 = 1.0, p1 = x, dp = 1.0;
        for (int it = 0; it < 100; ++it) {
            p0 = 1.0;
            p1 = x;
            for (int k = 2; k <= n; ++k) {
                const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dp = (n == 1) ? 1.0 : n * (x * p1 - p0) / (x * x - 1.0);
            const double dx = p1 / dp;
            x -= dx;
            if (fabs(dx) < 1e-16) break;
        }
        p0 = 1.0;
        p1 = x;
        for (int k = 2; k <= n; ++k) {
            const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        dp = (n == 1) ? 1.0 : n * (x * p1 - p0) / (x * x - 1.0);
        t[n - 1 - i] = x;
        w[n - 1 - i] = 2.0 / ((1.0 - x * x) * dp * dp);
    }
}

/* bmatu1b (caseFiles/tgv/BackstepProblem-stoch.f90:345-361): the tensor
 * quadrature weights of the Chebyshev-Gauss interpolant on [0,1]^3,
 * bmat(i,j,l) = sum_{k,m,n} 1/8 w_k w_m w_n h_i(x_k) h_j(x_m) h_l(x_n),
 * x = (t+1)/2 over the ncg-point Gauss-Legendre rule; index i + ncg(j + ncg l). */
int or_ke_weights(int p, double *bmat)
{
    ops_t o;
    if (make_ops(p, &o)) return -1;
    const int ncg = o.ncg;
    double t[16], w[16];
    or_gauss_legendre(ncg, t, w);
    for (int i = 1; i <= ncg; ++i)
        for (int j = 1; j <= ncg; ++j)
            for (int l = 1; l <= ncg; ++l) {
                double sum = 0.0;
                for (int k = 1; k <= ncg; ++k)
                    for (int m = 1; m <= ncg; ++m)
                        for (int n = 1; n <= ncg; ++n) {
                            const double hl = polyn(l, 0.5 * (t[n - 1] + 1.0), ncg, o.cxg);
                            const double hj = polyn(j, 0.5 * (t[m - 1] + 1.0), ncg, o.cxg);
                            const double hi = polyn(i, 0.5 * (t[k - 1] + 1.0), ncg, o.cxg);
                            sum = sum + 0.1250 * w[k - 1] * w[m - 1] * w[n - 1] * hi * hj * hl;
                        }
                bmat[(i - 1) + ncg * ((j - 1) + ncg * (l - 1))] = sum;
            }
    return 0;
}

/* Turbulent kinetic energy of do_user_stuff
 * (caseFiles/tgv/BackstepProblem-stoch.f90:629-650,653-657,679):
 * turbk = 1/2 sum_e sum_ijl (u^2 + v^2 + w^2) bmat(i,j,l) / jacob_1(1,1,1) / (8 pi^3),
 * u = (JQ)_2/(JQ)_1.  The per-point Jacobian is that of the first element's
 * first point, as in the reference (exact on a uniform mesh).  Loop order
 * e, i, j, l (l innermost), one running sum per component. */
int or_kinetic_energy(const or_mesh *mesh, const double *JQ, double *Ek)
{
    const int ncg = mesh->p + 1, np = ncg * ncg * ncg;
    double bmat[729];
    if (or_ke_weights(mesh->p, bmat)) return -1;
    double ur = 0.0, vr = 0.0, wr = 0.0;
    for (int e = 0; e < mesh->K; ++e) {
        const double *q = JQ + (size_t)e * NEQ * np;
        for (int i = 0; i < ncg; ++i)
            for (int j = 0; j < ncg; ++j)
                for (int l = 0; l < ncg; ++l) {
                    const int pt = i + ncg * (j + ncg * l);
                    const double u = q[1 * np + pt] / q[0 * np + pt];
                    const double v = q[2 * np + pt] / q[0 * np + pt];
                    const double w = q[3 * np + pt] / q[0 * np + pt];
                    ur = ur + (u * u) * bmat[pt];
                    vr = vr + (v * v) * bmat[pt];
                    wr = wr + (w * w) * bmat[pt];
                }
    }
    const double j0 = mesh->jacob[0];
    ur = ur / j0;
    vr = vr / j0;
    wr = wr / j0;
    *Ek = 0.5 * (ur + vr + wr) / (8.0 * (PI_REF * PI_REF * PI_REF));
    return 0;
}

/* energydissipation (src/Energydissipation/energydissipation.f90:2-89):
 * eps = 100 (sum |omega|^2 / jacob) / (sum 1/jacob) / Re over every Gauss
 * point (unweighted), with the velocity gradients from Gauss_Deriv of the
 * lagged staggered states stag_lag [K][3][5][ns] and u = (JQ)_2/(JQ)_1 of the
 * current state:  du/dy = (Qy_2 - u Qy_1)/(JQ)_1, omega = curl/jacob.
 * Loop order e, l, m, n over Q(l,m,n) (n innermost). */
int or_enstrophy(const or_mesh *mesh, const or_phys *ph, const double *JQ, const double *stag_lag, double *eps)
{
    ops_t o;
    if (make_ops(mesh->p, &o)) return -1;
    const int K = mesh->K, ncg = o.ncg, np = o.np, ns = o.ns;
    double *part = (double *)malloc(sizeof(double) * 2 * (size_t)K);
    if (!part) return -2;
#pragma omp parallel for schedule(static)
    for (int e = 0; e < K; ++e) {
        double G[4][3][729];
        const double *S = stag_lag + (size_t)e * 3 * NEQ * ns;
        const double *gmg = mesh->gmetg + (size_t)e * np * 9;
        for (int v = 0; v < 4; ++v)
            gauss_deriv_field(&o, gmg, S + (size_t)(0 * NEQ + v) * ns, S + (size_t)(1 * NEQ + v) * ns,
                              S + (size_t)(2 * NEQ + v) * ns, G[v][0], G[v][1], G[v][2]);
        const double *q = JQ + (size_t)e * NEQ * np;
        const double *jac = mesh->jacob + (size_t)e * np;
        double sum = 0.0, sum2 = 0.0;
        for (int l = 0; l < ncg; ++l)
            for (int m = 0; m < ncg; ++m)
                for (int n = 0; n < ncg; ++n) {
                    const int pt = l + ncg * (m + ncg * n);
                    const double q1 = q[0 * np + pt];
                    const double udery = ((G[1][1][pt] - (q[1 * np + pt] / q1) * G[0][1][pt])) / q1;
                    const double uderz = ((G[1][2][pt] - (q[1 * np + pt] / q1) * G[0][2][pt])) / q1;
                    const double vderx = ((G[2][0][pt] - (q[2 * np + pt] / q1) * G[0][0][pt])) / q1;
                    const double vderz = ((G[2][2][pt] - (q[2 * np + pt] / q1) * G[0][2][pt])) / q1;
                    const double wderx = ((G[3][0][pt] - (q[3 * np + pt] / q1) * G[0][0][pt])) / q1;
                    const double wdery = ((G[3][1][pt] - (q[3 * np + pt] / q1) * G[0][1][pt])) / q1;
                    const double c1 = (wdery - vderz) / jac[pt];
                    const double c2 = (uderz - wderx) / jac[pt];
                    const double c3 = (vderx - udery) / jac[pt];
                    double vort = c1 * c1 + c2 * c2 + c3 * c3;
                    vort = vort / jac[pt];
                    sum2 = sum2 + 1.0 / jac[pt];
                    sum = sum + vort;
                }
        part[2 * e] = sum;
        part[2 * e + 1] = sum2;
    }
    /* the reference's running sum crosses elements in order; summing per
     * element first changes only the rounding */
    double sum = 0.0, sum2 = 0.0;
    for (int e = 0; e < K; ++e) {
        sum += part[2 * e];
        sum2 += part[2 * e + 1];
    }
    free(part);
    *eps = 100.0 * (sum / sum2) * (1.0 / ph->re);
    return 0;
}


/* ------------------------------------------------------------------ */
/* Static Smagorinsky model                                            */
/* ------------------------------------------------------------------ */
/* simpleSmag (src/turbulence/staticSmago/staticSmag.f90:32-103) with
 * filteredStrain (:126-179), delta_g (:191-201) and nut_prolong (:206-261):
 * nu_t = Re (C_s Delta)^2 S at the Gauss points, S = (2 S_ij S_ij)^(1/2) from
 * Gauss_Deriv of the lagged staggered states, Delta = |1/jacob|^(1/3) /
 * (ncg^3)^(1/3), clamped to [0, numax]; prolonged to the three staggered
 * grids with negative values set to zero.  filteredStrain takes rho =
 * (JQ)_1 jacob at the element's first point and u = (JQ)_2 / rho, as the
 * reference does (exact on uniform meshes).  The shock and density sensors
 * (shockSensor, rhoSensor) are not restated.  Outputs (nullable):
 * nut_g [K][np], nut_s [K][3][ns]. */
int or_smagorinsky(const or_mesh *mesh, const or_phys *ph, double cs, double numax, const double *JQ,
                   const double *stag_lag, double *nut_g, double *nut_s)
{
    return or_smagorinsky_sensors(mesh, ph, cs, numax, JQ, stag_lag, 0, NULL, NULL, nut_g, nut_s);
}

/* simpleSmag with its shockSensor (nu_t *= 1 - ducros) and rhoSensor
 * (nu_t *= rhoSensor) switches, staticSmag.f90:67-69, before the clamp; the
 * sensors are the Gauss-point fields that RHS_ent_trans left (nullable when
 * the switch is off). */
int or_smagorinsky_sensors(const or_mesh *mesh, const or_phys *ph, double cs, double numax, const double *JQ,
                           const double *stag_lag, int sensors, const double *ducros, const double *rhosens,
                           double *nut_g, double *nut_s)
{
    ops_t o;
    if (make_ops(mesh->p, &o)) return -1;
    const int K = mesh->K, ncg = o.ncg, nc = o.nc, np = o.np, ns = o.ns;
#pragma omp parallel for schedule(static)
    for (int e = 0; e < K; ++e) {
        double G[4][3][729], nu[729];
        const double *S = stag_lag + (size_t)e * 3 * NEQ * ns;
        const double *gmg = mesh->gmetg + (size_t)e * np * 9;
        for (int v = 0; v < 4; ++v)
            gauss_deriv_field(&o, gmg, S + (size_t)(0 * NEQ + v) * ns, S + (size_t)(1 * NEQ + v) * ns,
                              S + (size_t)(2 * NEQ + v) * ns, G[v][0], G[v][1], G[v][2]);
        const double *q = JQ + (size_t)e * NEQ * np;
        const double *jac = mesh->jacob + (size_t)e * np;
        const double c3 = pow((double)(ncg * ncg * ncg), 1.0 / 3.0);
        for (int pt = 0; pt < np; ++pt) {
            const double rho = q[0 * np + pt] * jac[0];
            const double u = q[1 * np + pt] / rho, v = q[2 * np + pt] / rho, w = q[3 * np + pt] / rho;
            const double rx = G[0][0][pt], ry = G[0][1][pt], rz = G[0][2][pt];
            const double ux = (G[1][0][pt] - rx * u) / rho, uy = (G[1][1][pt] - ry * u) / rho, uz = (G[1][2][pt] - rz * u) / rho;
            const double vx = (G[2][0][pt] - rx * v) / rho, vy = (G[2][1][pt] - ry * v) / rho, vz = (G[2][2][pt] - rz * v) / rho;
            const double wx = (G[3][0][pt] - rx * w) / rho, wy = (G[3][1][pt] - ry * w) / rho, wz = (G[3][2][pt] - rz * w) / rho;
            const double S11 = ux, S12 = 0.50 * (uy + vx), S13 = 0.50 * (uz + wx);
            const double S21 = 0.50 * (vx + uy), S22 = vy, S23 = 0.50 * (vz + wy);
            const double S31 = 0.50 * (wx + uz), S32 = 0.50 * (wy + vz), S33 = wz;
            const double Sbar = sqrt(2 * S11 * S11 + 2 * S12 * S12 + 2 * S13 * S13 + 2 * S21 * S21 + 2 * S22 * S22
                                     + 2 * S23 * S23 + 2 * S31 * S31 + 2 * S32 * S32 + 2 * S33 * S33);
            const double delta = pow(fabs(1.0 / jac[pt]), 1.0 / 3.0) / c3;
            double nt = ph->re * (cs * delta) * (cs * delta) * Sbar;
            if (sensors & OR_LES_SHOCK_SENSOR) nt = nt * (1.0 - ducros[(size_t)e * np + pt]);
            if (sensors & OR_LES_RHO_SENSOR) nt = nt * rhosens[(size_t)e * np + pt];
            nu[pt] = fmin(fmax(nt, 0.0), numax);
        }
        if (nut_g)
            for (int pt = 0; pt < np; ++pt) nut_g[(size_t)e * np + pt] = nu[pt];
        if (nut_s)
            for (int d = 0; d < 3; ++d)
                for (int b = 0; b < ncg; ++b)
                    for (int a = 0; a < ncg; ++a)
                        for (int j = 0; j < nc; ++j) {
                            double sum = 0.0;
                            for (int k = 0; k < ncg; ++k) {
                                double t = o.bx[j * ncg + k] * nu[gidx(d, k, a, b, ncg)];
                                sum = (k == 0) ? t : sum + t;
                            }
                            if (sum < 0.0) sum = 0.0;
                            nut_s[((size_t)e * 3 + d) * ns + sidx(d, j, a, b, ncg, nc)] = sum;
                        }
    }
    return 0;
}

/* tstep with the static Smagorinsky model (src/tstep.f90:40-43): nu_t from
 * stag_lag at the start of the step (zero when have_lag = 0, as for the
 * entropy viscosity), five stages with amu = 1 + nu_t rho, and stag_lag
 * replaced by the last stage's staggered states. */
int or_rk_step_les(const or_mesh *mesh, const or_phys *ph, double cs, double numax, double *JQ, double *G,
                   double dt, double *stag_lag, int have_lag, double *nut_s, double *resid, int nthreads)
{
    double ark[5], brk[5], crk[5];
    or_rkcoefs(ark, brk, crk);
    const int np = (mesh->p + 1) * (mesh->p + 1) * (mesh->p + 1);
    const int ns = (mesh->p + 2) * (mesh->p + 1) * (mesh->p + 1);
    const size_t n = (size_t)mesh->K * NEQ * np;
    if (have_lag) {
        int rc = or_smagorinsky(mesh, ph, cs, numax, JQ, stag_lag, NULL, nut_s);
        if (rc) return rc;
    } else {
        memset(nut_s, 0, sizeof(double) * (size_t)mesh->K * 3 * ns);
    }
    double *Qt = (double *)malloc(sizeof(double) * n);
    if (!Qt) return -2;
    memset(G, 0, sizeof(double) * n);
    double last = 0.0;
    for (int kk = 0; kk < 5; ++kk) {
        or_debug dbg = {kk == 4 ? stag_lag : NULL, NULL, NULL, NULL, NULL};
        int rc = or_rhs_les(mesh, ph, JQ, NULL, nut_s, Qt, kk == 4 ? &dbg : NULL, nthreads);
        if (rc) { free(Qt); return rc; }
        const double a = ark[kk], c = crk[kk];
#pragma omp parallel for schedule(static)
        for (int e = 0; e < mesh->K; ++e)
            for (int i = 0; i < NEQ * np; ++i) {
                const size_t g = (size_t)e * NEQ * np + i;
                G[g] = a * G[g] + Qt[g];
                JQ[g] = JQ[g] + c * dt * G[g];
            }
        last = 0.0;
        const double *q = Qt + (size_t)(mesh->K - 1) * NEQ * np;
        for (int i = 0; i < NEQ * np; ++i) last = fmax(fabs(q[i]), last);
    }
    if (resid) *resid = last;
    free(Qt);
    return 0;
}

/* modalfilter (src/Energydissipation/energydissipation.f90:720-959) on the
 * whole state, in place: per element and momentum field, in = Q jacob(1,1,1);
 * mid = REDFT10 in all three directions (FFTW's unnormalised DCT-II,
 * Y_k = 2 sum_j X_j cos(pi (j + 1/2) k / N)); each mode divided by 2N (index
 * 1) or N (index >= 2) per direction in the reference's loop order; modes
 * with any 0-based index >= cutoff zeroed (the commented-out edit at
 * :778-783 with its 8,9 generalised; cutoff = N keeps all, as shipped);
 * divided by 2 per direction for index >= 2; out = REDFT01 in all three
 * directions (Y_k = X_0 + 2 sum_{j>=1} X_j cos(pi j (k + 1/2) / N));
 * Q = out / jacob(1,1,1).  Direct O(N) sums (FFTW's fast algorithm is a
 * different rounding of the same transform). */
static void dct_line(const double *in, double *out, int N, int stride, int type)
{
    for (int k = 0; k < N; ++k) {
        double s = 0.0;
        for (int j = 0; j < N; ++j) {
            double c;
            if (type == 2) c = 2.0 * cos(PI_REF * (j + 0.5) * k / N);
            else c = j == 0 ? 1.0 : 2.0 * cos(PI_REF * j * (k + 0.5) / N);
            s += c * in[j * stride];
        }
        out[k * stride] = s;
    }
}

static void dct3d(double *a, double *tmp, int N, int type)
{
    /* along n (fastest), then m, then l; a is [l][m][n] */
    for (int d = 0; d < 3; ++d) {
        const int stride = d == 0 ? 1 : (d == 1 ? N : N * N);
        for (int x = 0; x < N; ++x)
            for (int y = 0; y < N; ++y) {
                int base;
                if (d == 0) base = N * (x + N * y);
                else if (d == 1) base = x + N * N * y;
                else base = x + N * y;
                double lin[9], lout[9];
                for (int j = 0; j < N; ++j) lin[j] = a[base + j * stride];
                dct_line(lin, lout, N, 1, type);
                for (int j = 0; j < N; ++j) tmp[base + j * stride] = lout[j];
            }
        memcpy(a, tmp, sizeof(double) * N * N * N);
    }
}

int or_modal_filter(const or_mesh *mesh, double *JQ, int cutoff)
{
    const int N = mesh->p + 1, np = N * N * N;
    if (cutoff < 0 || cutoff > N) return -1;
#pragma omp parallel for schedule(static)
    for (int e = 0; e < mesh->K; ++e) {
        const double j0 = mesh->jacob[(size_t)e * np];
        for (int v = 1; v <= 3; ++v) {
            double a[729], t[729];
            double *q = JQ + ((size_t)e * NEQ + v) * np;
            for (int i = 0; i < np; ++i) a[i] = q[i] * j0;
            dct3d(a, t, N, 2);
#define MID(i, j, k) a[((i) - 1) + N * (((j) - 1) + N * ((k) - 1))]
            for (int i = 1; i <= N; ++i)
                for (int j = 1; j <= N; ++j) {
                    MID(1, i, j) = MID(1, i, j) / (2.0 * N);
                    MID(i, 1, j) = MID(i, 1, j) / (2.0 * N);
                    MID(i, j, 1) = MID(i, j, 1) / (2.0 * N);
                }
            for (int i = 1; i <= N; ++i)
                for (int j = 1; j <= N; ++j)
                    for (int k = 2; k <= N; ++k) {
                        MID(i, j, k) = MID(i, j, k) / N;
                        MID(j, k, i) = MID(j, k, i) / N;
                        MID(k, i, j) = MID(k, i, j) / N;
                    }
            for (int i = 1; i <= N; ++i)
                for (int j = 1; j <= N; ++j)
                    for (int k = 1; k <= N; ++k)
                        if (i > cutoff || j > cutoff || k > cutoff) MID(i, j, k) = 0.0;
            for (int i = 1; i <= N; ++i)
                for (int j = 1; j <= N; ++j)
                    for (int k = 2; k <= N; ++k) {
                        MID(i, j, k) = MID(i, j, k) / 2;
                        MID(j, k, i) = MID(j, k, i) / 2;
                        MID(k, i, j) = MID(k, i, j) / 2;
                    }
#undef MID
            dct3d(a, t, N, 3);
            for (int i = 0; i < np; ++i) q[i] = a[i] / j0;
        }
    }
    return 0;
}
