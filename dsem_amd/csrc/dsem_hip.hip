// dsem_hip.hip -- MI355X (gfx950) kernels and C ABI for the DSEM
// staggered-grid spectral-element RHS and the 5-stage RK4 2N step.
//
// Pipeline of one RHS evaluation (one workgroup per element for volume work,
// one work-item per mortar point for face work; all FP64):
//
//   k_traces        J*Q -> Q, Gauss->Lobatto interpolation evaluated at the
//                   six faces only (Int_fluxes' face values,
//                   fluxesH_newa.f90:262-352; send_q_to_mortars, send_q.f90:491-764)
//   k_face_inv      per mortar point: Roe flux with entropy fix scaled by the
//                   face norm and the averaged state (interface_flux,
//                   mortar_operations.f90:151-238; riemann, PhysLib.f90:231-320;
//                   V_Avg_Soln, mortar_operations.f90:781-837)
//   k_volume        Gauss-point gradients of Q from the staggered states whose
//                   face values are the mortar averages (Gauss_Deriv,
//                   SolnDerivsc.f90:14-83, with the combined operator D_int*B
//                   for the interior Lobatto points), kept in LDS; then per
//                   staggered grid: interpolated state, prolonged
//                   gradients, Euler + viscous contravariant fluxes at
//                   interior Lobatto points, viscous face traces r/norm, and
//                   the interior part of the flux divergence
//                   (Int_fluxes fluxesH_newa.f90:274-352, Vis_Fluxes
//                   Vis_Fluxes_new.f90:84-228, send_flux_to_mortars
//                   send_flux.f90:243-457, update_interior updatea.f90:100-113)
//   k_face_visc     per mortar point: averaged viscous flux subtracted from the
//                   Roe flux (mortar_vis_flux ViscousBC_slide.f90:231-271,
//                   send_flux_to_faces_add send_flux.f90:699-921)
//   k_update        face part of the divergence, Qt, and the RK register
//                   update g = a g + Qt, JQ += c dt g (updatea.f90:213-227)
//
// Element-major structure-of-arrays layouts (ncg = p+1, nc = p+2):
//   JQ, G, Qtp [K][5][np]        point n + ncg*(m + ncg*l)
//   T, FV      [K][6][5][nf]     face traces, face f-1, point i + ncg*j
//   QA, FI, FT [M][5][nf]        per mortar (master face indexing)
//   metrics    jacob [K][npm], gmetg [K][9][npm], gmet_d [K][3][nsm],
//              nhat [M][3][nfm], norm [M][nfm]; npm = np (per point) or 1
//              (per element, affine meshes) -- selected by the stride ms.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/dsem_hip.h"
#include "geom.h"

namespace {

constexpr int NEQ = 5;
constexpr int BLOCK = 256;
// Fused volume kernel: one element per workgroup.  FP64 at p = 7, 8 holds
// 105-151 KB of LDS per element, so one workgroup fills a CU, and its
// matrix-core tiles (48, 63 per direction) need 1024 work-items (16 lanes
// each); at p <= 6 every direction has at most 26 tiles, so 512 work-items
// hold them all and two elements (<= 72 KB each) share a CU.  The FP32
// modes (75 KB at p = 8) fit two workgroups of 512 per CU.
template <int P, int PM>
constexpr int vblock() { return (PM == 0 && P >= 7) ? 1024 : 512; }

// operators for every order p = 1..8, indexed [p]
__constant__ double c_D[9][90];   // D[k*nc + j]
__constant__ double c_B[9][90];   // B[j*ncg + k]
__constant__ double c_M[9][81];   // M[k*ncg + k2]
__constant__ double c_cx[9][10];  // Lobatto points
// single-precision copies for the mixed contraction modes
__constant__ float c_Df[9][90];
__constant__ float c_Bf[9][90];
__constant__ float c_Mf[9][81];
__constant__ double c_Dint[9][9];  // sum over the interior Lobatto j of D[k][j]
// modal filter: unnormalised DCT-II (FFTW REDFT10) and DCT-III (REDFT01) of
// length N = ncg, C2[k*N + j] = 2 cos(pi (j + 1/2) k / N),
// C3[k*N + j] = (j == 0 ? 1 : 2 cos(pi j (k + 1/2) / N))
__constant__ double c_dct2[9][81];
__constant__ double c_dct3[9][81];

// Intermediates of dsem_rhs_device_noalloc: one copy per GPU, held until the
// process ends.  Host code uses the device address that hipGetSymbolAddress
// gives for the current device (kept per device by scratch_setup), never the
// name itself.
__device__ __align__(256) double scratch[DSEM_NOALLOC_SCRATCH_BYTES / sizeof(double)];

// Contraction arithmetic of the volume kernel, by precision mode:
//   PM 0  FP64 (the drop-in path);
//   PM 1  FP32 multiply-add, FP32 accumulation;
//   PM 2  BF16 split products: a*b ~ ah*bh + ah*bl + al*bh with
//         x = xh + xl, xh = bf16(x), xl = bf16(x - xh); each product of two
//         bf16 values is exact in FP32, accumulation in FP32.
// In the mixed modes the contracted fields are shifted by an element
// constant first (see k_volume), so that FP32 rounding acts on the
// variation within the element and not on the low-Mach background.
template <int PM>
struct Ctr {
    using T = float;
    using A = float;
};
template <>
struct Ctr<0> {
    using T = double;
    using A = double;
};

__device__ __forceinline__ float bf16_round(float x)
{
    unsigned u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return __uint_as_float(u & 0xffff0000u);
}

template <int PM>
__device__ __forceinline__ typename Ctr<PM>::A mac(typename Ctr<PM>::T a, typename Ctr<PM>::T b,
                                                    typename Ctr<PM>::A acc)
{
    if constexpr (PM == 0) {
        return fma(a, b, acc);
    } else if constexpr (PM == 1) {
        return fmaf(a, b, acc);
    } else {
        const float ah = bf16_round(a), al = bf16_round(a - ah);
        const float bh = bf16_round(b), bl = bf16_round(b - bh);
        acc = fmaf(ah, bh, acc);
        acc = fmaf(ah, bl, acc);
        return fmaf(al, bh, acc);
    }
}

template <int PM>
__device__ __forceinline__ typename Ctr<PM>::T opB(int p, int idx)
{
    if constexpr (PM == 0) return c_B[p][idx]; else return c_Bf[p][idx];
}
template <int PM>
__device__ __forceinline__ typename Ctr<PM>::T opD(int p, int idx)
{
    if constexpr (PM == 0) return c_D[p][idx]; else return c_Df[p][idx];
}
template <int PM>
__device__ __forceinline__ typename Ctr<PM>::T opM(int p, int idx)
{
    if constexpr (PM == 0) return c_M[p][idx]; else return c_Mf[p][idx];
}

template <int P>
struct Sz {
    static constexpr int NCG = P + 1, NC = P + 2;
    static constexpr int NP = NCG * NCG * NCG, NS = NC * NCG * NCG, NF = NCG * NCG;
    static constexpr int R = (NP + BLOCK - 1) / BLOCK;
};

struct DevMesh {
    int K, M, ms;                 // ms: 1 per-point metrics, 0 per-element
    int rot;                      // some mortar face is rotated (orientation != 0)
    const double *jacob, *gmetg, *gmet0, *gmet1, *gmet2;
    const int *e2m;               // [K][6] (mortar << 4) | (map << 1) | side, see e2m_mortar
    const int *m_elem, *m_face, *m_sign, *m_nsign;
    const int *m_orient;          // [M] slave orientation 0..7, or null when every mortar is aligned
    const double *nhat, *norm;
};

struct DevPhys {
    double gamma, re, pr, mach;
    double gm1;        // gamma - 1
    double tfac;       // mach^2 gamma          (T = tfac p / rho)
    double hfac1;      // 1 / (pr mach^2 (gamma-1))
    double kap;        // 0.2 / (gamma-1)       (kappa = 1 + (mu-1) kap)
    double inv_re;
};

// Lobatto index j of grid d at line (a,b) -> staggered index
template <int P>
__device__ __forceinline__ int sidx(int d, int j, int a, int b)
{
    using S = Sz<P>;
    return d == 0 ? j + S::NC * (a + S::NCG * b) : (d == 1 ? a + S::NCG * (j + S::NC * b) : a + S::NCG * (b + S::NCG * j));
}

// Gauss index k along direction d at line (a,b)
template <int P>
__device__ __forceinline__ int gidx(int d, int k, int a, int b)
{
    using S = Sz<P>;
    return d == 0 ? k + S::NCG * (a + S::NCG * b) : (d == 1 ? a + S::NCG * (k + S::NCG * b) : a + S::NCG * (b + S::NCG * k));
}

// face (1..6) reached at Lobatto end j of grid d
template <int P>
__device__ __forceinline__ int face_of(int d, int j)
{
    const bool hi = (j != 0);
    return d == 0 ? (hi ? 4 : 6) : (d == 1 ? (hi ? 2 : 1) : (hi ? 5 : 3));
}

// Face point maps of rotated conforming faces.  Orientation codes DFLT,
// B1F2, B1B2, F1B2, F2F1, B2F1, B2B1, F2B1 = 0..7 (element_library.f90:99-101):
// mortar point (I, J) takes the slave face's point orient_point(o, I + N J)
// (rotate_to_mortar, mortar_operations.f90:461-605); rotate_flux_to_face
// (:616-758) is the inverse map, which is the same code except that 5 and 7
// exchange.  Written as an optional swap of (I, J) and a reflection of each.
template <int N>
__device__ __forceinline__ int orient_point(int o, int q)
{
    const int I = q % N, J = q / N;
    const bool sw = o >= 4;
    const bool fi = (o == 1) | (o == 2) | (o == 6) | (o == 7);
    const bool fj = (o == 2) | (o == 3) | (o == 5) | (o == 6);
    const int a = sw ? J : I, b = sw ? I : J;
    return (fi ? N - 1 - a : a) + N * (fj ? N - 1 - b : b);
}

// e2m entry of an element face: its mortar, the side (0 master, 1 slave) and
// the map code that takes the element's own face point to the mortar point
// (0 on master sides and aligned mortars, the inverse orientation on
// rotated slave faces).
__device__ __forceinline__ int e2m_mortar(int c) { return c >> 4; }
__device__ __forceinline__ int e2m_map(int c) { return (c >> 1) & 7; }

// ------------------------------------------------------------ pointwise physics
// Pointwise physics.  Euler fluxes: fflux/gflux/hflux (PhysLib.f90:17-124).
// Viscous fluxes x/y/z_vis_flux (PhysLib.f90:419-583), split so that the
// conservative gradients of one direction are consumed as soon as they are
// formed: prim_gradient turns dQ/dx_c into (du/dx_c, dv/dx_c, dw/dx_c,
// dT/dx_c) (PhysLib.f90:448-462); the stress and heat flux are formed in
// the volume kernel.  Both families are contracted with one metric row
// J grad xi_d as Int_fluxes/Vis_Fluxes do.
struct Prim {
    double rho, rinv, u, v, w, q2, p;
};

// Divisions by rho are multiplications by 1/rho here (one rounding apart
// from the reference's quotients).
__device__ __forceinline__ Prim primitives(const DevPhys &ph, const double *Q)
{
    Prim s;
    s.rho = Q[0];
    s.rinv = 1.0 / s.rho;
    s.u = Q[1] * s.rinv;
    s.v = Q[2] * s.rinv;
    s.w = Q[3] * s.rinv;
    s.q2 = s.u * s.u + s.v * s.v + s.w * s.w;
    s.p = ph.gm1 * (Q[4] - 0.5 * s.rho * s.q2);
    return s;
}

__device__ __forceinline__ void prim_gradient(const DevPhys &ph, const Prim &s, const double *Qc, double *out)
{
    const double uc = (Qc[1] - s.u * Qc[0]) * s.rinv;
    const double vc = (Qc[2] - s.v * Qc[0]) * s.rinv;
    const double wc = (Qc[3] - s.w * Qc[0]) * s.rinv;
    const double pc = ph.gm1 * (Qc[4] - 0.5 * Qc[0] * s.q2 - s.rho * (s.u * uc + s.v * vc + s.w * wc));
    out[0] = uc;
    out[1] = vc;
    out[2] = wc;
    out[3] = (pc - s.p * Qc[0] * s.rinv) * ph.tfac * s.rinv;
}

// Euler flux contracted with one metric row, from primitives.
__device__ __forceinline__ void euler_from_primitive(const Prim &s, const double *Q, double g0, double g1, double g2,
                                                     double *out)
{
    const double qn = g0 * s.u + g1 * s.v + g2 * s.w;   // contravariant velocity
    out[0] = g0 * Q[1] + g1 * Q[2] + g2 * Q[3];
    out[1] = g0 * (s.p + Q[1] * s.u) + g1 * (Q[1] * s.v) + g2 * (Q[1] * s.w);
    out[2] = g0 * (Q[1] * s.v) + g1 * (s.p + Q[2] * s.v) + g2 * (Q[3] * s.v);
    out[3] = g0 * (Q[1] * s.w) + g1 * (Q[2] * s.w) + g2 * (s.p + Q[3] * s.w);
    out[4] = qn * (Q[4] + s.p);
}

// Roe flux with entropy fix (PhysLib.f90:231-320)
__device__ __forceinline__ void roe_flux(const DevPhys &ph, double n0, double n1, double n2, double ds, const double *Ql,
                                         const double *Qr, double *flux)
{
    const double gamma = ph.gamma;
    const double rho = Ql[0], rhon = Qr[0];
    const double ul = Ql[1] / rho, vl = Ql[2] / rho, wl = Ql[3] / rho;
    const double pleft = (gamma - 1.0) * (Ql[4] - 0.5 / rho * (Ql[1] * Ql[1] + Ql[2] * Ql[2] + Ql[3] * Ql[3]));
    const double ur = Qr[1] / rhon, vr = Qr[2] / rhon, wr = Qr[3] / rhon;
    const double pright = (gamma - 1.0) * (Qr[4] - 0.5 / rhon * (Qr[1] * Qr[1] + Qr[2] * Qr[2] + Qr[3] * Qr[3]));
    const double ql = n0 * ul + n1 * vl + n2 * wl;
    const double qr = n0 * ur + n1 * vr + n2 * wr;
    const double gg = gamma / (gamma - 1.0);
    const double hl = 0.5 * (ul * ul + vl * vl + wl * wl) + gg * pleft / rho;
    const double hr = 0.5 * (ur * ur + vr * vr + wr * wr) + gg * pright / rhon;
    const double rtd = sqrt(rho * rhon);
    const double betal = rho / (rho + rtd);
    const double betar = 1.0 - betal;
    const double utd = betal * ul + betar * ur, vtd = betal * vl + betar * vr, wtd = betal * wl + betar * wr;
    const double htd = betal * hl + betar * hr;
    const double atd2 = (gamma - 1.0) * (htd - 0.5 * (utd * utd + vtd * vtd + wtd * wtd));
    const double atd = sqrt(atd2);
    const double qtd = utd * n0 + vtd * n1 + wtd * n2;
    const double hcoef = (gamma + 1.0) / 4.0 * atd / rtd;
    if (qtd >= 0.0) {
        const double dw1 = 0.5 * ((pright - pleft) / atd2 - (qr - ql) * rtd / atd);
        const double sp1 = qtd - atd;
        const double sp1m = fmin(sp1, 0.0);
        const double eta1 = fmax(-fabs(sp1) - hcoef * dw1, 0.0);
        const double udw1 = dw1 * (sp1m - 0.5 * eta1);
        const double rql = rho * ql;
        flux[0] = ds * (rql + udw1);
        flux[1] = ds * (rql * ul + pleft * n0 + udw1 * (utd - atd * n0));
        flux[2] = ds * (rql * vl + pleft * n1 + udw1 * (vtd - atd * n1));
        flux[3] = ds * (rql * wl + pleft * n2 + udw1 * (wtd - atd * n2));
        flux[4] = ds * (rql * hl + udw1 * (htd - qtd * atd));
    } else {
        const double dw4 = 0.5 * ((pright - pleft) / atd2 + (qr - ql) * rtd / atd);
        const double sp4 = qtd + atd;
        const double sp4p = fmax(sp4, 0.0);
        const double eta4 = fmax(-fabs(sp4) + hcoef * dw4, 0.0);
        const double udw4 = dw4 * (sp4p + 0.5 * eta4);
        const double rqr = rhon * qr;
        flux[0] = ds * (rqr - udw4);
        flux[1] = ds * (rqr * ur + pright * n0 - udw4 * (utd + atd * n0));
        flux[2] = ds * (rqr * vr + pright * n1 - udw4 * (vtd + atd * n1));
        flux[3] = ds * (rqr * wr + pright * n2 - udw4 * (wtd + atd * n2));
        flux[4] = ds * (rqr * hr - udw4 * (htd + qtd * atd));
    }
}

// ------------------------------------------------------------ kernels
// Face traces of the physical state (Int_fluxes face values, send_q_to_mortars):
// the element's state is staged in LDS with batched loads, then each trace
// is a 9-term sum over LDS.
template <int P>
__global__ __launch_bounds__(BLOCK) void k_traces(const double *__restrict__ JQ, DevMesh m, double *__restrict__ T)
{
    using S = Sz<P>;
    constexpr int NQ = NEQ * S::NP, TB = 8;   // loads in flight per work-item
    __shared__ double sQ[NQ];
    __shared__ double sBf[2 * S::NCG];   // B[0][k], B[nc-1][k]
    const int e = blockIdx.x;
    const double *q = JQ + (size_t)e * NQ;
    const double *jac = m.jacob + (size_t)e * (m.ms ? S::NP : 1);
    if (threadIdx.x < S::NCG) {
        sBf[threadIdx.x] = c_B[P][threadIdx.x];
        sBf[S::NCG + threadIdx.x] = c_B[P][(S::NC - 1) * S::NCG + threadIdx.x];
    }
    for (int i0 = threadIdx.x; i0 < NQ; i0 += BLOCK * TB) {
        double qv[TB], jv[TB];
#pragma unroll
        for (int u = 0; u < TB; ++u) {
            const int i = i0 + u * BLOCK, ic = i < NQ ? i : NQ - 1;
            qv[u] = q[ic];
            jv[u] = jac[(ic % S::NP) * m.ms];
        }
#pragma unroll
        for (int u = 0; u < TB; ++u) {
            const int i = i0 + u * BLOCK;
            if (i < NQ) sQ[i] = qv[u] * jv[u];
        }
    }
    __syncthreads();
    double *t = T + (size_t)e * 6 * NEQ * S::NF;
    for (int i = threadIdx.x; i < 6 * NEQ * S::NF; i += BLOCK) {
        const int qf = i % S::NF, v = (i / S::NF) % NEQ, f = i / (NEQ * S::NF) + 1;
        const int a = qf % S::NCG, b = qf / S::NCG;
        const int d = (f == 4 || f == 6) ? 0 : ((f == 1 || f == 2) ? 1 : 2);
        const double *Bj = sBf + ((f == 2 || f == 4 || f == 5) ? S::NCG : 0);
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < S::NCG; ++k) s = fma(Bj[k], sQ[v * S::NP + gidx<P>(d, k, a, b)], s);
        t[i] = s;
    }
}

template <int P>
__global__ __launch_bounds__(BLOCK) void k_face_inv(const double *__restrict__ T, DevMesh m, DevPhys ph,
                                                     double *__restrict__ FI, double *__restrict__ QA)
{
    using S = Sz<P>;
    const int idx = blockIdx.x * BLOCK + threadIdx.x;
    if (idx >= m.M * S::NF) return;
    const int mm = idx / S::NF, q = idx % S::NF;
    const int em = m.m_elem[2 * mm], es = m.m_elem[2 * mm + 1];
    const int fm = m.m_face[2 * mm], fs = m.m_face[2 * mm + 1];
    double Qm[5], Qs[5], fl[5];
    const int qs = m.m_orient ? orient_point<S::NCG>(m.m_orient[mm], q) : q;   // rotate_to_mortar
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        Qm[v] = T[(((size_t)em * 6 + fm - 1) * NEQ + v) * S::NF + q];
        Qs[v] = T[(((size_t)es * 6 + fs - 1) * NEQ + v) * S::NF + qs];
    }
    const int nfm = m.ms ? S::NF : 1, qq = q * m.ms;
    const double n0 = m.nhat[((size_t)mm * 3 + 0) * nfm + qq];
    const double n1 = m.nhat[((size_t)mm * 3 + 1) * nfm + qq];
    const double n2 = m.nhat[((size_t)mm * 3 + 2) * nfm + qq];
    const double ds = m.norm[(size_t)mm * nfm + qq];
    if (fm == 1 || fm == 3 || fm == 6) roe_flux(ph, n0, n1, n2, ds, Qs, Qm, fl);
    else roe_flux(ph, n0, n1, n2, ds, Qm, Qs, fl);
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        FI[((size_t)mm * NEQ + v) * S::NF + q] = fl[v];
        QA[((size_t)mm * NEQ + v) * S::NF + q] = (Qm[v] + Qs[v]) * 0.5;
    }
}

// Line contractions of the FP64 volume kernel on the matrix cores.
// v_mfma_f64_4x4x4f64 runs four independent 4x4x4 products per wave.  Lane l
// (c = l%4, blk = (l/4)%4, r = l/16) supplies A_blk[c][r] and B_blk[r][c]
// and receives D_blk[r][c] (layout measured on gfx950).  With A = four rows
// of a ROWS x ncg line operator (the Gauss->Lobatto interpolation B, ROWS =
// nc, or the Gauss derivative M = D_int B, ROWS = ncg) and B = four Gauss
// values of four lines, D[r][c] is the result at row 4 jt + r of line 4 g + c,
// the point this lane owns: the pointwise work that follows runs on it
// without staging, and each Gauss value is read from LDS once per 4x4 tile
// instead of once per output point.  ROWS x ncg^2 points form 4x4 tiles over
// (row, line); tile t = 4 * wave + blk, t = g * JT + jt.
template <int P, int ROWS = Sz<P>::NC>
struct MfTiles {
    using S = Sz<P>;
    static constexpr int JT = (ROWS + 3) / 4;      // tiles along the row index
    static constexpr int KS = (S::NCG + 3) / 4;    // steps of 4 Gauss points
    static constexpr int NL = S::NCG * S::NCG;     // lines per direction
    static constexpr int NT = ((NL + 3) / 4) * JT; // tiles per direction
};

template <int P, int ROWS = Sz<P>::NC>
struct MfLane {
    double A[MfTiles<P, ROWS>::KS];   // op[4 jt + c][4 ks + r], zero outside the operator
    int r, j, line;                   // line is also the line this lane loads
    bool lvalid;                      // tile and line exist (the lane loads)
    bool valid;                       // and its row exists (the lane owns a point)
    bool wave_active;                 // some tile of this wave exists (uniform per wave)
};

// A value the compiler cannot see through (an empty asm that may change it):
// keeps index arithmetic derived from it inside the loop that makes the copy.
__device__ __forceinline__ int opaque(int x)
{
    asm volatile("" : "+v"(x));
    return x;
}

template <int P, int ROWS = Sz<P>::NC>
__device__ __forceinline__ MfLane<P, ROWS> mf_lane(int tid, const double *op)
{
    using S = Sz<P>;
    using T = MfTiles<P, ROWS>;
    MfLane<P, ROWS> ml;
    const int l = tid % 64, c = l % 4, blk = (l / 4) % 4;
    ml.r = l / 16;
    const int t = (tid / 64) * 4 + blk;
    const int jt = t % T::JT, g = t / T::JT;
    ml.j = 4 * jt + ml.r;
    ml.line = 4 * g + c;
    ml.lvalid = t < T::NT && ml.line < T::NL;
    ml.valid = ml.lvalid && ml.j < ROWS;
    ml.wave_active = (tid / 64) * 4 < T::NT;
    if (!ml.lvalid) ml.line = 0;
#pragma unroll
    for (int ks = 0; ks < T::KS; ++ks) {
        const int row = 4 * jt + c, col = 4 * ks + ml.r;
        ml.A[ks] = (row < ROWS && col < S::NCG) ? op[row * S::NCG + col] : 0.0;
    }
    return ml;
}

// sum_k op[j][k] fld[k along the lane's line], for the point the lane owns;
// off: LDS offsets of the Gauss values this lane supplies.  Loads are
// unconditional (a masked load becomes a branch with its own wait): a lane
// whose Gauss index k lies past the line reads an in-range value that meets
// a zero coefficient (A = 0 for k >= ncg), and a lane past the last line only
// feeds outputs that no lane keeps.
template <int P, int ROWS>
__device__ __forceinline__ double mf_line(const double *fld, const MfLane<P, ROWS> &ml, const int *off)
{
    constexpr int KS = MfTiles<P, ROWS>::KS;
    double bv[KS];
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) bv[ks] = fld[off[ks]];
    double t = 0.0;
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) t = __builtin_amdgcn_mfma_f64_4x4x4f64(ml.A[ks], bv[ks], t, 0, 0, 0);
    return t;
}

// One staggered grid D of the fused volume kernel: fluxes at the nc*ncg^2
// points of grid D into LDS (interior) or the viscous face traces (faces),
// then the interior part of the divergence along D added into acc.
// MF (FP64 only): the interpolation of Q and the prolongation of the 15
// gradients run on the matrix cores (mf_line), one point per lane; otherwise
// each work-item forms its point's 9-term sums on the vector ALUs.
// Flux of the lane's staggered point in the matrix-core form, held in
// registers until every wave is past the previous divergence pass.
struct LaneFlux {
    double f[5];
    int s;
    bool store;
};

template <int P, int D, int PM, bool MF>
__device__ __forceinline__ LaneFlux volume_grid(const DevMesh &m, const DevPhys &ph, const double *__restrict__ QA,
                                            const double *__restrict__ muh, const double *__restrict__ nut,
                                            double *__restrict__ FV, double *__restrict__ Slag,
                                            const typename Ctr<PM>::T *sQ,
                                            const typename Ctr<PM>::T *sG, typename Ctr<PM>::T *sF,
                                            const double *sBar, const typename Ctr<PM>::T *sD,
                                            const typename Ctr<PM>::T *sB, const double *sDint, int e,
                                            const int *e2m_e, const MfLane<P> &ml, int tid)
{
    using S = Sz<P>;
    using CT = typename Ctr<PM>::T;
    using CA = typename Ctr<PM>::A;
    static_assert(!MF || PM == 0, "matrix-core contractions are FP64 only");
    constexpr int VBLOCK = vblock<P, PM>();
    static_assert(!MF || MfTiles<P>::NT <= VBLOCK / 16, "one tile per 16 lanes in one pass");
    constexpr int KS = MfTiles<P>::KS;
    const int nsm = m.ms ? S::NS : 1, nfm = m.ms ? S::NF : 1;
    const double *gmd = (D == 0 ? m.gmet0 : (D == 1 ? m.gmet1 : m.gmet2)) + (size_t)e * 3 * nsm;
    // Matrix-core form: the lane's point is fixed, so its metric row and face
    // values are loaded before the barrier and arrive while the other waves
    // finish the previous pass.
    int mf_s = 0;
    double mf_g[3] = {0.0, 0.0, 0.0}, mf_qa[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    double mf_nrm = 1.0, mf_nut = 0.0, mf_muh = 0.0;
    if constexpr (MF) {
        const int a = ml.line % S::NCG, b = ml.line / S::NCG;
        mf_s = sidx<P>(D, ml.valid ? ml.j : 0, a, b);
        if (ml.valid) {
#pragma unroll
            for (int i = 0; i < 3; ++i) mf_g[i] = gmd[i * nsm + mf_s * m.ms];
            if (nut) mf_nut = nut[((size_t)e * 3 + D) * S::NS + mf_s];
            if (muh) mf_muh = muh[((size_t)e * 3 + D) * S::NS + mf_s];
            if (ml.j == 0 || ml.j == S::NC - 1) {
                const int c = e2m_e[face_of<P>(D, ml.j) - 1];
                const int mo = e2m_mortar(c);
                const int q = m.rot ? orient_point<S::NCG>(e2m_map(c), a + S::NCG * b) : a + S::NCG * b;
#pragma unroll
                for (int v = 0; v < 5; ++v) mf_qa[v] = QA[((size_t)mo * NEQ + v) * S::NF + q];
                mf_nrm = m.norm[(size_t)mo * nfm + q * m.ms];
            }
        }
    }
    // Matrix-core form, grids 1 and 2: a lane's flux waits in registers, and the
    // barrier that keeps the previous divergence pass's reads of sF ahead of
    // this grid's writes comes after the arithmetic, so the two passes overlap
    // across waves.  Grid 0 (and the vector form) waits here: sG complete /
    // previous divergence pass done with sF.
    double mf_f[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    int mf_fs = 0;
    bool mf_fstore = false;
    if constexpr (!MF || D == 0) __syncthreads();
    // Matrix-core form: one pass, every lane of an active wave takes part in
    // the contractions (the tiles need all 64 lanes); only lanes that own a
    // point store.  Vector form: work-items stride over the points.
    // FP32 mode (PM 1): the 20 contractions of a point (5 interpolated state
    // fields, 15 prolonged gradients) run on v_mfma_f32_4x4x1f32.  Each 4x4
    // block is one line and four consecutive Lobatto rows j; lane i of a
    // block supplies field (4 ch + i) at Gauss point k of the line (A) and
    // the operator entry B[j_i][k] of its own row (B), and receives fields
    // 4 ch .. 4 ch + 3 at its own point (D column i, rows in the four
    // result registers).  Five chains of ncg steps give all 20 fields at the
    // lane's point; passes cover the ncg^2 lines x ceil(nc/4) row groups.
    constexpr bool MF32 = (PM == 1);
    constexpr int JG = (S::NC + 3) / 4;                       // row groups per line
    constexpr int NB32 = S::NCG * S::NCG * JG;                // blocks per grid
    constexpr int NPASS32 = (NB32 + VBLOCK / 4 - 1) / (VBLOCK / 4);
    const int w0 = (MF || MF32) ? 0 : tid;
    const int wend = MF ? (ml.wave_active ? 1 : 0) : (MF32 ? NPASS32 : S::NS);
    const int wstep = (MF || MF32) ? 1 : VBLOCK;
#pragma unroll 1
    for (int w = w0; w < wend; w += wstep) {
        int a, b, j;
        bool valid = true;
        int off[KS];
        int line32 = 0;
        if constexpr (MF32) {
            const int blk = w * (VBLOCK / 4) + tid / 4;
            const int line = blk / JG;
            j = 4 * (blk % JG) + tid % 4;
            valid = line < S::NF && j < S::NC;
            line32 = line < S::NF ? line : 0;                 // past the last line: read line 0, keep nothing
            a = line32 % S::NCG;
            b = line32 / S::NCG;
        } else if constexpr (MF) {
            j = ml.j;
            a = ml.line % S::NCG;
            b = ml.line / S::NCG;
            valid = ml.valid;
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) {
                const int k = 4 * ks + ml.r;
                off[ks] = gidx<P>(D, k < S::NCG ? k : S::NCG - 1, a, b);   // in range (ml.line is 0 past the last line)
            }
        } else {
            const int s0 = w;
            if (D == 0) { j = s0 % S::NC; a = (s0 / S::NC) % S::NCG; b = s0 / (S::NC * S::NCG); }
            else if (D == 1) { a = s0 % S::NCG; j = (s0 / S::NCG) % S::NC; b = s0 / (S::NCG * S::NC); }
            else { a = s0 % S::NCG; b = (s0 / S::NCG) % S::NCG; j = s0 / (S::NCG * S::NCG); }
        }
        const int q = a + S::NCG * b;
        const int jc = valid ? j : 0;
        const int s = MF ? mf_s : sidx<P>(D, jc, a, b);
        const bool face = (jc == 0 || jc == S::NC - 1);
        const int fc = face_of<P>(D, jc);
        const int mo = e2m_mortar(e2m_e[fc - 1]);
        const int qm = m.rot ? orient_point<S::NCG>(e2m_map(e2m_e[fc - 1]), q) : q;   // the face point on the mortar
        // Vector form: the point's metric row, and its line's face average and
        // norm (every lane: face_of names a face for any j, and interior lanes
        // hit the face lane's cache lines), requested before the contractions
        // instead of one wait each at their use
        double vg[3] = {0.0, 0.0, 0.0}, vqa[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, vnrm = 1.0;
        if constexpr (!MF) {
#pragma unroll
            for (int i = 0; i < 3; ++i) vg[i] = gmd[i * nsm + s * m.ms];
#pragma unroll
            for (int v = 0; v < 5; ++v) vqa[v] = QA[((size_t)mo * NEQ + v) * S::NF + qm];
            vnrm = m.norm[(size_t)mo * nfm + qm * m.ms];
        }
        double Qv[5];
        CT Bj[MF ? 1 : S::NCG];
        float fld32[MF32 ? 20 : 1];
        if constexpr (MF32) {
            typedef float f4 __attribute__((ext_vector_type(4)));
            float opk[S::NCG];
            const int jr = j < S::NC ? j : S::NC - 1;
#pragma unroll
            for (int k = 0; k < S::NCG; ++k) opk[k] = j < S::NC ? sB[jr * S::NCG + k] : 0.0f;
            int goff[S::NCG];
#pragma unroll
            for (int k = 0; k < S::NCG; ++k) goff[k] = gidx<P>(D, k, a, b);
#pragma unroll
            for (int ch = 0; ch < 5; ++ch) {
                const int F = 4 * ch + tid % 4;
                const CT *src = F < 5 ? sQ + F * S::NP : sG + (F - 5) * S::NP;
                f4 t = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int k = 0; k < S::NCG; ++k) t = __builtin_amdgcn_mfma_f32_4x4x1f32(src[goff[k]], opk[k], t, 0, 0, 0);
#pragma unroll
                for (int r = 0; r < 4; ++r) fld32[4 * ch + r] = t[r];
            }
#pragma unroll
            for (int v = 0; v < 5; ++v) Qv[v] = face ? vqa[v] : sBar[v] + (double)fld32[v];
        } else if constexpr (MF) {
#pragma unroll
            for (int v = 0; v < 5; ++v) {
                const double t = mf_line<P>(sQ + v * S::NP, ml, off);
                Qv[v] = face ? mf_qa[v] : t;
            }
        } else {
#pragma unroll
            for (int k = 0; k < S::NCG; ++k) Bj[k] = sB[j * S::NCG + k];
#pragma unroll
            for (int v = 0; v < 5; ++v) {
                if (face) {
                    Qv[v] = vqa[v];
                } else {
                    CA u = 0;
#pragma unroll
                    for (int k = 0; k < S::NCG; ++k) u = mac<PM>(Bj[k], sQ[v * S::NP + gidx<P>(D, k, a, b)], u);
                    if constexpr (PM == 0) Qv[v] = u; else Qv[v] = sBar[v] + (double)u;
                }
                if (v == 1 || v == 3) __builtin_amdgcn_sched_barrier(0);
            }
        }
        if (Slag && valid) {
#pragma unroll
            for (int v = 0; v < 5; ++v) Slag[(((size_t)e * 3 + D) * NEQ + v) * S::NS + s] = Qv[v];
        }
        if constexpr (MF) {
            if (!valid) { Qv[0] = 1.0; Qv[1] = Qv[2] = Qv[3] = 0.0; Qv[4] = 1.0; }   // no point: keep the arithmetic finite
        }
        const Prim pr = primitives(ph, Qv);
        // Velocity and temperature gradients, one direction c at a time: the
        // prolonged conservative gradient dQ/dx_c (5 fields) is turned into
        // (du/dx_c, dv/dx_c, dw/dx_c, dT/dx_c) and folded into the symmetric
        // velocity-gradient sums.  Unrolled in the matrix-core form (the next
        // direction's loads overlap this one's arithmetic, measured 8 % faster);
        // one direction at a time in the vector form, whose registers would spill.
        double Sxx = 0.0, Syy = 0.0, Szz = 0.0, Sxy = 0.0, Sxz = 0.0, Syz = 0.0, div = 0.0;
        double Tx = 0.0, Ty = 0.0, Tz = 0.0;
        constexpr int CUNROLL = (MF || MF32) ? 3 : 1;   // MF32: the fields are in registers already
#pragma unroll CUNROLL
        for (int c = 0; c < 3; ++c) {
            double Gc[5];
#pragma unroll
            for (int v = 0; v < 5; ++v) {
                const CT *gl = sG + (c * 5 + v) * S::NP;
                if constexpr (MF) {
                    Gc[v] = mf_line<P>(gl, ml, off);
                } else if constexpr (MF32) {
                    (void)gl;
                    Gc[v] = (double)fld32[5 + c * 5 + v];
                } else {
                    CA u = 0;
#pragma unroll
                    for (int k = 0; k < S::NCG; ++k) u = mac<PM>(Bj[k], gl[gidx<P>(D, k, a, b)], u);
                    Gc[v] = (double)u;
                }
            }
            double o[4];
            prim_gradient(ph, pr, Gc, o);
            const double uc = o[0], vc = o[1], wc = o[2];
            div = div + (c == 0 ? uc : (c == 1 ? vc : wc));
            Sxx += (c == 0) ? uc : 0.0;
            Syy += (c == 1) ? vc : 0.0;
            Szz += (c == 2) ? wc : 0.0;
            Sxy += (c == 0) ? vc : ((c == 1) ? uc : 0.0);   // uy + vx
            Sxz += (c == 0) ? wc : ((c == 2) ? uc : 0.0);   // wx + uz
            Syz += (c == 1) ? wc : ((c == 2) ? vc : 0.0);   // wy + vz
            Tx = (c == 0) ? o[3] : Tx;
            Ty = (c == 1) ? o[3] : Ty;
            Tz = (c == 2) ? o[3] : Tz;
        }
        if (valid) {
            double amu = 1.0;
            if (nut) amu = amu + (MF ? mf_nut : nut[((size_t)e * 3 + D) * S::NS + s]) * Qv[0];   // static Smagorinsky
            if (muh) amu = amu + (MF ? mf_muh : muh[((size_t)e * 3 + D) * S::NS + s]);
            const double g0 = MF ? mf_g[0] : vg[0];
            const double g1 = MF ? mf_g[1] : vg[1];
            const double g2 = MF ? mf_g[2] : vg[2];
            double rv[5];
            {
                // Stokes stress and Fourier heat flux (x/y/z_vis_flux, PhysLib.f90:464-470)
                const double hfac = (1.0 + (amu - 1.0) * ph.kap) * ph.hfac1;
                const double d3 = div * (1.0 / 3.0);
                const double txx = 2.0 * amu * (Sxx - d3), tyy = 2.0 * amu * (Syy - d3), tzz = 2.0 * amu * (Szz - d3);
                const double txy = amu * Sxy, txz = amu * Sxz, tyz = amu * Syz;
                const double f4 = pr.u * txx + pr.v * txy + pr.w * txz + hfac * Tx;
                const double g4 = pr.v * tyy + pr.u * txy + pr.w * tyz + hfac * Ty;
                const double h4 = pr.v * tyz + pr.u * txz + pr.w * tzz + hfac * Tz;
                rv[0] = 0.0;
                rv[1] = g0 * txx + g1 * txy + g2 * txz;
                rv[2] = g0 * txy + g1 * tyy + g2 * tyz;
                rv[3] = g0 * txz + g1 * tyz + g2 * tzz;
                rv[4] = g0 * f4 + g1 * g4 + g2 * h4;
            }
            if (face) {
                // r / norm as one reciprocal and five products (a rounding apart
                // from five divisions, which every wave with a face lane paid)
                const double rn = 1.0 / (MF ? mf_nrm : vnrm);
#pragma unroll
                for (int v = 0; v < 5; ++v) FV[(((size_t)e * 6 + fc - 1) * NEQ + v) * S::NF + q] = rv[v] * rn;
            } else {
                double fi[5];
                euler_from_primitive(pr, Qv, g0, g1, g2, fi);
                if constexpr (MF) {
#pragma unroll
                    for (int v = 0; v < 5; ++v) mf_f[v] = fi[v] - rv[v] * ph.inv_re;
                    mf_fs = s;
                    mf_fstore = true;
                } else {
#pragma unroll
                    for (int v = 0; v < 5; ++v) {
                        if constexpr (PM == 0) sF[v * S::NS + s] = fi[v] - rv[v] * ph.inv_re;
                        else sF[v * S::NS + s] = (CT)((fi[v] - rv[v] * ph.inv_re) - sBar[5 + D * 5 + v]);
                    }
                }
            }
        }
    }
    LaneFlux lf;
#pragma unroll
    for (int v = 0; v < 5; ++v) lf.f[v] = mf_f[v];
    lf.s = mf_fs;
    lf.store = mf_fstore;
    return lf;
}

// Grid D's fluxes into sF.  The matrix-core form stores its lanes' fluxes
// once every wave is past the previous divergence pass.  On return every
// wave is past grid D's point pass (its reads of sQ and sG).
template <int P, int D, int PM, bool MF>
__device__ __forceinline__ void volume_publish(const LaneFlux &lf, typename Ctr<PM>::T *sF)
{
    using S = Sz<P>;
    if constexpr (MF) {
        if constexpr (D > 0) __syncthreads();   // every wave is past the previous divergence pass
        if (lf.store) {
#pragma unroll
            for (int v = 0; v < 5; ++v) sF[v * S::NS + lf.s] = lf.f[v];
        }
    }
    __syncthreads();   // fluxes of this grid complete
}

// Interior part of the divergence along grid D, with the grid's fluxes in
// sF: acc -= sum_{interior j} D[k][j] F_j.
template <int P, int D, int PM>
__device__ __forceinline__ void volume_div(const typename Ctr<PM>::T *sF, const typename Ctr<PM>::T *sD,
                                           const double *sBar, const double *sDint, int tid,
                                           double (&acc)[(Sz<P>::NP + vblock<P, PM>() - 1) / vblock<P, PM>()][5])
{
    using S = Sz<P>;
    using CT = typename Ctr<PM>::T;
    using CA = typename Ctr<PM>::A;
    constexpr int VBLOCK = vblock<P, PM>();
    constexpr int R = (S::NP + VBLOCK - 1) / VBLOCK;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int pt = tid + r * VBLOCK;
        if (pt < S::NP) {
            const int n = pt % S::NCG, mi = (pt / S::NCG) % S::NCG, l = pt / (S::NCG * S::NCG);
            const int k = D == 0 ? n : (D == 1 ? mi : l);
            const int a = D == 0 ? mi : n;
            const int b = D == 2 ? mi : l;
            CT Dk[S::NC];
#pragma unroll
            for (int j = 1; j < S::NC - 1; ++j) Dk[j] = sD[k * S::NC + j];
#pragma unroll
            for (int v = 0; v < 5; ++v) {
                CA t = 0;
#pragma unroll
                for (int j = 1; j < S::NC - 1; ++j) t = mac<PM>(Dk[j], sF[v * S::NS + sidx<P>(D, j, a, b)], t);
                if constexpr (PM == 0) acc[r][v] -= t;
                else acc[r][v] -= (double)t + sBar[5 + D * 5 + v] * sDint[k];
            }
        }
    }
}

// Direct global->LDS copy (global_load_lds_dword): lane l of a wave moves one
// dword from its own address src to the LDS byte address M0 + 4 l, through no
// register.  Written as inline asm on purpose: the compiler's own form makes
// every later LDS read wait for all such copies (it cannot tell the
// destinations apart), which would serialize the copies with the work they
// are meant to overlap.  The caller waits for them itself (lds_copies_done)
// before the barrier that publishes them.  lds_wave_base must be the same for
// every lane of the wave.
typedef __attribute__((address_space(3))) void *LdsPtr;

__device__ __forceinline__ void lds_dword(const void *src, void *lds_wave_base)
{
    const unsigned l = __builtin_amdgcn_readfirstlane((unsigned)(size_t)(LdsPtr)lds_wave_base);
    // The LDS base goes in through the {m0} operand constraint: the compiler
    // writes M0 itself and knows the copy reads it, so no value of its own is
    // kept there across the copy.  The s_nop separates that write from the
    // copy's use of M0.
    asm volatile("s_nop 0\n\tglobal_load_lds_dword %0, off" : : "v"(src), "{m0}"(l) : "memory");
}

// s_waitcnt immediate of gfx9 for vmcnt(0) alone (expcnt 7, lgkmcnt 15).
constexpr int VMCNT0 = 0x0F70;

// Every copy this wave issued has landed in LDS.
__device__ __forceinline__ void lds_copies_done()
{
    asm volatile("s_waitcnt vmcnt(0)" : : : "memory");
}

// The next element's volume inputs, copied straight into LDS while the
// current element finishes: raw J*Q into sQ ([v][pt]), 1/J at every Gauss
// point into sJ, and the six faces' averages (face f+1 at [f][v][q],
// gathered by the mortars of the element's e2m entries e2m_e) into sA.
// Every wave copies 64-dword chunks w, w + NW, ...; each wave waits for its
// copies (lds_copies_done) before the barrier that publishes them.
template <int P>
__device__ __forceinline__ void volume_inputs_lds(const double *__restrict__ JQ, const double *__restrict__ QA,
                                                  const DevMesh &m, int e, const int *e2m_e, int tid,
                                                  double *sQ, double *sJ, double *sA)
{
    using S = Sz<P>;
    constexpr int NW = vblock<P, 0>() / 64;
    constexpr int NQV = NEQ * S::NP, NAV = 6 * NEQ * S::NF;
    constexpr int DQ = 2 * NQV, DJ = 2 * S::NP, DA = 2 * NAV;   // dwords
    constexpr int CQ = (DQ + 63) / 64, CJ = (DJ + 63) / 64, CA = (DA + 63) / 64;   // 64-dword chunks
    const int lane = tid % 64;
    const int w = __builtin_amdgcn_readfirstlane(tid / 64);
    // mortars first, before any copy is in flight
    int mo6[6], map6[6];
#pragma unroll
    for (int g = 0; g < 6; ++g) {
        mo6[g] = e2m_mortar(e2m_e[g]);
        map6[g] = e2m_map(e2m_e[g]);
    }
    const char *q = (const char *)(JQ + (size_t)e * NQV);
    const char *jac = (const char *)(m.jacob + (size_t)e * (m.ms ? S::NP : 1));
#pragma unroll
    for (int u = 0; u < (CQ + NW - 1) / NW; ++u) {
        const int c = w + u * NW, d = c * 64 + lane;
        if (d < DQ) lds_dword(q + 4 * d, (char *)sQ + 256 * c);
    }
#pragma unroll
    for (int u = 0; u < (CJ + NW - 1) / NW; ++u) {
        const int c = w + u * NW, d = c * 64 + lane;
        if (d < DJ) lds_dword(jac + 8 * ((d >> 1) * m.ms) + 4 * (d & 1), (char *)sJ + 256 * c);
    }
#pragma unroll
    for (int u = 0; u < (CA + NW - 1) / NW; ++u) {
        const int c = w + u * NW, d = c * 64 + lane;
        const int i = d >> 1, f = i / (NEQ * S::NF), vq = i % (NEQ * S::NF);
        int mo = mo6[0], mp = map6[0];
#pragma unroll
        for (int g = 1; g < 6; ++g) {
            mo = f == g ? mo6[g] : mo;
            mp = f == g ? map6[g] : mp;
        }
        const int vqm = m.rot ? (vq / S::NF) * S::NF + orient_point<S::NCG>(mp, vq % S::NF) : vq;
        if (d < DA) lds_dword((const char *)(QA + (size_t)mo * NEQ * S::NF + vqm) + 4 * (d & 1), (char *)sA + 256 * c);
    }
}

// Global loads of one element's state and 1/J and of its six faces'
// averages, all issued before any use.
template <int P, int PM>
__device__ __forceinline__ void volume_inputs(const double *__restrict__ JQ, const double *__restrict__ QA,
                                              const DevMesh &m, int e, int tid,
                                              double (&qv)[(NEQ * Sz<P>::NP + vblock<P, PM>() - 1) / vblock<P, PM>()],
                                              double (&jv)[(NEQ * Sz<P>::NP + vblock<P, PM>() - 1) / vblock<P, PM>()],
                                              double (&av)[(6 * NEQ * Sz<P>::NF + vblock<P, PM>() - 1) / vblock<P, PM>()])
{
    using S = Sz<P>;
    constexpr int VBLOCK = vblock<P, PM>();
    constexpr int NQV = NEQ * S::NP, NAV = 6 * NEQ * S::NF;
    constexpr int NQI = (NQV + VBLOCK - 1) / VBLOCK, NAI = (NAV + VBLOCK - 1) / VBLOCK;
    const double *q = JQ + (size_t)e * NEQ * S::NP;
    const double *jac = m.jacob + (size_t)e * (m.ms ? S::NP : 1);
    const int *e2m = m.e2m + (size_t)e * 6;
    int mort[6], map6[6];
#pragma unroll
    for (int f = 0; f < 6; ++f) {
        mort[f] = e2m_mortar(e2m[f]);
        map6[f] = e2m_map(e2m[f]);
    }
#pragma unroll
    for (int u = 0; u < NQI; ++u) {
        const int i = tid + u * VBLOCK;
        qv[u] = i < NQV ? q[i] : 0.0;
        jv[u] = i < NQV ? jac[(i % S::NP) * m.ms] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < NAI; ++u) {
        const int i = tid + u * VBLOCK;
        const int f = i / (NEQ * S::NF), vq = i % (NEQ * S::NF);
        // mortar of face f+1 by selects (a dynamic index into mort[] would go to scratch)
        int mo = mort[0], mp = map6[0];
#pragma unroll
        for (int g = 1; g < 6; ++g) {
            mo = f == g ? mort[g] : mo;
            mp = f == g ? map6[g] : mp;
        }
        const int vqm = m.rot ? (vq / S::NF) * S::NF + orient_point<S::NCG>(mp, vq % S::NF) : vq;
        av[u] = i < NAV ? QA[(size_t)mo * NEQ * S::NF + vqm] : 0.0;
    }
}

// Fused volume kernel: Gauss-point gradients into LDS (Gauss_Deriv,
// SolnDerivsc.f90:14-83), then for each staggered grid the interpolated
// state, prolonged gradients, contravariant Euler + viscous fluxes, the
// viscous face traces r/norm and the interior part of the divergence.
//
// Mixed modes (PM 1, 2): the LDS tiles hold FP32 values and the
// contractions run in FP32 on element-shifted fields, with the shifts added
// back in FP64 -- exact in exact arithmetic because B reproduces constants
// and the rows of D annihilate them:
//   Q     - Qbar   (Qbar = Q at the element's central Gauss point);
//   faces QA - Qbar in the gradient (D Qbar = 0 is dropped);
//   flux  F - Fbar_d (Fbar_d = contravariant Euler flux of Qbar with the
//         metric row at staggered point 0 of grid d); the divergence adds
//         Fbar_d * sum_{interior j} D[k][j] back in FP64.
// Pointwise physics, the gradient metric combination inputs, the RK state
// and all face kernels stay FP64.
// PLAIN: aligned faces, no entropy or eddy viscosity, no lagged states saved.
// That instance compiles the optional paths out (at p = 8 FP64: 126 VGPRs
// and no spill against 128 with a 12-byte spill, a quarter fewer VALU
// instructions in the code); the other handles every case by run-time
// branches.  The mixed modes use the general instance only: their plain one
// spilled 28 bytes at p = 8 and ran slower.
template <int P, int PM, bool PLAIN>
__global__ __launch_bounds__((vblock<P, PM>()), 4) void k_volume(const double *__restrict__ JQ, DevMesh m, DevPhys ph,
                                                    const double *__restrict__ QA, const double *__restrict__ muh,
                                                    const double *__restrict__ nut, double *__restrict__ FV,
                                                    double *__restrict__ Qtp, double *__restrict__ Slag)
{
    if constexpr (PLAIN) {
        muh = nullptr;
        nut = nullptr;
        Slag = nullptr;
        m.rot = 0;
    }
    using S = Sz<P>;
    using CT = typename Ctr<PM>::T;
    using CA = typename Ctr<PM>::A;
    constexpr int VBLOCK = vblock<P, PM>();
    constexpr int R = (S::NP + VBLOCK - 1) / VBLOCK;
    __shared__ CT sQ[NEQ * S::NP];            // physical Q at Gauss points (minus Qbar when mixed)
    __shared__ CT sG[15 * S::NP];             // dQ_v/dx_c at Gauss points, (c*5+v)
    // interior total flux of the current grid (minus Fbar); during the
    // gradient phase it holds the face averages of the six faces
    constexpr int NSF = NEQ * S::NS > 6 * NEQ * S::NF ? NEQ * S::NS : 6 * NEQ * S::NF;
    __shared__ CT sF[NSF];
    __shared__ double sBar[5 + 15];           // mixed modes: Qbar[5], Fbar[3][5]
    // operator matrices of this order: M = D_int B (Gauss derivative), D, B
    constexpr int OM = 0, OD = S::NCG * S::NCG, OB = OD + S::NCG * S::NC, NOP = OB + S::NC * S::NCG;
    __shared__ CT sOp[NOP];
    __shared__ double sDint[S::NCG];          // mixed modes: sum over interior j of D[k][j]
    const int tid0 = threadIdx.x;
    constexpr bool MF = (PM == 0);
    constexpr int NQV = NEQ * S::NP, NAV = 6 * NEQ * S::NF;
    constexpr int NQI = (NQV + VBLOCK - 1) / VBLOCK, NAI = (NAV + VBLOCK - 1) / VBLOCK;
    constexpr int NOI = (NOP + VBLOCK - 1) / VBLOCK;
    static_assert(NAV <= 15 * S::NP, "the face averages of the next element fit in sG");
    // operator matrices, stored once per workgroup (the first element's
    // barrier after its staging publishes them)
#pragma unroll
    for (int u = 0; u < NOI; ++u) {
        const int i = tid0 + u * VBLOCK;
        if (i < NOP) {
            const CT *src;
            if constexpr (PM == 0) src = i < OD ? &c_M[P][i - OM] : (i < OB ? &c_D[P][i - OD] : &c_B[P][i - OB]);
            else src = i < OD ? &c_Mf[P][i - OM] : (i < OB ? &c_Df[P][i - OD] : &c_Bf[P][i - OB]);
            sOp[i] = *src;
        }
    }
    // Persistent workgroup: elements blockIdx.x, + gridDim.x, ...  At 1024
    // work-items (one workgroup per CU, nothing else to hide a memory wait)
    // the FP64 form copies the next element's state, 1/J and face averages
    // straight into sQ and sG (both free once the last grid's point pass is
    // done) during the last divergence pass, so that only a workgroup's first
    // element waits for its prologue.
    constexpr bool PRE = MF && VBLOCK == 1024;
    static_assert(!PRE || NAV + S::NP <= 15 * S::NP, "the next element's 1/J and face averages fit in sG");
    // PRE: e2m entries of this element and of the next, in alternating slots
    __shared__ int sE2m[2][6];
    const int npm = m.ms ? S::NP : 1;
    if constexpr (PRE) {
        // the workgroup's first element: its e2m entries and those of the
        // next, then its inputs copied into LDS like every later element's
        const int e0 = blockIdx.x, e1 = min(e0 + (int)gridDim.x, m.K - 1);
        if (tid0 < 6) sE2m[0][tid0] = m.e2m[(size_t)e0 * 6 + tid0];
        else if (tid0 < 12) sE2m[1][tid0 - 6] = m.e2m[(size_t)e1 * 6 + tid0 - 6];
        __syncthreads();
        volume_inputs_lds<P>(JQ, QA, m, e0, sE2m[0], tid0, sQ, sG + NAV, sG);
    }
    // Without PRE the body runs once (grid = elements) and there is no loop.
    int slot = 0;
    int e = blockIdx.x;
#pragma unroll 1
    while (true) {
        const bool first = (e == blockIdx.x);
        // The work-item index passes through an empty asm once per element:
        // left visible, every index the element's passes derive from it (and
        // the matrix-core operand fragments) is hoisted out of the loop and
        // held in registers for its whole length (past 128 VGPRs, and into
        // scratch, at 1024 work-items).
        const int tid = PRE ? opaque(tid0) : tid0;
        const MfLane<P> ml = mf_lane<P>(tid, c_B[P]);
        const MfLane<P, S::NCG> mg = mf_lane<P, S::NCG>(tid, c_M[P]);   // gradient tiles (FP64 form)
        CA gpre[R][9];
        const double *jac = m.jacob + (size_t)e * (m.ms ? S::NP : 1);
        const int nxt = e + gridDim.x;
        int e2r[6];
        const int *e2m_e = e2r;
        if constexpr (PRE) {
            e2m_e = sE2m[slot];
        } else {
#pragma unroll
            for (int f = 0; f < 6; ++f) e2r[f] = m.e2m[(size_t)e * 6 + f];
        }

        // ---- prologue.  Every global load of the element's inputs is issued
        // before the first wait: one exposed memory latency per element
        // instead of one per loop trip.
        double qv[NQI], jv[NQI], av[NAI];
        if constexpr (!PRE) volume_inputs<P, PM>(JQ, QA, m, e, tid, qv, jv, av);
        // the previous element is done with sQ, sG, sF; PRE: this element's
        // copies into LDS have landed
        if constexpr (PRE) lds_copies_done();
        if (PRE || !first) __syncthreads();
        {
            // metric rows of this work-item's Gauss points, first used after
            // the gradient contractions
            const double *gg = m.gmetg + (size_t)e * 9 * npm;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int pt = tid + r * VBLOCK;
                if constexpr (PM != 0) {
                    // mixed modes: unconditional loads from a clamped point,
                    // converted after the select (guarded, converting loads:
                    // 1.11 ms at 22^3 p=8 FP32, these: 0.98 ms; the FP64
                    // kernel is faster with the guarded form)
                    const int pc = pt < S::NP ? pt : S::NP - 1;
#pragma unroll
                    for (int i = 0; i < 9; ++i) {
                        const double g = gg[i * npm + pc * m.ms];
                        gpre[r][i] = (CA)(pt < S::NP ? g : 0.0);
                    }
                } else {
#pragma unroll
                    for (int i = 0; i < 9; ++i) gpre[r][i] = pt < S::NP ? (CA)gg[i * npm + pt * m.ms] : (CA)0;
                }
            }
        }
        if constexpr (PM != 0) {
            // Qbar = Q at the central Gauss point, Fbar_d = Euler flux of Qbar
            // with the metric row at staggered point 0 of grid d
            constexpr int pc = (S::NCG / 2) * (1 + S::NCG + S::NCG * S::NCG);
            const int nsm = m.ms ? S::NS : 1;
            double qc = 0.0, gm[3] = {0.0, 0.0, 0.0};
            if (tid < 5) qc = JQ[((size_t)e * NEQ + tid) * S::NP + pc] * jac[pc * m.ms];
            if (tid < 3) {
                const double *gmd = (tid == 0 ? m.gmet0 : (tid == 1 ? m.gmet1 : m.gmet2)) + (size_t)e * 3 * nsm;
#pragma unroll
                for (int i = 0; i < 3; ++i) gm[i] = gmd[i * nsm];
            }
            if (tid < S::NCG) sDint[tid] = c_Dint[P][tid];
            if (tid < 5) sBar[tid] = qc;
            __syncthreads();
            if (tid < 3) {
                double qb[5], fb[5];
#pragma unroll
                for (int v = 0; v < 5; ++v) qb[v] = sBar[v];
                const Prim pb = primitives(ph, qb);
                euler_from_primitive(pb, qb, gm[0], gm[1], gm[2], fb);
#pragma unroll
                for (int v = 0; v < 5; ++v) sBar[5 + tid * 5 + v] = fb[v];
            }
        }

        CT *sQA = sF;   // face averages of the six faces (face f+1 at [f][v][q]), minus Qbar when mixed
        if constexpr (PRE) {
            // sQ holds this element's raw J*Q, sG its face averages and (from
            // NAV on) 1/J: Q = (J*Q) (1/J) in place
            for (int i = tid; i < NQV; i += VBLOCK) sQ[i] = sQ[i] * sG[NAV + i % S::NP];
            for (int i = tid; i < NAV; i += VBLOCK) sQA[i] = sG[i];
        } else {
#pragma unroll
            for (int u = 0; u < NQI; ++u) {
                const int i = tid + u * VBLOCK;
                if (i < NQV) {
                    const double qp = qv[u] * jv[u];
                    if constexpr (PM == 0) sQ[i] = qp; else sQ[i] = (CT)(qp - sBar[i / S::NP]);
                }
            }
#pragma unroll
            for (int u = 0; u < NAI; ++u) {
                const int i = tid + u * VBLOCK;
                if (i < NAV) {
                    if constexpr (PM == 0) sQA[i] = av[u]; else sQA[i] = (CT)(av[u] - sBar[(i % (NEQ * S::NF)) / S::NF]);
                }
            }
        }
        __syncthreads();

        // ---- gradients: w_d = (D_int B) Q + D(:,0) QA_lo + D(:,nc-1) QA_hi, then metric combine
        if constexpr (MF) {
            // Matrix-core form.  Direction by direction, the tile lane of each
            // Gauss point writes w_d (5 fields) into slots d*5+v of sG; one
            // direction's tiles cover every point once, so the three directions
            // need no barrier between them.  Then each work-item combines the 15
            // w values of its point with the metric rows, in place, in the same
            // order as the vector form: g0 w0, then + g1 w1, then + g2 w2.
            constexpr int KSG = MfTiles<P, S::NCG>::KS;
            static_assert(MfTiles<P, S::NCG>::NT <= VBLOCK / 16, "one gradient tile per 16 lanes in one pass");
            if (mg.wave_active) {
                const int a = mg.line % S::NCG, b = mg.line / S::NCG;   // mg.line = a + ncg b
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    int off[KSG];
#pragma unroll
                    for (int ks = 0; ks < KSG; ++ks) {
                        const int k = 4 * ks + mg.r;
                        off[ks] = gidx<P>(d, k < S::NCG ? k : S::NCG - 1, a, b);
                    }
                    double w[5];
#pragma unroll
                    for (int v = 0; v < 5; ++v) w[v] = mf_line<P, S::NCG>(sQ + v * S::NP, mg, off);
                    if (mg.valid) {
                        const int k = mg.j;
                        const double d0 = sOp[OD + k * S::NC], d1 = sOp[OD + k * S::NC + S::NC - 1];
                        const double *qa0 = sQA + (face_of<P>(d, 0) - 1) * NEQ * S::NF + mg.line;
                        const double *qa1 = sQA + (face_of<P>(d, 1) - 1) * NEQ * S::NF + mg.line;
                        const int pt = gidx<P>(d, k, a, b);
#pragma unroll
                        for (int v = 0; v < 5; ++v)
                            sG[(d * 5 + v) * S::NP + pt] = fma(d1, qa1[v * S::NF], fma(d0, qa0[v * S::NF], w[v]));
                    }
                }
            }
            __syncthreads();
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int pt = tid + r * VBLOCK;
                if (pt >= S::NP) break;
                double w[15];
#pragma unroll
                for (int i = 0; i < 15; ++i) w[i] = sG[i * S::NP + pt];
                const CA *g = gpre[r];
#pragma unroll
                for (int c = 0; c < 3; ++c)
#pragma unroll
                    for (int v = 0; v < 5; ++v)
                        sG[(c * 5 + v) * S::NP + pt] = fma(g[6 + c], w[10 + v], fma(g[3 + c], w[5 + v], g[c] * w[v]));
            }
        } else {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int pt = tid + r * VBLOCK;
                if (pt >= S::NP) break;
                const int n = pt % S::NCG, mi = (pt / S::NCG) % S::NCG, l = pt / (S::NCG * S::NCG);
                CA out[3][5];
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    const int k = d == 0 ? n : (d == 1 ? mi : l);
                    const int a = d == 0 ? mi : n;
                    const int b = d == 2 ? mi : l;
                    const int q = a + S::NCG * b;
                    CT Mk[S::NCG];
#pragma unroll
                    for (int k2 = 0; k2 < S::NCG; ++k2) Mk[k2] = sOp[OM + k * S::NCG + k2];
                    const CT d0 = sOp[OD + k * S::NC], d1 = sOp[OD + k * S::NC + S::NC - 1];
                    const CT *qa0 = sQA + (face_of<P>(d, 0) - 1) * NEQ * S::NF + q;
                    const CT *qa1 = sQA + (face_of<P>(d, 1) - 1) * NEQ * S::NF + q;
                    const CA *gd = gpre[r] + d * 3;
#pragma unroll
                    for (int v = 0; v < 5; ++v) {
                        CA t = 0;
#pragma unroll
                        for (int k2 = 0; k2 < S::NCG; ++k2) t = mac<PM>(Mk[k2], sQ[v * S::NP + gidx<P>(d, k2, a, b)], t);
                        t = mac<PM>(d0, qa0[v * S::NF], t);
                        t = mac<PM>(d1, qa1[v * S::NF], t);
#pragma unroll
                        for (int c = 0; c < 3; ++c) out[c][v] = d == 0 ? gd[c] * t : fma(gd[c], t, out[c][v]);
                    }
                    __builtin_amdgcn_sched_barrier(0);
                }
#pragma unroll
                for (int c = 0; c < 3; ++c)
#pragma unroll
                    for (int v = 0; v < 5; ++v) sG[(c * 5 + v) * S::NP + pt] = (CT)out[c][v];
            }
        }


        double acc[R][5];
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int v = 0; v < 5; ++v) acc[r][v] = 0.0;
        LaneFlux lf;
        lf = volume_grid<P, 0, PM, MF>(m, ph, QA, muh, nut, FV, Slag, sQ, sG, sF, sBar, sOp + OD, sOp + OB, sDint, e, e2m_e, ml, tid);
        volume_publish<P, 0, PM, MF>(lf, sF);
        volume_div<P, 0, PM>(sF, sOp + OD, sBar, sDint, tid, acc);
        lf = volume_grid<P, 1, PM, MF>(m, ph, QA, muh, nut, FV, Slag, sQ, sG, sF, sBar, sOp + OD, sOp + OB, sDint, e, e2m_e, ml, tid);
        volume_publish<P, 1, PM, MF>(lf, sF);
        volume_div<P, 1, PM>(sF, sOp + OD, sBar, sDint, tid, acc);
        lf = volume_grid<P, 2, PM, MF>(m, ph, QA, muh, nut, FV, Slag, sQ, sG, sF, sBar, sOp + OD, sOp + OB, sDint, e, e2m_e, ml, tid);
        volume_publish<P, 2, PM, MF>(lf, sF);
        if constexpr (PRE) {
            // wait for this element's face-trace stores first, in a form the
            // compiler accounts for (on every path): a wait it placed later
            // for them would also wait for every copy issued after them
            __builtin_amdgcn_s_waitcnt(VMCNT0);
            if (nxt < m.K) {
                // every wave is past this element's point passes: the next
                // element's inputs go into sQ and sG, and the e2m entries of
                // the element after it into the slot this element is done with
                const int *e2m_n = sE2m[slot ^ 1];
                volume_inputs_lds<P>(JQ, QA, m, nxt, e2m_n, tid, sQ, sG + NAV, sG);
                const int n2 = min(nxt + (int)gridDim.x, m.K - 1);
                if (tid < 6) lds_dword(m.e2m + (size_t)n2 * 6 + tid, sE2m[slot]);
            }
        }
        volume_div<P, 2, PM>(sF, sOp + OD, sBar, sDint, tid, acc);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int pt = tid + r * VBLOCK;
            if (pt < S::NP)
#pragma unroll
                for (int v = 0; v < 5; ++v) Qtp[((size_t)e * NEQ + v) * S::NP + pt] = acc[r][v];
        }
        if constexpr (!PRE) break;
        e += gridDim.x;
        slot ^= 1;
        if (e >= m.K) break;
    }
}

template <int P>
__global__ __launch_bounds__(BLOCK) void k_face_visc(const double *__restrict__ FV, const double *__restrict__ FI,
                                                      DevMesh m, DevPhys ph, double *__restrict__ FT)
{
    using S = Sz<P>;
    const int idx = blockIdx.x * BLOCK + threadIdx.x;
    if (idx >= m.M * S::NF) return;
    const int mm = idx / S::NF, q = idx % S::NF;
    const int em = m.m_elem[2 * mm], es = m.m_elem[2 * mm + 1];
    const int fm = m.m_face[2 * mm], fs = m.m_face[2 * mm + 1];
    const int nfm = m.ms ? S::NF : 1;
    const double nrm = m.norm[(size_t)mm * nfm + q * m.ms];
    const bool flip = m.m_nsign[mm] == -1;
    const int qs = m.m_orient ? orient_point<S::NCG>(m.m_orient[mm], q) : q;   // rotate_fv_to_mortar
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        const double fvm = FV[(((size_t)em * 6 + fm - 1) * NEQ + v) * S::NF + q];
        double fvs = FV[(((size_t)es * 6 + fs - 1) * NEQ + v) * S::NF + qs];
        if (flip) fvs = -fvs;
        const double qb = 0.5 * (fvm + fvs) * nrm;
        FT[((size_t)mm * NEQ + v) * S::NF + q] = FI[((size_t)mm * NEQ + v) * S::NF + q] - qb / ph.re;
    }
}

struct StageArgs {
    double a, cdt;                 // RK coefficients for this stage (a unused when first)
    int first;                     // first stage: g = Qt
    int rk;                        // 0: write Qt only, 1: RK update
    unsigned long long *resid;     // max |Qt| of the last element (nullable)
    unsigned int *nonfinite;       // set when any Qt is NaN/Inf (nullable)
};

// Face part of the divergence, Qt, and (rk) the register update
// g = a g + Qt, JQ += c dt g (updatea.f90:213-227).  With TRACES the updated
// state is staged in LDS and the face traces of the next stage are computed
// here, so the state is not read again by k_traces.
// HBM-bound: 512 work-items per element and UBLK independent points per
// work-item in flight (all loads of a batch issued before any store), so that
// enough bytes are outstanding per CU to approach the HBM bandwidth.
constexpr int UBLOCK = 512;
constexpr int UBLK = 4;

// One batch of UBLK points of k_update: Qtp, and for an RK stage G, JQ and
// (with traces) 1/J.  Indices past the element are clamped in range; the
// caller skips those points.
struct UpdBatch {
    double q[UBLK], g[UBLK], j[UBLK], c[UBLK];
};

template <int P, bool TRACES>
__device__ __forceinline__ UpdBatch upd_load(const double *__restrict__ Qtp, const double *__restrict__ G,
                                             const double *__restrict__ JQ, const double *__restrict__ jac, int ms,
                                             const StageArgs &st, size_t e0, int base)
{
    using S = Sz<P>;
    constexpr int NQ = NEQ * S::NP;
    UpdBatch bt;
#pragma unroll
    for (int u = 0; u < UBLK; ++u) {
        const int i = base + u * UBLOCK;
        const int ic = i < NQ ? i : NQ - 1;
        bt.q[u] = Qtp[e0 + ic];
        bt.g[u] = (st.rk && !st.first) ? G[e0 + ic] : 0.0;
        bt.j[u] = st.rk ? JQ[e0 + ic] : 0.0;
        bt.c[u] = (st.rk && TRACES) ? jac[(ic % S::NP) * ms] : 0.0;
    }
    return bt;
}

template <int P, bool TRACES>
__global__ __launch_bounds__(UBLOCK) void k_update(const double *__restrict__ Qtp, const double *__restrict__ FT, DevMesh m,
                                                    StageArgs st, double *__restrict__ JQ, double *__restrict__ G,
                                                    double *__restrict__ Qt, double *__restrict__ T)
{
    using S = Sz<P>;
    constexpr int NQ = NEQ * S::NP, NFT = 6 * NEQ * S::NF;
    constexpr int NFI = (NFT + UBLOCK - 1) / UBLOCK;            // face values per work-item
    constexpr int NTRIP = (NQ + UBLOCK * UBLK - 1) / (UBLOCK * UBLK);
    __shared__ double sFT[NFT];
    __shared__ double sQ[TRACES ? NQ : 1];
    __shared__ double sDf[2 * S::NCG];   // D[k][0], D[k][nc-1]: face part of the divergence
    __shared__ double sBf[2 * S::NCG];   // B[0][k], B[nc-1][k]: face traces
    const int e = blockIdx.x, tid = threadIdx.x;
    // Mortar and orientation sign of the six faces: uniform per element, so
    // these are scalar loads; each work-item picks its face by selects.
    const int *e2m = m.e2m + (size_t)e * 6;
    int fmo[6], fmp[6];
    double fsg[6];
#pragma unroll
    for (int f = 0; f < 6; ++f) {
        const int c = e2m[f];
        fmo[f] = e2m_mortar(c);
        fmp[f] = e2m_map(c);
        fsg[f] = ((c & 1) && m.m_sign[e2m_mortar(c)] == -1) ? -1.0 : 1.0;
    }
    // Every global load of the face fluxes and of the first batch of points
    // is issued before the first wait.
    double fv[NFI];
#pragma unroll
    for (int u = 0; u < NFI; ++u) {
        const int i = tid + u * UBLOCK;
        const int f = i / (NEQ * S::NF), r = i % (NEQ * S::NF);
        int mo = fmo[0], mp = fmp[0];
#pragma unroll
        for (int g = 1; g < 6; ++g) {
            mo = f == g ? fmo[g] : mo;
            mp = f == g ? fmp[g] : mp;
        }
        const int rm = m.rot ? (r / S::NF) * S::NF + orient_point<S::NCG>(mp, r % S::NF) : r;   // rotate_flux_to_face
        fv[u] = i < NFT ? FT[(size_t)mo * NEQ * S::NF + rm] : 0.0;
    }
    const double *jac = m.jacob + (size_t)e * (m.ms ? S::NP : 1);
    const size_t e0 = (size_t)e * NQ;
    UpdBatch bt = upd_load<P, TRACES>(Qtp, G, JQ, jac, m.ms, st, e0, tid);
    if (tid < S::NCG) {
        sDf[tid] = c_D[P][tid * S::NC];
        sDf[S::NCG + tid] = c_D[P][tid * S::NC + S::NC - 1];
        sBf[tid] = c_B[P][tid];
        sBf[S::NCG + tid] = c_B[P][(S::NC - 1) * S::NCG + tid];
    }
#pragma unroll
    for (int u = 0; u < NFI; ++u) {
        const int i = tid + u * UBLOCK;
        const int f = i / (NEQ * S::NF);
        double sg = fsg[0];
#pragma unroll
        for (int g = 1; g < 6; ++g) sg = f == g ? fsg[g] : sg;
        if (i < NFT) sFT[i] = sg < 0.0 ? -fv[u] : fv[u];
    }
    __syncthreads();
    double mx = 0.0;
    bool bad = false;
#pragma unroll
    for (int trip = 0; trip < NTRIP; ++trip) {
        const int base = tid + trip * UBLOCK * UBLK;
        if (trip > 0) bt = upd_load<P, TRACES>(Qtp, G, JQ, jac, m.ms, st, e0, base);
#pragma unroll
        for (int u = 0; u < UBLK; ++u) {
            const int i = base + u * UBLOCK;
            if (i >= NQ) continue;
            const int v = i / S::NP, pt = i % S::NP;
            const int n = pt % S::NCG, mi = (pt / S::NCG) % S::NCG, l = pt / (S::NCG * S::NCG);
            double qt = bt.q[u];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const int k = d == 0 ? n : (d == 1 ? mi : l);
                const int a = d == 0 ? mi : n;
                const int b = d == 2 ? mi : l;
                const int q = a + S::NCG * b;
                const double F0 = sFT[((face_of<P>(d, 0) - 1) * NEQ + v) * S::NF + q];
                const double F1 = sFT[((face_of<P>(d, 1) - 1) * NEQ + v) * S::NF + q];
                qt = qt - fma(sDf[k], F0, sDf[S::NCG + k] * F1);
            }
            if (st.rk) {
                const double g = st.first ? qt : fma(st.a, bt.g[u], qt);
                G[e0 + i] = g;
                const double jq = fma(st.cdt, g, bt.j[u]);
                JQ[e0 + i] = jq;
                if (TRACES) sQ[i] = jq * bt.c[u];
            } else {
                Qt[e0 + i] = qt;
            }
            mx = fmax(mx, fabs(qt));
            bad = bad || !isfinite(qt);
        }
    }
    if (bad && st.nonfinite) atomicOr(st.nonfinite, 1u);
    if (st.resid && e == m.K - 1) atomicMax(st.resid, (unsigned long long)__double_as_longlong(mx));
    if (TRACES) {
        __syncthreads();
        double *t = T + (size_t)e * NFT;
        for (int i = tid; i < NFT; i += UBLOCK) {
            const int qf = i % S::NF, v = (i / S::NF) % NEQ, f = i / (NEQ * S::NF) + 1;
            const int a = qf % S::NCG, b = qf / S::NCG;
            const int d = (f == 4 || f == 6) ? 0 : ((f == 1 || f == 2) ? 1 : 2);
            const double *Bj = sBf + ((f == 2 || f == 4 || f == 5) ? S::NCG : 0);
            double s2 = 0.0;
#pragma unroll
            for (int k = 0; k < S::NCG; ++k) s2 = fma(Bj[k], sQ[v * S::NP + gidx<P>(d, k, a, b)], s2);
            t[i] = s2;
        }
    }
}

// ------------------------------------------------------------ entropy viscosity
struct EvArgs {
    double cmu, mumax, cmax;
    int ducros, smoothing, cmax_cap;
};

// One smoothing update of entropy_viscosity_smooting (viscosity_smoothing.f90:4-195)
// at Gauss point (n,m,l) (1-based), reproducing the reference's term order
// for the corner (kind 0), face (kind 1..6: n=1, m=1, l=1, n=N, m=N, l=N)
// and interior (kind 7) formulas.
template <int P>
__device__ __forceinline__ double ev_smooth_point(const double *mu, int kind, int n, int m, int l)
{
    using S = Sz<P>;
    constexpr int N = S::NCG;
    auto MU = [&](int a, int b, int c) { return mu[(a - 1) + N * ((b - 1) + N * (c - 1))]; };
    const double SIXTH = 1.0 / 6.0, TENTH = 0.10, TWELVETH = 1.0 / 12.0;
    if (kind == 0) {
        const int dn = (n == 1) ? 1 : -1, dm = (m == 1) ? 1 : -1, dl = (l == 1) ? 1 : -1;
        return SIXTH * (3.0 * MU(n, m, l) + MU(n + dn, m, l) + MU(n, m + dm, l) + MU(n, m, l + dl));
    }
    if (kind == 1) return TENTH * (5.0 * MU(n, m, l) + MU(n + 1, m, l) + MU(n, m - 1, l) + MU(n, m + 1, l) + MU(n, m, l + 1) + MU(n, m, l - 1));
    if (kind == 2) return TENTH * (5.0 * MU(n, m, l) + MU(n, m + 1, l) + MU(n - 1, m, l) + MU(n + 1, m, l) + MU(n, m, l + 1) + MU(n, m, l - 1));
    if (kind == 3) return TENTH * (5.0 * MU(n, m, l) + MU(n, m, l + 1) + MU(n, m - 1, l) + MU(n, m + 1, l) + MU(n - 1, m, l) + MU(n + 1, m, l));
    if (kind == 4) return TENTH * (5.0 * MU(n, m, l) + MU(n - 1, m, l) + MU(n, m - 1, l) + MU(n, m + 1, l) + MU(n, m, l + 1) + MU(n, m, l - 1));
    if (kind == 5) return TENTH * (5.0 * MU(n, m, l) + MU(n, m - 1, l) + MU(n - 1, m, l) + MU(n + 1, m, l) + MU(n, m, l + 1) + MU(n, m, l - 1));
    if (kind == 6) return TENTH * (5.0 * MU(n, m, l) + MU(n, m, l - 1) + MU(n, m - 1, l) + MU(n, m + 1, l) + MU(n - 1, m, l) + MU(n + 1, m, l));
    return TWELVETH * (6.0 * MU(n, m, l) + MU(n + 1, m, l) + MU(n - 1, m, l) + MU(n, m + 1, l) + MU(n, m - 1, l) + MU(n, m, l + 1) + MU(n, m, l - 1));
}

// The first NV fields of each staggered grid of one element's lagged-state
// block (src = [3][NEQ][ns]) into LDS as [3][NV][ns], with every load in
// flight before the first store (a load-store loop waits out one memory
// latency per trip).
// D (ncg x nc) and B (nc x ncg) of this order into LDS: the contractions of
// the per-point passes read them with per-work-item rows, which from the
// constant tables are vector memory loads with their own latency.
template <int P>
__device__ __forceinline__ void stage_ops(double *sDm, double *sBm, int tid)
{
    using S = Sz<P>;
    for (int i = tid; i < S::NCG * S::NC; i += BLOCK) {
        sDm[i] = c_D[P][i];
        sBm[i] = c_B[P][i];
    }
}

template <int P, int NV>
__device__ __forceinline__ void stage_staggered(const double *__restrict__ src, double *sS, int tid)
{
    using S = Sz<P>;
    constexpr int N = 3 * NV * S::NS, NI = (N + BLOCK - 1) / BLOCK;
    double t[NI];
#pragma unroll
    for (int u = 0; u < NI; ++u) {
        const int i = tid + u * BLOCK, ic = i < N ? i : N - 1;
        const int d = ic / (NV * S::NS), r = ic % (NV * S::NS);
        t[u] = src[d * NEQ * S::NS + r];
    }
#pragma unroll
    for (int u = 0; u < NI; ++u) {
        const int i = tid + u * BLOCK;
        if (i < N) sS[i] = t[u];
    }
}

// Entropy (EV_routines/Main/Driver.f90:21-60) for one element:
//   RHS_ent_trans (RHS.f90:21-190) from the lagged face-averaged staggered
//   states Slag and the current state: D = (Phi + Gamma)/(T Re) and the
//   Ducros sensor (calculate_terms, RHS.f90:207-278);
//   Entropy_Viscosity (EViscosity_GL.f90:19-101): mu = min(max(Cmu Re rho h^2 |D|, 0), mumax)
//   (times the Ducros sensor when enabled; with cmax_cap the cap is
//   cmax * max(rho) * h * max(|u| + c) over the element instead of mumax);
//   optional smoothing; EV_Prolong (Prolong_gtol.f90:7-51) to the three grids.
// The second-derivative term and the density sensor of RHS_ent_trans, and
// the Strain routine, feed no output and are not computed.
template <int P>
__global__ __launch_bounds__(BLOCK) void k_ev(const double *__restrict__ JQ, const double *__restrict__ Slag, DevMesh m,
                                               DevPhys ph, EvArgs ev, double *__restrict__ muh,
                                               double *__restrict__ muhg_out, double *__restrict__ resid_out,
                                               double *__restrict__ ducros_out, double *__restrict__ rhos_out)
{
    using S = Sz<P>;
    __shared__ double sS[3 * NEQ * S::NS];
    __shared__ double sT[3 * S::NS];
    __shared__ double sMu[S::NP];
    __shared__ double sH[S::NP];   // h of each point, reused by the cap
    __shared__ double sRed[2 * BLOCK];
    __shared__ double sDm[S::NCG * S::NC], sBm[S::NC * S::NCG];
    const int e = blockIdx.x, tid = threadIdx.x;
    const double gamma = ph.gamma;
    stage_staggered<P, NEQ>(Slag + (size_t)e * 3 * NEQ * S::NS, sS, tid);
    stage_ops<P>(sDm, sBm, tid);
    __syncthreads();
    for (int i = tid; i < 3 * S::NS; i += BLOCK) {
        const int d = i / S::NS, s2 = i % S::NS;
        const double *q = sS + d * NEQ * S::NS;
        const double q1 = q[s2];
        const double a1 = q[S::NS + s2] / q1, a2 = q[2 * S::NS + s2] / q1, a3 = q[3 * S::NS + s2] / q1;
        const double Pv = (gamma - 1.0) * (q[4 * S::NS + s2] - 0.5 * q1 * (a1 * a1 + a2 * a2 + a3 * a3));
        sT[i] = Pv * gamma / q1;
    }
    __syncthreads();
    const int npm = m.ms ? S::NP : 1;
    const double *gg = m.gmetg + (size_t)e * 9 * npm;
    const double *jq = JQ + (size_t)e * NEQ * S::NP;
    const double *jac = m.jacob + (size_t)e * npm;
    const double c3 = pow((double)(S::NCG * S::NCG * S::NCG), 1.0 / 3.0);
    double rhomax = 0.0, utmax = 0.0;
    for (int pt = tid; pt < S::NP; pt += BLOCK) {
        const int n = pt % S::NCG, mi = (pt / S::NCG) % S::NCG, l = pt / (S::NCG * S::NCG);
        const int pq = pt * m.ms;
        double gr[6][3];
#pragma unroll
        for (int f = 0; f < 6; ++f) {
            double w[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const int k = d == 0 ? n : (d == 1 ? mi : l);
                const int a = d == 0 ? mi : n;
                const int b = d == 2 ? mi : l;
                const double *fld = (f < 5) ? sS + (d * NEQ + f) * S::NS : sT + d * S::NS;
                double t = 0.0;
#pragma unroll
                for (int j = 0; j < S::NC; ++j) t = fma(sDm[k * S::NC + j], fld[sidx<P>(d, j, a, b)], t);
                w[d] = t;
            }
#pragma unroll
            for (int c = 0; c < 3; ++c)
                gr[f][c] = gg[(0 * 3 + c) * npm + pq] * w[0] + gg[(1 * 3 + c) * npm + pq] * w[1] + gg[(2 * 3 + c) * npm + pq] * w[2];
        }
        const double jc = jac[pq];
        double Q[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) Q[v] = jq[v * S::NP + pt] * jc;
        const double rho = Q[0];
        const double u = Q[1] / rho, v = Q[2] / rho, w = Q[3] / rho;
        const double p = (gamma - 1.0) * (Q[4] - 0.5 * rho * (u * u + v * v + w * w));
        const double T = p * gamma / rho;
        const double ux = (gr[1][0] - u * gr[0][0]) / rho, uy = (gr[1][1] - u * gr[0][1]) / rho, uz = (gr[1][2] - u * gr[0][2]) / rho;
        const double vx = (gr[2][0] - v * gr[0][0]) / rho, vy = (gr[2][1] - v * gr[0][1]) / rho, vz = (gr[2][2] - v * gr[0][2]) / rho;
        const double wx = (gr[3][0] - w * gr[0][0]) / rho, wy = (gr[3][1] - w * gr[0][1]) / rho, wz = (gr[3][2] - w * gr[0][2]) / rho;
        const double tx = gr[5][0], ty = gr[5][1], tz = gr[5][2];
        const double s1 = ux + vy + wz;
        const double phi = (2.0 * (ux * ux + vy * vy + wz * wz) + ((vx + uy) * (vx + uy) + (uz + wx) * (uz + wx) + (vz + wy) * (vz + wy))
                            - 2.0 * (s1 * s1) / 3.0);
        const double gm = (tx * tx + ty * ty + tz * tz) / (T * (gamma - 1.0) * ph.pr);
        const double cur = sqrt((wy - vz) * (wy - vz) + (uz - wx) * (uz - wx) + (vx - uy) * (vx - uy));
        const double rhs = (((phi + gm) / T) / ph.re);
        const double duc = fmax(0.0, fmin(1.0, (s1 * s1 / (s1 * s1 + cur * cur + 1.0e-9))));
        if (resid_out) resid_out[(size_t)e * S::NP + pt] = rhs;
        if (ducros_out) ducros_out[(size_t)e * S::NP + pt] = duc;
        if (rhos_out) {
            // density-based sensor of calculate_terms (RHS.f90:263-275)
            const double eps = 1.0e-9;
            const double rx2 = gr[0][0] * gr[0][0], ry2 = gr[0][1] * gr[0][1], rz2 = gr[0][2] * gr[0][2];
            rhos_out[(size_t)e * S::NP + pt] =
                sqrt((1.0 / ((1.0 / (rx2 + eps)) + (1.0 / (ry2 + eps)) + 1.0)) +
                     (1.0 / ((1.0 / (rx2 + eps)) + (1.0 / (rz2 + eps)) + 1.0)) +
                     (1.0 / ((1.0 / (ry2 + eps)) + (1.0 / (rz2 + eps)) + 1.0)));
        }
        const double h = pow(fabs(1.0 / jc), 1.0 / 3.0) / c3;
        sH[pt] = h;
        double mur = ev.cmu * ph.re * rho * (h * h) * fabs(rhs);
        if (ev.ducros) mur = mur * duc;
        sMu[pt] = mur;
        if (ev.cmax_cap) {
            rhomax = fmax(rhomax, rho);
            utmax = fmax(utmax, sqrt(u * u + v * v + w * w) + sqrt(gamma * p / rho));
        }
    }
    double cap = ev.mumax;
    if (ev.cmax_cap) {
        sRed[tid] = rhomax;
        sRed[BLOCK + tid] = utmax;
        __syncthreads();
        for (int st = BLOCK / 2; st > 0; st >>= 1) {
            if (tid < st) {
                sRed[tid] = fmax(sRed[tid], sRed[tid + st]);
                sRed[BLOCK + tid] = fmax(sRed[BLOCK + tid], sRed[BLOCK + tid + st]);
            }
            __syncthreads();
        }
    }
    __syncthreads();
    for (int pt = tid; pt < S::NP; pt += BLOCK) {
        if (ev.cmax_cap) {
            cap = ev.cmax * sRed[0] * sH[pt] * sRed[BLOCK];
        }
        sMu[pt] = fmin(fmax(sMu[pt], 0.0), cap);
    }
    __syncthreads();
    if (ev.smoothing) {
        constexpr int N = S::NCG;
        if constexpr (N == 2) {
            // p = 1: every point is a corner and each corner neighbours three
            // others, so one work-item updates them in the reference's order
            // (viscosity_smoothing.f90:31-93)
            if (tid == 0) {
                constexpr int cn[8][3] = {{1, 1, 1}, {N, 1, 1}, {1, N, 1}, {1, 1, N},
                                          {1, N, N}, {N, 1, N}, {N, N, 1}, {N, N, N}};
#pragma unroll 1
                for (int c = 0; c < 8; ++c) {
                    const int n = cn[c][0], mm = cn[c][1], l = cn[c][2];
                    sMu[(n - 1) + N * ((mm - 1) + N * (l - 1))] = ev_smooth_point<P>(sMu, 0, n, mm, l);
                }
            }
            __syncthreads();
        } else {
            // corners (no two neighbour each other when N > 2): all eight
            // read, then write
            double cval = 0.0;
            int cidx = -1;
            if (tid < 8) {
                const int n = (tid & 1) ? N : 1, mm = (tid & 2) ? N : 1, l = (tid & 4) ? N : 1;
                cval = ev_smooth_point<P>(sMu, 0, n, mm, l);
                cidx = (n - 1) + N * ((mm - 1) + N * (l - 1));
            }
            __syncthreads();
            if (cidx >= 0) sMu[cidx] = cval;
            __syncthreads();
        }
        // faces: the six faces are independent; within a face, lexicographic
        // Gauss-Seidel equals a sweep over the diagonals u+v = h
        if (N > 2) {
            for (int hh = 0; hh <= 2 * (N - 3); ++hh) {
                for (int i = tid; i < 6 * (N - 2); i += BLOCK) {
                    const int f = i / (N - 2), uu = 2 + i % (N - 2), vv = 2 + hh - (uu - 2);
                    if (vv >= 2 && vv <= N - 1) {
                        int n, mm, l;
                        if (f == 0) { n = 1; mm = uu; l = vv; }
                        else if (f == 1) { n = uu; mm = 1; l = vv; }
                        else if (f == 2) { n = uu; mm = vv; l = 1; }
                        else if (f == 3) { n = N; mm = uu; l = vv; }
                        else if (f == 4) { n = uu; mm = N; l = vv; }
                        else { n = uu; mm = vv; l = N; }
                        sMu[(n - 1) + N * ((mm - 1) + N * (l - 1))] = ev_smooth_point<P>(sMu, 1 + f, n, mm, l);
                    }
                }
                __syncthreads();
            }
            // interior: sweep over the planes n+m+l = const
            for (int hh = 0; hh <= 3 * (N - 3); ++hh) {
                for (int i = tid; i < (N - 2) * (N - 2); i += BLOCK) {
                    const int n = 2 + i % (N - 2), mm = 2 + i / (N - 2), l = 2 + hh - (n - 2) - (mm - 2);
                    if (l >= 2 && l <= N - 1)
                        sMu[(n - 1) + N * ((mm - 1) + N * (l - 1))] = ev_smooth_point<P>(sMu, 7, n, mm, l);
                }
                __syncthreads();
            }
        }
    }
    if (muhg_out)
        for (int pt = tid; pt < S::NP; pt += BLOCK) muhg_out[(size_t)e * S::NP + pt] = sMu[pt];
    for (int i = tid; i < 3 * S::NS; i += BLOCK) {
        const int d = i / S::NS, s2 = i % S::NS;
        int j, a, b;
        if (d == 0) { j = s2 % S::NC; a = (s2 / S::NC) % S::NCG; b = s2 / (S::NC * S::NCG); }
        else if (d == 1) { a = s2 % S::NCG; j = (s2 / S::NCG) % S::NC; b = s2 / (S::NCG * S::NC); }
        else { a = s2 % S::NCG; b = (s2 / S::NCG) % S::NCG; j = s2 / (S::NCG * S::NCG); }
        double t = 0.0;
#pragma unroll
        for (int k = 0; k < S::NCG; ++k) t = fma(sBm[j * S::NCG + k], sMu[gidx<P>(d, k, a, b)], t);
        muh[(size_t)e * 3 * S::NS + i] = t;
    }
}

// Static Smagorinsky eddy viscosity (simpleSmag, staticSmag.f90:32-103;
// filteredStrain :126-179; delta_g :191-201; nut_prolong :206-261) for one
// element: Gauss-point gradients of the lagged staggered states,
// S = (2 S_ij S_ij)^(1/2), nu_t = Re (C_s Delta)^2 S clamped to [0, numax],
// prolonged to the three staggered grids with negative values set to zero.
// The shock and density sensors (Gauss-point fields left by k_ev) scale
// nu_t before the clamp when their switches are set.  As in filteredStrain, rho = (JQ)_1 jacob of the element's first point and
// u = (JQ)_2 / rho.
template <int P>
__global__ __launch_bounds__(BLOCK) void k_smag(const double *__restrict__ JQ, const double *__restrict__ Slag,
                                                 DevMesh m, DevPhys ph, double cs, double numax, int sensors,
                                                 const double *__restrict__ ducros, const double *__restrict__ rhos,
                                                 double *__restrict__ nut, double *__restrict__ nut_g)
{
    using S = Sz<P>;
    __shared__ double sS[3 * 4 * S::NS];
    __shared__ double sNu[S::NP];
    __shared__ double sDm[S::NCG * S::NC], sBm[S::NC * S::NCG];
    const int e = blockIdx.x, tid = threadIdx.x;
    stage_staggered<P, 4>(Slag + (size_t)e * 3 * NEQ * S::NS, sS, tid);
    stage_ops<P>(sDm, sBm, tid);
    __syncthreads();
    const int npm = m.ms ? S::NP : 1;
    const double *gg = m.gmetg + (size_t)e * 9 * npm;
    const double *jq = JQ + (size_t)e * NEQ * S::NP;
    const double *jac = m.jacob + (size_t)e * npm;
    const double c3 = pow((double)(S::NCG * S::NCG * S::NCG), 1.0 / 3.0);
    for (int pt = tid; pt < S::NP; pt += BLOCK) {
        const int n = pt % S::NCG, mi = (pt / S::NCG) % S::NCG, l = pt / (S::NCG * S::NCG);
        const int pq = pt * m.ms;
        double gr[4][3];
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            double wd[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const int k = d == 0 ? n : (d == 1 ? mi : l);
                const int a = d == 0 ? mi : n;
                const int b = d == 2 ? mi : l;
                const double *fld = sS + (d * 4 + f) * S::NS;
                double t = 0.0;
#pragma unroll
                for (int j = 0; j < S::NC; ++j) t = fma(sDm[k * S::NC + j], fld[sidx<P>(d, j, a, b)], t);
                wd[d] = t;
            }
#pragma unroll
            for (int c = 0; c < 3; ++c)
                gr[f][c] = gg[(0 * 3 + c) * npm + pq] * wd[0] + gg[(1 * 3 + c) * npm + pq] * wd[1] +
                           gg[(2 * 3 + c) * npm + pq] * wd[2];
        }
        const double rho = jq[pt] * jac[0];
        const double u = jq[S::NP + pt] / rho, v = jq[2 * S::NP + pt] / rho, w = jq[3 * S::NP + pt] / rho;
        const double ux = (gr[1][0] - gr[0][0] * u) / rho, uy = (gr[1][1] - gr[0][1] * u) / rho, uz = (gr[1][2] - gr[0][2] * u) / rho;
        const double vx = (gr[2][0] - gr[0][0] * v) / rho, vy = (gr[2][1] - gr[0][1] * v) / rho, vz = (gr[2][2] - gr[0][2] * v) / rho;
        const double wx = (gr[3][0] - gr[0][0] * w) / rho, wy = (gr[3][1] - gr[0][1] * w) / rho, wz = (gr[3][2] - gr[0][2] * w) / rho;
        const double S12 = 0.50 * (uy + vx), S13 = 0.50 * (uz + wx), S23 = 0.50 * (vz + wy);
        const double Sbar = sqrt(2.0 * ux * ux + 2.0 * S12 * S12 + 2.0 * S13 * S13 + 2.0 * S12 * S12 + 2.0 * vy * vy +
                                 2.0 * S23 * S23 + 2.0 * S13 * S13 + 2.0 * S23 * S23 + 2.0 * wz * wz);
        const double delta = pow(fabs(1.0 / jac[pq]), 1.0 / 3.0) / c3;
        double nt = ph.re * (cs * delta) * (cs * delta) * Sbar;
        // shockSensor / rhoSensor switches (staticSmag.f90:68-69), before the clamp
        if (sensors & DSEM_LES_SHOCK_SENSOR) nt = nt * (1.0 - ducros[(size_t)e * S::NP + pt]);
        if (sensors & DSEM_LES_RHO_SENSOR) nt = nt * rhos[(size_t)e * S::NP + pt];
        sNu[pt] = fmin(fmax(nt, 0.0), numax);
        if (nut_g) nut_g[(size_t)e * S::NP + pt] = sNu[pt];
    }
    __syncthreads();
    for (int i = tid; i < 3 * S::NS; i += BLOCK) {
        const int d = i / S::NS, s2 = i % S::NS;
        int j, a, b;
        if (d == 0) { j = s2 % S::NC; a = (s2 / S::NC) % S::NCG; b = s2 / (S::NC * S::NCG); }
        else if (d == 1) { a = s2 % S::NCG; j = (s2 / S::NCG) % S::NC; b = s2 / (S::NCG * S::NC); }
        else { a = s2 % S::NCG; b = (s2 / S::NCG) % S::NCG; j = s2 / (S::NCG * S::NCG); }
        double t = 0.0;
#pragma unroll
        for (int k = 0; k < S::NCG; ++k) t = fma(sBm[j * S::NCG + k], sNu[gidx<P>(d, k, a, b)], t);
        nut[(size_t)e * 3 * S::NS + i] = t < 0.0 ? 0.0 : t;
    }
}

// Per-element partial sums of the flow diagnostics, part[e][0..4]:
//   0..2  sum (u^2, v^2, w^2) bmat over the Gauss points, u = (JQ)_2/(JQ)_1
//         (turbulent kinetic energy of do_user_stuff,
//         caseFiles/tgv/BackstepProblem-stoch.f90:629-650)
//   3, 4  sum |omega|^2/jacob and sum 1/jacob (energydissipation,
//         src/Energydissipation/energydissipation.f90:19-85), with the
//         velocity gradients from Gauss_Deriv of the lagged staggered states
//         Slag (fields rho, rho u, rho v, rho w); skipped when Slag is null.
template <int P>
__global__ __launch_bounds__(BLOCK) void k_diag(const double *__restrict__ JQ, const double *__restrict__ Slag,
                                                 DevMesh m, const double *__restrict__ bmat, double *__restrict__ part)
{
    using S = Sz<P>;
    __shared__ double sS[3 * 4 * S::NS];
    __shared__ double sRed[5][BLOCK];
    __shared__ double sDm[S::NCG * S::NC], sBm[S::NC * S::NCG];
    const int e = blockIdx.x, tid = threadIdx.x;
    const bool vort = Slag != nullptr;
    if (vort) {
        stage_staggered<P, 4>(Slag + (size_t)e * 3 * NEQ * S::NS, sS, tid);
        stage_ops<P>(sDm, sBm, tid);
    }
    __syncthreads();
    const int npm = m.ms ? S::NP : 1;
    const double *gg = m.gmetg + (size_t)e * 9 * npm;
    const double *jq = JQ + (size_t)e * NEQ * S::NP;
    const double *jac = m.jacob + (size_t)e * npm;
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int pt = tid; pt < S::NP; pt += BLOCK) {
        const double q1 = jq[pt];
        const double u = jq[S::NP + pt] / q1, v = jq[2 * S::NP + pt] / q1, w = jq[3 * S::NP + pt] / q1;
        const double bw = bmat[pt];
        acc[0] += (u * u) * bw;
        acc[1] += (v * v) * bw;
        acc[2] += (w * w) * bw;
        if (vort) {
            const int n = pt % S::NCG, mi = (pt / S::NCG) % S::NCG, l = pt / (S::NCG * S::NCG);
            const int pq = pt * m.ms;
            double gr[4][3];
#pragma unroll
            for (int f = 0; f < 4; ++f) {
                double wd[3];
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    const int k = d == 0 ? n : (d == 1 ? mi : l);
                    const int a = d == 0 ? mi : n;
                    const int b = d == 2 ? mi : l;
                    const double *fld = sS + (d * 4 + f) * S::NS;
                    double t = 0.0;
#pragma unroll
                    for (int j = 0; j < S::NC; ++j) t = fma(sDm[k * S::NC + j], fld[sidx<P>(d, j, a, b)], t);
                    wd[d] = t;
                }
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    gr[f][c] = gg[(0 * 3 + c) * npm + pq] * wd[0] + gg[(1 * 3 + c) * npm + pq] * wd[1] +
                               gg[(2 * 3 + c) * npm + pq] * wd[2];
            }
            const double udery = (gr[1][1] - u * gr[0][1]) / q1, uderz = (gr[1][2] - u * gr[0][2]) / q1;
            const double vderx = (gr[2][0] - v * gr[0][0]) / q1, vderz = (gr[2][2] - v * gr[0][2]) / q1;
            const double wderx = (gr[3][0] - w * gr[0][0]) / q1, wdery = (gr[3][1] - w * gr[0][1]) / q1;
            const double jc = jac[pq];
            const double c1 = (wdery - vderz) / jc, c2 = (uderz - wderx) / jc, c3 = (vderx - udery) / jc;
            acc[3] += (c1 * c1 + c2 * c2 + c3 * c3) / jc;
            acc[4] += 1.0 / jc;
        }
    }
#pragma unroll
    for (int i = 0; i < 5; ++i) sRed[i][tid] = acc[i];
    __syncthreads();
    for (int st = BLOCK / 2; st > 0; st >>= 1) {
        if (tid < st)
#pragma unroll
            for (int i = 0; i < 5; ++i) sRed[i][tid] += sRed[i][tid + st];
        __syncthreads();
    }
    if (tid < 5) part[(size_t)e * 5 + tid] = sRed[tid][0];
}

// Sum of the per-element partials in a fixed order (one workgroup), so the
// diagnostics are reproducible from run to run.
__global__ __launch_bounds__(BLOCK) void k_diag_sum(const double *__restrict__ part, int K, double *__restrict__ out)
{
    __shared__ double sRed[5][BLOCK];
    const int tid = threadIdx.x;
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int e = tid; e < K; e += BLOCK)
#pragma unroll
        for (int i = 0; i < 5; ++i) acc[i] += part[(size_t)e * 5 + i];
#pragma unroll
    for (int i = 0; i < 5; ++i) sRed[i][tid] = acc[i];
    __syncthreads();
    for (int st = BLOCK / 2; st > 0; st >>= 1) {
        if (tid < st)
#pragma unroll
            for (int i = 0; i < 5; ++i) sRed[i][tid] += sRed[i][tid + st];
        __syncthreads();
    }
    if (tid < 5) out[tid] = sRed[tid][0];
}

// dtcalc (dtcalc.f90:9-58): per point 1/dti, reduced to the minimum
template <int P>
__global__ __launch_bounds__(BLOCK) void k_dtcalc(const double *__restrict__ JQ, DevMesh m, DevPhys ph,
                                                   unsigned long long *out)
{
    using S = Sz<P>;
    const int e = blockIdx.x;
    const double *q = JQ + (size_t)e * NEQ * S::NP;
    const int nsm = m.ms ? S::NS : 1;
    double best = 1.0e308;
    for (int pt = threadIdx.x; pt < S::NP; pt += BLOCK) {
        const int n = pt % S::NCG, mi = (pt / S::NCG) % S::NCG, l = pt / (S::NCG * S::NCG);
        const double dcx = c_cx[P][n + 1] - c_cx[P][n];
        const double dcy = c_cx[P][mi + 1] - c_cx[P][mi];
        const double dcz = c_cx[P][l + 1] - c_cx[P][l];
        const double r = q[0 * S::NP + pt];
        const double u0 = fabs(q[1 * S::NP + pt] / r), v0 = fabs(q[2 * S::NP + pt] / r), w0 = fabs(q[3 * S::NP + pt] / r);
        const double p = (ph.gamma - 1.0) * (q[4 * S::NP + pt] - 0.5 * r * (u0 * u0 + v0 * v0 + w0 * w0));
        const double a = sqrt(ph.gamma * p / r);
        const int s0 = (n + S::NC * (mi + S::NCG * l)) * m.ms;
        const int s1 = (n + S::NCG * (mi + S::NC * l)) * m.ms;
        const int s2 = (n + S::NCG * (mi + S::NCG * l)) * m.ms;
        const double *g0 = m.gmet0 + (size_t)e * 3 * nsm, *g1 = m.gmet1 + (size_t)e * 3 * nsm, *g2 = m.gmet2 + (size_t)e * 3 * nsm;
        const double uu = (u0 + a) * g0[s0] + (v0 + a) * g0[nsm + s0] + (w0 + a) * g0[2 * nsm + s0];
        const double vv = (u0 + a) * g1[s1] + (v0 + a) * g1[nsm + s1] + (w0 + a) * g1[2 * nsm + s1];
        const double ww = (u0 + a) * g2[s2] + (v0 + a) * g2[nsm + s2] + (w0 + a) * g2[2 * nsm + s2];
        const double jac = m.jacob[(size_t)e * (m.ms ? S::NP : 1) + pt * m.ms];
        const double dti = fabs(uu * jac) / dcx + fabs(vv * jac) / dcy + fabs(ww * jac) / dcz + 1.0e-14;
        best = fmin(best, 1.0 / dti);
    }
    atomicMin(out, (unsigned long long)__double_as_longlong(best));
}

// One 1-D transform of the modal filter along direction d, for the three
// momentum fields of the element in LDS: out = C (in) along d.
template <int P>
__device__ __forceinline__ void modal_pass(const double *in, double *out, const double *C, int d, int tid)
{
    using S = Sz<P>;
    for (int i = tid; i < 3 * S::NP; i += BLOCK) {
        const int v = i / S::NP, pt = i % S::NP;
        const int n = pt % S::NCG, mi = (pt / S::NCG) % S::NCG, l = pt / (S::NCG * S::NCG);
        const int k = d == 0 ? n : (d == 1 ? mi : l);
        const int a = d == 0 ? mi : n;
        const int b = d == 2 ? mi : l;
        double t = 0.0;
#pragma unroll
        for (int j = 0; j < S::NCG; ++j) t = fma(C[k * S::NCG + j], in[v * S::NP + gidx<P>(d, j, a, b)], t);
        out[i] = t;
    }
}

// modalfilter (src/Energydissipation/energydissipation.f90:720-959), called
// every 50 steps before tstep (src/3DDriver_parfmdf.f90:462-463), for one
// element: the momentum fields Q = (J Q) jacob(1,1,1) go to modal space by a
// 3-D DCT-II, are divided by 2N (mode 0) or N (modes >= 1) per direction,
// modes with any index >= cutoff are zeroed (the reference keeps them all:
// its mode edit at :778-783 is commented out, cutoff = N), then divided by 2
// per direction for modes >= 1 and brought back by a 3-D DCT-III, and
// divided by jacob(1,1,1).  With cutoff = N this is the identity up to
// round-off.
template <int P>
__global__ __launch_bounds__(BLOCK) void k_modal(double *__restrict__ JQ, DevMesh m, int cutoff)
{
    using S = Sz<P>;
    constexpr int N = S::NCG, NQ = 3 * S::NP;
    __shared__ double sA[NQ], sB[NQ];
    __shared__ double sC2[N * N], sC3[N * N];
    const int e = blockIdx.x, tid = threadIdx.x;
    const double j0 = m.jacob[(size_t)e * (m.ms ? S::NP : 1)];
    double *q = JQ + ((size_t)e * NEQ + 1) * S::NP;   // rho u, rho v, rho w
    for (int i = tid; i < NQ; i += BLOCK) sA[i] = q[i] * j0;
    for (int i = tid; i < N * N; i += BLOCK) {
        sC2[i] = c_dct2[P][i];
        sC3[i] = c_dct3[P][i];
    }
    __syncthreads();
    modal_pass<P>(sA, sB, sC2, 0, tid);
    __syncthreads();
    modal_pass<P>(sB, sA, sC2, 1, tid);
    __syncthreads();
    modal_pass<P>(sA, sB, sC2, 2, tid);
    __syncthreads();
    const double n2 = 2.0 * N, n1 = (double)N;
    for (int i = tid; i < NQ; i += BLOCK) {
        const int pt = i % S::NP;
        const int a = pt % N, b = (pt / N) % N, c = pt / (N * N);
        double x = sB[i];
        x = x / (a == 0 ? n2 : n1);
        x = x / (b == 0 ? n2 : n1);
        x = x / (c == 0 ? n2 : n1);
        if (a >= cutoff || b >= cutoff || c >= cutoff) x = 0.0;
        if (a > 0) x = x / 2.0;
        if (b > 0) x = x / 2.0;
        if (c > 0) x = x / 2.0;
        sB[i] = x;
    }
    __syncthreads();
    modal_pass<P>(sB, sA, sC3, 0, tid);
    __syncthreads();
    modal_pass<P>(sA, sB, sC3, 1, tid);
    __syncthreads();
    modal_pass<P>(sB, sA, sC3, 2, tid);
    __syncthreads();
    for (int i = tid; i < NQ; i += BLOCK) q[i] = sA[i] / j0;
}

#include "adjoint.inc"

}  // namespace

// ====================================================================== host
namespace {

// Carpenter-Kennedy 2N RK4(5) coefficients as rkcoefs(4) writes them
// (src/Independentroutines/baselib.f90:135-152): ark = A_k, crk = B_k.
const double kArk[5] = {0.0, -0.4178904745, -1.192151694643, -1.697784692471, -1.514183444257};
const double kCrk[5] = {0.1496590219993, 0.3792103129999, 0.8229550293869, 0.6994504559488, 0.1530572479681};

enum KernelId { KID_TRACES = 0, KID_FACE_INV, KID_VOLUME, KID_FACE_VISC, KID_UPDATE, KID_DTCALC, KID_EV, KID_DIAG, KID_ADJ, KID_LES, KID_MODAL, KID_COUNT };
const char *const kKernelNames[KID_COUNT] = {"k_traces", "k_face_inv", "k_volume", "k_face_visc",
                                             "k_update", "k_dtcalc", "k_ev", "k_diag", "k_adj", "k_smag",
                                             "k_modal"};

template <typename T>
struct DevBuf {
    T *ptr = nullptr;
    size_t n = 0;
};

}  // namespace

struct dsem_ctx {
    int device = 0;
    int K = 0, p = 0, M = 0, ncg = 0, nc = 0, np = 0, ns = 0, nf = 0;
    int ms = 1;
    dsem_params params{};
    std::string err;
    hipStream_t stream = nullptr;
    dsem::Ops ops;
    std::vector<double> h_xg, h_jacob;   // Gauss points and 1/J on the host
    std::vector<double> h_corners;       // [K][8][3], for host-side geometry queries
    std::vector<int32_t> h_melem, h_mface;
    // device buffers
    double *JQ = nullptr, *G = nullptr, *Qtp = nullptr, *Qt = nullptr;
    double *T = nullptr, *FV = nullptr, *QA = nullptr, *FI = nullptr, *FT = nullptr;
    double *jacob = nullptr, *gmetg = nullptr, *gmet[3] = {nullptr, nullptr, nullptr};
    double *nhat = nullptr, *norm = nullptr, *muh = nullptr;
    double *Slag = nullptr, *muhg = nullptr, *ev_resid = nullptr, *ev_ducros = nullptr;   // entropy viscosity
    bool have_lag = false;
    bool keep_lag = false;                     // save the last stage's staggered states without EV
    bool traces_cur = false;                   // T holds the face traces of the current JQ
    int vol_grid[3] = {0, 0, 0};               // resident volume-kernel workgroups per precision mode
    bool les = false;                          // static Smagorinsky model
    double les_cs = 0.0, les_numax = 0.0;
    double *nut = nullptr, *nut_g = nullptr;   // eddy viscosity on the staggered grids / Gauss points
    int les_sensors = 0;                       // DSEM_LES_SHOCK_SENSOR | DSEM_LES_RHO_SENSOR
    double *ev_rhos = nullptr;                 // density sensor at the Gauss points (k_ev)
    double *bmat = nullptr, *diag_part = nullptr, *diag_out = nullptr;   // diagnostics
    // adjoint buffers (allocated by the first VJP call)
    double *aQVb = nullptr, *aGVb = nullptr, *aFTbe = nullptr, *aFVb = nullptr, *aQAbe = nullptr;
    double *aTb = nullptr, *aFIb = nullptr, *aJQvb = nullptr, *aNuPart = nullptr, *aNu = nullptr;
    double *aGb = nullptr, *aLam = nullptr, *aOut = nullptr;
    double *aMuB = nullptr, *aWB = nullptr, *aSlagB = nullptr;   // entropy-viscosity adjoint (shock = 1)
    double *aNutB = nullptr;                   // Smagorinsky adjoint: that of the staggered nu_t
    // per tape slot: bit 0 the recorded step had lagged states, bit 1 it ran
    // the Smagorinsky model (with the constants in tape_cs, tape_numax)
    std::vector<char> tape_lag;
    std::vector<double> tape_cs, tape_numax;
    std::vector<int> tape_sens;                // per tape slot: the Smagorinsky sensor switches
    bool tape_les = false;                     // the tape's slots hold nu_t (the layout of tape_reserve)
    // step tape: per slot and stage the stage input state and the forward
    // intermediates the adjoint reads (face traces T, mortar averages QA,
    // viscous face traces FV)
    double *tape = nullptr;
    int tape_slots = 0;
    int *e2m = nullptr, *m_elem = nullptr, *m_face = nullptr, *m_sign = nullptr, *m_nsign = nullptr;
    int *m_orient = nullptr;                   // [M] orientation codes, allocated only when some face is rotated
    bool rotated = false;                      // some mortar has orientation != 0
    unsigned long long *d_scalars = nullptr;   // [0] resid bits, [1] dt bits, [2] nonfinite flag
    double *scratch = nullptr;                 // device address of this device's `scratch` array
    std::vector<void *> allocations;
    // timing
    bool timing = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double t_ms[KID_COUNT] = {};
    int64_t t_n[KID_COUNT] = {};
};

namespace {

int set_err(dsem_ctx *c, int code, const std::string &msg)
{
    if (c) c->err = msg;
    return code;
}

#define HIP_TRY(ctx, call)                                                                                  \
    do {                                                                                                    \
        hipError_t e_ = (call);                                                                             \
        if (e_ != hipSuccess)                                                                               \
            return set_err((ctx), DSEM_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));         \
    } while (0)

template <typename T>
int dev_alloc(dsem_ctx *c, T **p, size_t n)
{
    void *ptr = nullptr;
    hipError_t e = hipMalloc(&ptr, sizeof(T) * (n ? n : 1));
    if (e != hipSuccess) return set_err(c, DSEM_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    c->allocations.push_back(ptr);
    *p = static_cast<T *>(ptr);
    return DSEM_OK;
}

template <typename T>
int upload(dsem_ctx *c, T **p, const std::vector<T> &h)
{
    int rc = dev_alloc(c, p, h.size());
    if (rc) return rc;
    HIP_TRY(c, hipMemcpy(*p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
    return DSEM_OK;
}

DevMesh dev_mesh(const dsem_ctx *c)
{
    DevMesh m;
    m.K = c->K; m.M = c->M; m.ms = c->ms;
    m.jacob = c->jacob; m.gmetg = c->gmetg; m.gmet0 = c->gmet[0]; m.gmet1 = c->gmet[1]; m.gmet2 = c->gmet[2];
    m.e2m = c->e2m; m.m_elem = c->m_elem; m.m_face = c->m_face; m.m_sign = c->m_sign; m.m_nsign = c->m_nsign;
    m.m_orient = c->m_orient;
    m.rot = c->rotated ? 1 : 0;
    m.nhat = c->nhat; m.norm = c->norm;
    return m;
}

DevPhys dev_phys(const dsem_ctx *c)
{
    const dsem_params &p = c->params;
    DevPhys d;
    d.gamma = p.gamma;
    d.re = p.re;
    d.pr = p.pr;
    d.mach = p.mach;
    d.gm1 = p.gamma - 1.0;
    d.tfac = p.mach * p.mach * p.gamma;
    d.hfac1 = 1.0 / (p.pr * p.mach * p.mach * (p.gamma - 1.0));
    d.kap = 2.0e-1 / (p.gamma - 1.0);
    d.inv_re = 1.0 / p.re;
    return d;
}

// Launch helper: brackets the launch with events when timing is on.
template <typename F>
int launch(dsem_ctx *c, int kid, F &&fn)
{
    if (c->timing) HIP_TRY(c, hipEventRecord(c->ev0, c->stream));
    fn();
    HIP_TRY(c, hipGetLastError());
    if (c->timing) {
        HIP_TRY(c, hipEventRecord(c->ev1, c->stream));
        HIP_TRY(c, hipEventSynchronize(c->ev1));
        float ms = 0.f;
        HIP_TRY(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
        c->t_ms[kid] += ms;
        c->t_n[kid] += 1;
    }
    return DSEM_OK;
}

#define TRY(x)                 \
    do {                       \
        int rc_ = (x);         \
        if (rc_) return rc_;   \
    } while (0)

// Workgroups of the volume kernel.  The FP64 kernels of 1024 work-items
// (p >= 7) are persistent: as many workgroups as are resident at once, each
// looping over elements with a stride of the grid (never more than the
// elements; queried once per context).  The others take one element each.
template <int P, int PM>
int volume_grid_size(dsem_ctx *c, int &grid)
{
    if (!(PM == 0 && vblock<P, PM>() == 1024)) {
        grid = c->K;
        return DSEM_OK;
    }
    int &g = c->vol_grid[PM];
    if (g == 0) {
        int per_cu = 0, ncu = 0;
        // both instances hold the same LDS, which allows one workgroup per CU
        HIP_TRY(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_volume<P, PM, false>, vblock<P, PM>(), 0));
        HIP_TRY(c, hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device));
        if (per_cu < 1 || ncu < 1) return set_err(c, DSEM_ERR_HIP, "volume kernel: no resident work-group fits");
        g = per_cu * ncu;
    }
    grid = g < c->K ? g : c->K;
    return DSEM_OK;
}

// One RHS evaluation on state JQ; the final kernel either writes Qt or
// performs an RK register update (st.rk).
// traces_ready: T already holds the face traces of JQ (computed by the
// previous stage's update); make_traces: this stage's update also computes
// the traces of the updated state for the next stage.
// Per-call intermediates of one RHS evaluation: face traces T, mortar
// averages QA, Roe fluxes FI, viscous face traces FV, total face fluxes FT
// and the partial residual Qtp.  Every entry a call reads is written earlier
// in the same call, so none needs clearing.
struct RhsBufs {
    double *T, *QA, *FI, *FV, *FT, *Qtp;
};

RhsBufs ctx_bufs(const dsem_ctx *c) { return RhsBufs{c->T, c->QA, c->FI, c->FV, c->FT, c->Qtp}; }

template <int P>
int run_rhs(dsem_ctx *c, double *JQ, const StageArgs &st, double *Qt_out, bool traces_ready, bool make_traces,
            bool save_lag, const RhsBufs &b)
{
    using S = Sz<P>;
    const DevMesh m = dev_mesh(c);
    const DevPhys ph = dev_phys(c);
    const int K = c->K;
    const int face_blocks = (c->M * S::NF + BLOCK - 1) / BLOCK;
    const double *muh = c->params.shock ? c->muh : nullptr;
    const double *nut = c->les ? c->nut : nullptr;
    hipStream_t s = c->stream;
    c->traces_cur = false;   // only dsem_rk_steps vouches for T after its last stage
    if (!traces_ready) TRY(launch(c, KID_TRACES, [&] { k_traces<P><<<K, BLOCK, 0, s>>>(JQ, m, b.T); }));
    TRY(launch(c, KID_FACE_INV, [&] { k_face_inv<P><<<face_blocks, BLOCK, 0, s>>>(b.T, m, ph, b.FI, b.QA); }));
    double *slag = save_lag ? c->Slag : nullptr;
    int gv0 = 0, gv1 = 0, gv2 = 0;
    if (c->params.precision == DSEM_PREC_FP32) TRY((volume_grid_size<P, 1>(c, gv1)));
    else if (c->params.precision == DSEM_PREC_BF16X3) TRY((volume_grid_size<P, 2>(c, gv2)));
    else TRY((volume_grid_size<P, 0>(c, gv0)));
    const bool plain = !c->rotated && !muh && !nut && !slag;
    auto vol = [&](auto pm, auto pl, int grid) {
        constexpr int PM = decltype(pm)::value;
        k_volume<P, PM, decltype(pl)::value><<<grid, vblock<P, PM>(), 0, s>>>(JQ, m, ph, b.QA, muh, nut, b.FV, b.Qtp,
                                                                             slag);
    };
    using T_ = std::true_type;
    using F_ = std::false_type;
    TRY(launch(c, KID_VOLUME, [&] {
        if (c->params.precision == DSEM_PREC_FP32)
            vol(std::integral_constant<int, 1>{}, F_{}, gv1);
        else if (c->params.precision == DSEM_PREC_BF16X3)
            vol(std::integral_constant<int, 2>{}, F_{}, gv2);
        else
            plain ? vol(std::integral_constant<int, 0>{}, T_{}, gv0) : vol(std::integral_constant<int, 0>{}, F_{}, gv0);
    }));
    TRY(launch(c, KID_FACE_VISC, [&] { k_face_visc<P><<<face_blocks, BLOCK, 0, s>>>(b.FV, b.FI, m, ph, b.FT); }));
    if (make_traces)
        TRY(launch(c, KID_UPDATE, [&] { k_update<P, true><<<K, UBLOCK, 0, s>>>(b.Qtp, b.FT, m, st, JQ, c->G, Qt_out, b.T); }));
    else
        TRY(launch(c, KID_UPDATE, [&] { k_update<P, false><<<K, UBLOCK, 0, s>>>(b.Qtp, b.FT, m, st, JQ, c->G, Qt_out, b.T); }));
    return DSEM_OK;
}

template <int P>
int run_dtcalc(dsem_ctx *c, unsigned long long *out)
{
    const DevMesh m = dev_mesh(c);
    const DevPhys ph = dev_phys(c);
    return launch(c, KID_DTCALC, [&] { k_dtcalc<P><<<c->K, BLOCK, 0, c->stream>>>(c->JQ, m, ph, out); });
}

#define DISPATCH_P(ctx, fn, ...)                        \
    switch ((ctx)->p) {                                 \
    case 1: return fn<1>(ctx, __VA_ARGS__);             \
    case 2: return fn<2>(ctx, __VA_ARGS__);             \
    case 3: return fn<3>(ctx, __VA_ARGS__);             \
    case 4: return fn<4>(ctx, __VA_ARGS__);             \
    case 5: return fn<5>(ctx, __VA_ARGS__);             \
    case 6: return fn<6>(ctx, __VA_ARGS__);             \
    case 7: return fn<7>(ctx, __VA_ARGS__);             \
    case 8: return fn<8>(ctx, __VA_ARGS__);             \
    default: return set_err(ctx, DSEM_ERR_ARG, "p");    \
    }

int rhs_dispatch(dsem_ctx *c, double *JQ, const StageArgs &st, double *Qt_out, bool traces_ready = false,
                 bool make_traces = false, bool save_lag = false, const RhsBufs *bufs = nullptr)
{
    const RhsBufs b = bufs ? *bufs : ctx_bufs(c);
    DISPATCH_P(c, run_rhs, JQ, st, Qt_out, traces_ready, make_traces, save_lag, b);
}

template <int P>
int run_ev(dsem_ctx *c, bool debug)
{
    const DevMesh m = dev_mesh(c);
    const DevPhys ph = dev_phys(c);
    EvArgs ev;
    ev.cmu = c->params.cmu;
    ev.mumax = c->params.mumax;
    ev.cmax = c->params.cmax;
    ev.ducros = (c->params.ev_flags & DSEM_EV_DUCROS) != 0;
    ev.smoothing = (c->params.ev_flags & DSEM_EV_SMOOTHING) != 0;
    ev.cmax_cap = (c->params.ev_flags & DSEM_EV_CMAX_CAP) != 0;
    // the Smagorinsky sensors read the Ducros and density sensors of this call
    const bool sens = c->les && c->les_sensors;
    return launch(c, KID_EV, [&] {
        k_ev<P><<<c->K, BLOCK, 0, c->stream>>>(c->JQ, c->Slag, m, ph, ev, c->muh, c->muhg, debug ? c->ev_resid : nullptr,
                                               (debug || sens) ? c->ev_ducros : nullptr,
                                               (debug || sens) ? c->ev_rhos : nullptr);
    });
}

int ev_dispatch(dsem_ctx *c, bool debug)
{
    DISPATCH_P(c, run_ev, debug);
}

// Entropy at the start of a step: mu_h from the lagged states, or zero on a
// fresh start (the reference divides 0/0 there and its max(NaN,0) yields 0).
int ev_step(dsem_ctx *c, bool debug)
{
    if (!c->params.shock) return DSEM_OK;
    if (!c->have_lag) {
        HIP_TRY(c, hipMemsetAsync(c->muh, 0, sizeof(double) * (size_t)c->K * 3 * c->ns, c->stream));
        HIP_TRY(c, hipMemsetAsync(c->muhg, 0, sizeof(double) * (size_t)c->K * c->np, c->stream));
        return DSEM_OK;
    }
    return ev_dispatch(c, debug);
}

template <int P>
int run_les(dsem_ctx *c, bool gauss)
{
    const DevMesh m = dev_mesh(c);
    const DevPhys ph = dev_phys(c);
    return launch(c, KID_LES, [&] {
        k_smag<P><<<c->K, BLOCK, 0, c->stream>>>(c->JQ, c->Slag, m, ph, c->les_cs, c->les_numax, c->les_sensors,
                                                 c->ev_ducros, c->ev_rhos, c->nut, gauss ? c->nut_g : nullptr);
    });
}

int les_dispatch(dsem_ctx *c, bool gauss)
{
    DISPATCH_P(c, run_les, gauss);
}

// Static Smagorinsky at the start of a step (tstep.f90:40-43): nu_t from the
// lagged states, or zero on a fresh start (the reference's zero lagged
// density gives 0/0 there, which its max(.,0) turns into 0).
int les_step(dsem_ctx *c, bool gauss)
{
    if (!c->les) return DSEM_OK;
    if (!c->have_lag) {
        HIP_TRY(c, hipMemsetAsync(c->nut, 0, sizeof(double) * (size_t)c->K * 3 * c->ns, c->stream));
        HIP_TRY(c, hipMemsetAsync(c->nut_g, 0, sizeof(double) * (size_t)c->K * c->np, c->stream));
        return DSEM_OK;
    }
    return les_dispatch(c, gauss);
}

int dtcalc_dispatch(dsem_ctx *c, unsigned long long *out)
{
    DISPATCH_P(c, run_dtcalc, out);
}

template <int P>
int run_diag(dsem_ctx *c, bool vorticity)
{
    const DevMesh m = dev_mesh(c);
    const double *slag = vorticity ? c->Slag : nullptr;
    return launch(c, KID_DIAG, [&] {
        k_diag<P><<<c->K, BLOCK, 0, c->stream>>>(c->JQ, slag, m, c->bmat, c->diag_part);
        k_diag_sum<<<1, BLOCK, 0, c->stream>>>(c->diag_part, c->K, c->diag_out);
    });
}

int diag_dispatch(dsem_ctx *c, bool vorticity)
{
    DISPATCH_P(c, run_diag, vorticity);
}

template <int P>
int run_modal(dsem_ctx *c, int cutoff)
{
    const DevMesh m = dev_mesh(c);
    return launch(c, KID_MODAL, [&] { k_modal<P><<<c->K, BLOCK, 0, c->stream>>>(c->JQ, m, cutoff); });
}

int modal_dispatch(dsem_ctx *c, int cutoff)
{
    DISPATCH_P(c, run_modal, cutoff);
}

int adjoint_alloc(dsem_ctx *c)
{
    if (c->aQVb) return DSEM_OK;
    const size_t K = c->K, n = K * NEQ * c->np, nfe = K * 6 * NEQ * c->nf;
    const size_t face_blocks = ((size_t)c->M * c->nf + BLOCK - 1) / BLOCK;
    TRY(dev_alloc(c, &c->aQVb, K * 3 * NEQ * c->ns));
    TRY(dev_alloc(c, &c->aGVb, K * 3 * 15 * c->ns));
    TRY(dev_alloc(c, &c->aFTbe, nfe));
    TRY(dev_alloc(c, &c->aFVb, nfe));
    TRY(dev_alloc(c, &c->aQAbe, nfe));
    TRY(dev_alloc(c, &c->aTb, nfe));
    TRY(dev_alloc(c, &c->aFIb, (size_t)c->M * NEQ * c->nf));
    TRY(dev_alloc(c, &c->aJQvb, n));
    TRY(dev_alloc(c, &c->aNuPart, K + face_blocks));
    TRY(dev_alloc(c, &c->aNu, 16));
    TRY(dev_alloc(c, &c->aGb, n));
    TRY(dev_alloc(c, &c->aLam, n));
    TRY(dev_alloc(c, &c->aOut, n));
    return DSEM_OK;
}

// Buffers of the lagged-state adjoints: entropy viscosity (fixed at
// creation) and the Smagorinsky model (switched on at any time).
int adjoint_lag_alloc(dsem_ctx *c)
{
    const size_t K = c->K;
    if (c->params.shock && !c->aMuB) TRY(dev_alloc(c, &c->aMuB, K * 3 * c->ns));
    if ((c->params.shock || c->les) && !c->aWB) TRY(dev_alloc(c, &c->aWB, K * (c->params.shock ? 18 : 12) * c->np));
    if ((c->params.shock || c->les) && !c->aSlagB) TRY(dev_alloc(c, &c->aSlagB, K * 3 * NEQ * c->ns));
    if (c->les && !c->aNutB) TRY(dev_alloc(c, &c->aNutB, K * 3 * c->ns));
    return DSEM_OK;
}

// Forward intermediates of one RHS evaluation that its adjoint reads.
struct FwdTape {
    double *T, *QA, *FV;
};

// Adjoint of one RHS evaluation at state JQ (device) for the cotangent in
// Gb, with the forward intermediates in f (recomputed first when fwd).
// Plain VJP (Qb == nullptr): Out = (dR/dJQ)^T Gb.  RK stage reverse
// (Qb != nullptr): Gb <- Gb + cdt Qb, Qb += (dR/dJQ)^T Gb, Gb <- a Gb.
// nu_slot receives <Gb, dR/dnu> (device, one double).
// With entropy viscosity: muh is the step's mu_h (staggered), MuB
// accumulates its adjoint, and SlagB (the adjoint of the lagged states this
// evaluation leaves, last stage only) joins the staggered-state adjoints.
// With the Smagorinsky model: nut is the step's staggered nu_t and NutB
// accumulates its adjoint.
template <int P>
int run_rhs_vjp(dsem_ctx *c, double *JQ, FwdTape f, bool fwd, double *Gb, double *Qb, double cdt, double a,
                double *Out, double *nu_slot, const double *muh, double *MuB, const double *SlagB,
                const double *nut, double *NutB)
{
    using S = Sz<P>;
    const DevMesh m = dev_mesh(c);
    const DevPhys ph = dev_phys(c);
    const int K = c->K;
    const int face_blocks = (c->M * S::NF + BLOCK - 1) / BLOCK;
    if (fwd) {
        StageArgs st{0.0, 0.0, 0, 0, nullptr, nullptr};
        TRY(run_rhs<P>(c, JQ, st, c->Qt, false, false, false, ctx_bufs(c)));   // forward intermediates in c->T, QA, FV
        f = FwdTape{c->T, c->QA, c->FV};
    }
    hipStream_t s = c->stream;
    double *nu_f = c->aNuPart, *nu_v = c->aNuPart + face_blocks;
    return launch(c, KID_ADJ, [&] {
        k_adj_update<P><<<K, BLOCK, 0, s>>>(Gb, Qb, cdt, c->aFTbe);
        k_adj_face_visc<P><<<face_blocks, BLOCK, 0, s>>>(c->aFTbe, f.FV, m, ph, c->aFIb, c->aFVb, nu_f);
        auto vol = [&](auto rot, auto lag) {
            k_adj_volume<P, (P >= 4), decltype(rot)::value, decltype(lag)::value><<<K, ABLOCK, 0, s>>>(
                JQ, m, ph, f.QA, muh, Gb, c->aFVb, c->aQVb, c->aGVb, nu_v, MuB, SlagB, nut, NutB);
        };
        auto lagged = [&](auto rot) {
            if (nut) vol(rot, std::integral_constant<int, 2>{});
            else if (muh || MuB || SlagB) vol(rot, std::integral_constant<int, 1>{});
            else vol(rot, std::integral_constant<int, 0>{});
        };
        if (c->rotated) lagged(std::true_type{});
        else lagged(std::false_type{});
        if constexpr (P >= 5) k_adj_gather_dma<P><<<K, GBLOCK, 0, s>>>(m, c->aQVb, c->aGVb, c->aJQvb, c->aQAbe);
        else k_adj_gather<P><<<K, GBLOCK, 0, s>>>(m, c->aQVb, c->aGVb, c->aJQvb, c->aQAbe);
        k_adj_face_inv<P><<<face_blocks, BLOCK, 0, s>>>(f.T, m, ph, c->aFIb, c->aQAbe, c->aTb);
        k_adj_traces<P><<<K, BLOCK, 0, s>>>(c->aTb, c->aJQvb, m, Out, Qb, Gb, a, c->aNuPart, face_blocks + K, nu_slot);
    });
}

int rhs_vjp_dispatch(dsem_ctx *c, double *JQ, FwdTape f, bool fwd, double *Gb, double *Qb, double cdt, double a,
                     double *Out, double *nu_slot, const double *muh = nullptr, double *MuB = nullptr,
                     const double *SlagB = nullptr, const double *nut = nullptr, double *NutB = nullptr)
{
    DISPATCH_P(c, run_rhs_vjp, JQ, f, fwd, Gb, Qb, cdt, a, Out, nu_slot, muh, MuB, SlagB, nut, NutB);
}

EvArgs ev_args(const dsem_ctx *c)
{
    EvArgs ev;
    ev.cmu = c->params.cmu;
    ev.mumax = c->params.mumax;
    ev.cmax = c->params.cmax;
    ev.ducros = (c->params.ev_flags & DSEM_EV_DUCROS) != 0;
    ev.smoothing = (c->params.ev_flags & DSEM_EV_SMOOTHING) != 0;
    ev.cmax_cap = (c->params.ev_flags & DSEM_EV_CMAX_CAP) != 0;
    return ev;
}

// Adjoint of the entropy-viscosity update (k_ev) at state JQ and lagged
// states Slag: JQb += (d mu_h / d JQ)^T MuB, SlagB = (d mu_h / d Slag)^T MuB.
template <int P>
int run_ev_vjp(dsem_ctx *c, const double *JQ, const double *Slag, const double *MuB, double *JQb, double *SlagB)
{
    const DevMesh m = dev_mesh(c);
    const DevPhys ph = dev_phys(c);
    const EvArgs ev = ev_args(c);
    hipStream_t s = c->stream;
    return launch(c, KID_ADJ, [&] {
        k_adj_ev<P><<<c->K, BLOCK, 0, s>>>(JQ, Slag, m, ph, ev, MuB, JQb, c->aWB);
        k_adj_ev_gather<P><<<c->K, BLOCK, 0, s>>>(Slag, c->aWB, ph, SlagB);
    });
}

int ev_vjp_dispatch(dsem_ctx *c, const double *JQ, const double *Slag, const double *MuB, double *JQb, double *SlagB)
{
    DISPATCH_P(c, run_ev_vjp, JQ, Slag, MuB, JQb, SlagB);
}

// Adjoint of the Smagorinsky update (k_smag, with its sensor switches) at
// state JQ and lagged states Slag: JQb += (d nu_t / d JQ)^T NutB, SlagB +=
// (d nu_t / d Slag)^T NutB, nu_slot = <NutB, d nu_t / d nu> (nu_t is linear
// in Re).
template <int P>
int run_les_vjp(dsem_ctx *c, const double *JQ, const double *Slag, double cs, double numax, int sensors,
                const double *NutB, double *JQb, double *SlagB, double *nu_slot)
{
    const DevMesh m = dev_mesh(c);
    const DevPhys ph = dev_phys(c);
    hipStream_t s = c->stream;
    const int ev_on = c->params.shock ? 1 : 0;
    return launch(c, KID_ADJ, [&] {
        k_adj_smag<P><<<c->K, BLOCK, 0, s>>>(JQ, Slag, m, ph, cs, numax, sensors, ev_on, NutB, JQb, c->aWB,
                                             c->aNuPart);
        k_adj_smag_gather<P><<<c->K, BLOCK, 0, s>>>(c->aWB, SlagB, c->aNuPart, c->K, nu_slot);
    });
}

int les_vjp_dispatch(dsem_ctx *c, const double *JQ, const double *Slag, double cs, double numax, int sensors,
                     const double *NutB, double *JQb, double *SlagB, double *nu_slot)
{
    DISPATCH_P(c, run_les_vjp, JQ, Slag, cs, numax, sensors, NutB, JQb, SlagB, nu_slot);
}

// Tape layout per slot and stage: state [n], T [nT], QA [nQA], FV [nT].
size_t tape_stage_size(const dsem_ctx *c)
{
    const size_t n = (size_t)c->K * NEQ * c->np, nT = (size_t)c->K * 6 * NEQ * c->nf;
    const size_t nQA = (size_t)c->M * NEQ * c->nf;
    return n + 2 * nT + nQA;
}

// After a slot's five stages: with entropy viscosity the step's mu_h
// [K][3][ns]; with entropy viscosity or the Smagorinsky model the lagged
// states the step read [K][3][5][ns]; with the Smagorinsky model its nu_t
// [K][3][ns].
size_t tape_ev_size(const dsem_ctx *c)
{
    const size_t nm = (size_t)c->K * 3 * c->ns;
    return (c->params.shock ? nm : 0) + ((c->params.shock || c->tape_les) ? nm * NEQ : 0) + (c->tape_les ? nm : 0);
}

size_t tape_slot_size(const dsem_ctx *c)
{
    return 5 * tape_stage_size(c) + tape_ev_size(c);
}

void tape_ev(const dsem_ctx *c, int slot, double **muh, double **slag, double **nut)
{
    const size_t nm = (size_t)c->K * 3 * c->ns;
    double *b = c->tape + (size_t)slot * tape_slot_size(c) + 5 * tape_stage_size(c);
    *muh = c->params.shock ? b : nullptr;
    if (c->params.shock) b += nm;
    *slag = (c->params.shock || c->tape_les) ? b : nullptr;
    b += (c->params.shock || c->tape_les) ? nm * NEQ : 0;
    *nut = c->tape_les ? b : nullptr;
}

void tape_stage(const dsem_ctx *c, int slot, int k, double **S, FwdTape *f)
{
    const size_t n = (size_t)c->K * NEQ * c->np, nT = (size_t)c->K * 6 * NEQ * c->nf;
    const size_t nQA = (size_t)c->M * NEQ * c->nf;
    double *b = c->tape + (size_t)slot * tape_slot_size(c) + (size_t)k * tape_stage_size(c);
    *S = b;
    f->T = b + n;
    f->QA = b + n + nT;
    f->FV = b + n + nT + nQA;
}

int tape_reserve(dsem_ctx *c, int nslots)
{
    if (nslots <= c->tape_slots && c->tape_les == c->les) return DSEM_OK;
    nslots = std::max(nslots, c->tape_slots);   // a layout change keeps the slot count
    if (c->tape) {
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        HIP_TRY(c, hipFree(c->tape));
        for (auto &p : c->allocations)
            if (p == c->tape) p = nullptr;
        c->tape = nullptr;
        c->tape_slots = 0;
    }
    c->tape_les = c->les;
    TRY(dev_alloc(c, &c->tape, (size_t)nslots * tape_slot_size(c)));
    c->tape_slots = nslots;
    c->tape_lag.assign(nslots, 0);
    c->tape_cs.assign(nslots, 0.0);
    c->tape_numax.assign(nslots, 0.0);
    c->tape_sens.assign(nslots, 0);
    return DSEM_OK;
}


// One RK step of c->JQ (advanced in place) that records each stage's input
// state and forward intermediates into tape slot `slot`.
int record_step(dsem_ctx *c, double dt, int slot)
{
    const size_t n = (size_t)c->K * NEQ * c->np;
    double *T0 = c->T, *QA0 = c->QA, *FV0 = c->FV;
    int rc = DSEM_OK;
    if (c->les != c->tape_les)
        return set_err(c, DSEM_ERR_ARG, "the Smagorinsky switch changed since dsem_tape_reserve: reserve again");
    const bool lag_model = c->params.shock || c->les;
    c->tape_lag[slot] = 0;
    if (lag_model) {
        // the step's entropy viscosity (tstep's Entropy call), eddy viscosity
        // (its Smagorinsky call) and the lagged states they read, kept for the
        // adjoint
        double *tm, *ts, *tn;
        tape_ev(c, slot, &tm, &ts, &tn);
        c->tape_lag[slot] = (c->have_lag ? 1 : 0) | (c->les ? 2 : 0);
        c->tape_cs[slot] = c->les_cs;
        c->tape_numax[slot] = c->les_numax;
        c->tape_sens[slot] = c->les_sensors;
        if (c->have_lag)
            HIP_TRY(c, hipMemcpyAsync(ts, c->Slag, sizeof(double) * (size_t)c->K * 3 * NEQ * c->ns,
                                      hipMemcpyDeviceToDevice, c->stream));
        if (c->params.shock) {
            TRY(ev_step(c, false));
            HIP_TRY(c, hipMemcpyAsync(tm, c->muh, sizeof(double) * (size_t)c->K * 3 * c->ns,
                                      hipMemcpyDeviceToDevice, c->stream));
        }
        if (c->les) {
            TRY(les_step(c, false));
            HIP_TRY(c, hipMemcpyAsync(tn, c->nut, sizeof(double) * (size_t)c->K * 3 * c->ns,
                                      hipMemcpyDeviceToDevice, c->stream));
        }
    }
    for (int k = 0; k < 5 && rc == DSEM_OK; ++k) {
        double *S;
        FwdTape f;
        tape_stage(c, slot, k, &S, &f);
        hipError_t e = hipMemcpyAsync(S, c->JQ, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream);
        if (e != hipSuccess) { rc = set_err(c, DSEM_ERR_HIP, hipGetErrorString(e)); break; }
        StageArgs st;
        st.a = kArk[k];
        st.cdt = kCrk[k] * dt;
        st.first = (k == 0);
        st.rk = 1;
        st.resid = nullptr;
        st.nonfinite = nullptr;
        c->T = f.T;
        c->QA = f.QA;
        c->FV = f.FV;
        // the last stage leaves the next step's lagged states, as dsem_rk_steps does
        rc = rhs_dispatch(c, c->JQ, st, nullptr, false, false, (lag_model || c->keep_lag) && k == 4);
        c->T = T0;
        c->QA = QA0;
        c->FV = FV0;
    }
    if (rc == DSEM_OK && (lag_model || c->keep_lag)) c->have_lag = true;
    c->traces_cur = false;
    return rc;
}

// Adjoint of the step recorded in `slot`: dQbar (device) from the adjoint of
// the state after the step to that of the state before it; nu_host += the
// step's <., dR/dnu>.
// With entropy viscosity or the Smagorinsky model, dSlagbar (device,
// nullable) holds on entry the adjoint of the lagged states the step leaves
// (its last stage's face-averaged staggered states) and on return that of
// the lagged states the step's mu_h and nu_t were computed from; their own
// adjoints, accumulated over the five stages, go back through k_ev's and
// k_smag's adjoints into dQbar (both are functions of the step's start
// state) and dSlagbar.
int vjp_recorded_step(dsem_ctx *c, double dt, int slot, double *dQbar, double *nu_host, double *dSlagbar = nullptr)
{
    const size_t n = (size_t)c->K * NEQ * c->np, nm = (size_t)c->K * 3 * c->ns;
    const bool ev = c->params.shock;
    const bool les = (c->tape_lag[slot] & 2) != 0;
    const bool lagged = (c->tape_lag[slot] & 1) != 0;
    double *tm = nullptr, *ts = nullptr, *tn = nullptr;
    if (ev || les) tape_ev(c, slot, &tm, &ts, &tn);
    if (ev) HIP_TRY(c, hipMemsetAsync(c->aMuB, 0, sizeof(double) * nm, c->stream));
    if (les) HIP_TRY(c, hipMemsetAsync(c->aNutB, 0, sizeof(double) * nm, c->stream));
    HIP_TRY(c, hipMemsetAsync(c->aGb, 0, sizeof(double) * n, c->stream));
    for (int k = 4; k >= 0; --k) {
        double *S;
        FwdTape f;
        tape_stage(c, slot, k, &S, &f);
        TRY(rhs_vjp_dispatch(c, S, f, false, c->aGb, dQbar, kCrk[k] * dt, kArk[k], nullptr, c->aNu + 1 + k,
                             ev ? tm : nullptr, ev ? c->aMuB : nullptr, ((ev || les) && k == 4) ? dSlagbar : nullptr,
                             les ? tn : nullptr, les ? c->aNutB : nullptr));
    }
    if ((ev || les) && lagged) {
        double *S0;
        FwdTape f0;
        tape_stage(c, slot, 0, &S0, &f0);
        double *slb = dSlagbar ? dSlagbar : c->aSlagB;
        if (ev) TRY(ev_vjp_dispatch(c, S0, ts, c->aMuB, dQbar, slb));   // sets slb
        else HIP_TRY(c, hipMemsetAsync(slb, 0, sizeof(double) * nm * NEQ, c->stream));
        if (les)
            TRY(les_vjp_dispatch(c, S0, ts, c->tape_cs[slot], c->tape_numax[slot], c->tape_sens[slot], c->aNutB, dQbar,
                                 slb, c->aNu + 6));
    } else if (dSlagbar) {
        // a fresh start: mu_h = nu_t = 0 whatever the state
        HIP_TRY(c, hipMemsetAsync(dSlagbar, 0, sizeof(double) * nm * NEQ, c->stream));
    }
    double h[6];
    HIP_TRY(c, hipMemcpyAsync(h, c->aNu + 1, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (nu_host) *nu_host = h[4] + h[3] + h[2] + h[1] + h[0] + ((les && lagged) ? h[5] : 0.0);
    return DSEM_OK;
}

int upload_operators(dsem_ctx *c)
{
    double D[9][90] = {}, B[9][90] = {}, Mo[9][81] = {}, cx[9][10] = {};
    for (int p = 1; p <= 8; ++p) {
        dsem::Ops o;
        dsem::make_ops(p, o);
        std::memcpy(D[p], o.D, sizeof(o.D));
        std::memcpy(B[p], o.B, sizeof(o.B));
        std::memcpy(Mo[p], o.M, sizeof(o.M));
        std::memcpy(cx[p], o.cx, sizeof(o.cx));
    }
    HIP_TRY(c, hipMemcpyToSymbol(HIP_SYMBOL(c_D), D, sizeof(D)));
    HIP_TRY(c, hipMemcpyToSymbol(HIP_SYMBOL(c_B), B, sizeof(B)));
    HIP_TRY(c, hipMemcpyToSymbol(HIP_SYMBOL(c_M), Mo, sizeof(Mo)));
    HIP_TRY(c, hipMemcpyToSymbol(HIP_SYMBOL(c_cx), cx, sizeof(cx)));
    float Df[9][90] = {}, Bf[9][90] = {}, Mf[9][81] = {};
    double Dint[9][9] = {};
    for (int p = 1; p <= 8; ++p) {
        const int ncg = p + 1, nc = p + 2;
        for (int i = 0; i < 90; ++i) { Df[p][i] = (float)D[p][i]; Bf[p][i] = (float)B[p][i]; }
        for (int i = 0; i < 81; ++i) Mf[p][i] = (float)Mo[p][i];
        for (int k = 0; k < ncg; ++k) {
            double sum = 0.0;
            for (int j = 1; j < nc - 1; ++j) sum += D[p][k * nc + j];
            Dint[p][k] = sum;
        }
    }
    HIP_TRY(c, hipMemcpyToSymbol(HIP_SYMBOL(c_Df), Df, sizeof(Df)));
    HIP_TRY(c, hipMemcpyToSymbol(HIP_SYMBOL(c_Bf), Bf, sizeof(Bf)));
    HIP_TRY(c, hipMemcpyToSymbol(HIP_SYMBOL(c_Mf), Mf, sizeof(Mf)));
    HIP_TRY(c, hipMemcpyToSymbol(HIP_SYMBOL(c_Dint), Dint, sizeof(Dint)));
    double C2[9][81] = {}, C3[9][81] = {};
    const double pi = 3.141592653589793238462643;
    for (int p = 1; p <= 8; ++p) {
        const int N = p + 1;
        for (int k = 0; k < N; ++k)
            for (int j = 0; j < N; ++j) {
                C2[p][k * N + j] = 2.0 * cos(pi * (j + 0.5) * k / N);
                C3[p][k * N + j] = j == 0 ? 1.0 : 2.0 * cos(pi * j * (k + 0.5) / N);
            }
    }
    HIP_TRY(c, hipMemcpyToSymbol(HIP_SYMBOL(c_dct2), C2, sizeof(C2)));
    HIP_TRY(c, hipMemcpyToSymbol(HIP_SYMBOL(c_dct3), C3, sizeof(C3)));
    return DSEM_OK;
}

// Device address of `scratch` on the context's device, looked up once per
// device with that device current and shared by every context on it.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_dev[kMaxDevices] = {};

int scratch_setup(dsem_ctx *c)
{
    if (c->device < 0 || c->device >= kMaxDevices) return set_err(c, DSEM_ERR_ARG, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[c->device]) {
        void *ptr = nullptr;
        HIP_TRY(c, hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) return set_err(c, DSEM_ERR_HIP, "hipGetSymbolAddress(scratch) returned no address");
        g_scratch_dev[c->device] = static_cast<double *>(ptr);
    }
    c->scratch = g_scratch_dev[c->device];
    return DSEM_OK;
}

int create_impl(const dsem_mesh_desc *mesh, const dsem_params *prm, int device, dsem_ctx *c)
{
    if (!mesh || !prm) return set_err(c, DSEM_ERR_ARG, "null mesh or params");
    if (mesh->p < 1 || mesh->p > 8) return set_err(c, DSEM_ERR_ARG, "p must be in 1..8 (nc <= 10 as in size.f90)");
    if (mesh->K < 1 || mesh->nmortar < 1 || !mesh->mortar_elem || !mesh->mortar_face || !mesh->corners)
        return set_err(c, DSEM_ERR_ARG, "empty mesh");
    if (prm->precision != DSEM_PREC_FP64 && prm->precision != DSEM_PREC_FP32 && prm->precision != DSEM_PREC_BF16X3)
        return set_err(c, DSEM_ERR_ARG, "precision must be DSEM_PREC_FP64, _FP32 or _BF16X3");
    if (!(prm->gamma > 1.0) || !(prm->re > 0.0) || !(prm->pr > 0.0) || !(prm->mach > 0.0))
        return set_err(c, DSEM_ERR_ARG, "physics parameters must be positive (gamma > 1)");
    c->device = device;
    c->K = mesh->K; c->p = mesh->p; c->M = mesh->nmortar;
    c->ncg = c->p + 1; c->nc = c->p + 2;
    c->np = c->ncg * c->ncg * c->ncg; c->ns = c->nc * c->ncg * c->ncg; c->nf = c->ncg * c->ncg;
    c->params = *prm;
    // topology checks: every element face in exactly one mortar, orientation 0
    std::vector<int> e2m((size_t)c->K * 6, -1);
    std::vector<int> me(2 * (size_t)c->M), mf(2 * (size_t)c->M), msg(c->M, 1), mns(c->M, 1), mor(c->M, 0);
    if (c->M >= (1 << 27)) return set_err(c, DSEM_ERR_ARG, "too many mortars");
    for (int k = 0; k < c->M; ++k) {
        const int o = mesh->mortar_orient ? mesh->mortar_orient[k] : 0;
        if (o < 0 || o > 7) return set_err(c, DSEM_ERR_ARG, "mortar orientation must be 0..7");
        mor[k] = o;
        if (o) c->rotated = true;
        for (int s = 0; s < 2; ++s) {
            const int e = mesh->mortar_elem[2 * k + s], f = mesh->mortar_face[2 * k + s];
            if (e < 0 || e >= c->K || f < 1 || f > 6) return set_err(c, DSEM_ERR_ARG, "mortar element/face out of range");
            if (e2m[(size_t)e * 6 + f - 1] != -1) return set_err(c, DSEM_ERR_ARG, "element face in two mortars");
            // the slave reads mortar data through the inverse of its
            // orientation (rotate_flux_to_face): the same code with 5 and 7
            // exchanged
            const int map = s == 0 ? 0 : (o == 5 ? 7 : (o == 7 ? 5 : o));
            e2m[(size_t)e * 6 + f - 1] = (k << 4) | (map << 1) | s;
            me[2 * k + s] = e;
            mf[2 * k + s] = f;
        }
        if (mesh->mortar_sign) msg[k] = mesh->mortar_sign[k];
        if (mesh->mortar_nsign) mns[k] = mesh->mortar_nsign[k];
        if ((msg[k] != 1 && msg[k] != -1) || (mns[k] != 1 && mns[k] != -1))
            return set_err(c, DSEM_ERR_ARG, "mortar sign/nsign must be +-1");
    }
    for (size_t i = 0; i < e2m.size(); ++i)
        if (e2m[i] < 0) return set_err(c, DSEM_ERR_UNSUPPORTED, "open element face: boundary conditions are not part of this path");

    HIP_TRY(c, hipSetDevice(device));
    HIP_TRY(c, hipStreamCreate(&c->stream));
    TRY(scratch_setup(c));
    HIP_TRY(c, hipEventCreate(&c->ev0));
    HIP_TRY(c, hipEventCreate(&c->ev1));
    TRY(upload_operators(c));
    if (!dsem::make_ops(c->p, c->ops)) return set_err(c, DSEM_ERR_ARG, "operators");

    dsem::Geometry g;
    dsem::element_geometry(c->ops, c->K, mesh->corners, g);
    std::vector<double> nhat, norm;
    dsem::mortar_geometry(c->ops, c->M, mesh->corners, mesh->mortar_elem, mesh->mortar_face, nhat, norm);
    c->h_xg = g.xg;
    c->h_corners.assign(mesh->corners, mesh->corners + (size_t)c->K * 24);
    c->h_melem.assign(mesh->mortar_elem, mesh->mortar_elem + 2 * (size_t)c->M);
    c->h_mface.assign(mesh->mortar_face, mesh->mortar_face + 2 * (size_t)c->M);
    c->h_jacob = g.jacob;
    int mode = prm->metrics;
    if (mode == DSEM_METRICS_AUTO) mode = g.affine ? DSEM_METRICS_PER_ELEMENT : DSEM_METRICS_PER_POINT;
    if (mode == DSEM_METRICS_PER_ELEMENT && !g.affine)
        return set_err(c, DSEM_ERR_ARG, "per-element metrics need parallelepiped elements");
    c->ms = (mode == DSEM_METRICS_PER_ELEMENT) ? 0 : 1;
    if (c->ms == 0) {
        // keep the value at the first point of every element / mortar
        auto compact = [](std::vector<double> &v, size_t groups, size_t stride) {
            std::vector<double> o(groups);
            for (size_t i = 0; i < groups; ++i) o[i] = v[i * stride];
            v.swap(o);
        };
        compact(g.jacob, (size_t)c->K, c->np);
        compact(g.gmetg, (size_t)c->K * 9, c->np);
        for (int d = 0; d < 3; ++d) compact(g.gmet[d], (size_t)c->K * 3, c->ns);
        compact(nhat, (size_t)c->M * 3, c->nf);
        compact(norm, (size_t)c->M, c->nf);
    }
    TRY(upload(c, &c->jacob, g.jacob));
    TRY(upload(c, &c->gmetg, g.gmetg));
    for (int d = 0; d < 3; ++d) TRY(upload(c, &c->gmet[d], g.gmet[d]));
    TRY(upload(c, &c->nhat, nhat));
    TRY(upload(c, &c->norm, norm));
    TRY(upload(c, &c->e2m, e2m));
    TRY(upload(c, &c->m_elem, me));
    TRY(upload(c, &c->m_face, mf));
    TRY(upload(c, &c->m_sign, msg));
    TRY(upload(c, &c->m_nsign, mns));
    if (c->rotated) TRY(upload(c, &c->m_orient, mor));
    const size_t nst = (size_t)c->K * NEQ * c->np;
    TRY(dev_alloc(c, &c->JQ, nst));
    TRY(dev_alloc(c, &c->G, nst));
    TRY(dev_alloc(c, &c->Qtp, nst));
    TRY(dev_alloc(c, &c->Qt, nst));
    TRY(dev_alloc(c, &c->T, (size_t)c->K * 6 * NEQ * c->nf));
    TRY(dev_alloc(c, &c->FV, (size_t)c->K * 6 * NEQ * c->nf));
    TRY(dev_alloc(c, &c->QA, (size_t)c->M * NEQ * c->nf));
    TRY(dev_alloc(c, &c->FI, (size_t)c->M * NEQ * c->nf));
    TRY(dev_alloc(c, &c->FT, (size_t)c->M * NEQ * c->nf));
    TRY(dev_alloc(c, &c->muh, (size_t)c->K * 3 * c->ns));
    TRY(dev_alloc(c, &c->muhg, (size_t)c->K * c->np));
    TRY(upload(c, &c->bmat, dsem::ke_weights(c->ops)));
    TRY(dev_alloc(c, &c->diag_part, (size_t)c->K * 5));
    TRY(dev_alloc(c, &c->diag_out, 8));
    if (c->params.shock) {
        TRY(dev_alloc(c, &c->Slag, (size_t)c->K * 3 * NEQ * c->ns));
        TRY(dev_alloc(c, &c->ev_resid, (size_t)c->K * c->np));
        TRY(dev_alloc(c, &c->ev_ducros, (size_t)c->K * c->np));
        TRY(dev_alloc(c, &c->ev_rhos, (size_t)c->K * c->np));
    }
    TRY(dev_alloc(c, &c->d_scalars, 4));
    HIP_TRY(c, hipMemset(c->JQ, 0, sizeof(double) * nst));
    HIP_TRY(c, hipMemset(c->G, 0, sizeof(double) * nst));
    HIP_TRY(c, hipMemset(c->muh, 0, sizeof(double) * (size_t)c->K * 3 * c->ns));
    HIP_TRY(c, hipMemset(c->muhg, 0, sizeof(double) * (size_t)c->K * c->np));
    return DSEM_OK;
}

void pack_state(const dsem_ctx *c, const double *in, int layout, double *out)
{
    // layout 1 (reference padded 10^3 per variable) -> compact
    const int ncg = c->ncg, np = c->np;
    for (int e = 0; e < c->K; ++e)
        for (int v = 0; v < NEQ; ++v)
            for (int l = 0; l < ncg; ++l)
                for (int m = 0; m < ncg; ++m)
                    for (int n = 0; n < ncg; ++n) {
                        const size_t src = (((size_t)e * NEQ + v) * 10 + l) * 100 + (size_t)m * 10 + n;
                        const size_t dst = ((size_t)e * NEQ + v) * np + n + ncg * (m + ncg * l);
                        if (layout == 1) out[dst] = in[src];
                        else out[src] = in[dst];
                    }
}

}  // namespace

// ====================================================================== C ABI
extern "C" {

int dsem_create(const dsem_mesh_desc *mesh, const dsem_params *params, int device, dsem_ctx **out)
{
    if (!out) return DSEM_ERR_ARG;
    *out = nullptr;
    dsem_ctx *c = new (std::nothrow) dsem_ctx();
    if (!c) return DSEM_ERR_NOMEM;
    int rc = create_impl(mesh, params, device, c);
    *out = c;   // returned even on failure so that dsem_last_error can report; caller destroys
    return rc;
}

int dsem_destroy(dsem_ctx *c)
{
    if (!c) return DSEM_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (void *p : c->allocations) (void)hipFree(p);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return DSEM_OK;
}

const char *dsem_last_error(const dsem_ctx *c)
{
    return c ? c->err.c_str() : "null context";
}

int dsem_sizes(const dsem_ctx *c, int32_t *K, int32_t *p, int32_t *nmortar, int64_t *dof)
{
    if (!c) return DSEM_ERR_ARG;
    if (K) *K = c->K;
    if (p) *p = c->p;
    if (nmortar) *nmortar = c->M;
    if (dof) *dof = (int64_t)NEQ * c->K * c->np;
    return DSEM_OK;
}

int dsem_gauss_points(dsem_ctx *c, double *xg, double *jacob)
{
    if (!c || !c->stream) return DSEM_ERR_ARG;
    if (xg) std::memcpy(xg, c->h_xg.data(), sizeof(double) * c->h_xg.size());
    if (jacob) std::memcpy(jacob, c->h_jacob.data(), sizeof(double) * c->h_jacob.size());
    return DSEM_OK;
}

int dsem_points(const dsem_ctx *c, double *cx, double *cxg)
{
    if (!c || !c->stream) return DSEM_ERR_ARG;
    if (cx) std::memcpy(cx, c->ops.cx, sizeof(double) * c->nc);
    if (cxg) std::memcpy(cxg, c->ops.cxg, sizeof(double) * c->ncg);
    return DSEM_OK;
}

int dsem_operators(const dsem_ctx *c, double *dmx, double *bx)
{
    if (!c || !c->stream) return DSEM_ERR_ARG;
    if (dmx) std::memcpy(dmx, c->ops.D, sizeof(double) * c->ncg * c->nc);
    if (bx) std::memcpy(bx, c->ops.B, sizeof(double) * c->nc * c->ncg);
    return DSEM_OK;
}

int dsem_get_metrics(dsem_ctx *c, double *gmetg, double *gmet, double *nhat, double *norm)
{
    if (!c || !c->stream) return DSEM_ERR_ARG;
    // recomputed on the host from the corners: the same routine and values
    // as the per-point arrays on the device
    if (gmetg || gmet) {
        dsem::Geometry g;
        dsem::element_geometry(c->ops, c->K, c->h_corners.data(), g);
        const size_t np = c->np, ns = c->ns;
        for (size_t e = 0; e < (size_t)c->K; ++e) {
            if (gmetg)
                for (size_t pt = 0; pt < np; ++pt)
                    for (int i = 0; i < 9; ++i) gmetg[(e * np + pt) * 9 + i] = g.gmetg[(e * 9 + i) * np + pt];
            if (gmet)
                for (int d = 0; d < 3; ++d)
                    for (size_t s = 0; s < ns; ++s)
                        for (int i = 0; i < 3; ++i)
                            gmet[((e * 3 + d) * ns + s) * 3 + i] = g.gmet[d][(e * 3 + i) * ns + s];
        }
    }
    if (nhat || norm) {
        std::vector<double> nh, nr;
        dsem::mortar_geometry(c->ops, c->M, c->h_corners.data(), c->h_melem.data(), c->h_mface.data(), nh, nr);
        const size_t nf = c->nf;
        for (size_t mm = 0; mm < (size_t)c->M; ++mm)
            for (size_t q = 0; q < nf; ++q) {
                if (nhat)
                    for (int i = 0; i < 3; ++i) nhat[(mm * nf + q) * 3 + i] = nh[(mm * 3 + i) * nf + q];
                if (norm) norm[mm * nf + q] = nr[mm * nf + q];
            }
    }
    return DSEM_OK;
}

int dsem_set_state(dsem_ctx *c, const double *JQ, int layout)
{
    if (!c || !c->stream || !JQ || (layout != 0 && layout != 1)) return set_err(c, DSEM_ERR_ARG, "set_state arguments");
    const size_t n = (size_t)c->K * NEQ * c->np;
    std::vector<double> tmp;
    const double *src = JQ;
    if (layout == 1) {
        tmp.resize(n);
        pack_state(c, JQ, 1, tmp.data());
        src = tmp.data();
    }
    HIP_TRY(c, hipSetDevice(c->device));
    c->traces_cur = false;
    HIP_TRY(c, hipMemcpyAsync(c->JQ, src, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->have_lag = false;   // a new state starts without lagged staggered states
    return DSEM_OK;
}

int dsem_get_state(dsem_ctx *c, double *JQ, int layout)
{
    if (!c || !c->stream || !JQ || (layout != 0 && layout != 1)) return set_err(c, DSEM_ERR_ARG, "get_state arguments");
    const size_t n = (size_t)c->K * NEQ * c->np;
    HIP_TRY(c, hipSetDevice(c->device));
    if (layout == 0) {
        HIP_TRY(c, hipMemcpyAsync(JQ, c->JQ, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        return DSEM_OK;
    }
    std::vector<double> tmp(n);
    HIP_TRY(c, hipMemcpyAsync(tmp.data(), c->JQ, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    std::memset(JQ, 0, sizeof(double) * (size_t)c->K * NEQ * 1000);
    pack_state(c, tmp.data(), 0, JQ);
    return DSEM_OK;
}

int dsem_set_viscosity(dsem_ctx *c, const double *muh)
{
    if (!c || !c->stream || !muh) return set_err(c, DSEM_ERR_ARG, "set_viscosity arguments");
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipMemcpyAsync(c->muh, muh, sizeof(double) * (size_t)c->K * 3 * c->ns, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return DSEM_OK;
}

int dsem_get_viscosity(dsem_ctx *c, double *muhg, double *muh)
{
    if (!c || !c->stream) return set_err(c, DSEM_ERR_ARG, "get_viscosity arguments");
    HIP_TRY(c, hipSetDevice(c->device));
    if (muhg)
        HIP_TRY(c, hipMemcpyAsync(muhg, c->muhg, sizeof(double) * (size_t)c->K * c->np, hipMemcpyDeviceToHost, c->stream));
    if (muh)
        HIP_TRY(c, hipMemcpyAsync(muh, c->muh, sizeof(double) * (size_t)c->K * 3 * c->ns, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return DSEM_OK;
}

int dsem_ev_update(dsem_ctx *c, double *muhg, double *resid, double *ducros)
{
    if (!c || !c->stream) return set_err(c, DSEM_ERR_ARG, "ev_update arguments");
    if (!c->params.shock) return set_err(c, DSEM_ERR_ARG, "entropy viscosity needs params.shock = 1");
    HIP_TRY(c, hipSetDevice(c->device));
    TRY(ev_step(c, true));
    const size_t n = (size_t)c->K * c->np;
    if (muhg) HIP_TRY(c, hipMemcpyAsync(muhg, c->muhg, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    if (resid && c->have_lag) HIP_TRY(c, hipMemcpyAsync(resid, c->ev_resid, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    if (ducros && c->have_lag) HIP_TRY(c, hipMemcpyAsync(ducros, c->ev_ducros, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (!c->have_lag) {
        if (resid) std::memset(resid, 0, sizeof(double) * n);
        if (ducros) std::memset(ducros, 0, sizeof(double) * n);
    }
    return DSEM_OK;
}

int dsem_set_lagged_states(dsem_ctx *c, const double *stag)
{
    if (!c || !c->stream || !stag || !c->Slag) return set_err(c, DSEM_ERR_ARG, "lagged states need shock = 1 or dsem_keep_lagged_states");
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipMemcpyAsync(c->Slag, stag, sizeof(double) * (size_t)c->K * 3 * NEQ * c->ns, hipMemcpyHostToDevice,
                              c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->have_lag = true;
    return DSEM_OK;
}

int dsem_get_lagged_states(dsem_ctx *c, double *stag)
{
    if (!c || !c->stream || !stag || !c->Slag) return set_err(c, DSEM_ERR_ARG, "lagged states need shock = 1 or dsem_keep_lagged_states");
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipMemcpyAsync(stag, c->Slag, sizeof(double) * (size_t)c->K * 3 * NEQ * c->ns, hipMemcpyDeviceToHost,
                              c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return DSEM_OK;
}

int dsem_keep_lagged_states(dsem_ctx *c, int enable)
{
    if (!c || !c->stream) return set_err(c, DSEM_ERR_ARG, "keep_lagged_states arguments");
    HIP_TRY(c, hipSetDevice(c->device));
    if (enable && !c->Slag) TRY(dev_alloc(c, &c->Slag, (size_t)c->K * 3 * NEQ * c->ns));
    c->keep_lag = enable != 0;
    return DSEM_OK;
}

int dsem_set_les(dsem_ctx *c, int enable, double cs, double numax)
{
    if (!c || !c->stream) return set_err(c, DSEM_ERR_ARG, "set_les arguments");
    if (enable && !(cs >= 0.0 && numax >= 0.0)) return set_err(c, DSEM_ERR_ARG, "set_les: cs and numax must be >= 0");
    HIP_TRY(c, hipSetDevice(c->device));
    if (enable) {
        if (!c->Slag) TRY(dev_alloc(c, &c->Slag, (size_t)c->K * 3 * NEQ * c->ns));
        if (!c->nut) {
            TRY(dev_alloc(c, &c->nut, (size_t)c->K * 3 * c->ns));
            TRY(dev_alloc(c, &c->nut_g, (size_t)c->K * c->np));
            HIP_TRY(c, hipMemsetAsync(c->nut, 0, sizeof(double) * (size_t)c->K * 3 * c->ns, c->stream));
            HIP_TRY(c, hipMemsetAsync(c->nut_g, 0, sizeof(double) * (size_t)c->K * c->np, c->stream));
            HIP_TRY(c, hipStreamSynchronize(c->stream));
        }
        c->les_cs = cs;
        c->les_numax = numax;
    }
    c->les = enable != 0;
    return DSEM_OK;
}

int dsem_set_les_sensors(dsem_ctx *c, int sensors)
{
    if (!c || !c->stream || (sensors & ~(DSEM_LES_SHOCK_SENSOR | DSEM_LES_RHO_SENSOR)))
        return set_err(c, DSEM_ERR_ARG, "set_les_sensors arguments");
    HIP_TRY(c, hipSetDevice(c->device));
    if (sensors && !c->ev_ducros) {
        // without entropy viscosity nothing computes the sensors: they stay
        // zero, as the reference's never-written d%ducros and d%rhoSensor do
        TRY(dev_alloc(c, &c->ev_ducros, (size_t)c->K * c->np));
        TRY(dev_alloc(c, &c->ev_rhos, (size_t)c->K * c->np));
        HIP_TRY(c, hipMemsetAsync(c->ev_ducros, 0, sizeof(double) * (size_t)c->K * c->np, c->stream));
        HIP_TRY(c, hipMemsetAsync(c->ev_rhos, 0, sizeof(double) * (size_t)c->K * c->np, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
    }
    c->les_sensors = sensors;
    return DSEM_OK;
}

int dsem_get_ev_sensors(dsem_ctx *c, double *ducros, double *rho_sensor)
{
    if (!c || !c->stream || !c->ev_ducros) return set_err(c, DSEM_ERR_ARG, "get_ev_sensors: no sensors computed");
    HIP_TRY(c, hipSetDevice(c->device));
    const size_t n = (size_t)c->K * c->np;
    if (ducros) HIP_TRY(c, hipMemcpyAsync(ducros, c->ev_ducros, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    if (rho_sensor) HIP_TRY(c, hipMemcpyAsync(rho_sensor, c->ev_rhos, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return DSEM_OK;
}

int dsem_les_update(dsem_ctx *c, double *nut_g, double *nut_s)
{
    if (!c || !c->stream || !c->les) return set_err(c, DSEM_ERR_ARG, "les_update needs dsem_set_les(ctx, 1, ...)");
    HIP_TRY(c, hipSetDevice(c->device));
    TRY(les_step(c, true));
    if (nut_g) HIP_TRY(c, hipMemcpyAsync(nut_g, c->nut_g, sizeof(double) * (size_t)c->K * c->np, hipMemcpyDeviceToHost,
                                         c->stream));
    if (nut_s) HIP_TRY(c, hipMemcpyAsync(nut_s, c->nut, sizeof(double) * (size_t)c->K * 3 * c->ns,
                                         hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return DSEM_OK;
}

int dsem_set_les_viscosity(dsem_ctx *c, const double *nut_s)
{
    if (!c || !c->stream || !c->les || !nut_s) return set_err(c, DSEM_ERR_ARG, "set_les_viscosity needs the model on");
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipMemcpyAsync(c->nut, nut_s, sizeof(double) * (size_t)c->K * 3 * c->ns, hipMemcpyHostToDevice,
                              c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return DSEM_OK;
}

int dsem_reduce_ke(dsem_ctx *c, double *Ek, double *eps)
{
    if (!c || !c->stream || (!Ek && !eps)) return set_err(c, DSEM_ERR_ARG, "reduce_ke arguments");
    if (eps && !(c->Slag && c->have_lag))
        return set_err(c, DSEM_ERR_ARG, "the enstrophy dissipation needs the lagged staggered states of a step taken "
                                        "with shock = 1 or dsem_keep_lagged_states");
    HIP_TRY(c, hipSetDevice(c->device));
    TRY(diag_dispatch(c, eps != nullptr));
    double h[5];
    HIP_TRY(c, hipMemcpyAsync(h, c->diag_out, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    const double pi = 3.141592653589793238462643;
    if (Ek) {
        const double j0 = c->h_jacob[0];
        *Ek = 0.5 * (h[0] / j0 + h[1] / j0 + h[2] / j0) / (8.0 * (pi * pi * pi));
    }
    if (eps) *eps = 100.0 * (h[3] / h[4]) * (1.0 / c->params.re);
    return DSEM_OK;
}

int dsem_rhs_device(dsem_ctx *c, const double *dJQ, double *dQt)
{
    if (!c || !c->stream || !dJQ || !dQt) return set_err(c, DSEM_ERR_ARG, "rhs_device arguments");
    HIP_TRY(c, hipSetDevice(c->device));
    StageArgs st{0.0, 0.0, 0, 0, nullptr, nullptr};
    return rhs_dispatch(c, const_cast<double *>(dJQ), st, dQt);
}

int dsem_rhs_device_noalloc(dsem_ctx *c, const double *dJQ, double *dQt)
{
    if (!c || !c->stream || !dJQ || !dQt) return set_err(c, DSEM_ERR_ARG, "rhs_device_noalloc arguments");
    if (!c->scratch) return set_err(c, DSEM_ERR_ARG, "no scratch address for this device");
    // T, FV [K][6][5][nf]; QA, FI, FT [M][5][nf]; Qtp [K][5][np]; each part
    // starts on a 256-byte boundary
    size_t used = 0;
    auto take = [&](size_t n) {
        double *p = c->scratch + used;
        used += (n + 31) / 32 * 32;
        return p;
    };
    const size_t nface = (size_t)c->K * 6 * NEQ * c->nf, nmort = (size_t)c->M * NEQ * c->nf;
    RhsBufs b;
    b.T = take(nface);
    b.FV = take(nface);
    b.QA = take(nmort);
    b.FI = take(nmort);
    b.FT = take(nmort);
    b.Qtp = take((size_t)c->K * NEQ * c->np);
    if (used * sizeof(double) > DSEM_NOALLOC_SCRATCH_BYTES)
        return set_err(c, DSEM_ERR_ARG, "rhs_device_noalloc: this mesh needs " + std::to_string(used * sizeof(double)) +
                                            " bytes of scratch, more than the " +
                                            std::to_string((unsigned long long)DSEM_NOALLOC_SCRATCH_BYTES) +
                                            " bytes of the static array (use dsem_rhs_device)");
    HIP_TRY(c, hipSetDevice(c->device));
    StageArgs st{0.0, 0.0, 0, 0, nullptr, nullptr};
    return rhs_dispatch(c, const_cast<double *>(dJQ), st, dQt, false, false, false, &b);
}

int dsem_rhs(dsem_ctx *c, double t, double *Qt_out)
{
    (void)t;   // no time-dependent source terms on this path (source terms = none)
    if (!c || !c->stream) return set_err(c, DSEM_ERR_ARG, "rhs arguments");
    HIP_TRY(c, hipSetDevice(c->device));
    StageArgs st{0.0, 0.0, 0, 0, nullptr, nullptr};
    TRY(rhs_dispatch(c, c->JQ, st, c->Qt));
    if (Qt_out) {
        HIP_TRY(c, hipMemcpyAsync(Qt_out, c->Qt, sizeof(double) * (size_t)c->K * NEQ * c->np, hipMemcpyDeviceToHost,
                                  c->stream));
    }
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return DSEM_OK;
}

int dsem_rk_steps(dsem_ctx *c, double t, double dt, int nsteps, double *resid_out)
{
    (void)t;
    if (!c || !c->stream || nsteps < 1) return set_err(c, DSEM_ERR_ARG, "rk_steps arguments");
    HIP_TRY(c, hipSetDevice(c->device));
    const bool check = resid_out != nullptr;
    if (check) HIP_TRY(c, hipMemsetAsync(c->d_scalars, 0, sizeof(unsigned long long) * 4, c->stream));
    for (int s = 0; s < nsteps; ++s) {
        TRY(ev_step(c, false));
        TRY(les_step(c, false));
        for (int k = 0; k < 5; ++k) {
            StageArgs st;
            st.a = kArk[k];
            st.cdt = kCrk[k] * dt;
            st.first = (k == 0);
            st.rk = 1;
            const bool last = check && s == nsteps - 1 && k == 4;
            st.resid = last ? &c->d_scalars[0] : nullptr;
            st.nonfinite = check ? reinterpret_cast<unsigned int *>(&c->d_scalars[2]) : nullptr;
            // The first stage needs the traces of the state: left in T by the
            // previous call's last stage unless the state or T changed since.
            // Every stage's update, the last included, leaves the traces of the
            // updated state for the next stage or call.
            const bool ready = (s == 0 && k == 0) ? c->traces_cur : true;
            TRY(rhs_dispatch(c, c->JQ, st, nullptr, ready, true, (c->params.shock || c->keep_lag || c->les) && k == 4));
        }
        if (c->params.shock || c->keep_lag || c->les) c->have_lag = true;
    }
    c->traces_cur = true;
    if (check) {
        unsigned long long h[4];
        HIP_TRY(c, hipMemcpyAsync(h, c->d_scalars, sizeof(h), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        double r;
        std::memcpy(&r, &h[0], sizeof(double));
        *resid_out = r;
        if (h[2] & 0xffffffffull) return set_err(c, DSEM_ERR_NONFINITE, "NaN or Inf in the residual");
    } else {
        HIP_TRY(c, hipStreamSynchronize(c->stream));
    }
    return DSEM_OK;
}

int dsem_rk_step(dsem_ctx *c, double t, double dt, double *resid_out)
{
    return dsem_rk_steps(c, t, dt, 1, resid_out);
}

// allow_lagged: the tape entry points that carry the adjoint of the lagged
// states (entropy viscosity and the Smagorinsky model read them).
static int vjp_check(dsem_ctx *c, bool allow_lagged = false)
{
    if (!c || !c->stream) return set_err(c, DSEM_ERR_ARG, "vjp arguments");
    if ((c->params.shock || c->les) && !allow_lagged)
        return set_err(c, DSEM_ERR_UNSUPPORTED,
                       "with entropy viscosity or the Smagorinsky model use the tape entry points "
                       "(dsem_rk_step_record, dsem_rk_step_vjp_tape_lagged) or dsem_ev_vjp");
    if (c->params.precision != DSEM_PREC_FP64)
        return set_err(c, DSEM_ERR_UNSUPPORTED, "the adjoint is built for the FP64 path");
    HIP_TRY(c, hipSetDevice(c->device));
    TRY(adjoint_alloc(c));
    return adjoint_lag_alloc(c);
}

int dsem_rhs_vjp_device(dsem_ctx *c, const double *dJQ, const double *dLam, double *dJQbar, double *nubar)
{
    TRY(vjp_check(c));
    if (!dJQ || !dLam || !dJQbar) return set_err(c, DSEM_ERR_ARG, "rhs_vjp_device pointers");
    TRY(rhs_vjp_dispatch(c, const_cast<double *>(dJQ), FwdTape{nullptr, nullptr, nullptr}, true,
                         const_cast<double *>(dLam), nullptr, 0.0, 0.0, dJQbar, c->aNu));
    double h = 0.0;
    HIP_TRY(c, hipMemcpyAsync(&h, c->aNu, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (nubar) *nubar = h;
    return DSEM_OK;
}

int dsem_rhs_vjp(dsem_ctx *c, const double *lam, double *JQbar, double *nubar)
{
    TRY(vjp_check(c));
    if (!lam || !JQbar) return set_err(c, DSEM_ERR_ARG, "rhs_vjp pointers");
    const size_t n = (size_t)c->K * NEQ * c->np;
    HIP_TRY(c, hipMemcpyAsync(c->aLam, lam, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    TRY(dsem_rhs_vjp_device(c, c->JQ, c->aLam, c->aOut, nubar));
    HIP_TRY(c, hipMemcpyAsync(JQbar, c->aOut, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return DSEM_OK;
}

int dsem_rk_step_vjp_device(dsem_ctx *c, double dt, double *dQbar, double *nubar)
{
    TRY(vjp_check(c));
    if (!dQbar) return set_err(c, DSEM_ERR_ARG, "rk_step_vjp_device pointer");
    TRY(tape_reserve(c, 1));
    TRY(record_step(c, dt, 0));
    double *S0;
    FwdTape f;
    tape_stage(c, 0, 0, &S0, &f);
    c->traces_cur = false;
    HIP_TRY(c, hipMemcpyAsync(c->JQ, S0, sizeof(double) * (size_t)c->K * NEQ * c->np, hipMemcpyDeviceToDevice,
                              c->stream));   // the start state stays the context's state
    return vjp_recorded_step(c, dt, 0, dQbar, nubar);
}

int dsem_tape_reserve(dsem_ctx *c, int nslots)
{
    TRY(vjp_check(c, true));
    if (nslots < 1) return set_err(c, DSEM_ERR_ARG, "tape_reserve: nslots >= 1");
    return tape_reserve(c, nslots);
}

int dsem_rk_step_record(dsem_ctx *c, double dt, int slot)
{
    TRY(vjp_check(c, true));
    if (slot < 0 || slot >= c->tape_slots) return set_err(c, DSEM_ERR_ARG, "rk_step_record: slot out of range");
    TRY(record_step(c, dt, slot));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return DSEM_OK;
}

int dsem_rk_step_vjp_tape(dsem_ctx *c, double dt, int slot, double *dQbar, double *nubar)
{
    TRY(vjp_check(c));
    if (slot < 0 || slot >= c->tape_slots || !dQbar) return set_err(c, DSEM_ERR_ARG, "rk_step_vjp_tape arguments");
    if (c->tape_lag[slot] & 2)
        return set_err(c, DSEM_ERR_UNSUPPORTED, "slot recorded with the Smagorinsky model: use dsem_rk_step_vjp_tape_lagged");
    return vjp_recorded_step(c, dt, slot, dQbar, nubar);
}

int dsem_rk_step_vjp_tape_lagged(dsem_ctx *c, double dt, int slot, double *dQbar, double *dSlagbar, double *nubar)
{
    TRY(vjp_check(c, true));
    if (slot < 0 || slot >= c->tape_slots || !dQbar) return set_err(c, DSEM_ERR_ARG, "rk_step_vjp_tape_lagged arguments");
    const bool lag_model = c->params.shock || (c->tape_lag[slot] & 2);
    if (lag_model && !dSlagbar)
        return set_err(c, DSEM_ERR_ARG, "rk_step_vjp_tape_lagged: dSlagbar needed with shock = 1 or the Smagorinsky model");
    return vjp_recorded_step(c, dt, slot, dQbar, nubar, lag_model ? dSlagbar : nullptr);
}

int dsem_ev_vjp(dsem_ctx *c, const double *muhbar, double *JQbar, double *Slagbar)
{
    TRY(vjp_check(c, true));
    if (!c->params.shock || !c->have_lag) return set_err(c, DSEM_ERR_ARG, "ev_vjp needs shock = 1 and lagged states");
    if (!muhbar || !JQbar || !Slagbar) return set_err(c, DSEM_ERR_ARG, "ev_vjp pointers");
    const size_t n = (size_t)c->K * NEQ * c->np, nm = (size_t)c->K * 3 * c->ns, nl = nm * NEQ;
    HIP_TRY(c, hipMemcpyAsync(c->aMuB, muhbar, sizeof(double) * nm, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemsetAsync(c->aOut, 0, sizeof(double) * n, c->stream));
    TRY(ev_vjp_dispatch(c, c->JQ, c->Slag, c->aMuB, c->aOut, c->aSlagB));
    HIP_TRY(c, hipMemcpyAsync(JQbar, c->aOut, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipMemcpyAsync(Slagbar, c->aSlagB, sizeof(double) * nl, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return DSEM_OK;
}

int dsem_get_lagged_states_device(dsem_ctx *c, double *dstag, int32_t *valid)
{
    if (!c || !c->stream || !c->Slag) return set_err(c, DSEM_ERR_ARG, "lagged states need shock = 1 or dsem_keep_lagged_states");
    HIP_TRY(c, hipSetDevice(c->device));
    if (valid) *valid = c->have_lag ? 1 : 0;
    if (dstag && c->have_lag)
        HIP_TRY(c, hipMemcpyAsync(dstag, c->Slag, sizeof(double) * (size_t)c->K * 3 * NEQ * c->ns,
                                  hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return DSEM_OK;
}

int dsem_set_lagged_states_device(dsem_ctx *c, const double *dstag)
{
    if (!c || !c->stream || !c->Slag) return set_err(c, DSEM_ERR_ARG, "lagged states need shock = 1 or dsem_keep_lagged_states");
    HIP_TRY(c, hipSetDevice(c->device));
    if (!dstag) {
        c->have_lag = false;   // as at a fresh start: the next step's mu_h is zero
        return DSEM_OK;
    }
    HIP_TRY(c, hipMemcpyAsync(c->Slag, dstag, sizeof(double) * (size_t)c->K * 3 * NEQ * c->ns,
                              hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->have_lag = true;
    return DSEM_OK;
}

int dsem_rk_step_vjp(dsem_ctx *c, double dt, double *Qbar, double *nubar)
{
    TRY(vjp_check(c));
    if (!Qbar) return set_err(c, DSEM_ERR_ARG, "rk_step_vjp pointer");
    const size_t n = (size_t)c->K * NEQ * c->np;
    HIP_TRY(c, hipMemcpyAsync(c->aOut, Qbar, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    TRY(dsem_rk_step_vjp_device(c, dt, c->aOut, nubar));
    HIP_TRY(c, hipMemcpyAsync(Qbar, c->aOut, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return DSEM_OK;
}

int dsem_modal_filter(dsem_ctx *c, int cutoff)
{
    if (!c || !c->stream) return set_err(c, DSEM_ERR_ARG, "modal_filter arguments");
    if (cutoff < 0 || cutoff > c->ncg) return set_err(c, DSEM_ERR_ARG, "modal_filter: cutoff must be in 0..p+1");
    HIP_TRY(c, hipSetDevice(c->device));
    c->traces_cur = false;   // the state changes: the next step recomputes its face traces
    TRY(modal_dispatch(c, cutoff));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return DSEM_OK;
}

int dsem_set_state_device(dsem_ctx *c, const double *dJQ)
{
    if (!c || !c->stream || !dJQ) return set_err(c, DSEM_ERR_ARG, "set_state_device arguments");
    HIP_TRY(c, hipSetDevice(c->device));
    c->traces_cur = false;
    HIP_TRY(c, hipMemcpyAsync(c->JQ, dJQ, sizeof(double) * (size_t)c->K * NEQ * c->np, hipMemcpyDeviceToDevice,
                              c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->have_lag = false;
    return DSEM_OK;
}

int dsem_get_state_device(dsem_ctx *c, double *dJQ)
{
    if (!c || !c->stream || !dJQ) return set_err(c, DSEM_ERR_ARG, "get_state_device arguments");
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipMemcpyAsync(dJQ, c->JQ, sizeof(double) * (size_t)c->K * NEQ * c->np, hipMemcpyDeviceToDevice,
                              c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return DSEM_OK;
}

int dsem_quadrature_weights(const dsem_ctx *c, double *bmat)
{
    if (!c || !bmat) return DSEM_ERR_ARG;
    const std::vector<double> w = dsem::ke_weights(c->ops);
    std::memcpy(bmat, w.data(), sizeof(double) * w.size());
    return DSEM_OK;
}

int dsem_dtcalc(dsem_ctx *c, double *dt_out)
{
    if (!c || !c->stream || !dt_out) return set_err(c, DSEM_ERR_ARG, "dtcalc arguments");
    HIP_TRY(c, hipSetDevice(c->device));
    const double big = 1.0e308;
    unsigned long long init;
    std::memcpy(&init, &big, sizeof(init));
    HIP_TRY(c, hipMemcpyAsync(&c->d_scalars[1], &init, sizeof(init), hipMemcpyHostToDevice, c->stream));
    TRY(dtcalc_dispatch(c, &c->d_scalars[1]));
    unsigned long long h;
    HIP_TRY(c, hipMemcpyAsync(&h, &c->d_scalars[1], sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    std::memcpy(dt_out, &h, sizeof(double));
    return DSEM_OK;
}

int dsem_state_pointers(dsem_ctx *c, double **dJQ, double **dG)
{
    if (!c) return DSEM_ERR_ARG;
    if (dJQ) *dJQ = c->JQ;
    if (dG) *dG = c->G;
    return DSEM_OK;
}

int dsem_stream(dsem_ctx *c, void **s)
{
    if (!c || !s) return DSEM_ERR_ARG;
    *s = c->stream;
    return DSEM_OK;
}

int dsem_synchronize(dsem_ctx *c)
{
    if (!c || !c->stream) return DSEM_ERR_ARG;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return DSEM_OK;
}

int dsem_timing_enable(dsem_ctx *c, int enable)
{
    if (!c) return DSEM_ERR_ARG;
    c->timing = enable != 0;
    return DSEM_OK;
}

int dsem_timing_reset(dsem_ctx *c)
{
    if (!c) return DSEM_ERR_ARG;
    for (int i = 0; i < KID_COUNT; ++i) { c->t_ms[i] = 0.0; c->t_n[i] = 0; }
    return DSEM_OK;
}

int dsem_timing_count(void) { return KID_COUNT; }

int dsem_timing_query(const dsem_ctx *c, int kid, const char **name, double *total_ms, int64_t *launches)
{
    if (!c || kid < 0 || kid >= KID_COUNT) return DSEM_ERR_ARG;
    if (name) *name = kKernelNames[kid];
    if (total_ms) *total_ms = c->t_ms[kid];
    if (launches) *launches = c->t_n[kid];
    return DSEM_OK;
}

}  // extern "C"
