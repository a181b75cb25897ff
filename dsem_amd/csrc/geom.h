// Host-side set-up of the DSEM operators and hex8 geometry (run once per
// context).  Restates the reference set-up routines in FP64 with their
// operation order; see geom.cpp for the citations.
#pragma once
#include <cstdint>
#include <vector>

namespace dsem {

struct Ops {
    int p = 0, ncg = 0, nc = 0;
    double cx[10] = {}, cxg[9] = {};
    double D[9 * 10] = {};   // D[k*nc + j]: derivative at Gauss k of Lobatto basis j
    double B[10 * 9] = {};   // B[j*ncg + k]: Gauss basis k at Lobatto point j
    double M[9 * 9] = {};    // M[k*ncg + k2] = sum_{interior j} D[k][j] B[j][k2]
};

bool make_ops(int p, Ops &o);

// Per-point metrics of K hex8 elements (layouts documented in geom.cpp).
struct Geometry {
    int K = 0, p = 0;
    std::vector<double> xg;      // [K][np][3]
    std::vector<double> jacob;   // [K][np]
    std::vector<double> gmetg;   // [K][9][np]   component i*3+j = d xi_i / d x_j
    std::vector<double> gmet[3]; // [K][3][ns]   J * grad xi_d on staggered grid d
    bool affine = false;         // every element a parallelepiped
};

bool element_geometry(const Ops &o, int K, const double *corners, Geometry &g);

// Mortar normals from the master face: nhat [M][3][nf], norm [M][nf].
void mortar_geometry(const Ops &o, int M, const double *corners, const int32_t *m_elem,
                     const int32_t *m_face, std::vector<double> &nhat, std::vector<double> &norm);

// Tensor quadrature weights of the Gauss-point interpolant on [0,1]^3
// (bmatu1b of the TGV statistics), [ncg^3], index i + ncg*(j + ncg*l).
std::vector<double> ke_weights(const Ops &o);

}  // namespace dsem
