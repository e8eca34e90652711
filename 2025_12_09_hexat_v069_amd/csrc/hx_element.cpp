// hx_element.cpp -- element matrices of the structured box mesh (host, FP64).
//
// Element.jl:64-102 builds Ke with B^T D B over 2x2x2 Gauss points for an
// arbitrary hexahedron.  Every element of the structured grid is the same
// dx*dy*dz box, so here the Jacobian is diagonal and the integrand is written
// in Lame form,  a(u,v) = lam div u div v + 2 mu eps(u):eps(v),  which gives
// the same matrix (tests/test_element_host.py pins it against the oracle and
// against exact integration).
#include <cmath>
#include <cstring>

namespace hx {

// reference node order n1..n8 (Element.jl:8-17)
static const int kNat[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                               {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};

// shape functions and physical gradients of a box of size h at natural point q
static void box_shape(const double h[3], const double q[3], double N[8], double G[8][3]) {
    for (int a = 0; a < 8; ++a) {
        double f[3], df[3];
        for (int d = 0; d < 3; ++d) {
            f[d] = 0.5 * (1.0 + kNat[a][d] * q[d]);
            df[d] = 0.5 * kNat[a][d] * 2.0 / h[d];  // d/dx of f = (1/2) s * (2/h)
        }
        N[a] = f[0] * f[1] * f[2];
        G[a][0] = df[0] * f[1] * f[2];
        G[a][1] = f[0] * df[1] * f[2];
        G[a][2] = f[0] * f[1] * df[2];
    }
}

void element_stiffness(double dx, double dy, double dz, double nu, double E, double* Ke) {
    const double lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
    const double mu = E / (2.0 * (1.0 + nu));
    const double h[3] = {dx, dy, dz};
    const double w = dx * dy * dz / 8.0;  // det J * Gauss weight 1
    const double g = 1.0 / std::sqrt(3.0);
    std::memset(Ke, 0, 576 * sizeof(double));
    for (int p = 0; p < 8; ++p) {
        const double q[3] = {(p & 1) ? g : -g, (p & 2) ? g : -g, (p & 4) ? g : -g};
        double N[8], G[8][3];
        box_shape(h, q, N, G);
        for (int a = 0; a < 8; ++a)
            for (int b = 0; b < 8; ++b) {
                const double gg = G[a][0] * G[b][0] + G[a][1] * G[b][1] + G[a][2] * G[b][2];
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) {
                        double v = lam * G[a][i] * G[b][j] + mu * G[a][j] * G[b][i];
                        if (i == j) v += mu * gg;
                        Ke[24 * (3 * a + i) + 3 * b + j] += w * v;
                    }
            }
    }
}

void scalar_matrices(double dx, double dy, double dz, int quirk, double* Ks, double* Ms) {
    // quirk: the reference's scalar element puts its top face at z=+dz
    // (Element.jl:107), i.e. a box 1.5*dz tall.
    const double h[3] = {dx, dy, quirk ? 1.5 * dz : dz};
    const double w = h[0] * h[1] * h[2] / 8.0;
    const double g = 1.0 / std::sqrt(3.0);
    std::memset(Ks, 0, 64 * sizeof(double));
    std::memset(Ms, 0, 64 * sizeof(double));
    for (int p = 0; p < 8; ++p) {
        const double q[3] = {(p & 1) ? g : -g, (p & 2) ? g : -g, (p & 4) ? g : -g};
        double N[8], G[8][3];
        box_shape(h, q, N, G);
        for (int a = 0; a < 8; ++a)
            for (int b = 0; b < 8; ++b) {
                Ks[8 * a + b] += w * (G[a][0] * G[b][0] + G[a][1] * G[b][1] + G[a][2] * G[b][2]);
                Ms[8 * a + b] += w * N[a] * N[b];
            }
    }
}

}  // namespace hx
