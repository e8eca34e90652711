/*
 * hexat_oracle.h -- CPU FP64 restatement of the HEXAT topology-optimisation
 * hot path.  TEST INFRASTRUCTURE ONLY: this library is what the HIP kernels
 * are compared against (tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py).  The product never links or calls it.
 *
 * Conventions (0-based; the reference is 1-based Julia):
 *   grid of nx*ny*nz unit-box elements of size dx*dy*dz,
 *   node(i,j,k) = i + NX*(j + NY*k), NX = nx+1, NY = ny+1      (Mesh.jl:13-15)
 *   elem(i,j,k) = i + nx*(j + ny*k)                             (Mesh.jl:44-54)
 *   element-local node order n1..n8 = (0,0,0),(1,0,0),(1,1,0),(0,1,0),
 *                                     (0,0,1),(1,0,1),(1,1,1),(0,1,1)
 *                                                     (Element.jl:8-17)
 *   DOF = 3*node + d, node-major                    (GPUSolver.jl:50-57)
 *   Ke is 24x24 row-major (symmetric).
 */
#ifndef HEXAT_ORACLE_H
#define HEXAT_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int nx, ny, nz;
    double dx, dy, dz;
} ox_grid;

/* Element.jl:45-62 (Voigt order xx,yy,zz,xy,yz,xz), row-major 6x6 */
void ox_material_matrix(double E, double nu, double* D);
/* Element.jl:64-102: 2x2x2 Gauss H8 stiffness for a dx*dy*dz box */
void ox_ke_hex(double dx, double dy, double dz, double nu, double E, double* Ke);
/* Element.jl:104-125: scalar Laplacian and consistent mass.  quirk=1
 * reproduces the reference's top-face z=+dz (element 1.5*dz tall). */
void ox_scalar_ke_me(double dx, double dy, double dz, int quirk, double* Ks, double* Ms);

/* GPUSolver.jl:37-79 / CPUSolver.jl:50-94: y = sum_e Ee[e] * Ke x_e */
void ox_matvec(const ox_grid* g, const double* Ke, const double* Ee,
               const double* x, double* y);
/* GPUSolver.jl:301-315: diag = sum_e Ee[e] * diag(Ke) */
void ox_diag(const ox_grid* g, const double* Ke, const double* Ee, double* diag);

/* GPUSolver.jl:274-609 (CG loop :452-526) without the retry/AMG branches:
 * Jacobi-PCG on (K_ff + shift*I), free DOFs are those with fixed[d]==0.
 * M^-1 = 1/(diag + shift) (GPUSolver.jl:402).  x holds the warm start on
 * entry (fixed DOFs ignored) and the solution on exit (fixed DOFs = 0).
 * Convergence test ||r||/||b|| < tol after the r update (:472-511).
 * Returns 0 converged, 1 max_iter, 2 breakdown, 3 zero rhs. */
int ox_pcg(const ox_grid* g, const double* Ke, const double* Ee,
           const uint8_t* fixed, const double* b, double* x,
           double shift, double tol, int max_iter,
           double* rel_out, int* iters_out);

/* GPUHelmholtz.jl:39-102 & :165-229: nodal Helmholtz filter solve
 * (R^2 Ks + Ms) xn = b, b_n = sum_e rho_e*dx*dy*dz/8, out_e = mean of nodes.
 * Jacobi-PCG from x0 = 0; the residual is tested every check_every
 * iterations (reference: 10) up to max_iter (reference: 200). */
int ox_helmholtz(const ox_grid* g, double R, int quirk, const double* rho,
                 double* out, double tol, int max_iter, int check_every,
                 double* rel_out, int* iters_out);
/* nodal Helmholtz operator y = sum_e Ae x_e, Ae = R^2 Ks + Ms */
void ox_helm_matvec(const ox_grid* g, const double* Ae, const double* x, double* y);

/* Stress.jl:10-100: centroid stress of E_e-scaled material (E_e = E*rho_e),
 * zero for elements with rho_e <= 1e-6 (Stress.jl:85).  Any output may be
 * NULL.  principal (3/elem, descending), vm, voigt (6/elem), l1, dir (3/elem,
 * eigenvector of the largest principal stress, unit length, sign arbitrary) */
void ox_stress(const ox_grid* g, double nu, double E, const double* rho,
               const double* u, double* principal, double* vm, double* voigt,
               double* l1, double* dir);

/* ce_e = u_e^T Ke u_e (unit-E element energy) */
void ox_element_energy(const ox_grid* g, const double* Ke, const double* u, double* ce);

/* Jacobi eigen-decomposition of a symmetric 3x3 (row-major); values sorted
 * descending, vecs columns (row-major 3x3, column j = vector j). */
void ox_eig3(const double* A, double* vals, double* vecs);

/* OC update (Andreassen et al. 2011, top88 lines 55-64): bisection on the
 * Lagrange multiplier, new x_e = clamp(x*sqrt(-dc/(dv*lmid)), x-move, x+move)
 * within [xmin,1]; volume tested on mean(x_new) (mass-preserving filter).
 * Stops when (l2-l1)/(l1+l2) <= rel_tol.  Returns the number of steps. */
int ox_oc(int n, const double* x, const double* dc, const double* dv,
          double volfrac, double move, double xmin, double rel_tol,
          double* xnew, double* lmid_out);

/* FP32 mirror of the reference's GPU solver arithmetic (GPUSolver.jl runs in
 * Float32): Float32 storage, EbE products and dot products. */
void ox_matvec_f32(const ox_grid* g, const float* Ke, const float* Ee, const float* x, float* y);
int ox_pcg_f32(const ox_grid* g, const float* Ke, const float* Ee, const uint8_t* fixed, const float* b,
               float* x, float shift, double tol, int max_iter, double* rel_out, int* iters_out);

#ifdef __cplusplus
}
#endif
#endif
