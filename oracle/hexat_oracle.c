/*
 * hexat_oracle.c -- CPU FP64 restatement of the HEXAT hot path.
 *
 * TEST INFRASTRUCTURE: the HIP kernels are compared against it (used only by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg).  It
 * follows the reference Julia files cited per function, in FP64 (the
 * reference computes in FP32).  OpenMP parallelism uses static partitioning
 * and an 8-colour element schedule (elements of one (i%2,j%2,k%2) colour
 * share no node), so scatter needs neither atomics nor per-thread buffers.
 */
#include "hexat_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Element.jl:8-17 */
static const double NAT[8][3] = {
    {-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
    {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
/* (i,j,k) offsets of local nodes n1..n8, Mesh.jl:45-52 */
static const int NOFF[8][3] = {
    {0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
    {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};

static inline long node_id(const ox_grid* g, int i, int j, int k) {
    const long NX = g->nx + 1, NY = g->ny + 1;
    return (long)i + NX * ((long)j + NY * (long)k);
}

static void elem_nodes(const ox_grid* g, int i, int j, int k, long* n) {
    for (int a = 0; a < 8; ++a)
        n[a] = node_id(g, i + NOFF[a][0], j + NOFF[a][1], k + NOFF[a][2]);
}

/* Element.jl:23-43 */
static void shape_functions(double xi, double eta, double zeta, double* N, double dN[8][3]) {
    for (int i = 0; i < 8; ++i) {
        double tx = 1.0 + xi * NAT[i][0];
        double ty = 1.0 + eta * NAT[i][1];
        double tz = 1.0 + zeta * NAT[i][2];
        if (N) N[i] = 0.125 * tx * ty * tz;
        dN[i][0] = 0.125 * NAT[i][0] * ty * tz;
        dN[i][1] = 0.125 * tx * NAT[i][1] * tz;
        dN[i][2] = 0.125 * tx * ty * NAT[i][2];
    }
}

/* Element.jl:45-62 */
void ox_material_matrix(double E, double nu, double* D) {
    double f = E / ((1.0 + nu) * (1.0 - 2.0 * nu));
    double c1 = (1.0 - nu) * f, c2 = nu * f, c3 = 0.5 * (1.0 - 2.0 * nu) * f;
    memset(D, 0, 36 * sizeof(double));
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) D[6 * i + j] = (i == j) ? c1 : c2;
    D[6 * 3 + 3] = D[6 * 4 + 4] = D[6 * 5 + 5] = c3;
}

static double inv3(const double J[3][3], double Ji[3][3]) {
    double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                 J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                 J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    double id = 1.0 / det;
    Ji[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) * id;
    Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
    Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
    Ji[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * id;
    Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
    Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
    Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * id;
    Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
    Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
    return det;
}

/* dN/dx at a point for an element with the given 8x3 node coordinates */
static double dN_dx_at(const double X[8][3], double xi, double eta, double zeta,
                       double* N, double dNx[8][3]) {
    double dN[8][3], J[3][3], Ji[3][3];
    shape_functions(xi, eta, zeta, N, dN);
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            double s = 0.0;
            for (int a = 0; a < 8; ++a) s += dN[a][r] * X[a][c];
            J[r][c] = s; /* J = dN_dxi' * nodes (Element.jl:73) */
        }
    double det = inv3(J, Ji);
    /* dN_dx = dN_dxi * inv(J)' (Element.jl:79) */
    for (int a = 0; a < 8; ++a)
        for (int c = 0; c < 3; ++c)
            dNx[a][c] = dN[a][0] * Ji[c][0] + dN[a][1] * Ji[c][1] + dN[a][2] * Ji[c][2];
    return det;
}

/* Element.jl:82-89 */
static void fill_B(const double dNx[8][3], double B[6][24]) {
    memset(B, 0, sizeof(double) * 6 * 24);
    for (int i = 0; i < 8; ++i) {
        int c = 3 * i;
        B[0][c] = dNx[i][0];
        B[1][c + 1] = dNx[i][1];
        B[2][c + 2] = dNx[i][2];
        B[3][c] = dNx[i][1];
        B[3][c + 1] = dNx[i][0];
        B[4][c + 1] = dNx[i][2];
        B[4][c + 2] = dNx[i][1];
        B[5][c] = dNx[i][2];
        B[5][c + 2] = dNx[i][0];
    }
}

static void box_nodes(double dx, double dy, double dz, double ztop, double X[8][3]) {
    /* Element.jl:95-100: box centred at the origin; ztop overrides the top z */
    for (int a = 0; a < 8; ++a) {
        X[a][0] = 0.5 * NAT[a][0] * dx;
        X[a][1] = 0.5 * NAT[a][1] * dy;
        X[a][2] = NAT[a][2] < 0 ? -0.5 * dz : ztop;
    }
}

/* Element.jl:64-102 */
void ox_ke_hex(double dx, double dy, double dz, double nu, double E, double* Ke) {
    double X[8][3], D[36], B[6][24], dNx[8][3], DB[6][24];
    box_nodes(dx, dy, dz, 0.5 * dz, X);
    ox_material_matrix(E, nu, D);
    memset(Ke, 0, 576 * sizeof(double));
    const double g = 1.0 / sqrt(3.0), gp[2] = {-g, g};
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int c = 0; c < 2; ++c) {
                double det = dN_dx_at(X, gp[a], gp[b], gp[c], NULL, dNx);
                fill_B(dNx, B);
                for (int r = 0; r < 6; ++r)
                    for (int j = 0; j < 24; ++j) {
                        double s = 0.0;
                        for (int k = 0; k < 6; ++k) s += D[6 * r + k] * B[k][j];
                        DB[r][j] = s;
                    }
                for (int i = 0; i < 24; ++i)
                    for (int j = 0; j < 24; ++j) {
                        double s = 0.0;
                        for (int k = 0; k < 6; ++k) s += B[k][i] * DB[k][j];
                        Ke[24 * i + j] += s * det;
                    }
            }
}

/* Element.jl:104-125 (quirk: top nodes at z=+dz, Element.jl:107) */
void ox_scalar_ke_me(double dx, double dy, double dz, int quirk, double* Ks, double* Ms) {
    double X[8][3], dNx[8][3], N[8];
    box_nodes(dx, dy, dz, quirk ? dz : 0.5 * dz, X);
    memset(Ks, 0, 64 * sizeof(double));
    memset(Ms, 0, 64 * sizeof(double));
    const double g = 1.0 / sqrt(3.0), gp[2] = {-g, g};
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int c = 0; c < 2; ++c) {
                double det = dN_dx_at(X, gp[a], gp[b], gp[c], N, dNx);
                for (int i = 0; i < 8; ++i)
                    for (int j = 0; j < 8; ++j) {
                        Ks[8 * i + j] += (dNx[i][0] * dNx[j][0] + dNx[i][1] * dNx[j][1] +
                                          dNx[i][2] * dNx[j][2]) * det;
                        Ms[8 * i + j] += N[i] * N[j] * det;
                    }
            }
}

/* Element loop over one colour: elements (i,j,k) with i%2==ci etc. */
#define FOR_COLOUR_BEGIN(g)                                                          \
    for (int colour = 0; colour < 8; ++colour) {                                     \
        const int ci = colour & 1, cj = (colour >> 1) & 1, ck = (colour >> 2) & 1;   \
        const int hx = ((g)->nx - ci + 1) / 2, hy = ((g)->ny - cj + 1) / 2,          \
                  hz = ((g)->nz - ck + 1) / 2;                                        \
        const long ncol = (long)hx * hy * hz;                                        \
        _Pragma("omp parallel for schedule(static)")                                 \
        for (long t = 0; t < ncol; ++t) {                                            \
            const int i = ci + 2 * (int)(t % hx);                                    \
            const int j = cj + 2 * (int)((t / hx) % hy);                             \
            const int k = ck + 2 * (int)(t / ((long)hx * hy));                       \
            const long e = (long)i + (long)(g)->nx * ((long)j + (long)(g)->ny * k);
#define FOR_COLOUR_END \
    }                  \
    }

/* GPUSolver.jl:37-79, CPUSolver.jl:50-94 */
void ox_matvec(const ox_grid* g, const double* Ke, const double* Ee,
               const double* x, double* y) {
    const long nn = (long)(g->nx + 1) * (g->ny + 1) * (g->nz + 1);
    memset(y, 0, 3 * nn * sizeof(double));
    FOR_COLOUR_BEGIN(g)
    const double f = Ee[e];
    if (f != 0.0) {
        long n[8];
        double xe[24], ye[24];
        elem_nodes(g, i, j, k, n);
        for (int a = 0; a < 8; ++a)
            for (int d = 0; d < 3; ++d) xe[3 * a + d] = x[3 * n[a] + d];
        for (int r = 0; r < 24; ++r) {
            double s = 0.0;
            for (int c = 0; c < 24; ++c) s += Ke[24 * r + c] * xe[c];
            ye[r] = f * s;
        }
        for (int a = 0; a < 8; ++a)
            for (int d = 0; d < 3; ++d) y[3 * n[a] + d] += ye[3 * a + d];
    }
    FOR_COLOUR_END
}

/* GPUSolver.jl:301-315 */
void ox_diag(const ox_grid* g, const double* Ke, const double* Ee, double* diag) {
    const long nn = (long)(g->nx + 1) * (g->ny + 1) * (g->nz + 1);
    memset(diag, 0, 3 * nn * sizeof(double));
    FOR_COLOUR_BEGIN(g)
    const double f = Ee[e];
    if (f != 0.0) {
        long n[8];
        elem_nodes(g, i, j, k, n);
        for (int a = 0; a < 8; ++a)
            for (int d = 0; d < 3; ++d) diag[3 * n[a] + d] += Ke[24 * (3 * a + d) + 3 * a + d] * f;
    }
    FOR_COLOUR_END
}

static double dot(long n, const double* a, const double* b) {
    double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (long i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

/* GPUSolver.jl:274-609 (CG loop :452-526) */
int ox_pcg(const ox_grid* g, const double* Ke, const double* Ee,
           const uint8_t* fixed, const double* b, double* x,
           double shift, double tol, int max_iter,
           double* rel_out, int* iters_out) {
    const long nn = (long)(g->nx + 1) * (g->ny + 1) * (g->nz + 1), nd = 3 * nn;
    double* minv = (double*)malloc(nd * sizeof(double));
    double* r = (double*)malloc(nd * sizeof(double));
    double* p = (double*)malloc(nd * sizeof(double));
    double* q = (double*)malloc(nd * sizeof(double));
    double* bf = (double*)malloc(nd * sizeof(double));
    int status = 1, it = 0;
    double rel = 0.0;
    ox_diag(g, Ke, Ee, minv);
    for (long d = 0; d < nd; ++d) {
        minv[d] = fixed[d] ? 0.0 : 1.0 / (minv[d] + shift); /* :402 */
        bf[d] = fixed[d] ? 0.0 : b[d];                          /* b_gpu = f[free] :361 */
        if (fixed[d]) x[d] = 0.0;
    }
    const double normb = sqrt(dot(nd, bf, bf));
    if (normb == 0.0) { /* :373 */
        memset(x, 0, nd * sizeof(double));
        status = 3;
        goto done;
    }
    /* initial residual (:415-422) */
    ox_matvec(g, Ke, Ee, x, q);
    for (long d = 0; d < nd; ++d) r[d] = fixed[d] ? 0.0 : bf[d] - (q[d] + shift * x[d]);
    for (long d = 0; d < nd; ++d) p[d] = minv[d] * r[d];
    double rz = dot(nd, r, p);
    for (it = 1; it <= max_iter; ++it) {
        ox_matvec(g, Ke, Ee, p, q);
        for (long d = 0; d < nd; ++d) q[d] = fixed[d] ? 0.0 : q[d] + shift * p[d];
        const double pq = dot(nd, p, q);
        if (fabs(pq) < 1e-300 || isnan(pq)) { /* :462-466 (threshold rescaled for FP64) */
            status = 2;
            break;
        }
        const double alpha = rz / pq;
        for (long d = 0; d < nd; ++d) {
            x[d] += alpha * p[d];
            r[d] -= alpha * q[d];
        }
        rel = sqrt(dot(nd, r, r)) / normb;
        if (rel < tol) {
            status = 0;
            break;
        }
        double rz_new = 0.0;
#pragma omp parallel for reduction(+ : rz_new) schedule(static)
        for (long d = 0; d < nd; ++d) rz_new += r[d] * (minv[d] * r[d]);
        const double beta = rz_new / rz;
        for (long d = 0; d < nd; ++d) p[d] = minv[d] * r[d] + beta * p[d];
        rz = rz_new;
    }
    if (it > max_iter) it = max_iter;
done:
    if (rel_out) *rel_out = rel;
    if (iters_out) *iters_out = it;
    free(minv); free(r); free(p); free(q); free(bf);
    return status;
}

/* GPUHelmholtz.jl:52-76 */
void ox_helm_matvec(const ox_grid* g, const double* Ae, const double* x, double* y) {
    const long nn = (long)(g->nx + 1) * (g->ny + 1) * (g->nz + 1);
    memset(y, 0, nn * sizeof(double));
    FOR_COLOUR_BEGIN(g)
    long n[8];
    (void)e;
    elem_nodes(g, i, j, k, n);
    for (int r = 0; r < 8; ++r) {
        double s = 0.0;
        for (int c = 0; c < 8; ++c) s += Ae[8 * r + c] * x[n[c]];
        y[n[r]] += s;
    }
    FOR_COLOUR_END
}

/* GPUHelmholtz.jl:108-229 */
int ox_helmholtz(const ox_grid* g, double R, int quirk, const double* rho,
                 double* out, double tol, int max_iter, int check_every,
                 double* rel_out, int* iters_out) {
    const long nn = (long)(g->nx + 1) * (g->ny + 1) * (g->nz + 1);
    const long ne = (long)g->nx * g->ny * g->nz;
    double Ks[64], Ms[64], Ae[64];
    ox_scalar_ke_me(g->dx, g->dy, g->dz, quirk, Ks, Ms);
    for (int t = 0; t < 64; ++t) Ae[t] = R * R * Ks[t] + Ms[t]; /* :128-129 */
    double* dg = (double*)calloc(nn, sizeof(double));
    double* b = (double*)calloc(nn, sizeof(double));
    double* x = (double*)calloc(nn, sizeof(double));
    double* r = (double*)malloc(nn * sizeof(double));
    double* z = (double*)malloc(nn * sizeof(double));
    double* p = (double*)malloc(nn * sizeof(double));
    double* q = (double*)malloc(nn * sizeof(double));
    const double vol8 = g->dx * g->dy * g->dz / 8.0; /* :175-176 */
    int it = 0, status = 1;
    double rel = 0.0;
    FOR_COLOUR_BEGIN(g)
    long n[8];
    elem_nodes(g, i, j, k, n);
    for (int a = 0; a < 8; ++a) {
        dg[n[a]] += Ae[9 * a]; /* :91-102 */
        b[n[a]] += rho[e] * vol8; /* :39-50 */
    }
    FOR_COLOUR_END
    const double normb = sqrt(dot(nn, b, b));
    if (normb == 0.0) { /* :179-182 */
        memcpy(out, rho, ne * sizeof(double));
        status = 3;
        goto done;
    }
    for (long n = 0; n < nn; ++n) {
        dg[n] = 1.0 / dg[n];
        r[n] = b[n];
        z[n] = r[n] * dg[n];
        p[n] = z[n];
    }
    double rho_old = dot(nn, r, z);
    for (it = 1; it <= max_iter; ++it) {
        ox_helm_matvec(g, Ae, p, q);
        const double alpha = rho_old / dot(nn, p, q);
        for (long n = 0; n < nn; ++n) {
            x[n] += alpha * p[n];
            r[n] -= alpha * q[n];
        }
        if (it % check_every == 0) { /* :208-214 */
            rel = sqrt(dot(nn, r, r)) / normb;
            if (rel < tol) {
                status = 0;
                break;
            }
        }
        for (long n = 0; n < nn; ++n) z[n] = r[n] * dg[n];
        const double rho_new = dot(nn, r, z);
        const double beta = rho_new / rho_old;
        for (long n = 0; n < nn; ++n) p[n] = z[n] + beta * p[n];
        rho_old = rho_new;
    }
    if (it > max_iter) it = max_iter;
    /* :78-89 extract: element mean of its 8 nodal values */
#pragma omp parallel for schedule(static)
    for (long e = 0; e < ne; ++e) {
        const int i = (int)(e % g->nx), j = (int)((e / g->nx) % g->ny), k = (int)(e / ((long)g->nx * g->ny));
        long n[8];
        elem_nodes(g, i, j, k, n);
        double s = 0.0;
        for (int a = 0; a < 8; ++a) s += x[n[a]];
        out[e] = s / 8.0;
    }
done:
    if (rel_out) *rel_out = rel;
    if (iters_out) *iters_out = it;
    free(dg); free(b); free(x); free(r); free(z); free(p); free(q);
    return status;
}

/* cyclic Jacobi eigenvalue iteration for a symmetric 3x3 */
void ox_eig3(const double* Ain, double* vals, double* vecs) {
    double A[3][3], V[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) A[i][j] = Ain[3 * i + j];
    for (int sweep = 0; sweep < 50; ++sweep) {
        double off = fabs(A[0][1]) + fabs(A[0][2]) + fabs(A[1][2]);
        double scale = fabs(A[0][0]) + fabs(A[1][1]) + fabs(A[2][2]) + off;
        if (off <= 1e-300 || off <= 1e-17 * scale) break;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                if (A[p][q] == 0.0) continue;
                double theta = (A[q][q] - A[p][p]) / (2.0 * A[p][q]);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < 3; ++k) { /* A = A*G */
                    double akp = A[k][p], akq = A[k][q];
                    A[k][p] = c * akp - s * akq;
                    A[k][q] = s * akp + c * akq;
                }
                for (int k = 0; k < 3; ++k) { /* A = G'*A */
                    double apk = A[p][k], aqk = A[q][k];
                    A[p][k] = c * apk - s * aqk;
                    A[q][k] = s * apk + c * aqk;
                }
                for (int k = 0; k < 3; ++k) {
                    double vkp = V[k][p], vkq = V[k][q];
                    V[k][p] = c * vkp - s * vkq;
                    V[k][q] = s * vkp + c * vkq;
                }
            }
    }
    int idx[3] = {0, 1, 2};
    double ev[3] = {A[0][0], A[1][1], A[2][2]};
    for (int a = 0; a < 3; ++a) /* descending (Stress.jl:45) */
        for (int b = a + 1; b < 3; ++b)
            if (ev[idx[b]] > ev[idx[a]]) {
                int t = idx[a];
                idx[a] = idx[b];
                idx[b] = t;
            }
    for (int a = 0; a < 3; ++a) {
        vals[a] = ev[idx[a]];
        for (int k = 0; k < 3; ++k) vecs[3 * k + a] = V[k][idx[a]];
    }
}

/* Stress.jl:10-100 */
void ox_stress(const ox_grid* g, double nu, double E, const double* rho,
               const double* u, double* principal, double* vm, double* voigt,
               double* l1, double* dir) {
    const long ne = (long)g->nx * g->ny * g->nz;
    double X[8][3], dNx[8][3], B[6][24];
    box_nodes(g->dx, g->dy, g->dz, 0.5 * g->dz, X);
    dN_dx_at(X, 0.0, 0.0, 0.0, NULL, dNx); /* centroid (Stress.jl:14-20) */
    fill_B(dNx, B);
#pragma omp parallel for schedule(static)
    for (long e = 0; e < ne; ++e) {
        const int i = (int)(e % g->nx), j = (int)((e / g->nx) % g->ny), k = (int)(e / ((long)g->nx * g->ny));
        double pr[3] = {0, 0, 0}, dv[3] = {0, 0, 0}, sv[6] = {0, 0, 0, 0, 0, 0}, v = 0.0, l = 0.0;
        if (rho[e] > 1e-6) { /* Stress.jl:85 */
            long n[8];
            double ue[24], eps[6], D[36], S[9], vec[9];
            elem_nodes(g, i, j, k, n);
            for (int a = 0; a < 8; ++a)
                for (int d = 0; d < 3; ++d) ue[3 * a + d] = u[3 * n[a] + d];
            for (int r = 0; r < 6; ++r) {
                double s = 0.0;
                for (int c = 0; c < 24; ++c) s += B[r][c] * ue[c];
                eps[r] = s;
            }
            ox_material_matrix(E * rho[e], nu, D);
            for (int r = 0; r < 6; ++r) {
                double s = 0.0;
                for (int c = 0; c < 6; ++c) s += D[6 * r + c] * eps[c];
                sv[r] = s;
            }
            S[0] = sv[0]; S[4] = sv[1]; S[8] = sv[2];
            S[1] = S[3] = sv[3];
            S[5] = S[7] = sv[4];
            S[2] = S[6] = sv[5];
            ox_eig3(S, pr, vec);
            for (int d = 0; d < 3; ++d) dv[d] = vec[3 * d + 0];
            /* Stress.jl:52 */
            v = sqrt(0.5 * ((sv[0] - sv[1]) * (sv[0] - sv[1]) + (sv[1] - sv[2]) * (sv[1] - sv[2]) +
                            (sv[2] - sv[0]) * (sv[2] - sv[0])) +
                     3.0 * (sv[3] * sv[3] + sv[4] * sv[4] + sv[5] * sv[5]));
            l = fabs(pr[0]) + fabs(pr[1]) + fabs(pr[2]); /* Stress.jl:90 */
        }
        if (principal) for (int d = 0; d < 3; ++d) principal[3 * e + d] = pr[d];
        if (dir) for (int d = 0; d < 3; ++d) dir[3 * e + d] = dv[d];
        if (voigt) for (int d = 0; d < 6; ++d) voigt[6 * e + d] = sv[d];
        if (vm) vm[e] = v;
        if (l1) l1[e] = l;
    }
}

void ox_element_energy(const ox_grid* g, const double* Ke, const double* u, double* ce) {
    const long ne = (long)g->nx * g->ny * g->nz;
#pragma omp parallel for schedule(static)
    for (long e = 0; e < ne; ++e) {
        const int i = (int)(e % g->nx), j = (int)((e / g->nx) % g->ny), k = (int)(e / ((long)g->nx * g->ny));
        long n[8];
        double ue[24], s = 0.0;
        elem_nodes(g, i, j, k, n);
        for (int a = 0; a < 8; ++a)
            for (int d = 0; d < 3; ++d) ue[3 * a + d] = u[3 * n[a] + d];
        for (int r = 0; r < 24; ++r) {
            double t = 0.0;
            for (int c = 0; c < 24; ++c) t += Ke[24 * r + c] * ue[c];
            s += ue[r] * t;
        }
        ce[e] = s;
    }
}

/* top88 OC (Andreassen et al. 2011, lines 55-64) */
int ox_oc(int n, const double* x, const double* dc, const double* dv,
          double volfrac, double move, double xmin, double rel_tol,
          double* xnew, double* lmid_out) {
    double l1 = 0.0, l2 = 1e9, lmid = 0.0;
    int steps = 0;
    while ((l2 - l1) / (l1 + l2) > rel_tol && steps < 200) {
        lmid = 0.5 * (l2 + l1);
        double sum = 0.0;
        for (int e = 0; e < n; ++e) {
            double be = -dc[e] / (dv[e] * lmid);
            double cand = x[e] * sqrt(be > 0.0 ? be : 0.0);
            double lo = x[e] - move > xmin ? x[e] - move : xmin;
            double hi = x[e] + move < 1.0 ? x[e] + move : 1.0;
            double v = cand < lo ? lo : (cand > hi ? hi : cand);
            xnew[e] = v;
            sum += v;
        }
        if (sum > volfrac * n) l1 = lmid;
        else l2 = lmid;
        ++steps;
    }
    if (lmid_out) *lmid_out = lmid;
    return steps;
}

/* ---- FP32 mirror of the reference's GPU solver arithmetic ----------------
 * GPUSolver.jl runs everything in Float32: Ke, factors, the EbE matvec
 * (GPUSolver.jl:60-75, four products per statement), CUBLAS Sdot for the dot
 * products and Float32 Jacobi.  This mirror keeps Float32 storage and
 * arithmetic (dot products as per-thread Float32 partials), so it follows
 * the reference's rounding behaviour rather than the exact FP64 answer. */
static float dotf(long n, const float* a, const float* b) {
    float s = 0.0f;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (long i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

void ox_matvec_f32(const ox_grid* g, const float* Ke, const float* Ee, const float* x, float* y) {
    const long nn = (long)(g->nx + 1) * (g->ny + 1) * (g->nz + 1);
    memset(y, 0, 3 * nn * sizeof(float));
    FOR_COLOUR_BEGIN(g)
    const float f = Ee[e];
    if (f != 0.0f) {
        long n[8];
        float xe[24];
        elem_nodes(g, i, j, k, n);
        for (int a = 0; a < 8; ++a)
            for (int d = 0; d < 3; ++d) xe[3 * a + d] = x[3 * n[a] + d];
        for (int r = 0; r < 24; ++r) {
            float v = 0.0f;
            for (int c = 0; c < 24; c += 4)
                v += Ke[24 * r + c] * xe[c] + Ke[24 * r + c + 1] * xe[c + 1] + Ke[24 * r + c + 2] * xe[c + 2] +
                     Ke[24 * r + c + 3] * xe[c + 3];
            y[3 * n[r / 3] + r % 3] += f * v;
        }
    }
    FOR_COLOUR_END
}

int ox_pcg_f32(const ox_grid* g, const float* Ke, const float* Ee, const uint8_t* fixed, const float* b,
               float* x, float shift, double tol, int max_iter, double* rel_out, int* iters_out) {
    const long nn = (long)(g->nx + 1) * (g->ny + 1) * (g->nz + 1), nd = 3 * nn;
    float* minv = (float*)malloc(nd * sizeof(float));
    float* r = (float*)malloc(nd * sizeof(float));
    float* p = (float*)malloc(nd * sizeof(float));
    float* q = (float*)malloc(nd * sizeof(float));
    float* z = (float*)malloc(nd * sizeof(float));
    float* bf = (float*)malloc(nd * sizeof(float));
    int status = 1, it = 0;
    double rel = 0.0;
    memset(minv, 0, nd * sizeof(float));
    FOR_COLOUR_BEGIN(g)
    if (Ee[e] != 0.0f) {
        long n[8];
        elem_nodes(g, i, j, k, n);
        for (int a = 0; a < 8; ++a)
            for (int d = 0; d < 3; ++d) minv[3 * n[a] + d] += Ke[24 * (3 * a + d) + 3 * a + d] * Ee[e];
    }
    FOR_COLOUR_END
    for (long d = 0; d < nd; ++d) {
        minv[d] = fixed[d] ? 0.0f : 1.0f / (minv[d] + shift);
        bf[d] = fixed[d] ? 0.0f : b[d];
        if (fixed[d]) x[d] = 0.0f;
    }
    const float normb = sqrtf(dotf(nd, bf, bf));
    if (normb == 0.0f) {
        memset(x, 0, nd * sizeof(float));
        status = 3;
        goto done;
    }
    ox_matvec_f32(g, Ke, Ee, x, q);
    for (long d = 0; d < nd; ++d) r[d] = fixed[d] ? 0.0f : bf[d] - (q[d] + shift * x[d]);
    for (long d = 0; d < nd; ++d) p[d] = minv[d] * r[d];
    float rz = dotf(nd, r, p);
    for (it = 1; it <= max_iter; ++it) {
        ox_matvec_f32(g, Ke, Ee, p, q);
        for (long d = 0; d < nd; ++d) q[d] = fixed[d] ? 0.0f : q[d] + shift * p[d];
        const float pq = dotf(nd, p, q);
        if (fabsf(pq) < 1e-20f || isnan(pq)) {
            status = 2;
            break;
        }
        const float alpha = rz / pq;
        for (long d = 0; d < nd; ++d) {
            x[d] += alpha * p[d];
            r[d] -= alpha * q[d];
        }
        rel = sqrtf(dotf(nd, r, r)) / normb;
        if (rel < tol) {
            status = 0;
            break;
        }
        for (long d = 0; d < nd; ++d) z[d] = r[d] * minv[d];
        const float rz_new = dotf(nd, r, z);
        const float beta = rz_new / rz;
        for (long d = 0; d < nd; ++d) p[d] = z[d] + beta * p[d];
        rz = rz_new;
    }
    if (it > max_iter) it = max_iter;
done:
    if (rel_out) *rel_out = rel;
    if (iters_out) *iters_out = it;
    free(minv); free(r); free(p); free(q); free(z); free(bf);
    return status;
}
