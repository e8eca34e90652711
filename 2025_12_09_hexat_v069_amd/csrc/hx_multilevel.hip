// hx_multilevel.hip -- additive multilevel diagonal scaling, the linear and
// symmetric form of a U-Net preconditioner:
//
//   z = D0^-1 r + sum_{l=1..L} P_l D_l^-1 P_l^T r
//
// P_l is trilinear interpolation from the node grid of level l (element size
// 2^l, every dimension halved per level) to the fine grid, composed from 2:1
// steps; P_l^T is the matching full-weighting restriction.  D_l is the Jacobi
// diagonal of the level-l grid with the moduli averaged over 2x2x2 children
// and the element stiffness scaled by the element size (an H8 elasticity
// stiffness grows linearly with it).  A coarse DOF is fixed when the fine DOF
// at the same place is.  M^-1 is symmetric positive definite, so the solve
// stays plain PCG.  The encoder is the restriction cascade r_l = R r_{l-1}, the
// bottleneck of every level is its diagonal, and the decoder is the
// prolongation cascade y_{l-1} = D_{l-1}^-1 r_{l-1} + P y_l, ending at
// z = D0^-1 r + mask(P y_1) on the fine grid.  All gathers; no atomics, so the
// result is deterministic.
//
// The reference has no counterpart; its strong-preconditioner slot is the
// learned one (north-star path) and its CPU AMG branch (GPUSolver.jl:168-268).
#include "hx_launch.h"

namespace hx {

namespace {

constexpr int NT = 256;

__device__ __forceinline__ void node_ijk(const GridArgs& g, long n, int& i, int& j, int& k) {
    i = (int)(n % g.NX);
    j = (int)((n / g.NX) % g.NY);
    k = (int)(n / ((long)g.NX * g.NY));
}

// Ec[I,J,K] = mean of the 8 children (2I+a, 2J+b, 2K+c) of the finer level
__global__ __launch_bounds__(NT) void k_ml_coarsen_E(GridArgs gf, const double* __restrict__ Ef, GridArgs gc,
                                                     double* __restrict__ Ec, float* __restrict__ Ec32) {
    const long e = (long)blockIdx.x * NT + threadIdx.x;
    if (e >= gc.ne) return;
    const int I = (int)(e % gc.nx), J = (int)((e / gc.nx) % gc.ny), K = (int)(e / ((long)gc.nx * gc.ny));
    double s = 0.0;
    for (int c = 0; c < 2; ++c)
        for (int b = 0; b < 2; ++b)
            for (int a = 0; a < 2; ++a)
                s += Ef[(long)(2 * K + c) * gf.nx * gf.ny + (long)(2 * J + b) * gf.nx + (2 * I + a)];
    s *= 0.125;
    Ec[e] = s;
    if (Ec32) Ec32[e] = (float)s;
}

// coarse Jacobi inverse diagonal; the fine node of coarse node (I,J,K) is
// (stride I, stride J, stride K) on the fine grid (NXf x NYf nodes)
template <typename T>
__global__ __launch_bounds__(NT) void k_ml_minv(GridArgs gc, double kd0, double kd1, double kd2,
                                                const double* __restrict__ Ec, const uint8_t* __restrict__ fixed_f,
                                                int stride, int NXf, int NYf, T* __restrict__ minv) {
    const long n = (long)blockIdx.x * NT + threadIdx.x;
    if (n >= gc.nn) return;
    int i, j, k;
    node_ijk(gc, n, i, j, k);
    double S = 0.0;
    for (int dz = 0; dz < 2; ++dz)
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) {
                const int ie = i - 1 + dx, je = j - 1 + dy, ke = k - 1 + dz;
                if (ie < 0 || je < 0 || ke < 0 || ie >= gc.nx || je >= gc.ny || ke >= gc.nz) continue;
                S += Ec[(long)ke * gc.nx * gc.ny + (long)je * gc.nx + ie];
            }
    const long f = (long)stride * i + (long)NXf * ((long)stride * j + (long)NYf * ((long)stride * k));
    const double kd[3] = {kd0, kd1, kd2};
    for (int c = 0; c < 3; ++c) {
        const bool fr = (fixed_f == nullptr || fixed_f[3 * f + c] == 0) && S > 0.0;
        minv[3 * n + c] = fr ? T(1.0 / (kd[c] * S)) : T(0);
    }
}

// rc = R rf (full weighting: weight 1/2 per axis offset); on the coarsest
// level also yc = w Dc^-1 rc
template <typename T, bool SCALE>
__global__ __launch_bounds__(NT) void k_ml_restrict(GridArgs gf, const T* __restrict__ rf, GridArgs gc,
                                                    T* __restrict__ rc, const T* __restrict__ minvc,
                                                    T* __restrict__ yc, T w, const PcgState* st) {
    if (st != nullptr && st->done) return;
    const long n = (long)blockIdx.x * NT + threadIdx.x;
    if (n >= gc.nn) return;
    int I, J, K;
    node_ijk(gc, n, I, J, K);
    T acc[3] = {T(0), T(0), T(0)};
    for (int dk = -1; dk <= 1; ++dk) {
        const int k = 2 * K + dk;
        if (k < 0 || k >= gf.NZ) continue;
        const T wk = dk ? T(0.5) : T(1);
        for (int dj = -1; dj <= 1; ++dj) {
            const int j = 2 * J + dj;
            if (j < 0 || j >= gf.NY) continue;
            const T wj = dj ? wk * T(0.5) : wk;
            const long row = 3 * ((long)k * gf.NY * gf.NX + (long)j * gf.NX);
            for (int di = -1; di <= 1; ++di) {
                const int i = 2 * I + di;
                if (i < 0 || i >= gf.NX) continue;
                const T wt = di ? wj * T(0.5) : wj;
                const T* v = rf + row + 3L * i;
                acc[0] += wt * v[0];
                acc[1] += wt * v[1];
                acc[2] += wt * v[2];
            }
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        rc[3 * n + c] = acc[c];
        if constexpr (SCALE) yc[3 * n + c] = w * minvc[3 * n + c] * acc[c];
    }
}

// y_f = w D_f^-1 r_f + P y_c (intermediate level), or on the fine grid
// z = D0^-1 r + mask(P y_1) with the per-block partial sums of r'z
template <typename T, bool FINE>
__global__ __launch_bounds__(NT) void k_ml_prolong(GridArgs gc, const T* __restrict__ yc, GridArgs gf,
                                                   const T* __restrict__ rf, const T* __restrict__ minvf, T w,
                                                   T* __restrict__ yf, const PcgState* st, double* partials) {
    __shared__ double red[NT / 64];
    if (st != nullptr && st->done) return;
    double rz = 0.0;
    for (long n = (long)blockIdx.x * NT + threadIdx.x; n < gf.nn; n += (long)gridDim.x * NT) {
        int i, j, k;
        node_ijk(gf, n, i, j, k);
        // per axis: even index -> one coarse node (weight 1), odd -> two (1/2 each)
        const int I0 = i >> 1, J0 = j >> 1, K0 = k >> 1;
        const int nI = (i & 1) ? 2 : 1, nJ = (j & 1) ? 2 : 1, nK = (k & 1) ? 2 : 1;
        const T wI = (i & 1) ? T(0.5) : T(1), wJ = (j & 1) ? T(0.5) : T(1), wK = (k & 1) ? T(0.5) : T(1);
        T acc[3] = {T(0), T(0), T(0)};
        if (yc != nullptr)
        for (int c3 = 0; c3 < nK; ++c3)
            for (int b = 0; b < nJ; ++b)
                for (int a = 0; a < nI; ++a) {
                    const T* v = yc + 3 * ((long)(K0 + c3) * gc.NY * gc.NX + (long)(J0 + b) * gc.NX + (I0 + a));
                    const T wt = wI * wJ * wK;
                    acc[0] += wt * v[0];
                    acc[1] += wt * v[1];
                    acc[2] += wt * v[2];
                }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const T m = minvf[3 * n + c], r = rf[3 * n + c];
            if constexpr (FINE) {
                const T z = m != T(0) ? m * r + acc[c] : T(0);
                yf[3 * n + c] = z;
                rz += (double)r * (double)z;
            } else {
                yf[3 * n + c] = w * m * r + acc[c];
            }
        }
    }
    if constexpr (FINE) {
        const double v = block_sum<NT>(rz, red);
        if (threadIdx.x == 0) partials[blockIdx.x] = v;
    }
}

// r'z of the new residual: mode 0 beta = rz/rz_old, rz = rz (an iteration);
// 1 rz = rz, beta = 0 (start); 2 park rz in scal[5] (reliable update)
__global__ __launch_bounds__(NT) void k_fin_rz(PcgState* st, const double* partials, unsigned nb, int mode) {
    __shared__ double red[NT / 64];
    if (mode != 2 && st->done) return;  // parking follows a finished phase
    const double s = reduce_partials<NT>(partials, 1, 0, nb, red);
    if (threadIdx.x == 0) {
        if (mode == 0) {  // the update recorded ||r||; decide here, after beta
            if (!(s > 0.0) || !(st->rz > 0.0)) {  // M^-1 is SPD: r'z <= 0 only at r = 0 or in breakdown
                st->done = 1;
                st->status = st->rel == 0.0 ? 0 : 4;
            } else {
                st->beta = s / st->rz;
                st->rz = s;
                pcg_decide(st);
            }
        } else if (mode == 1) {
            st->rz = s;
            st->beta = 0.0;
        } else {
            st->scal[5] = s;
        }
    }
}

constexpr int kProlongBlocks = 1024;  // fine-level blocks (multiple of 8)

inline unsigned blocks_for(long n) { return (unsigned)((n + NT - 1) / NT); }

}  // namespace

int ml_levels_for(const GridArgs& g, int max_levels) {
    int L = 0;
    int nx = g.nx, ny = g.ny, nz = g.nz;
    while (L < max_levels && nx % 2 == 0 && ny % 2 == 0 && nz % 2 == 0) {
        nx /= 2;
        ny /= 2;
        nz /= 2;
        ++L;
    }
    return L;
}

GridArgs ml_coarse_grid(const GridArgs& g) {
    GridArgs c = g;
    c.nx = g.nx / 2;
    c.ny = g.ny / 2;
    c.nz = g.nz / 2;
    c.NX = c.nx + 1;
    c.NY = c.ny + 1;
    c.NZ = c.nz + 1;
    c.nn = (long)c.NX * c.NY * c.NZ;
    c.ne = (long)c.nx * c.ny * c.nz;
    c.dx = 2 * g.dx;
    c.dy = 2 * g.dy;
    c.dz = 2 * g.dz;
    return c;
}

void launch_ml_coarsen_E(const GridArgs& gf, const double* Ef, const GridArgs& gc, double* Ec, float* Ec32,
                         hipStream_t s) {
    hipLaunchKernelGGL(k_ml_coarsen_E, dim3(blocks_for(gc.ne)), dim3(NT), 0, s, gf, Ef, gc, Ec, Ec32);
}

template <typename T>
void launch_ml_minv(const GridArgs& gc, const double kd[3], const double* Ec, const uint8_t* fixed_f, int stride,
                    const GridArgs& gfine, T* minv, hipStream_t s) {
    hipLaunchKernelGGL((k_ml_minv<T>), dim3(blocks_for(gc.nn)), dim3(NT), 0, s, gc, kd[0], kd[1], kd[2], Ec, fixed_f,
                       stride, gfine.NX, gfine.NY, minv);
}

template <typename T>
void launch_ml_apply(const MlHierarchy<T>& h, const T* r, const T* minv0, T* z, PcgState* st_, double* partials,
                     int fin_mode, hipStream_t s) {
    const int L = h.levels;
    // an iteration stops once the solve is done; parking (after a finished
    // phase) and the start always run
    const PcgState* st = fin_mode == 0 ? st_ : nullptr;
    if (L == 0) {  // no coarse level: plain Jacobi through the same finaliser
        hipLaunchKernelGGL((k_ml_prolong<T, true>), dim3(kProlongBlocks), dim3(NT), 0, s, h.g[0], (const T*)nullptr,
                           h.g[0], r, minv0, T(1), z, st, partials);
    } else {
        const T* rf = r;
        for (int l = 1; l <= L; ++l) {
            if (l == L)
                hipLaunchKernelGGL((k_ml_restrict<T, true>), dim3(blocks_for(h.g[l].nn)), dim3(NT), 0, s, h.g[l - 1],
                                   rf, h.g[l], h.r[l], h.minv[l], h.y[l], T(h.w[l]), st);
            else
                hipLaunchKernelGGL((k_ml_restrict<T, false>), dim3(blocks_for(h.g[l].nn)), dim3(NT), 0, s,
                                   h.g[l - 1], rf, h.g[l], h.r[l], (const T*)nullptr, (T*)nullptr, T(1), st);
            rf = h.r[l];
        }
        for (int l = L - 1; l >= 1; --l)
            hipLaunchKernelGGL((k_ml_prolong<T, false>), dim3(blocks_for(h.g[l].nn)), dim3(NT), 0, s, h.g[l + 1],
                               h.y[l + 1], h.g[l], h.r[l], h.minv[l], T(h.w[l]), h.y[l], st, partials);
        hipLaunchKernelGGL((k_ml_prolong<T, true>), dim3(kProlongBlocks), dim3(NT), 0, s, h.g[1], h.y[1], h.g[0], r,
                           minv0, T(1), z, st, partials);
    }
    hipLaunchKernelGGL(k_fin_rz, dim3(1), dim3(NT), 0, s, st_, partials, (unsigned)kProlongBlocks, fin_mode);
}

int ml_partial_blocks() { return kProlongBlocks; }

template void launch_ml_minv<double>(const GridArgs&, const double*, const double*, const uint8_t*, int,
                                     const GridArgs&, double*, hipStream_t);
template void launch_ml_minv<float>(const GridArgs&, const double*, const double*, const uint8_t*, int,
                                    const GridArgs&, float*, hipStream_t);
template void launch_ml_apply<double>(const MlHierarchy<double>&, const double*, const double*, double*, PcgState*,
                                      double*, int, hipStream_t);
template void launch_ml_apply<float>(const MlHierarchy<float>&, const float*, const float*, float*, PcgState*,
                                     double*, int, hipStream_t);

}  // namespace hx
