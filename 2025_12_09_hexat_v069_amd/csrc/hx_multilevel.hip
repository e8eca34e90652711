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
// The V-cycle form (launch_vc_apply) smooths every level above the coarsest,
// which it solves exactly (or by its diagonal): per level l < L, with S_l =
// omega / (absolute row sums of the level's stiffness), rediscretised
// (averaged moduli, element stiffness scaled by the element size) and
// restricted to the DOFs solved for,
//   x = S r;  res = mask(r - A x);  e = V_{l+1}(R res);
//   x += mask(P e);  z = x + S (r - A x)
// The row-sum diagonal bounds A from above (x'Ax <= x'D x for the absolute
// row sums D), so every omega < 2 keeps the smoothing convergent and the
// V-cycle symmetric positive definite whatever the moduli contrast.
//
// The reference has no counterpart; its strong-preconditioner slot is the
// learned U-Net preconditioner and its CPU AMG branch (GPUSolver.jl:168-268).
#include "hx_launch.h"

#include <algorithm>

#include "../../include/hexat.h"

namespace hx {

namespace {

constexpr int NT = 256;

// node counts stay below 2^31 (3 nn fits an int, checked when the grid is set)
__device__ __forceinline__ void node_ijk(const GridArgs& g, long n, int& i, int& j, int& k) {
    const int m = (int)n, q = m / g.NX;
    i = m - q * g.NX;
    k = q / g.NY;
    j = q - k * g.NY;
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

// rc = R rf (full weighting: weight 1/2 per axis offset; on the coarsest
// level also yc = w Dc^-1 rc), marching in z: a workgroup owns a CX x CY
// column of coarse nodes over kc coarse planes.  Coarse plane K sums fine planes 2K-1,
// 2K, 2K+1 (weights 1/2, 1, 1/2), each first restricted in xy from a
// (2CX+1) x (2CY+1) fine tile staged in LDS; plane 2K+1's xy sum is carried
// to K+1.  The next pair of fine planes is loaded into registers while the
// current pair is summed, so one launch streams the fine level once (the
// chunk's first plane and the xy halo aside) with every workgroup resident.
constexpr int kRzX = 32, kRzY = 8;
template <typename T, bool SCALE, int CX = kRzX, int CY = kRzY>
__global__ __launch_bounds__(CX * CY) void k_ml_restrict_zm(GridArgs gf, const T* __restrict__ rf, GridArgs gc,
                                                            T* __restrict__ rc, const T* __restrict__ minvc,
                                                            T* __restrict__ yc, T w, const PcgState* st, int ntx,
                                                            int nty, int kc) {
    constexpr int NTH = CX * CY, ROW = 3 * (2 * CX + 1), PLANE = (2 * CY + 1) * ROW, PER = (PLANE + NTH - 1) / NTH;
    __shared__ T sf[2][PLANE];
    {  // column blockIdx.y
        rf += (long)blockIdx.y * 3 * gf.nn;
        rc += (long)blockIdx.y * 3 * gc.nn;
        if constexpr (SCALE) yc += (long)blockIdx.y * 3 * gc.nn;
        if (st != nullptr) st += blockIdx.y;
    }
    if (st != nullptr && st->done) return;
    const int t = threadIdx.x;
    const int bx = blockIdx.x % ntx, by = (blockIdx.x / ntx) % nty, bz = blockIdx.x / (ntx * nty);
    const int I0 = bx * CX, J0 = by * CY, K0 = bz * kc, K1 = min(K0 + kc, gc.NZ);
    const int i0 = 2 * I0 - 1, j0 = 2 * J0 - 1;
    const long fplane = 3L * gf.NX * gf.NY;
    // this thread's slots of a fine plane tile: (y, x, component) with
    // clamped in-plane offsets; out-of-grid slots read as 0 through a select
    int off[PER];
    unsigned okm = 0;
#pragma unroll
    for (int m = 0; m < PER; ++m) {
        const int idx = t + m * NTH, y = idx / ROW, x3 = idx - y * ROW, x = x3 / 3, i = i0 + x, j = j0 + y;
        if (idx < PLANE && i >= 0 && i < gf.NX && j >= 0 && j < gf.NY) okm |= 1u << m;
        off[m] = 3 * (min(max(j, 0), gf.NY - 1) * gf.NX + min(max(i, 0), gf.NX - 1)) + (x3 - 3 * x);
    }
    const T* const zero = reinterpret_cast<const T*>(kZeroWord);
    auto load = [&](int k, T (&v)[PER]) __attribute__((always_inline)) {
        const unsigned ok = (k >= 0 && k < gf.NZ) ? okm : 0u;
        const T* p = rf + (long)min(max(k, 0), gf.NZ - 1) * fplane;
#pragma unroll
        for (int m = 0; m < PER; ++m) v[m] = *(((ok >> m) & 1u) ? p + off[m] : zero);
    };
    auto put = [&](int b, const T (&v)[PER]) __attribute__((always_inline)) {
#pragma unroll
        for (int m = 0; m < PER; ++m)
            if (m < PER - 1 || t + m * NTH < PLANE) sf[b][t + m * NTH] = v[m];
    };
    const int tx = t % CX, ty = t / CX;
    auto xyr = [&](int b, T (&u)[3]) __attribute__((always_inline)) {
        u[0] = u[1] = u[2] = T(0);
#pragma unroll
        for (int dy = 0; dy < 3; ++dy)
#pragma unroll
            for (int dx = 0; dx < 3; ++dx) {
                const T wt = T((dx == 1 ? 1.0 : 0.5) * (dy == 1 ? 1.0 : 0.5));
                const T* v = &sf[b][(2 * ty + dy) * ROW + 3 * (2 * tx + dx)];
                u[0] += wt * v[0];
                u[1] += wt * v[1];
                u[2] += wt * v[2];
            }
    };
    const int I = I0 + tx, J = J0 + ty;
    const bool own = I < gc.NX && J < gc.NY;
    T va[PER], vb[PER], ulo[3];
    load(2 * K0 - 1, va);
    put(0, va);
    load(2 * K0, va);
    load(2 * K0 + 1, vb);
    __syncthreads();
    xyr(0, ulo);
    __syncthreads();
    for (int K = K0; K < K1; ++K) {
        put(0, va);
        put(1, vb);
        if (K + 1 < K1) {
            load(2 * K + 2, va);
            load(2 * K + 3, vb);
        }
        __syncthreads();
        T um[3], uh[3];
        xyr(0, um);
        xyr(1, uh);
        if (own) {
            const long n = (long)K * gc.NY * gc.NX + (long)J * gc.NX + I;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const T acc = T(0.5) * ulo[c] + um[c] + T(0.5) * uh[c];
                rc[3 * n + c] = acc;
                if constexpr (SCALE) yc[3 * n + c] = w * minvc[3 * n + c] * acc;
            }
        }
        ulo[0] = uh[0];
        ulo[1] = uh[1];
        ulo[2] = uh[2];
        __syncthreads();
    }
}

// coarse planes per workgroup of the restriction onto gc: the deepest chunk
// (at most kRzMaxKc) that still gives min_blocks workgroups (the hierarchy's
// launch width, 512 by default: two per compute unit).  On the 256x128x128 hierarchy's transitions (FP32) this is
// 5 planes for the first (17-18 us against 23 us for one plane) and one
// plane for the smaller ones (4.8-6.6 us; each further plane costs a
// memory round trip there).
constexpr int kRzMaxKc = 8;
inline int restrict_zm_kc(const GridArgs& gc, int ntx, int nty, int min_blocks) {
    int kc = kRzMaxKc;
    while (kc > 1 && (long)ntx * nty * ((gc.NZ + kc - 1) / kc) < min_blocks) --kc;
    return kc;
}

// rc = R rf onto gc (and, with minvc, yc = w Dc^-1 rc); nrhs columns
// (blockIdx.y) of 3 nn values on each grid, one state per column
template <typename T>
void launch_restrict(const GridArgs& gf, const T* rf, const GridArgs& gc, T* rc, const T* minvc, T* yc, T w,
                     const PcgState* st, int min_blocks, hipStream_t s, int nrhs = 1) {
    const int ntx = (gc.NX + kRzX - 1) / kRzX, nty = (gc.NY + kRzY - 1) / kRzY;
    const int kc = restrict_zm_kc(gc, ntx, nty, (min_blocks + nrhs - 1) / nrhs);
    const dim3 grid(ntx * nty * ((gc.NZ + kc - 1) / kc), nrhs);
    if (minvc != nullptr)
        hipLaunchKernelGGL((k_ml_restrict_zm<T, true>), grid, dim3(kRzX * kRzY), 0, s, gf, rf, gc, rc, minvc, yc, w,
                           st, ntx, nty, kc);
    else
        hipLaunchKernelGGL((k_ml_restrict_zm<T, false>), grid, dim3(kRzX * kRzY), 0, s, gf, rf, gc, rc, minvc, yc, w,
                           st, ntx, nty, kc);
}

// interpolation weights of fine index i from the coarse line: corners I0 and
// I0+1 (clamped; weight 0 when i is even), so every thread runs 8 corners
__device__ __forceinline__ void lin_corners(int i, int Nc, int& a0, int& a1, float& w0, float& w1) {
    a0 = i >> 1;
    a1 = min(a0 + 1, Nc - 1);
    w0 = (i & 1) ? 0.5f : 1.0f;
    w1 = (i & 1) ? 0.5f : 0.0f;
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
        T acc[3] = {T(0), T(0), T(0)};
        if (yc != nullptr) {
            int xa[2], ya[2], za[2];
            float xw[2], yw[2], zw[2];
            lin_corners(i, gc.NX, xa[0], xa[1], xw[0], xw[1]);
            lin_corners(j, gc.NY, ya[0], ya[1], yw[0], yw[1]);
            lin_corners(k, gc.NZ, za[0], za[1], zw[0], zw[1]);
#pragma unroll
            for (int c3 = 0; c3 < 2; ++c3)
#pragma unroll
                for (int b = 0; b < 2; ++b)
#pragma unroll
                    for (int a = 0; a < 2; ++a) {
                        const T* v = yc + 3 * ((long)za[c3] * gc.NY * gc.NX + (long)ya[b] * gc.NX + xa[a]);
                        const T wt = T(xw[a] * yw[b] * zw[c3]);
                        acc[0] += wt * v[0];
                        acc[1] += wt * v[1];
                        acc[2] += wt * v[2];
                    }
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

// Levels s..L in one workgroup: restriction from level s-1 down to L, the
// coarsest scaling, prolongation back up to y_s.  These grids are small
// (a few thousand nodes at most), where one launch per level would cost more
// than the work.  Their residuals and outputs live in LDS (dynamic, laid out
// level by level: r_s, y_s, r_{s+1}, ...), so only the first restriction
// reads memory and only y_s is written back; passes are separated by
// workgroup barriers.
constexpr int kCoarseThreads = 1024;

// restriction of node n of grid gc from the finer grid gf held at rf: 27
// taps with clamped addresses and weight 0 outside the grid
template <typename T>
__device__ __forceinline__ void restrict_node(const GridArgs& gf, const T* rf, const GridArgs& gc, long n, T* acc) {
    int I, J, K;
    node_ijk(gc, n, I, J, K);
    int ia[3], ja[3], ka[3];
    float wi[3], wj[3], wk[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const int i = 2 * I + d - 1, j = 2 * J + d - 1, k = 2 * K + d - 1;
        ia[d] = min(max(i, 0), gf.NX - 1);
        ja[d] = min(max(j, 0), gf.NY - 1);
        ka[d] = min(max(k, 0), gf.NZ - 1);
        const float wd = d == 1 ? 1.0f : 0.5f;
        wi[d] = (i >= 0 && i < gf.NX) ? wd : 0.0f;
        wj[d] = (j >= 0 && j < gf.NY) ? wd : 0.0f;
        wk[d] = (k >= 0 && k < gf.NZ) ? wd : 0.0f;
    }
    acc[0] = acc[1] = acc[2] = T(0);
#pragma unroll 1
    for (int dk = 0; dk < 3; ++dk)  // 27 loads in flight per plane (register budget of 1024 threads)
#pragma unroll
        for (int dj = 0; dj < 3; ++dj)
#pragma unroll
            for (int di = 0; di < 3; ++di) {
                const T wt = T(wk[dk] * wj[dj] * wi[di]);
                const T* v = rf + 3 * ((long)ka[dk] * gf.NY * gf.NX + (long)ja[dj] * gf.NX + ia[di]);
                acc[0] += wt * v[0];
                acc[1] += wt * v[1];
                acc[2] += wt * v[2];
            }
}

// trilinear interpolation at node n of grid gf from the coarser grid gc at yc
template <typename T>
__device__ __forceinline__ void prolong_node(const GridArgs& gc, const T* yc, const GridArgs& gf, long n, T* acc) {
    int i, j, k;
    node_ijk(gf, n, i, j, k);
    int xa[2], ya[2], za[2];
    float xw[2], yw[2], zw[2];
    lin_corners(i, gc.NX, xa[0], xa[1], xw[0], xw[1]);
    lin_corners(j, gc.NY, ya[0], ya[1], yw[0], yw[1]);
    lin_corners(k, gc.NZ, za[0], za[1], zw[0], zw[1]);
    acc[0] = acc[1] = acc[2] = T(0);
#pragma unroll
    for (int c3 = 0; c3 < 2; ++c3)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                const T* v = yc + 3 * ((long)za[c3] * gc.NY * gc.NX + (long)ya[b] * gc.NX + xa[a]);
                const T wt = T(xw[a] * yw[b] * zw[c3]);
                acc[0] += wt * v[0];
                acc[1] += wt * v[1];
                acc[2] += wt * v[2];
            }
}

// LDS values of levels s..L (r and y of each)
template <typename T>
long coarse_lds_values(const MlHierarchy<T>& h, int s0) {
    long v = 0;
    for (int l = s0; l <= h.levels; ++l) v += 6 * h.g[l].nn;
    return v;
}

template <typename T>
__global__ __launch_bounds__(kCoarseThreads) void k_ml_coarse(MlHierarchy<T> h, int s0, const PcgState* st) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    T* lds = reinterpret_cast<T*>(lds_raw);
    if (st != nullptr && st->done) return;
    const int L = h.levels;
    auto R = [&](int l) {  // r_l in LDS; y_l follows it
        long o = 0;
        for (int m = s0; m < l; ++m) o += 6 * h.g[m].nn;
        return lds + o;
    };
    auto Y = [&](int l) { return R(l) + 3 * h.g[l].nn; };
    for (int l = s0; l <= L; ++l) {
        const GridArgs gf = h.g[l - 1], gc = h.g[l];
        for (long n = threadIdx.x; n < gc.nn; n += kCoarseThreads) {
            T acc[3];
            if (l == s0) restrict_node(gf, (const T*)h.r[l - 1], gc, n, acc);  // from memory
            else restrict_node(gf, (const T*)R(l - 1), gc, n, acc);           // from LDS
            for (int c = 0; c < 3; ++c) {
                R(l)[3 * n + c] = acc[c];
                if (l == L && h.ainv == nullptr) Y(l)[3 * n + c] = T(h.w[l]) * h.minv[l][3 * n + c] * acc[c];
            }
        }
        __syncthreads();
    }
    if (h.ainv != nullptr) {  // exact coarsest level: y_L = A_L^-1 r_L
        const int N = 3 * (int)h.g[L].nn;
        const T* rl = R(L);
        for (int d = threadIdx.x; d < N; d += kCoarseThreads) {
            T acc = T(0);
            for (int j = 0; j < N; ++j) acc += h.ainv[d * N + j] * rl[j];
            Y(L)[d] = acc;
        }
        __syncthreads();
    }
    for (int l = L - 1; l >= s0; --l) {
        const GridArgs gf = h.g[l], gc = h.g[l + 1];
        for (long n = threadIdx.x; n < gf.nn; n += kCoarseThreads) {
            T acc[3];
            prolong_node(gc, (const T*)Y(l + 1), gf, n, acc);
            for (int c = 0; c < 3; ++c) Y(l)[3 * n + c] = T(h.w[l]) * h.minv[l][3 * n + c] * R(l)[3 * n + c] + acc[c];
        }
        __syncthreads();
    }
    for (long i = threadIdx.x; i < 3 * h.g[s0].nn; i += kCoarseThreads) h.y[s0][i] = Y(s0)[i];
}

// V-cycle levels s..L in one workgroup per column (the grids too small for a
// launch per level): from the residual of level s-1 (memory) down to the
// coarsest level and back, every level's r, x and a work vector in LDS
// (level by level r_l, x_l, t_l), passes separated by barriers; only
// e_s = x_s is written back.  Level l < L: x = S r; t = mask(r - A x);
// r_{l+1} = R t; ...; x += mask(P x_{l+1}); t = A x; x += S (r - t).  Level L:
// x = A_L^-1 r (exact: one wave per row, lanes over the columns) or D_L^-1 r.
// The level operator runs in two passes: every element's 24 outputs E_e Ke x_e
// in the parity basis (one thread per element, the coefficients in registers)
// into a work buffer W, component-major, then every node sums the outputs of
// its elements in a fixed order (deterministic).  The first loads -- the
// exact inverse (when it fits the budget, stage_ainv), the smoothed levels'
// moduli h.Eop[l] and smoother diagonals h.sinv[l], and level s-1's residual
// (restricted from LDS: a gather of its taps from memory by one workgroup is
// bound by the workgroup's address throughput) -- are all issued, coalesced,
// before any is used.
template <typename T>
__host__ __device__ long vc_coarse_work_values(const MlHierarchy<T>& h, int s0) {
    long w = 3 * h.g[s0 - 1].nn;  // level s0-1's residual, staged for the first restriction
    for (int l = s0; l < h.levels; ++l) w = 24 * h.g[l].ne > w ? 24 * h.g[l].ne : w;
    return w;
}
template <typename T>
long vc_coarse_lds_values(const MlHierarchy<T>& h, int s0, bool stage_ainv = false) {
    long v = vc_coarse_work_values(h, s0);
    for (int l = s0; l < h.levels; ++l) v += h.g[l].ne + 3 * h.g[l].nn;
    for (int l = s0; l <= h.levels; ++l) v += 9 * h.g[l].nn;
    if (stage_ainv) v += 9 * h.g[h.levels].nn * h.g[h.levels].nn;
    return v;
}

// node (i, j, k) of node index n < 2^31 on grid g, unsigned arithmetic
__device__ __forceinline__ void node_ijk_u(const GridArgs& g, unsigned n, int& i, int& j, int& k) {
    const unsigned q = n / (unsigned)g.NX;
    i = (int)(n - q * (unsigned)g.NX);
    k = (int)(q / (unsigned)g.NY);
    j = (int)(q - (unsigned)k * (unsigned)g.NY);
}

// component c of the full-weighting restriction at node n of gc from src on
// gf (LDS; 32-bit offsets): separable weights, clamped taps with weight 0
// outside the grid
template <typename T>
__device__ __forceinline__ T restrict_comp(const GridArgs& gf, const T* src, const GridArgs& gc, unsigned n, int c) {
    int I, J, K;
    node_ijk_u(gc, n, I, J, K);
    int ia[3], ja[3], ka[3];
    T wi[3], wj[3], wk[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const int i = 2 * I + d - 1, j = 2 * J + d - 1, k = 2 * K + d - 1;
        const T wd = d == 1 ? T(1) : T(0.5);
        wi[d] = (i >= 0 && i < gf.NX) ? wd : T(0);
        wj[d] = (j >= 0 && j < gf.NY) ? wd : T(0);
        wk[d] = (k >= 0 && k < gf.NZ) ? wd : T(0);
        ia[d] = 3 * min(max(i, 0), gf.NX - 1) + c;
        ja[d] = 3 * gf.NX * min(max(j, 0), gf.NY - 1);
        ka[d] = 3 * gf.NX * gf.NY * min(max(k, 0), gf.NZ - 1);
    }
    T acc = T(0);
#pragma unroll
    for (int dk = 0; dk < 3; ++dk) {
        T pk = T(0);
#pragma unroll
        for (int dj = 0; dj < 3; ++dj) {
            const T* row = src + ka[dk] + ja[dj];
            pk += wj[dj] * (wi[0] * row[ia[0]] + wi[1] * row[ia[1]] + wi[2] * row[ia[2]]);
        }
        acc += wk[dk] * pk;
    }
    return acc;
}
// component c of the trilinear interpolation at node n of gf from src on gc
// (LDS; 32-bit offsets)
template <typename T>
__device__ __forceinline__ T prolong_comp(const GridArgs& gc, const T* src, const GridArgs& gf, unsigned n, int c) {
    int i, j, k;
    node_ijk_u(gf, n, i, j, k);
    int xa[2], ya[2], za[2];
    float xw[2], yw[2], zw[2];
    lin_corners(i, gc.NX, xa[0], xa[1], xw[0], xw[1]);
    lin_corners(j, gc.NY, ya[0], ya[1], yw[0], yw[1]);
    lin_corners(k, gc.NZ, za[0], za[1], zw[0], zw[1]);
    T acc = T(0);
#pragma unroll
    for (int c3 = 0; c3 < 2; ++c3)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int a = 0; a < 2; ++a)
                acc += T(xw[a] * yw[b] * zw[c3]) * src[3 * ((za[c3] * gc.NY + ya[b]) * gc.NX + xa[a]) + c];
    return acc;
}

constexpr int kCoarsePre = 18;       // exact-inverse values per thread fetched with the first loads
constexpr int kCoarseStagePre = 4;   // level-data values per thread fetched with them
constexpr int kCoarseResPre = 8;     // values of the finer level's residual per thread fetched with them

// spare workgroups of the one-workgroup pass's launch (blockIdx.x >= 1): the
// PCG update's x half, x += alpha p over column col, alpha as the update
// kernel left it in the state (the bits of k_pcg_update_free)
constexpr int kCoarseXBlocks = 255;

template <typename T>
__global__ __launch_bounds__(kCoarseThreads) void k_vc_coarse(MlHierarchy<T> h, int s0, ParityK<T> K, int stage_ainv,
                                                              const PcgState* st, T* xu, const T* pu, long xn) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    T* lds = reinterpret_cast<T*>(lds_raw);
    const int col = blockIdx.y;
    if (st != nullptr && st[col].done) return;
    if (blockIdx.x > 0) {
        if (xu == nullptr || st == nullptr) return;
        const double alpha = st[col].alpha;
        T* x = xu + (long)col * xn;
        const T* p = pu + (long)col * xn;
        const long nthr = (long)(gridDim.x - 1) * kCoarseThreads;
        const long tid = (long)(blockIdx.x - 1) * kCoarseThreads + threadIdx.x;
        constexpr int V = 16 / (int)sizeof(T);
        typedef T VT __attribute__((ext_vector_type(16 / sizeof(T))));
        const long nv = xn / V;
        if ((((uintptr_t)x | (uintptr_t)p) & 15u) == 0) {
            for (long v = tid; v < nv; v += nthr) {
                VT xv = ((const VT*)x)[v];
                const VT pv = ((const VT*)p)[v];
#pragma unroll
                for (int l = 0; l < V; ++l) xv[l] = T((double)xv[l] + alpha * (double)pv[l]);
                ((VT*)x)[v] = xv;
            }
        } else {
            for (long i = tid; i < nv * V; i += nthr) x[i] = T((double)x[i] + alpha * (double)p[i]);
        }
        for (long i = nv * V + tid; i < xn; i += nthr) x[i] = T((double)x[i] + alpha * (double)p[i]);
        return;
    }
    const int L = h.levels, t = threadIdx.x;
    auto Ev = [&](int l) {  // moduli and smoother diagonal of smoothed level l
        long o = 0;
        for (int m = s0; m < l; ++m) o += h.g[m].ne + 3 * h.g[m].nn;
        return lds + o;
    };
    auto Sv = [&](int l) { return Ev(l) + h.g[l].ne; };
    auto Rv = [&](int l) {
        long o = 0;
        for (int m = s0; m < L; ++m) o += h.g[m].ne + 3 * h.g[m].nn;
        for (int m = s0; m < l; ++m) o += 9 * h.g[m].nn;
        return lds + o;
    };
    auto Xv = [&](int l) { return Rv(l) + 3 * h.g[l].nn; };
    auto Tv = [&](int l) { return Rv(l) + 6 * h.g[l].nn; };
    T* const W = Rv(L) + 9 * h.g[L].nn;
    T* const Av = W + vc_coarse_work_values(h, s0);  // the exact inverse, when staged

    // ---- first loads, all unconditional (clamped 32-bit indices, a valid
    // stand-in pointer), all issued before any is parked in LDS: level s0-1's
    // residual, the smoothed levels' moduli and smoother diagonals (E_l, S_l
    // level by level, in LDS order from the start, from their contiguous
    // copy) and the exact inverse
    const GridArgs gr = h.g[s0 - 1];
    const unsigned nres = 3u * (unsigned)gr.nn;
    const T* rin = h.r[s0 - 1] + (long)col * nres;
    const unsigned nstage = (unsigned)h.cstage_n;
    const T* ssrc = nstage > 0 ? h.cstage : rin;
    const int N = 3 * h.g[L].nn;
    const unsigned na = stage_ainv ? (unsigned)(N * N) : 0u;
    const T* asrc = na > 0 ? h.ainv : rin;
    auto clampi = [](unsigned i, unsigned n) { return n > 0 ? min(i, n - 1) : 0u; };
    T rpre[kCoarseResPre], spre[kCoarseStagePre], apre[kCoarsePre];
#pragma unroll
    for (int b = 0; b < kCoarseResPre; ++b) rpre[b] = rin[clampi(b * kCoarseThreads + t, nres)];
#pragma unroll
    for (int b = 0; b < kCoarseStagePre; ++b) spre[b] = ssrc[clampi(b * kCoarseThreads + t, nstage)];
#pragma unroll
    for (int b = 0; b < kCoarsePre; ++b) apre[b] = asrc[clampi(b * kCoarseThreads + t, na)];
#pragma unroll
    for (int b = 0; b < kCoarseResPre; ++b) {
        const unsigned i = b * kCoarseThreads + t;
        if (i < nres) W[i] = rpre[b];
    }
#pragma unroll
    for (int b = 0; b < kCoarseStagePre; ++b) {
        const unsigned i = b * kCoarseThreads + t;
        if (i < nstage) lds[i] = spre[b];
    }
#pragma unroll
    for (int b = 0; b < kCoarsePre; ++b) {
        const unsigned i = b * kCoarseThreads + t;
        if (i < na) Av[i] = apre[b];
    }
    for (unsigned i = kCoarseResPre * kCoarseThreads + t; i < nres; i += kCoarseThreads) W[i] = rin[i];
    for (unsigned i = kCoarseStagePre * kCoarseThreads + t; i < nstage; i += kCoarseThreads) lds[i] = h.cstage[i];
    for (unsigned i = kCoarsePre * kCoarseThreads + t; i < na; i += kCoarseThreads) Av[i] = h.ainv[i];
    __syncthreads();
    for (int u = t; u < 3 * (int)h.g[s0].nn; u += kCoarseThreads)  // restriction onto level s0
        Rv(s0)[u] = restrict_comp(gr, (const T*)W, h.g[s0], (unsigned)u / 3u, (int)((unsigned)u % 3u));
    __syncthreads();

    // y = A_l x over the level's nodes; with res, y = mask(r - A_l x)
    auto apply = [&](int l, const T* x, T* y, bool res) {
        const GridArgs g = h.g[l];
        const T* E = Ev(l);
        const int ne = (int)g.ne, nx = g.nx, ny = g.ny, NX = g.NX, NXY = g.NX * g.NY;
        for (int e = t; e < ne; e += kCoarseThreads) {  // element outputs, component-major in W
            const unsigned q = (unsigned)e / (unsigned)nx;
            const int ie = e - (int)q * nx, ke = (int)(q / (unsigned)ny), je = (int)q - ke * ny;
            const int n0 = ke * NXY + je * NX + ie;
            T u[24], yv[24];
#pragma unroll
            for (int Lb = 0; Lb < 8; ++Lb) {
                const int nd = n0 + (Lb >> 2) * NXY + ((Lb >> 1) & 1) * NX + (Lb & 1);
#pragma unroll
                for (int c = 0; c < 3; ++c) u[3 * Lb + c] = x[3 * nd + c];
            }
            element_apply_parity(u, yv, K, E[e]);
#pragma unroll
            for (int i = 0; i < 24; ++i) W[i * ne + e] = yv[i];
        }
        __syncthreads();
        const T* S = Sv(l);
        const T* r = Rv(l);
        for (int n = t; n < (int)g.nn; n += kCoarseThreads) {  // node sums, elements in a fixed order
            int i, j, k;
            node_ijk_u(g, (unsigned)n, i, j, k);
            T acc[3] = {T(0), T(0), T(0)};
#pragma unroll
            for (int e8 = 0; e8 < 8; ++e8) {
                const int ex = e8 & 1, ey = (e8 >> 1) & 1, ez = e8 >> 2;
                const int ie = i - 1 + ex, je = j - 1 + ey, ke = k - 1 + ez;
                if (ie < 0 || je < 0 || ke < 0 || ie >= nx || je >= ny || ke >= g.nz) continue;
                const int eidx = (ke * ny + je) * nx + ie;
                const int La = (1 - ex) + 2 * (1 - ey) + 4 * (1 - ez);
#pragma unroll
                for (int c = 0; c < 3; ++c) acc[c] += W[(3 * La + c) * ne + eidx];
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const int d = 3 * n + c;
                y[d] = res ? (S[d] != T(0) ? r[d] - acc[c] : T(0)) : acc[c];
            }
        }
    };
    // down: pre-smoothing, residual, restriction; the coarsest level; up: correction,
    // post-smoothing.  The level above an exact coarsest level visits it twice
    // (smoothing, correction, smoothing, correction, smoothing: symmetric): after its
    // first post-smoothing it takes the down step again from its current iterate.
    // One loop with each operation at one place (a second copy of the level
    // operator's code spills the per-level arrays of h to scratch).
    const int wl = h.ainv != nullptr ? L - 1 : -1;
    int phase = s0 < L ? 0 : 1;  // 0: down step at level l, 1: the coarsest level, 2: up step at level l
    bool again = false;
    for (int l = s0;;) {
        if (phase == 0) {
            const long nl = 3 * h.g[l].nn;
            if (!again)
                for (long d = t; d < nl; d += kCoarseThreads) Xv(l)[d] = Sv(l)[d] * Rv(l)[d];
            __syncthreads();
            apply(l, Xv(l), Tv(l), true);
            __syncthreads();
            const GridArgs gf = h.g[l], gc = h.g[l + 1];
            for (int u = t; u < 3 * (int)gc.nn; u += kCoarseThreads)
                Rv(l + 1)[u] = restrict_comp(gf, (const T*)Tv(l), gc, (unsigned)u / 3u, (int)((unsigned)u % 3u));
            __syncthreads();
            if (l + 1 < L) {
                ++l;
                continue;
            }
            phase = 1;
        }
        if (phase == 1) {
            {  // the coarsest level
                const long nL = 3 * h.g[L].nn;
                const T* rl = Rv(L);
                if (h.ainv != nullptr) {  // 8 lanes per row (columns q, q + 8, ...), then an 8-lane sum
                    const int q = t & 7;
                    const T* A = stage_ainv ? Av : h.ainv;
                    for (int d0 = 0; d0 < N; d0 += kCoarseThreads / 8) {  // uniform trip count: every lane shuffles
                        const int d = d0 + (t >> 3);
                        T acc = T(0);
                        if (d < N)
                            for (int j = q; j < N; j += 8) acc += A[(long)d * N + j] * rl[j];
#pragma unroll
                        for (int off = 4; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
                        if (q == 0 && d < N) Xv(L)[d] = acc;
                    }
                } else {
                    for (long d = t; d < nL; d += kCoarseThreads) Xv(L)[d] = h.minv[L][d] * rl[d];
                }
                __syncthreads();
            }
            if (L - 1 < s0) break;
            l = L - 1;
            phase = 2;
        }
        {  // up step at level l: correction, post-smoothing
            const GridArgs gf = h.g[l], gc = h.g[l + 1];
            for (int u = t; u < 3 * (int)gf.nn; u += kCoarseThreads)
                if (Sv(l)[u] != T(0))
                    Xv(l)[u] += prolong_comp(gc, (const T*)Xv(l + 1), gf, (unsigned)u / 3u, (int)((unsigned)u % 3u));
            __syncthreads();
            apply(l, Xv(l), Tv(l), false);
            __syncthreads();
            const long nl = 3 * h.g[l].nn;
            for (long d = t; d < nl; d += kCoarseThreads) Xv(l)[d] += Sv(l)[d] * (Rv(l)[d] - Tv(l)[d]);
            __syncthreads();
        }
        if (l == wl && !again) {  // the second visit: down from the current iterate
            again = true;
            phase = 0;
            continue;
        }
        if (--l < s0) break;
    }
    T* const yout = h.y[s0] + (long)col * 3 * h.g[s0].nn;
    for (long i = t; i < 3 * h.g[s0].nn; i += kCoarseThreads) yout[i] = Xv(s0)[i];
}

// r'z of the new residual.  mode 0 (an iteration): record ||r|| from the
// update's partials (r'r at index 2b+1 of nb_upd blocks), beta = rz/rz_old,
// rz = rz, then the convergence decision; 1 (start): rz = rz, beta = 0;
// 2: park rz in scal[5] (reliable update)
// r'r (mode 0): rr_partials[j * rr_stride], j < nb_rr
// Column blockIdx.x: state st + column, partials at column * pcol and
// rr_partials at column * rrcol.
__global__ __launch_bounds__(NT) void k_fin_rz(PcgState* st, const double* partials, unsigned nb,
                                               const double* rr_partials, int rr_stride, unsigned nb_rr, int mode,
                                               long pcol, long rrcol) {
    __shared__ double red[NT / 64];
    st += blockIdx.x;
    partials += (long)blockIdx.x * pcol;
    rr_partials += (long)blockIdx.x * rrcol;
    if (mode != 2 && st->done) return;  // parking follows a finished phase
    const double s = reduce_partials<NT>(partials, 1, 0, nb, red);
    double rr = 0.0;
    if (mode == 0) rr = reduce_partials<NT>(rr_partials, rr_stride, 0, nb_rr, red);
    if (threadIdx.x == 0) {
        if (mode == 0) {
            pcg_record(st, rr);
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

// The coarsest level solved exactly: A assembled element by element (an
// element's 24 DOFs are distinct, so the entries of one element go in
// parallel) into the lower triangle of a symmetric matrix held in LDS,
// boundary DOFs and DOFs without stiffness replaced by identity rows and
// columns, then swept on every pivot with the symmetric sweep operator
//   a_ij -= a_ik a_jk / a_kk (i, j != k);  a_ik /= a_kk;  a_kk = -1 / a_kk
// which leaves -A^-1 (A is symmetric positive definite: no pivoting); those
// rows and columns of the inverse zeroed.  One workgroup; run once per solve.
constexpr int kExactThreads = 1024;
__device__ __forceinline__ int tri(int i, int j) { return i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i; }
template <typename T>
__global__ __launch_bounds__(kExactThreads) void k_ml_exact_setup(GridArgs gc, const double* __restrict__ Ec,
                                                                  const double* __restrict__ Kbin, double h,
                                                                  const uint8_t* __restrict__ fixed_f, int stride,
                                                                  int NXf, int NYf, T* __restrict__ ainv) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    double* P = reinterpret_cast<double*>(lds_raw);  // (i, j <= i) at i (i + 1) / 2 + j
    __shared__ int keep[kMlExactMax];
    const int N = 3 * (int)gc.nn, t = threadIdx.x, NP = N * (N + 1) / 2;
    for (int q = t; q < NP; q += kExactThreads) P[q] = 0.0;
    __syncthreads();
    for (int e = 0; e < (int)gc.ne; ++e) {
        const int I = e % gc.nx, J = (e / gc.nx) % gc.ny, K = e / (gc.nx * gc.ny);
        const double Ee = h * Ec[e];
        for (int ab = t; ab < 576; ab += kExactThreads) {
            const int ra = ab / 24, cb = ab - 24 * ra, La = ra / 3, Lb = cb / 3;
            const int na = (I + (La & 1)) + gc.NX * ((J + ((La >> 1) & 1)) + gc.NY * (K + (La >> 2)));
            const int nb = (I + (Lb & 1)) + gc.NX * ((J + ((Lb >> 1) & 1)) + gc.NY * (K + (Lb >> 2)));
            const int row = 3 * na + ra - 3 * La, col = 3 * nb + cb - 3 * Lb;
            if (row >= col) P[row * (row + 1) / 2 + col] += Ee * Kbin[ab];
        }
        __syncthreads();
    }
    for (int d = t; d < N; d += kExactThreads) {
        int i, j, k;
        node_ijk(gc, d / 3, i, j, k);
        const long f = (long)stride * i + (long)NXf * ((long)stride * j + (long)NYf * ((long)stride * k));
        keep[d] = (fixed_f == nullptr || fixed_f[3 * f + d % 3] == 0) && P[tri(d, d)] > 0.0;
    }
    __syncthreads();
    for (int idx = t; idx < N * N; idx += kExactThreads) {
        const int i = idx / N, j = idx - i * N;
        if (j <= i && (!keep[i] || !keep[j])) P[tri(i, j)] = i == j ? 1.0 : 0.0;
    }
    __syncthreads();
    // sweeps: thread (ti, tj) of a 32 x 32 grid updates the entries (ti + 32 a, tj + 32 b) of the lower
    // triangle, with its pivot-column values read once per pivot
    constexpr int G = 32, MB = (kMlExactMax + G - 1) / G;
    const int ti = t / G, tj = t % G;
    for (int k = 0; k < N; ++k) {
        const double rd = 1.0 / P[tri(k, k)];
        double ci[MB], cj[MB];
#pragma unroll
        for (int a = 0; a < MB; ++a) {
            const int i = ti + G * a, j = tj + G * a;
            ci[a] = i < N ? P[tri(i, k)] * rd : 0.0;
            cj[a] = j < N ? P[tri(j, k)] : 0.0;
        }
#pragma unroll
        for (int a = 0; a < MB; ++a) {
            const int i = ti + G * a;
            if (i >= N || i == k) continue;
            const int row = i * (i + 1) / 2;
#pragma unroll
            for (int b = 0; b <= a; ++b) {
                const int j = tj + G * b;
                if (j > i || j == k) continue;
                P[row + j] -= ci[a] * cj[b];
            }
        }
        __syncthreads();
        for (int i = t; i < N; i += kExactThreads) P[tri(i, k)] = i == k ? -rd : P[tri(i, k)] * rd;
        __syncthreads();
    }
    for (int idx = t; idx < N * N; idx += kExactThreads) {
        const int i = idx / N, j = idx - i * N;
        ainv[idx] = (keep[i] && keep[j]) ? T(-P[tri(i, j)]) : T(0);
    }
}

// V-cycle smoother diagonal of level grid gc (moduli Ec, element stiffness h
// Kbin in binary node order): S = omega / sum_j |A_ij| over the DOFs j solved
// for (not fixed, nonzero diagonal), 0 where DOF i is not solved for.  The
// coupling block of node n with neighbour n + (di,dj,dk) sums, over the
// elements around n that hold both, E_e Kbin[La][Lb]; the absolute value is
// taken of the sum.  Every loop is unrolled, so the Kbin indices are
// constants (uniform loads).  The fine node of coarse node (I,J,K) is
// stride * (I,J,K) on the fine grid (NXf x NYf nodes).
template <typename T>
__global__ __launch_bounds__(NT) void k_vc_l1(GridArgs gc, const double* __restrict__ Ec, double h,
                                              const double* __restrict__ Kbin, const uint8_t* __restrict__ fixed_f,
                                              int stride, int NXf, int NYf, double omega, T* __restrict__ sinv) {
    const long n = (long)blockIdx.x * NT + threadIdx.x;
    if (n >= gc.nn) return;
    int i, j, k;
    node_ijk(gc, n, i, j, k);
    auto fixed_at = [&](int a, int b, int cz, int comp) {
        if (fixed_f == nullptr) return false;
        const long f = (long)stride * a + (long)NXf * ((long)stride * b + (long)NYf * ((long)stride * cz));
        return fixed_f[3 * f + comp] != 0;
    };
    double Ee[8];  // element (i-1+ex, j-1+ey, k-1+ez) at ex + 2 ey + 4 ez, 0 outside the grid
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int ie = i - 1 + (e & 1), je = j - 1 + ((e >> 1) & 1), ke = k - 1 + (e >> 2);
        const bool in = ie >= 0 && je >= 0 && ke >= 0 && ie < gc.nx && je < gc.ny && ke < gc.nz;
        Ee[e] = in ? h * Ec[(long)ke * gc.nx * gc.ny + (long)je * gc.nx + ie] : 0.0;
    }
    double l1[3] = {0.0, 0.0, 0.0}, diag[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int d = 0; d < 27; ++d) {
        const int di = d % 3 - 1, dj = (d / 3) % 3 - 1, dk = d / 9 - 1;
        double B[3][3] = {};
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int ex = e & 1, ey = (e >> 1) & 1, ez = e >> 2;
            const int bx = di + 1 - ex, by = dj + 1 - ey, bz = dk + 1 - ez;  // neighbour's place in the element
            if (bx < 0 || bx > 1 || by < 0 || by > 1 || bz < 0 || bz > 1) continue;
            const int La = (1 - ex) + 2 * (1 - ey) + 4 * (1 - ez), Lb = bx + 2 * by + 4 * bz;
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = 0; b < 3; ++b) B[a][b] += Ee[e] * Kbin[(3 * La + a) * 24 + 3 * Lb + b];
        }
        const int mi = i + di, mj = j + dj, mk = k + dk;
        const bool in = mi >= 0 && mj >= 0 && mk >= 0 && mi < gc.NX && mj < gc.NY && mk < gc.NZ;
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            if (!in || fixed_at(mi, mj, mk, b)) continue;
#pragma unroll
            for (int a = 0; a < 3; ++a) l1[a] += fabs(B[a][b]);
        }
        if (d == 13)
#pragma unroll
            for (int a = 0; a < 3; ++a) diag[a] = B[a][a];
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const bool solved = !fixed_at(i, j, k, a) && diag[a] > 0.0;
        sinv[3 * n + a] = solved ? T(omega / l1[a]) : T(0);
    }
}

// operator moduli of a level: h * Ec in the operator's precision
template <typename T>
__global__ __launch_bounds__(NT) void k_vc_scale_E(long ne, const double* __restrict__ Ec, double h,
                                                   T* __restrict__ Eop) {
    const long e = (long)blockIdx.x * NT + threadIdx.x;
    if (e < ne) Eop[e] = T(h * Ec[e]);
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

// first level handled by the one-workgroup kernel: the first whose grid has
// at most kCoarseNodes nodes and whose levels fit the LDS budget (levels above
// it run one multi-block launch each); levels + 1 when none qualifies
constexpr long kCoarseNodes = 2048;
constexpr long kCoarseLdsBytes = 128 * 1024;
template <typename T>
int ml_coarse_start(const MlHierarchy<T>& h) {
    int s0 = 1;
    while (s0 <= h.levels &&
           (h.g[s0].nn > kCoarseNodes || coarse_lds_values(h, s0) * (long)sizeof(T) > kCoarseLdsBytes))
        ++s0;
    return s0;
}

template <typename T>
void launch_ml_apply(const MlHierarchy<T>& h, const T* r, const T* minv0, T* z, PcgState* st_, double* partials,
                     int fin_mode, hipStream_t s) {
    const int L = h.levels;
    // an iteration stops once the solve is done; parking (after a finished
    // phase) and the start always run
    const PcgState* st = fin_mode == 0 ? st_ : nullptr;
    double* const upd_partials = partials;
    partials += 2 * kVecBlocks;  // after the update's (read by the mode-0 finaliser)
    if (L == 0) {  // no coarse level: plain Jacobi through the same finaliser
        hipLaunchKernelGGL((k_ml_prolong<T, true>), dim3(kProlongBlocks), dim3(NT), 0, s, h.g[0], (const T*)nullptr,
                           h.g[0], r, minv0, T(1), z, st, partials);
    } else {
        const int s0 = ml_coarse_start(h);
        // levels 1..s0-1 (all of them when no level fits the LDS kernel): tiled restriction
        const int last_tiled = s0 > L ? L : s0 - 1;
        for (int l = 1; l <= last_tiled; ++l) {
            const T* rf = l == 1 ? r : h.r[l - 1];
            if (l == L)
                launch_restrict<T>(h.g[l - 1], rf, h.g[l], h.r[l], h.minv[l], h.y[l], T(h.w[l]), st, h.min_blocks, s);
            else
                launch_restrict<T>(h.g[l - 1], rf, h.g[l], h.r[l], nullptr, nullptr, T(1), st, h.min_blocks, s);
        }
        if (last_tiled < L) {
            MlHierarchy<T> hc = h;
            hc.r[0] = const_cast<T*>(r);  // level 0's residual, when s0 = 1
            hipLaunchKernelGGL((k_ml_coarse<T>), dim3(1), dim3(kCoarseThreads),
                               coarse_lds_values(h, s0) * sizeof(T), s, hc, s0, st);
        }
        const int top = s0 <= L ? s0 - 1 : L - 1;  // finest level whose y the coarse pass did not produce
        for (int l = top; l >= 1; --l)
            hipLaunchKernelGGL((k_ml_prolong<T, false>), dim3(blocks_for(h.g[l].nn)), dim3(NT), 0, s, h.g[l + 1],
                               h.y[l + 1], h.g[l], h.r[l], h.minv[l], T(h.w[l]), h.y[l], st, partials);
        hipLaunchKernelGGL((k_ml_prolong<T, true>), dim3(kProlongBlocks), dim3(NT), 0, s, h.g[1], h.y[1], h.g[0], r,
                           minv0, T(1), z, st, partials);
    }
    hipLaunchKernelGGL(k_fin_rz, dim3(1), dim3(NT), 0, s, st_, partials, (unsigned)kProlongBlocks,
                       (const double*)upd_partials + 1, 2, (unsigned)kVecBlocks, fin_mode, 0L, 0L);
}

int ml_partial_blocks() { return kProlongBlocks; }

// every level above the coarsest is smoothed
template <typename T>
int vc_levels(const MlHierarchy<T>& h) {
    return h.levels;
}

// first level of the one-workgroup V-cycle pass: at most kVcCoarseNodes nodes
// (its level operator runs on one compute unit, so larger levels are faster
// as multi-block launches) and its levels' data within the LDS budget;
// levels + 1 when none fits
constexpr long kVcCoarseNodes = 512;
template <typename T>
int vc_coarse_start(const MlHierarchy<T>& h) {
    int s0 = 1;
    while (s0 <= h.levels &&
           (h.g[s0].nn > kVcCoarseNodes || vc_coarse_lds_values(h, s0) * (long)sizeof(T) > kCoarseLdsBytes))
        ++s0;
    return s0;
}

template <typename T>
long vc_coarse_stage_values(const MlHierarchy<T>& h) {
    if (h.levels == 0) return 0;
    long n = 0;
    for (int l = vc_coarse_start(h); l < h.levels; ++l) n += h.g[l].ne + 3 * h.g[l].nn;
    return n;
}
template <typename T>
int launch_vc_coarse_stage(MlHierarchy<T>& h, T* buf, hipStream_t s) {
    h.cstage = nullptr;
    h.cstage_n = vc_coarse_stage_values(h);
    if (h.cstage_n == 0) return HX_OK;
    long o = 0;
    for (int l = vc_coarse_start(h); l < h.levels; ++l) {
        if (hipMemcpyAsync(buf + o, h.Eop[l], h.g[l].ne * sizeof(T), hipMemcpyDeviceToDevice, s) != hipSuccess ||
            hipMemcpyAsync(buf + o + h.g[l].ne, h.sinv[l], 3 * h.g[l].nn * sizeof(T), hipMemcpyDeviceToDevice, s) !=
                hipSuccess)
            return HX_ERR_HIP;
        o += h.g[l].ne + 3 * h.g[l].nn;
    }
    h.cstage = buf;
    return HX_OK;
}

template <typename T>
void launch_vc_smoother(const GridArgs& gc, const double* Ec, double h, const double* Kbin, const uint8_t* fixed_f,
                        int stride, const GridArgs& gfine, double omega, T* sinv, hipStream_t s) {
    hipLaunchKernelGGL((k_vc_l1<T>), dim3(blocks_for(gc.nn)), dim3(NT), 0, s, gc, Ec, h, Kbin, fixed_f, stride,
                       gfine.NX, gfine.NY, omega, sinv);
}

template <typename T>
void launch_vc_scale_E(long ne, const double* Ec, double h, T* Eop, hipStream_t s) {
    hipLaunchKernelGGL((k_vc_scale_E<T>), dim3(blocks_for(ne)), dim3(NT), 0, s, ne, Ec, h, Eop);
}

template <typename T>
bool vc_coarse_pass(const MlHierarchy<T>& h) {
    return h.levels >= 1 && vc_coarse_start(h) <= h.levels;
}

template <typename T>
int launch_vc_apply(const MlHierarchy<T>& h, const ParityK<T>& K, const MvShape& sh_pre, const MvShape& sh_post,
                    const T* r, T* z, PcgState* st_, double* partials, long max_partials, int fin_mode, hipStream_t s,
                    bool rr_post, int nrhs, T* xup, const T* pup) {
    const int L = h.levels;
    if (rr_post && L == 0) return HX_ERR_BAD_ARG;  // no post-smoothing operator to form r'r
    if (L == 0) {  // no coarse level: plain Jacobi (the additive form's path)
        if (nrhs != 1) return HX_ERR_BAD_ARG;
        launch_ml_apply<T>(h, r, h.minv[0], z, st_, partials, fin_mode, s);
        return HX_OK;
    }
    const int s0 = vc_coarse_start(h);
    const int c = std::min(s0, L);  // levels smoothed by multi-block launches
    PcgState* const st = fin_mode == 0 ? st_ : nullptr;
    // the update's partials (2 kVecBlocks per column), then per column the
    // fine post-smoothing operator's r'z and r'r partials (nb0 each)
    double* const upd_partials = partials;
    double* const rz_partials = partials + 2L * kVecBlocks * nrhs;
    const int kc0 = smoothing_post_kchunk(sh_post.tile, sh_post.kchunk, (int)sizeof(T));  // as the launch runs it
    if (2L * nrhs * (kVecBlocks + matvec_parity_blocks(h.g[0], sh_post.tile, kc0)) > max_partials)
        return HX_ERR_BAD_ARG;
    auto level_r = [&](int l) { return l == 0 ? r : (const T*)h.r[l]; };
    // coarse levels: shallower z chunks until the launch has h.min_blocks
    // workgroups over all columns (or one plane each), so the small grids are
    // not latency-bound
    const int min_blocks = (h.min_blocks + nrhs - 1) / nrhs;
    auto level_shape = [&](int l, const MvShape& sh) {
        MvShape s2 = sh;
        if (l > 0)
            while (s2.kchunk > 1 && matvec_parity_blocks(h.g[l], s2.tile, s2.kchunk) < min_blocks) s2.kchunk /= 2;
        return s2;
    };
    auto restrict_to = [&](int l, bool scale) {  // R vres[l-1] -> r[l] (and y[l] = D_l^-1 r[l])
        launch_restrict<T>(h.g[l - 1], (const T*)h.vres[l - 1], h.g[l], h.r[l], scale ? (const T*)h.minv[l] : nullptr,
                           scale ? h.y[l] : nullptr, T(scale ? h.w[l] : 1), st, h.min_blocks, s, nrhs);
    };
    for (int l = 0; l < c; ++l) {  // pre-smoothing x = S r, residual, restriction
        MvArgs<T> a{nullptr, h.vres[l], h.Eop[l], T(0), MV_PLAIN, st, nullptr, 3 * h.g[l].nn, level_r(l), h.sinv[l],
                    nullptr};
        a.form = 3;
        launch_matvec_parity<T>(h.g[l], K, a, level_shape(l, sh_pre), s, nrhs);
        if (l + 1 < c) restrict_to(l + 1, false);
    }
    // levels s0..L in one workgroup per column, from vres[s0-1] (the first launch also does x += alpha p)
    auto coarse_pass = [&](bool with_x) {
        MlHierarchy<T> hc = h;
        hc.r[s0 - 1] = h.vres[s0 - 1];
        const bool stage = h.ainv != nullptr &&
                           vc_coarse_lds_values(h, s0, true) * (long)sizeof(T) <= kCoarseLdsBytes;
        const bool xupd = with_x && xup != nullptr && pup != nullptr && st != nullptr;
        hipLaunchKernelGGL((k_vc_coarse<T>), dim3(xupd ? 1 + kCoarseXBlocks : 1, nrhs), dim3(kCoarseThreads),
                           vc_coarse_lds_values(h, s0, stage) * sizeof(T), s, hc, s0, K, stage ? 1 : 0, st,
                           xupd ? xup : nullptr, xupd ? pup : nullptr, 3 * h.g[0].nn);
    };
    if (s0 <= L) coarse_pass(true);
    else restrict_to(L, true);  // no level fits it: the coarsest level by its diagonal
    // The last multi-block level visits the pass twice when the pass holds the
    // level above an exact coarsest level (which visits that one twice itself)
    // and it is not the fine level: after its post-smoothing, its residual
    // again (form 6), the pass, and a post-smoothing from its iterate (form 7)
    // into vres, which the level below then prolongates.
    const int w2 = (h.ainv != nullptr && s0 == L - 1 && c - 1 >= 1) ? c - 1 : -1;
    int nb0 = 0;
    for (int l = c - 1; l >= 0; --l) {  // post-smoothing with the prolongation of the correction fused
        MvArgs<T> a{nullptr, l == 0 ? z : h.y[l], h.Eop[l], T(0), MV_PLAIN, st, l == 0 ? rz_partials : nullptr,
                    3 * h.g[l].nn, level_r(l), h.sinv[l], nullptr};
        a.form = 4;
        a.ec = l + 1 == w2 ? h.vres[w2] : h.y[l + 1];
        a.cNX = h.g[l + 1].NX;
        a.cNY = h.g[l + 1].NY;
        a.cNZ = h.g[l + 1].NZ;
        a.ccs = 3 * h.g[l + 1].nn;
        const int nb = launch_matvec_parity<T>(h.g[l], K, a, level_shape(l, sh_post), s, nrhs);
        if (l == 0) nb0 = nb;
        if (l == w2) {
            MvArgs<T> b{h.y[l], h.vres[l], h.Eop[l], T(0), MV_PLAIN, st, nullptr, 3 * h.g[l].nn, level_r(l),
                        h.sinv[l], nullptr};
            b.form = 6;  // vres = mask(r - A y)
            launch_matvec_parity<T>(h.g[l], K, b, level_shape(l, sh_pre), s, nrhs);
            coarse_pass(false);  // y[l+1]: the second correction
            MvArgs<T> d{h.y[l], h.vres[l], h.Eop[l], T(0), MV_PLAIN, st, nullptr, 3 * h.g[l].nn, level_r(l),
                        h.sinv[l], nullptr};
            d.form = 7;  // vres = y + mask(P y[l+1]) + S (r - A (...))
            d.ec = h.y[l + 1];
            d.cNX = h.g[l + 1].NX;
            d.cNY = h.g[l + 1].NY;
            d.cNZ = h.g[l + 1].NZ;
            d.ccs = 3 * h.g[l + 1].nn;
            launch_matvec_parity<T>(h.g[l], K, d, level_shape(l, sh_post), s, nrhs);
        }
    }
    // r'r from the update's partials (unmasked update) or from the fine
    // post-smoothing operator's, after its r'z (free update)
    if (rr_post)
        hipLaunchKernelGGL(k_fin_rz, dim3(nrhs), dim3(NT), 0, s, st_, rz_partials, (unsigned)nb0,
                           (const double*)rz_partials + nb0, 1, (unsigned)nb0, fin_mode, 2L * nb0, 2L * nb0);
    else
        hipLaunchKernelGGL(k_fin_rz, dim3(nrhs), dim3(NT), 0, s, st_, rz_partials, (unsigned)nb0,
                           (const double*)upd_partials + 1, 2, (unsigned)kVecBlocks, fin_mode, 2L * nb0,
                           2L * kVecBlocks);
    return HX_OK;
}

template <typename T>
void launch_ml_exact_setup(const GridArgs& gc, const double* Ec, const double* Kbin, double h, const uint8_t* fixed_f,
                           int stride, const GridArgs& gfine, T* ainv, hipStream_t s) {
    const long N = 3 * gc.nn;
    hipLaunchKernelGGL((k_ml_exact_setup<T>), dim3(1), dim3(kExactThreads), N * (N + 1) / 2 * sizeof(double), s, gc,
                       Ec, Kbin, h, fixed_f, stride, gfine.NX, gfine.NY, ainv);
}
template void launch_ml_exact_setup<double>(const GridArgs&, const double*, const double*, double, const uint8_t*, int,
                                            const GridArgs&, double*, hipStream_t);
template void launch_ml_exact_setup<float>(const GridArgs&, const double*, const double*, double, const uint8_t*, int,
                                           const GridArgs&, float*, hipStream_t);

// allow the coarse kernel's dynamic LDS above the default 64 KiB
void ml_prepare_kernels() {
    static bool done = false;
    if (done) return;
    (void)hipFuncSetAttribute((const void*)k_ml_coarse<double>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)kCoarseLdsBytes);
    (void)hipFuncSetAttribute((const void*)k_ml_coarse<float>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)kCoarseLdsBytes);
    (void)hipFuncSetAttribute((const void*)k_vc_coarse<double>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)kCoarseLdsBytes);
    (void)hipFuncSetAttribute((const void*)k_vc_coarse<float>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)kCoarseLdsBytes);
    constexpr int kExactLds = kMlExactMax * (kMlExactMax + 1) / 2 * (int)sizeof(double);
    (void)hipFuncSetAttribute((const void*)k_ml_exact_setup<double>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              kExactLds);
    (void)hipFuncSetAttribute((const void*)k_ml_exact_setup<float>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              kExactLds);
    done = true;
}

template void launch_ml_minv<double>(const GridArgs&, const double*, const double*, const uint8_t*, int,
                                     const GridArgs&, double*, hipStream_t);
template void launch_ml_minv<float>(const GridArgs&, const double*, const double*, const uint8_t*, int,
                                    const GridArgs&, float*, hipStream_t);
template void launch_ml_apply<double>(const MlHierarchy<double>&, const double*, const double*, double*, PcgState*,
                                      double*, int, hipStream_t);
template void launch_ml_apply<float>(const MlHierarchy<float>&, const float*, const float*, float*, PcgState*,
                                     double*, int, hipStream_t);
template int vc_levels<double>(const MlHierarchy<double>&);
template int vc_levels<float>(const MlHierarchy<float>&);
template void launch_vc_smoother<double>(const GridArgs&, const double*, double, const double*, const uint8_t*, int,
                                         const GridArgs&, double, double*, hipStream_t);
template void launch_vc_smoother<float>(const GridArgs&, const double*, double, const double*, const uint8_t*, int,
                                        const GridArgs&, double, float*, hipStream_t);
template void launch_vc_scale_E<double>(long, const double*, double, double*, hipStream_t);
template void launch_vc_scale_E<float>(long, const double*, double, float*, hipStream_t);
template long vc_coarse_stage_values<double>(const MlHierarchy<double>&);
template long vc_coarse_stage_values<float>(const MlHierarchy<float>&);
template int launch_vc_coarse_stage<double>(MlHierarchy<double>&, double*, hipStream_t);
template int launch_vc_coarse_stage<float>(MlHierarchy<float>&, float*, hipStream_t);
template int launch_vc_apply<double>(const MlHierarchy<double>&, const ParityK<double>&, const MvShape&,
                                     const MvShape&, const double*, double*, PcgState*, double*, long, int,
                                     hipStream_t, bool, int, double*, const double*);
template int launch_vc_apply<float>(const MlHierarchy<float>&, const ParityK<float>&, const MvShape&, const MvShape&,
                                    const float*, float*, PcgState*, double*, long, int, hipStream_t, bool, int,
                                    float*, const float*);
template bool vc_coarse_pass<double>(const MlHierarchy<double>&);
template bool vc_coarse_pass<float>(const MlHierarchy<float>&);

}  // namespace hx
