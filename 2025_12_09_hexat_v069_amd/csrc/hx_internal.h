// hx_internal.h -- shared definitions of libhexat's HIP kernels and host glue.
//
// Element operator in the per-axis parity ("monomial") basis.  With the
// element's nodes in binary order L = bx + 2*by + 4*bz and the transform
//   xh[s] = sum_L prod_{axis a in s} xi_a(L) * x[L]      (xi = -1 / +1)
// every box-element stiffness (2x2x2 Gauss, any dx,dy,dz,nu) has exactly the
// 45 non-zeros listed in kParityPattern; the scalar Laplacian and mass are
// diagonal.  y = Ke x becomes  xh = V^T x,  yh = Kp xh,  y = V yh  with
// Kp = V^T Ke V / 64: 72 + 90 + 72 flops instead of 1152 for dense Ke.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>

namespace hx {

// (row, col) of the 45 non-zeros of Kp; row/col = 3*s + component.
struct RC {
    int8_t r, c;
};
constexpr int kParityNnz = 45;
constexpr RC kParityPattern[kParityNnz] = {
    // linear modes s in {1,2,4}: constant-strain block
    {3, 3}, {3, 7}, {3, 14},
    {4, 4}, {4, 6},
    {5, 5}, {5, 12},
    {6, 4}, {6, 6},
    {7, 3}, {7, 7}, {7, 14},
    {8, 8}, {8, 13},
    {12, 5}, {12, 12},
    {13, 8}, {13, 13},
    {14, 3}, {14, 7}, {14, 14},
    // bilinear modes s in {3,5,6}
    {9, 9}, {9, 20},
    {10, 10}, {10, 17},
    {11, 11}, {11, 16}, {11, 18},
    {15, 15}, {15, 19},
    {16, 11}, {16, 16}, {16, 18},
    {17, 10}, {17, 17},
    {18, 11}, {18, 16}, {18, 18},
    {19, 15}, {19, 19},
    {20, 9}, {20, 20},
    // trilinear mode s = 7
    {21, 21}, {22, 22}, {23, 23}};

// Kp is symmetric: entry n and the entry at (c, r) hold the same value.  The
// lower of the two indices, so kernels touch 33 distinct values, not 45.
constexpr int parity_sym(int n) {
    for (int m = 0; m < n; ++m)
        if (kParityPattern[m].r == kParityPattern[n].c && kParityPattern[m].c == kParityPattern[n].r) return m;
    return n;
}

struct ParitySymTable {
    int v[kParityNnz];
};
constexpr ParitySymTable make_parity_sym() {
    ParitySymTable t{};
    for (int n = 0; n < kParityNnz; ++n) t.v[n] = parity_sym(n);
    return t;
}
constexpr ParitySymTable kParitySym = make_parity_sym();  // compile-time table

template <typename T>
struct ParityK {
    T v[kParityNnz];
};

// reference local node n1..n8 -> binary local index L
constexpr int kRefToBin[8] = {0, 1, 3, 2, 4, 5, 7, 6};

// In-register transforms on 8 values of one component.
template <typename T>
__device__ __forceinline__ void fwd8(T* a) {  // (lo,hi) -> (lo+hi, hi-lo)
#pragma unroll
    for (int b = 1; b < 8; b <<= 1)
#pragma unroll
        for (int L = 0; L < 8; ++L)
            if (!(L & b)) {
                T lo = a[L], hi = a[L | b];
                a[L] = lo + hi;
                a[L | b] = hi - lo;
            }
}
template <typename T>
__device__ __forceinline__ void inv8(T* a) {  // (lo,hi) -> (lo-hi, lo+hi)
#pragma unroll
    for (int b = 1; b < 8; b <<= 1)
#pragma unroll
        for (int L = 0; L < 8; ++L)
            if (!(L & b)) {
                T lo = a[L], hi = a[L | b];
                a[L] = lo - hi;
                a[L | b] = lo + hi;
            }
}

// y (24, index 3L+c) = E * Ke x for one element, parity basis.
template <typename T>
__device__ __forceinline__ void element_apply_parity(const T* __restrict__ x, T* __restrict__ y,
                                                     const ParityK<T>& K, T E) {
    T xh[24], yh[24];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        T t[8];
#pragma unroll
        for (int L = 0; L < 8; ++L) t[L] = x[3 * L + c];
        fwd8(t);
#pragma unroll
        for (int s = 0; s < 8; ++s) xh[3 * s + c] = t[s];
    }
#pragma unroll
    for (int i = 0; i < 24; ++i) yh[i] = T(0);
#pragma unroll
    for (int n = 0; n < kParityNnz; ++n) yh[kParityPattern[n].r] += K.v[n] * xh[kParityPattern[n].c];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        T t[8];
#pragma unroll
        for (int s = 0; s < 8; ++s) t[s] = E * yh[3 * s + c];
        inv8(t);
#pragma unroll
        for (int L = 0; L < 8; ++L) y[3 * L + c] = t[L];
    }
}

// Device-side state of one PCG solve (all scalars stay on the device).
struct PcgState {
    double rz;       // r'z of the current iterate
    double pq;       // p'Ap
    double alpha, beta;
    double rr;       // r'r
    double normb;    // ||b||
    double rel;      // ||r||/||b|| (recursive)
    double best_rel;
    double tol;
    double shift;
    double scal[6];  // spare scalar slots (compliance, OC bracket, change ...)
    double snap_rel; // residual of the last best-iterate snapshot (infinite: none)
    int iter;
    int max_iter;
    int done;        // 0 running, 1 finished
    int status;      // hx_status
    int check_every;
    // best-iterate snapshot (GPUSolver.jl:492-499): mode 1 = the finaliser of
    // each iteration flags a new best iterate with a residual below 1e-2
    // (snap_pending), which the next update kernel copies as it reads x
    // anyway; mode 2 = copies at the FP32 form's check points only; 0 = none
    int snap_pending;
    int snap_mode;
    int oc_widen;    // 8-point OC search: its first pass starts from a guessed bracket
    unsigned ticket[4];
};

// Convergence bookkeeping after an r update (GPUSolver.jl:472-511): record
// ||r|| and the iteration, then decide: instability (:480-484), explosion
// (:486-490), convergence at a check iteration, iteration cap.
__device__ inline void pcg_record(PcgState* st, double rr) {
    st->iter += 1;
    st->rr = rr;
    st->rel = sqrt(rr) / st->normb;
}
__device__ inline void pcg_decide(PcgState* st) {
    const int it = st->iter;
    const double rel = st->rel;
    st->snap_pending = 0;
    if (it > 50 && rel > st->best_rel * 100.0) {
        st->done = 1;
        st->status = 5;
    } else if (!(rel == rel) || isinf(rel) || rel > 1000.0) {
        st->done = 1;
        st->status = 5;
    } else {
        if (rel < st->best_rel) st->best_rel = rel;
        if (st->snap_mode == 1 && rel < 1e-2 && rel < st->snap_rel) {
            st->snap_rel = rel;
            st->snap_pending = 1;
        }
        if (rel < st->tol && (it % st->check_every) == 0) {
            st->done = 1;
            st->status = 0;
        } else if (it >= st->max_iter) {
            st->done = 1;
            st->status = 3;
        }
    }
}

// Grid description passed by value to every kernel.
// A zero word in global memory: a gather reads it in place of an
// out-of-grid value through a selected address, so the load is unconditional
// (a select on the loaded value, or a load under a branch of its own, is
// compiled into a wait per load).  Never written; not const, so it shares
// the global address space with the arrays (the loads stay global loads).
static __device__ double kZeroWord[2] = {0.0, 0.0};

struct GridArgs {
    int nx, ny, nz;
    int NX, NY, NZ;
    long nn, ne;
    double dx, dy, dz;
};

constexpr int kRedThreads = 256;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// Deterministic block sum (fixed tree).  Result valid in thread 0.
template <int NT>
__device__ __forceinline__ double block_sum(double v, double* sh) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x < 64) {
        r = (threadIdx.x < NT / 64) ? sh[threadIdx.x] : 0.0;
        r = wave_sum(r);
    }
    return r;
}
template <int NT>
__device__ __forceinline__ double block_max(double v, double* sh) {
    v = wave_max(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double r = -INFINITY;
    if (threadIdx.x < 64) {
        r = (threadIdx.x < NT / 64) ? sh[threadIdx.x] : -INFINITY;
        r = wave_max(r);
    }
    return r;
}

// Sum partials[nv*b + j] over b, j = 0..NV-1, in a fixed order; results valid
// in thread 0.  Partials come from an earlier kernel on the same stream, so
// plain loads see them (the kernel boundary orders them); no fences or tickets
// are used -- an agent-scope fence writes back the XCD's L2 on gfx950 and
// costs far more than the extra one-block launch.  Each thread issues the
// loads of kUnroll blocks before adding any of them, so the pass costs about
// one memory latency per kUnroll*NT blocks instead of one per NT; the per-
// thread order (b = tid, tid+NT, ...) is the plain loop's.
constexpr int kUnroll = 8;
template <int NT, int NV>
__device__ __forceinline__ void reduce_partials_n(const double* partials, int nv, unsigned nblocks, double* sh,
                                                  double (&out)[NV]) {
    double s[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) s[j] = 0.0;
    for (unsigned b0 = 0; b0 < nblocks; b0 += kUnroll * NT) {
        double t[kUnroll][NV];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            const unsigned b = b0 + u * NT + threadIdx.x;
#pragma unroll
            for (int j = 0; j < NV; ++j) t[u][j] = b < nblocks ? partials[(size_t)nv * b + j] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kUnroll; ++u)
#pragma unroll
            for (int j = 0; j < NV; ++j) s[j] += t[u][j];
    }
#pragma unroll
    for (int j = 0; j < NV; ++j) out[j] = block_sum<NT>(s[j], sh);
}
template <int NT>
__device__ __forceinline__ double reduce_partials(const double* partials, int nv, int j, unsigned nblocks,
                                                  double* sh) {
    double r[1];
    reduce_partials_n<NT, 1>(partials + j, nv, nblocks, sh, r);
    return r[0];
}
template <int NT>
__device__ __forceinline__ double reduce_partials_max(const double* partials, unsigned nblocks, double* sh) {
    double s = -INFINITY;
    for (unsigned b0 = 0; b0 < nblocks; b0 += kUnroll * NT) {
        double t[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            const unsigned b = b0 + u * NT + threadIdx.x;
            t[u] = b < nblocks ? partials[b] : -INFINITY;
        }
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) s = fmax(s, t[u]);
    }
    return block_max<NT>(s, sh);
}

}  // namespace hx
