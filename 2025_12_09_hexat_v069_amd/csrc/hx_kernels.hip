// hx_kernels.hip -- streaming kernels of the PCG recurrences, the Helmholtz
// filter, the element post-processing (stress, energies, sensitivities) and
// the OC update.  All reductions are two-stage and deterministic: per-block
// partials in a fixed tree, then a one-block finaliser kernel sums the
// partials in block order and updates the device-side PcgState, so no scalar
// ever returns to the host inside a solve (the reference pays three
// blocking CUBLAS dots per iteration, GPUSolver.jl:460,472,522).
#include "hx_launch.h"

namespace hx {

namespace {

constexpr int NT = kRedThreads;

__device__ __forceinline__ long gtid() { return (long)blockIdx.x * NT + threadIdx.x; }
__device__ __forceinline__ long gstride() { return (long)gridDim.x * NT; }

// 16-byte vectors of the streaming kernels: 4 floats or 2 doubles per load,
// so one wave keeps 1 KiB per array in flight instead of 256 B.
template <typename T>
struct Vec16 {
    static constexpr int n = 16 / sizeof(T);
    typedef T type __attribute__((ext_vector_type(16 / sizeof(T))));
};
__device__ __forceinline__ bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

__global__ void k_pcg_reset(PcgState* st0, double tol, double shift, int max_iter, int check_every, int nrhs,
                            int snap_mode) {
    if (blockIdx.x == 0 && (int)threadIdx.x < nrhs) {
        PcgState* st = st0 + threadIdx.x;
        st->rz = st->pq = st->alpha = st->beta = st->rr = st->normb = st->rel = 0.0;
        st->best_rel = INFINITY;
        st->snap_rel = INFINITY;
        st->tol = tol;
        st->shift = shift;
        for (int i = 0; i < 6; ++i) st->scal[i] = 0.0;
        st->iter = 0;
        st->max_iter = max_iter;
        st->done = 0;
        st->status = 0;
        st->check_every = check_every > 0 ? check_every : 1;
        st->snap_pending = 0;
        st->snap_mode = snap_mode;
        for (int i = 0; i < 4; ++i) st->ticket[i] = 0;
    }
}

// x = free ? x : 0 ; r = free ? b - (q + shift x) : 0 ; p = minv r
template <typename T>
__global__ __launch_bounds__(NT) void k_pcg_init(long n, const T* __restrict__ b, T* __restrict__ x,
                                                 const T* __restrict__ q, const T* __restrict__ minv,
                                                 T* __restrict__ r, T* __restrict__ p, double shift, PcgState* st,
                                                 double* partials) {
    __shared__ double red[NT / 64];
    {
        const long off = (long)blockIdx.y * n;
        b += off; x += off; q += off; r += off; p += off;
        st += blockIdx.y;
        partials += (long)blockIdx.y * gridDim.x * 3;
    }
    double rz = 0.0, bb = 0.0, rr = 0.0;
    for (long i = gtid(); i < n; i += gstride()) {
        const double m = (double)minv[i];
        const bool fr = m != 0.0;
        const double bi = (double)b[i];
        const double xi = fr ? (double)x[i] : 0.0;
        const T rt = fr ? T(bi - ((double)q[i] + shift * xi)) : T(0);
        const T zt = T(m * (double)rt);
        x[i] = T(xi);
        r[i] = rt;
        p[i] = zt;
        rz += (double)rt * (double)zt;
        rr += (double)rt * (double)rt;
        bb += fr ? bi * bi : 0.0;
    }
    const double v0 = block_sum<NT>(rz, red);
    const double v1 = block_sum<NT>(bb, red);
    const double v2 = block_sum<NT>(rr, red);
    if (threadIdx.x == 0) {
        partials[3 * blockIdx.x + 0] = v0;
        partials[3 * blockIdx.x + 1] = v1;
        partials[3 * blockIdx.x + 2] = v2;
    }
}

__global__ __launch_bounds__(NT) void k_fin_init(PcgState* st0, const double* partials0, unsigned nb) {
    __shared__ double red[NT / 64];
    PcgState* st = st0 + blockIdx.x;
    const double* partials = partials0 + (size_t)blockIdx.x * nb * 3;
    double sv[3];
    reduce_partials_n<NT, 3>(partials, 3, nb, red, sv);
    const double s0 = sv[0], s1 = sv[1], s2 = sv[2];
    if (threadIdx.x == 0) {
        st->rz = s0;
        st->normb = sqrt(s1);
        st->rr = s2;
        st->rel = s1 > 0.0 ? sqrt(s2 / s1) : 0.0;
        if (s1 == 0.0) {  // GPUSolver.jl:373
            st->done = 1;
            st->status = 10;
        }
    }
}

// Convergence bookkeeping after the r update (pcg_record / pcg_decide in
// hx_internal.h); with set_beta the Jacobi r'z also gives beta.  An explicit
// preconditioner (set_beta = false) records only: its own finaliser sets beta
// from r'z and then decides, so a phase that stops leaves a consistent beta.
__device__ void pcg_after_update(PcgState* st, double rz_new, double rr, bool set_beta = true) {
    pcg_record(st, rr);
    if (!set_beta) return;
    pcg_decide(st);
    st->beta = rz_new / st->rz;
    st->rz = rz_new;
}

// alpha of an update: with npq > 0, rz / p'q from the operator's partials,
// reduced in every block (same order, same value), block 0 recording p'q and
// alpha or the breakdown (NaN returned); else the alpha already in the state
__device__ double update_alpha(PcgState* st, const double* pq_partials, unsigned npq, double* red, double& s_alpha) {
    if (npq == 0) return st->alpha;
    const double pq = reduce_partials<NT>(pq_partials, 1, 0, npq, red);
    if (threadIdx.x == 0) {
        const bool ok = fabs(pq) >= 1e-300;  // also catches NaN (GPUSolver.jl:462-466)
        s_alpha = ok ? st->rz / pq : 0.0;
        if (blockIdx.x == 0) {
            st->pq = pq;
            if (ok) st->alpha = s_alpha;
            else {
                st->done = 1;
                st->status = 4;
            }
        }
        if (!ok) s_alpha = NAN;
    }
    __syncthreads();
    return s_alpha;
}

// x += alpha p ; r = free ? r - alpha q : 0 ; rz = r'(minv r), rr = r'r
template <typename T>
__global__ __launch_bounds__(NT) void k_pcg_update(long n, T* __restrict__ x, const T* __restrict__ p,
                                                   T* __restrict__ r, const T* __restrict__ q,
                                                   const T* __restrict__ minv, PcgState* st, double* partials,
                                                   const double* __restrict__ pq_partials, unsigned npq,
                                                   T* __restrict__ xbest) {
    __shared__ double red[NT / 64];
    __shared__ double s_alpha;
    {
        const long off = (long)blockIdx.y * n;
        x += off; p += off; r += off; q += off;
        if (xbest) xbest += off;
        st += blockIdx.y;
        partials += (long)blockIdx.y * gridDim.x * 2;
    }
    if (st->done) return;
    // the last finaliser flagged x as the best iterate so far: keep it as it is read
    const bool snap = xbest != nullptr && st->snap_pending;
    const double alpha = update_alpha(st, pq_partials + (size_t)blockIdx.y * npq, npq, red, s_alpha);
    if (alpha != alpha) return;  // breakdown: block 0 has recorded it
    double rz = 0.0, rr = 0.0;
    auto one = [&](T& xi, T pi, T& ri, T qi, T mi) {
        const double m = (double)mi;
        xi = T((double)xi + alpha * (double)pi);
        ri = m != 0.0 ? T((double)ri - alpha * (double)qi) : T(0);
        const double rd = (double)ri;
        rz += rd * (m * rd);
        rr += rd * rd;
    };
    // Thread t takes the V-element groups t, t + stride, ... then the tail;
    // minv is shared by all columns, so a column offset can misalign the
    // others against it: such columns load the same groups element by
    // element, keeping the summation order (and so the bits) of the
    // single-column solve.
    typedef typename Vec16<T>::type VT;
    constexpr int V = Vec16<T>::n;
    const long nv = n / V;
    if (aligned16(x) && aligned16(p) && aligned16(r) && aligned16(q) && aligned16(minv) &&
        (!snap || aligned16(xbest))) {
        for (long v = gtid(); v < nv; v += gstride()) {
            VT xv = ((const VT*)x)[v], rv = ((const VT*)r)[v];
            const VT pv = ((const VT*)p)[v], qv = ((const VT*)q)[v], mv = ((const VT*)minv)[v];
            if (snap) ((VT*)xbest)[v] = xv;
#pragma unroll
            for (int l = 0; l < V; ++l) {
                T xi = xv[l], ri = rv[l];
                one(xi, pv[l], ri, qv[l], mv[l]);
                xv[l] = xi;
                rv[l] = ri;
            }
            ((VT*)x)[v] = xv;
            ((VT*)r)[v] = rv;
        }
    } else {
        for (long v = gtid(); v < nv; v += gstride())
#pragma unroll
            for (int l = 0; l < V; ++l) {
                const long i = v * V + l;
                T xi = x[i], ri = r[i];
                if (snap) xbest[i] = xi;
                one(xi, p[i], ri, q[i], minv[i]);
                x[i] = xi;
                r[i] = ri;
            }
    }
    for (long i = nv * V + gtid(); i < n; i += gstride()) {
        T xi = x[i], ri = r[i];
        if (snap) xbest[i] = xi;
        one(xi, p[i], ri, q[i], minv[i]);
        x[i] = xi;
        r[i] = ri;
    }
    const double v0 = block_sum<NT>(rz, red);
    const double v1 = block_sum<NT>(rr, red);
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x + 0] = v0;
        partials[2 * blockIdx.x + 1] = v1;
    }
}

__global__ __launch_bounds__(NT) void k_fin_update(PcgState* st0, const double* partials0, unsigned nb,
                                                   int set_beta) {
    __shared__ double red[NT / 64];
    PcgState* st = st0 + blockIdx.x;
    if (st->done) return;
    const double* partials = partials0 + (size_t)blockIdx.x * nb * 2;
    double sv[2];
    reduce_partials_n<NT, 2>(partials, 2, nb, red, sv);
    const double s0 = sv[0], s1 = sv[1];
    if (threadIdx.x == 0) pcg_after_update(st, s0, s1, set_beta != 0);
}

// x += alpha p ; r -= alpha q, unmasked and with no reductions, for the
// V-cycle preconditioner: its smoother S is zero exactly where minv is
// (fixed DOFs, DOFs without stiffness), it never reads r there, and its
// post-smoothing operator forms r'r over the other DOFs.  r at those DOFs
// then holds the change of the reaction, K (x - x0), bounded by the loads.
// Free DOFs get the bits of k_pcg_update.
template <typename T>
__global__ __launch_bounds__(NT) void k_pcg_update_free(long n, T* __restrict__ x, const T* __restrict__ p,
                                                        T* __restrict__ r, const T* __restrict__ q, PcgState* st,
                                                        const double* __restrict__ pq_partials, unsigned npq,
                                                        T* __restrict__ xbest) {
    __shared__ double red[NT / 64];
    __shared__ double s_alpha;
    {  // column blockIdx.y
        const long off = (long)blockIdx.y * n;
        x += off; p += off; r += off; q += off;
        if (xbest) xbest += off;
        st += blockIdx.y;
    }
    if (st->done) return;
    const bool snap = xbest != nullptr && st->snap_pending;  // as in k_pcg_update
    const double alpha = update_alpha(st, pq_partials + (size_t)blockIdx.y * npq, npq, red, s_alpha);
    if (alpha != alpha) return;
    auto one = [&](T& xi, T pi, T& ri, T qi) {
        xi = T((double)xi + alpha * (double)pi);
        ri = T((double)ri - alpha * (double)qi);
    };
    typedef typename Vec16<T>::type VT;
    constexpr int V = Vec16<T>::n;
    const long nv = n / V;
    if (aligned16(x) && aligned16(p) && aligned16(r) && aligned16(q) && (!snap || aligned16(xbest))) {
        for (long v = gtid(); v < nv; v += gstride()) {
            VT xv = ((const VT*)x)[v], rv = ((const VT*)r)[v];
            const VT pv = ((const VT*)p)[v], qv = ((const VT*)q)[v];
            if (snap) ((VT*)xbest)[v] = xv;
#pragma unroll
            for (int l = 0; l < V; ++l) {
                T xi = xv[l], ri = rv[l];
                one(xi, pv[l], ri, qv[l]);
                xv[l] = xi;
                rv[l] = ri;
            }
            ((VT*)x)[v] = xv;
            ((VT*)r)[v] = rv;
        }
    } else {
        for (long i = gtid(); i < nv * V; i += gstride()) {
            if (snap) xbest[i] = x[i];
            one(x[i], p[i], r[i], q[i]);
        }
    }
    for (long i = nv * V + gtid(); i < n; i += gstride()) {
        if (snap) xbest[i] = x[i];
        one(x[i], p[i], r[i], q[i]);
    }
}

// r -= alpha q alone (k_pcg_update_free's r half, the same bits): the V-cycle
// PCG leaves x += alpha p to the spare workgroups of its one-workgroup
// coarse pass, which runs latency-bound on one compute unit
template <typename T>
__global__ __launch_bounds__(NT) void k_pcg_update_r(long n, T* __restrict__ r, const T* __restrict__ q,
                                                     PcgState* st, const double* __restrict__ pq_partials,
                                                     unsigned npq) {
    __shared__ double red[NT / 64];
    __shared__ double s_alpha;
    {  // column blockIdx.y
        const long off = (long)blockIdx.y * n;
        r += off; q += off;
        st += blockIdx.y;
    }
    if (st->done) return;
    const double alpha = update_alpha(st, pq_partials + (size_t)blockIdx.y * npq, npq, red, s_alpha);
    if (alpha != alpha) return;
    typedef typename Vec16<T>::type VT;
    constexpr int V = Vec16<T>::n;
    const long nv = n / V;
    if (aligned16(r) && aligned16(q)) {
        for (long v = gtid(); v < nv; v += gstride()) {
            VT rv = ((const VT*)r)[v];
            const VT qv = ((const VT*)q)[v];
#pragma unroll
            for (int l = 0; l < V; ++l) rv[l] = T((double)rv[l] - alpha * (double)qv[l]);
            ((VT*)r)[v] = rv;
        }
    } else {
        for (long i = gtid(); i < nv * V; i += gstride()) r[i] = T((double)r[i] - alpha * (double)q[i]);
    }
    for (long i = nv * V + gtid(); i < n; i += gstride()) r[i] = T((double)r[i] - alpha * (double)q[i]);
}

// p = minv r + beta p
template <typename T>
__global__ __launch_bounds__(NT) void k_pcg_pupdate(long n, const T* __restrict__ r, const T* __restrict__ minv,
                                                    T* __restrict__ p, const PcgState* st) {
    {
        const long off = (long)blockIdx.y * n;
        r += off; p += off;
        st += blockIdx.y;
    }
    if (st->done) return;
    const double beta = st->beta;
    for (long i = gtid(); i < n; i += gstride())
        p[i] = T((double)minv[i] * (double)r[i] + beta * (double)p[i]);
}

__global__ __launch_bounds__(NT) void k_copy16(long n, const uint4* __restrict__ src, uint4* __restrict__ dst) {
    for (long i = gtid(); i < n; i += gstride()) dst[i] = src[i];
}

__global__ __launch_bounds__(NT) void k_dot(long n, const double* __restrict__ a, const double* __restrict__ b,
                                            double* partials) {
    __shared__ double red[NT / 64];
    double s = 0.0;
    for (long i = gtid(); i < n; i += gstride()) s += a[i] * b[i];
    const double v = block_sum<NT>(s, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = v;
}

// r'z per column (blockIdx.y) of FP32 vectors with FP64 sums, parked in
// scal[5] for k_ir_resume: a reliable update of an explicitly preconditioned
// FP32 phase keeps the last z (of the recursive residual, which the true one
// matches to FP32 rounding) instead of applying the preconditioner again
template <typename T>
__global__ __launch_bounds__(NT) void k_rz_part(long n, const T* __restrict__ r, const T* __restrict__ z,
                                                double* partials) {
    __shared__ double red[NT / 64];
    r += (long)blockIdx.y * n;
    z += (long)blockIdx.y * n;
    double s = 0.0;
    for (long i = gtid(); i < n; i += gstride()) s += (double)r[i] * (double)z[i];
    const double v = block_sum<NT>(s, red);
    if (threadIdx.x == 0) partials[(long)blockIdx.y * gridDim.x + blockIdx.x] = v;
}
__global__ __launch_bounds__(NT) void k_fin_rz_park(PcgState* st0, const double* partials, unsigned nb) {
    __shared__ double red[NT / 64];
    const double t = reduce_partials<NT>(partials + (size_t)blockIdx.x * nb, 1, 0, nb, red);
    if (threadIdx.x == 0) st0[blockIdx.x].scal[5] = t;
}

__global__ __launch_bounds__(NT) void k_fin_dot(double* out, const double* partials, unsigned nb) {
    __shared__ double red[NT / 64];
    const double t = reduce_partials<NT>(partials, 1, 0, nb, red);
    if (threadIdx.x == 0) *out = t;
}

__global__ __launch_bounds__(NT) void k_fill(long n, double v, double* a) {
    for (long i = gtid(); i < n; i += gstride()) a[i] = v;
}

__global__ __launch_bounds__(NT) void k_fill32(long n, float v, float* a) {
    for (long i = gtid(); i < n; i += gstride()) a[i] = v;
}

template <typename Tin, typename Tout>
__global__ __launch_bounds__(NT) void k_convert(long n, const Tin* __restrict__ a, Tout* __restrict__ b) {
    for (long i = gtid(); i < n; i += gstride()) b[i] = Tout(a[i]);
}

// r = free ? b - (q + shift x) : 0 in FP64; r32 copy; rz (with minv rounded
// to FP32), rr and optionally bb.  mode bit0: set normb = sqrt(bb);
// mode bit1: park rz in scal[5] (reliable update) instead of st->rz.
__global__ __launch_bounds__(NT) void k_true_residual(long n, const double* __restrict__ b,
                                                      const double* __restrict__ q, const double* __restrict__ x,
                                                      const double* __restrict__ minv, double shift,
                                                      double* __restrict__ r64, float* __restrict__ r32,
                                                      PcgState* st, double* partials, int mode) {
    __shared__ double red[NT / 64];
    {
        const long off = (long)blockIdx.y * n;
        b += off; q += off; x += off;
        if (r64) r64 += off;
        if (r32) r32 += off;
        st += blockIdx.y;
        partials += (long)blockIdx.y * gridDim.x * 3;
    }
    double rz = 0.0, rr = 0.0, bb = 0.0;
    for (long i = gtid(); i < n; i += gstride()) {
        const double m = minv[i];
        const bool fr = m != 0.0;
        const double bi = b[i];
        const double sx = shift != 0.0 ? shift * x[i] : 0.0;  // x is not read without a shift
        const double ri = fr ? bi - (q[i] + sx) : 0.0;
        if (r64) r64[i] = ri;
        const float rf = (float)ri;
        if (r32) r32[i] = rf;
        rz += (double)rf * ((double)(float)m * (double)rf);
        rr += ri * ri;
        bb += fr ? bi * bi : 0.0;
    }
    const double v0 = block_sum<NT>(rz, red);
    const double v1 = block_sum<NT>(rr, red);
    const double v2 = block_sum<NT>(bb, red);
    if (threadIdx.x == 0) {
        partials[3 * blockIdx.x + 0] = v0;
        partials[3 * blockIdx.x + 1] = v1;
        partials[3 * blockIdx.x + 2] = v2;
    }
}

__global__ __launch_bounds__(NT) void k_fin_true(PcgState* st0, const double* partials0, unsigned nb, int mode) {
    __shared__ double red[NT / 64];
    PcgState* st = st0 + blockIdx.x;
    const double* partials = partials0 + (size_t)blockIdx.x * nb * 3;
    double sv[3];
    reduce_partials_n<NT, 3>(partials, 3, nb, red, sv);
    const double s0 = sv[0], s1 = sv[1], s2 = sv[2];
    if (threadIdx.x == 0) {
        if (mode & 1) st->normb = sqrt(s2);
        if (mode & 2) st->scal[5] = s0;
        else st->rz = s0;
        st->scal[4] = s1;  // true ||r||^2
        st->rel = st->normb > 0.0 ? sqrt(s1) / st->normb : 0.0;
    }
}

// Resume the FP32 recurrence of each column after a reliable update: the
// true residual's r'z (scal[5]) replaces the recursive one and rescales beta.
// A column whose true residual is below tol_final, that failed, or that used
// up its iterations stays finished.
__global__ void k_ir_resume(PcgState* st0, double tol_final, double delta, int first, int nrhs) {
    if (blockIdx.x != 0 || (int)threadIdx.x >= nrhs) return;
    PcgState* st = st0 + threadIdx.x;
    if (st->normb == 0.0) {
        st->done = 1;
        st->status = 10;
        return;
    }
    if (st->status == 4 || st->status == 5) {
        st->done = 1;
        return;
    }
    if (st->rel < tol_final) {
        st->done = 1;
        st->status = 0;
        return;
    }
    if (!first && st->iter >= st->max_iter) {
        st->done = 1;
        st->status = 3;
        return;
    }
    const double rzt = st->scal[5];
    st->beta = first ? 0.0 : st->beta * (rzt / st->rz);
    st->rz = rzt;
    st->tol = fmax(tol_final, delta * st->rel);
    st->done = 0;
    st->status = 0;
    st->scal[3] = (double)st->iter;  // first iteration of the next FP32 phase
}

// The FP32 recurrences double-buffer p (p, pb): iteration j of a phase reads
// buffer j%2 and writes the other.  Before a phase resumes, move each
// column's live direction into p (it sits in pb when the phase ended on an
// even iteration count j = 0, 2, ...).
template <typename T>
__global__ __launch_bounds__(NT) void k_fix_p(long n, T* __restrict__ p, const T* __restrict__ pb,
                                              const PcgState* st) {
    st += blockIdx.y;
    const long off = (long)blockIdx.y * n;
    const long k = (long)st->iter - (long)st->scal[3];  // active iterations of the phase that just ended
    if (k <= 0 || ((k - 1) & 1)) return;                // last write went to p (odd j) or none
    for (long i = gtid(); i < n; i += gstride()) p[off + i] = pb[off + i];
}

template <typename T>
__global__ __launch_bounds__(NT) void k_mask(long n, const T* __restrict__ minv, T* __restrict__ x) {
    x += (long)blockIdx.y * n;
    for (long i = gtid(); i < n; i += gstride())
        if (minv[i] == T(0)) x[i] = T(0);
}

__global__ __launch_bounds__(NT) void k_accumulate(long n, double* __restrict__ x64, float* __restrict__ x32) {
    x64 += (long)blockIdx.y * n;
    x32 += (long)blockIdx.y * n;
    for (long i = gtid(); i < n; i += gstride()) {
        x64[i] += (double)x32[i];
        x32[i] = 0.0f;
    }
}

// ---------------------------------------------------------------- Helmholtz
constexpr int kBinToRef[8] = {0, 1, 3, 2, 4, 5, 7, 6};

__device__ __forceinline__ void node_ijk(const GridArgs& g, long n, int& i, int& j, int& k) {
    i = (int)(n % g.NX);
    j = (int)((n / g.NX) % g.NY);
    k = (int)(n / ((long)g.NX * g.NY));
}

// GPUHelmholtz.jl:39-50 as a node gather: b_n = vol/8 * sum_{e ni n} in_e
__global__ __launch_bounds__(NT) void k_helm_rhs(GridArgs g, const double* __restrict__ in_e, double vol8,
                                                 double* __restrict__ b) {
    const long n = gtid();
    if (n >= g.nn) return;
    int i, j, k;
    node_ijk(g, n, i, j, k);
    double s = 0.0;  // out-of-grid elements add the zero word (same sum, every load in flight)
#pragma unroll
    for (int dz = 0; dz < 2; ++dz)
#pragma unroll
        for (int dy = 0; dy < 2; ++dy)
#pragma unroll
            for (int dx = 0; dx < 2; ++dx) {
                const int ie = i - 1 + dx, je = j - 1 + dy, ke = k - 1 + dz;
                const bool in = ie >= 0 && je >= 0 && ke >= 0 && ie < g.nx && je < g.ny && ke < g.nz;
                s += *(in ? in_e + ((long)ke * g.nx * g.ny + (long)je * g.nx + ie) : kZeroWord);
            }
    b[n] = s * vol8;
}

// GPUHelmholtz.jl:91-102,144: inv of the assembled diagonal
__global__ __launch_bounds__(NT) void k_helm_minv(GridArgs g, HelmArgs A, double* __restrict__ minv) {
    const long n = gtid();
    if (n >= g.nn) return;
    int i, j, k;
    node_ijk(g, n, i, j, k);
    double s = 0.0;
    for (int dz = 0; dz < 2; ++dz)
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) {
                const int ie = i - 1 + dx, je = j - 1 + dy, ke = k - 1 + dz;
                if (ie < 0 || je < 0 || ke < 0 || ie >= g.nx || je >= g.ny || ke >= g.nz) continue;
                const int a = kBinToRef[(1 - dx) + 2 * (1 - dy) + 4 * (1 - dz)];
                s += A.Ae[9 * a];
            }
    minv[n] = 1.0 / s;
}

// GPUHelmholtz.jl:52-76 as a node gather (64 MACs per node, x via cache)
// y = A x for the scalar filter operator: the uniform element matrix Ae summed
// over the 8 elements around each node (GPUHelmholtz.jl:165-229 assembles the
// same sum).  A thread owns a node column (i, j) of a 64x4 tile and marches
// a chunk of z planes, keeping the 3x3 neighbourhoods of planes k-1, k, k+1 in
// registers: one new 3x3 plane (9 loads from clamped addresses, coalesced
// along i) per node instead of 64 gathers.  Elements outside the grid get
// weight 0.  In the PCG mode it also writes the per-block x'y partials.
// With BOX the element is a box (both filter forms are: the reference's is
// 1.5 dz tall), so an entry of Ae depends only on which coordinates its two
// nodes share: eight values (16 SGPRs) instead of 64, which the compiler
// otherwise spills to vector lanes and reads back (72 v_readlane per plane).
constexpr int kHx = 64, kHy = 4, kHkc = 16;
struct HelmBox {
    double c[8];  // at sx + 2 sy + 4 sz: s = 1 where the two nodes share that coordinate
};
template <bool BOX>
__global__ __launch_bounds__(NT) void k_helm_matvec(GridArgs g, HelmArgs A, HelmBox hb, const double* __restrict__ x,
                                                    double* __restrict__ y, int mode, PcgState* st,
                                                    double* partials, int ntx, int nty) {
    __shared__ double red[NT / 64];
    if (mode == MV_PCG && st->done) return;
    const int bx = blockIdx.x % ntx, by = (blockIdx.x / ntx) % nty, bz = blockIdx.x / (ntx * nty);
    const int i = bx * kHx + (threadIdx.x % kHx), j = by * kHy + (threadIdx.x / kHx);
    const int k0 = bz * kHkc, k1 = min(k0 + kHkc, g.NZ);
    const bool active = i < g.NX && j < g.NY;
    int ii[3], jj[3];
    double wx[2], wy[2];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        ii[d] = min(max(i - 1 + d, 0), g.NX - 1);
        jj[d] = min(max(j - 1 + d, 0), g.NY - 1);
    }
#pragma unroll
    for (int d = 0; d < 2; ++d) {  // element i-1+d / j-1+d inside the grid
        wx[d] = (i - 1 + d >= 0 && i - 1 + d < g.nx) ? 1.0 : 0.0;
        wy[d] = (j - 1 + d >= 0 && j - 1 + d < g.ny) ? 1.0 : 0.0;
    }
    const long plane = (long)g.NX * g.NY;
    double V[3][3][3];  // [z: k-1, k, k+1][y][x]
    auto load = [&](int kk, double (&P)[3][3]) {
        const double* xp = x + (long)min(max(kk, 0), g.NZ - 1) * plane;
#pragma unroll
        for (int dy = 0; dy < 3; ++dy)
#pragma unroll
            for (int dx = 0; dx < 3; ++dx) P[dy][dx] = xp[(long)jj[dy] * g.NX + ii[dx]];
    };
    load(k0 - 1, V[0]);
    load(k0, V[1]);
    double pq = 0.0;
    for (int k = k0; k < k1; ++k) {
        load(k + 1, V[2]);
        double acc = 0.0;
#pragma unroll
        for (int dz = 0; dz < 2; ++dz) {
            const double wz = (k - 1 + dz >= 0 && k - 1 + dz < g.nz) ? 1.0 : 0.0;
#pragma unroll
            for (int dy = 0; dy < 2; ++dy)
#pragma unroll
                for (int dx = 0; dx < 2; ++dx) {
                    const int a = kBinToRef[(1 - dx) + 2 * (1 - dy) + 4 * (1 - dz)];
                    double e = 0.0;
#pragma unroll
                    for (int L = 0; L < 8; ++L) {
                        double coef;
                        if constexpr (BOX)
                            coef = hb.c[((1 - dx) == (L & 1)) + 2 * ((1 - dy) == ((L >> 1) & 1)) +
                                        4 * ((1 - dz) == (L >> 2))];
                        else
                            coef = A.Ae[8 * a + kBinToRef[L]];
                        e += coef * V[dz + (L >> 2)][dy + ((L >> 1) & 1)][dx + (L & 1)];
                    }
                    acc += (wx[dx] * wy[dy] * wz) * e;
                }
        }
        if (active) {
            y[(long)k * plane + (long)j * g.NX + i] = acc;
            pq += V[1][1][1] * acc;
        }
#pragma unroll
        for (int dy = 0; dy < 3; ++dy)
#pragma unroll
            for (int dx = 0; dx < 3; ++dx) {
                V[0][dy][dx] = V[1][dy][dx];
                V[1][dy][dx] = V[2][dy][dx];
            }
    }
    if (mode == MV_PCG) {
        const double v = block_sum<NT>(pq, red);
        if (threadIdx.x == 0) partials[blockIdx.x] = v;
    }
}

// Sum-factorised form of the same operator.  Every element of the filter
// has the same box matrix R^2 Ks + Ms, and on a box Ks = Kx(x)My(x)Mz +
// Mx(x)Ky(x)Mz + Mx(x)My(x)Kz, Ms = Mx(x)My(x)Mz with the 1-D element
// matrices K = k [1 -1; -1 1], M = m [2 1; 1 2] (k = 1/h, m = h/6), so the
// assembled operator factorises per axis:
//   y = Kx U + Mx W,  U = R^2 My a,  W = (R^2 Ky + My) a + R^2 My b,
//   a = Mz x,  b = Kz x,
// each factor a 3-point stencil whose end entries vanish where the element
// between two nodes lies outside the grid.  A thread owns a node column and
// marches z (a, b from its own column's three planes); row neighbours'
// contributions P = a and Q = (m_y - R^2 k_y) a + R^2 m_y b come through LDS
// (double-buffered: one barrier per plane), the x neighbours' S = -k_x U +
// m_x W through a DPP lane shift (a tile row is one wave).  ~22 FMAs per node
// and plane instead of 64, one coalesced load per plane instead of 9.
// FUSE: the PCG's p = minv r + beta p_old is formed from the planes staged
// anyway and written to p_out (p_old in x), as the elasticity operator does.
struct HelmSF {
    double kx, mx, ky, my, kz, mz, R2;
};
constexpr int kHsfX = 64, kHsfY = 16, kHsfKc = 16;  // tile: one wave per row; 62 x 14 owned node columns
__device__ __forceinline__ double wave_from_left(double v) {  // lane l <- lane l-1 (lane 0 gets 0)
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x138, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x138, 0xF, 0xF, true);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double wave_from_right(double v) {  // lane l <- lane l+1 (lane 63 gets 0)
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x130, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x130, 0xF, 0xF, true);
    return __hiloint2double(hi, lo);
}
template <bool FUSE>
__global__ __launch_bounds__(kHsfX * kHsfY) void k_helm_sf(GridArgs g, HelmSF c, const double* __restrict__ x,
                                                          const double* __restrict__ r,
                                                          const double* __restrict__ minv, double* __restrict__ p_out,
                                                          double* __restrict__ y, int mode, PcgState* st,
                                                          double* partials, int ntx, int nty) {
    constexpr int NT2 = kHsfX * kHsfY;
    __shared__ double PQ[2][2][kHsfY][kHsfX];  // [buffer][P, Q][row][column]
    __shared__ double red[NT2 / 64];
    if (mode == MV_PCG && st->done) return;
    const double beta = FUSE ? st->beta : 0.0;
    const int tx = threadIdx.x % kHsfX, ty = threadIdx.x / kHsfX;
    const int bx = blockIdx.x % ntx, by = (blockIdx.x / ntx) % nty, bz = blockIdx.x / (ntx * nty);
    const int i = bx * (kHsfX - 2) - 1 + tx, j = by * (kHsfY - 2) - 1 + ty;
    const int k0 = bz * kHsfKc, k1 = min(k0 + kHsfKc, g.NZ);
    const bool own = tx >= 1 && tx <= kHsfX - 2 && ty >= 1 && ty <= kHsfY - 2 && i < g.NX && j < g.NY;
    // per-axis stencil flags: the element on the low / high side of the node exists
    const double nlx = (i >= 1 && i <= g.NX - 1) ? 1.0 : 0.0, nrx = (i >= 0 && i <= g.NX - 2) ? 1.0 : 0.0;
    const double nly = (j >= 1 && j <= g.NY - 1) ? 1.0 : 0.0, nry = (j >= 0 && j <= g.NY - 2) ? 1.0 : 0.0;
    const double kxc = c.kx * (nlx + nrx), mxc = 2.0 * c.mx * (nlx + nrx);
    const double kyc = c.ky * (nly + nry), myc = 2.0 * c.my * (nly + nry);
    const double qa = c.my - c.R2 * c.ky, qb = c.R2 * c.my;  // Q = qa a + qb b
    const double wa = c.R2 * kyc + myc, wb = c.R2 * myc;     // centre of W
    const double ua = c.R2 * myc, un = c.R2 * c.my;          // U = ua a + un (neighbours' a)
    const long col = (long)min(max(j, 0), g.NY - 1) * g.NX + min(max(i, 0), g.NX - 1);
    const long plane = (long)g.NX * g.NY;
    // the staged value of plane k at this column: x, or the fused p; the raw
    // values of a plane are requested two planes ahead of their use
    struct Raw {
        double x, r, m;
    };
    auto request = [&](int k) __attribute__((always_inline)) -> Raw {
        const long o = (long)min(max(k, 0), g.NZ - 1) * plane + col;
        Raw q;
        q.x = x[o];
        if constexpr (FUSE) {
            q.r = r[o];
            q.m = minv[o];
        } else {
            q.r = q.m = 0.0;
        }
        return q;
    };
    auto staged = [&](const Raw& q) __attribute__((always_inline)) -> double {
        if constexpr (FUSE) return q.m * q.r + beta * q.x;
        else return q.x;
    };
    double v0 = staged(request(k0 - 1)), v1 = staged(request(k0));
    Raw n1 = request(k0 + 1), n2 = request(k0 + 2);
    double pq = 0.0;
    int buf = 0;
    for (int k = k0; k < k1; ++k) {
        const double v2 = staged(n1);  // plane k+1, requested two planes ago
        n1 = n2;
        n2 = request(k + 3);  // past the grid: re-reads the last plane (a cache hit)
        if constexpr (FUSE) {
            if (own) p_out[(long)k * plane + col] = v1;
        }
        // z-pass at plane k (wave-uniform coefficients)
        const double nlz = k >= 1 ? 1.0 : 0.0, nrz = k <= g.NZ - 2 ? 1.0 : 0.0;
        const double a = c.mz * (nlz * v0 + nrz * v2) + 2.0 * c.mz * (nlz + nrz) * v1;
        const double b = -c.kz * (nlz * v0 + nrz * v2) + c.kz * (nlz + nrz) * v1;
        PQ[buf][0][ty][tx] = a;
        PQ[buf][1][ty][tx] = qa * a + qb * b;
        __syncthreads();
        const int ylo = ty >= 1 ? ty - 1 : ty, yhi = ty + 1 < kHsfY ? ty + 1 : ty;  // halo rows: unused results
        const double U = ua * a + un * (nly * PQ[buf][0][ylo][tx] + nry * PQ[buf][0][yhi][tx]);
        const double W = wa * a + wb * b + nly * PQ[buf][1][ylo][tx] + nry * PQ[buf][1][yhi][tx];
        // x-pass: this node's share of its x neighbours, moved by one lane
        const double S = -c.kx * U + c.mx * W;
        const double yv = kxc * U + mxc * W + nlx * wave_from_left(S) + nrx * wave_from_right(S);
        if (own) {
            y[(long)k * plane + col] = yv;
            pq += v1 * yv;
        }
        v0 = v1;
        v1 = v2;
        buf ^= 1;
    }
    if (mode == MV_PCG) {
        const double vsum = block_sum<NT2>(pq, red);
        if (threadIdx.x == 0) partials[blockIdx.x] = vsum;
    }
}

inline int helm_sf_tiles(const GridArgs& g, int* ntx, int* nty) {
    *ntx = (g.NX + kHsfX - 3) / (kHsfX - 2);
    *nty = (g.NY + kHsfY - 3) / (kHsfY - 2);
    return *ntx * *nty * ((g.NZ + kHsfKc - 1) / kHsfKc);
}

inline int helm_tiles(const GridArgs& g, int* ntx, int* nty) {
    *ntx = (g.NX + kHx - 1) / kHx;
    *nty = (g.NY + kHy - 1) / kHy;
    return *ntx * *nty * ((g.NZ + kHkc - 1) / kHkc);
}

// alpha = rz / p'q with the breakdown test (GPUSolver.jl:462-466)
__global__ __launch_bounds__(NT) void k_fin_alpha(PcgState* st0, const double* partials0, unsigned nb) {
    __shared__ double red[NT / 64];
    PcgState* st = st0 + blockIdx.x;
    if (st->done) return;
    const double s = reduce_partials<NT>(partials0 + (size_t)blockIdx.x * nb, 1, 0, nb, red);
    if (threadIdx.x == 0) {
        st->pq = s;
        if (!(fabs(s) >= 1e-300)) {  // also catches NaN
            st->done = 1;
            st->status = 4;
        } else {
            st->alpha = st->rz / s;
        }
    }
}

// GPUHelmholtz.jl:78-89
__global__ __launch_bounds__(NT) void k_helm_extract(GridArgs g, const double* __restrict__ xn,
                                                     double* __restrict__ out) {
    const long e = gtid();
    if (e >= g.ne) return;
    const int i = (int)(e % g.nx), j = (int)((e / g.nx) % g.ny), k = (int)(e / ((long)g.nx * g.ny));
    double s = 0.0;
#pragma unroll
    for (int L = 0; L < 8; ++L)
        s += xn[(long)(i + (L & 1)) + (long)g.NX * ((j + ((L >> 1) & 1)) + (long)g.NY * (k + (L >> 2)))];
    out[e] = s / 8.0;
}

// ------------------------------------------------------------ element post
// cyclic Jacobi for a symmetric 3x3 (A row-major); vals descending, V columns.
// The rotations of one sweep, (0,1), (0,2), (1,2), are written out with
// constant indices, so A and V stay in registers.
template <int P, int Q>
__device__ __forceinline__ void jacobi_rot(double (&A)[3][3], double (&V)[3][3]) {
    if (A[P][Q] == 0.0) return;
    const double th = (A[Q][Q] - A[P][P]) / (2.0 * A[P][Q]);
    const double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
    const double c = 1.0 / sqrt(t * t + 1.0), sn = t * c;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double akp = A[k][P], akq = A[k][Q];
        A[k][P] = c * akp - sn * akq;
        A[k][Q] = sn * akp + c * akq;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double apk = A[P][k], aqk = A[Q][k];
        A[P][k] = c * apk - sn * aqk;
        A[Q][k] = sn * apk + c * aqk;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double vkp = V[k][P], vkq = V[k][Q];
        V[k][P] = c * vkp - sn * vkq;
        V[k][Q] = sn * vkp + c * vkq;
    }
}
__device__ void eig3_jacobi(double (&A)[3][3], double vals[3], double vec[3]) {
    double V[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int sweep = 0; sweep < 50; ++sweep) {
        const double off = fabs(A[0][1]) + fabs(A[0][2]) + fabs(A[1][2]);
        const double scale = fabs(A[0][0]) + fabs(A[1][1]) + fabs(A[2][2]) + off;
        if (off <= 1e-300 || off <= 1e-17 * scale) break;
        jacobi_rot<0, 1>(A, V);
        jacobi_rot<0, 2>(A, V);
        jacobi_rot<1, 2>(A, V);
    }
    // descending order of the diagonal: a 3-element sorting network on
    // (value, column) pairs, the earlier column first on ties
    double ev[3] = {A[0][0], A[1][1], A[2][2]};
    int idx[3] = {0, 1, 2};
    auto cswap = [&](int a, int b) {
        if (ev[b] > ev[a]) {
            const double tv = ev[a];
            ev[a] = ev[b];
            ev[b] = tv;
            const int ti = idx[a];
            idx[a] = idx[b];
            idx[b] = ti;
        }
    };
    cswap(0, 1);
    cswap(0, 2);
    cswap(1, 2);
#pragma unroll
    for (int a = 0; a < 3; ++a) vals[a] = ev[a];
#pragma unroll
    for (int k = 0; k < 3; ++k) vec[k] = idx[0] == 0 ? V[k][0] : idx[0] == 1 ? V[k][1] : V[k][2];
}

// Stress.jl:10-100 (centroid stress, von Mises, principal, L1, direction) and
// the SIMP energy / sensitivity, all from one read of the element's 24 DOFs.
// EIG: principal stresses, L1 or direction requested (the eigen solver's
// registers are then part of the kernel; without it, as in the TO loop, the
// kernel keeps more waves resident).
template <bool EIG>
__global__ __launch_bounds__(NT) void k_element_post(GridArgs g, ParityK<double> K, ElemArgs a) {
    __shared__ double red[NT / 64];
    const long e = gtid();
    double vmv = 0.0;
    if (e < g.ne) {
        const int i = (int)(e % g.nx), j = (int)((e / g.nx) % g.ny), k = (int)(e / ((long)g.nx * g.ny));
        double xh[24];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double t[8];
#pragma unroll
            for (int L = 0; L < 8; ++L) {
                const long m = (long)(i + (L & 1)) + (long)g.NX * ((j + ((L >> 1) & 1)) + (long)g.NY * (k + (L >> 2)));
                t[L] = a.u[3 * m + c];
            }
            fwd8(t);
#pragma unroll
            for (int s = 0; s < 8; ++s) xh[3 * s + c] = t[s];
        }
        if (a.ce || a.dc) {
            double yh[24];
#pragma unroll
            for (int r = 0; r < 24; ++r) yh[r] = 0.0;
#pragma unroll
            for (int t = 0; t < kParityNnz; ++t)  // Kp symmetric: its 33 distinct values
                yh[kParityPattern[t].r] += K.v[kParitySym.v[t]] * xh[kParityPattern[t].c];
            double ce = 0.0;
#pragma unroll
            for (int r = 3; r < 24; ++r) ce += xh[r] * yh[r];
            if (a.ce) a.ce[e] = ce;
            if (a.dc) {
                const double xp = a.xphys[e], pm1 = a.penal - 1.0;
                // integer exponents 1 and 2 (penal 2, 3) by products: FP64 pow is a long routine
                const double xw = pm1 == 2.0 ? xp * xp : pm1 == 1.0 ? xp : pow(xp, pm1);
                a.dc[e] = -a.penal * xw * (a.E0 - a.Emin) * ce;
            }
        }
        if (a.Emod) {
            const double Em = a.Emod[e];
            double sv[6] = {0, 0, 0, 0, 0, 0}, pr[3] = {0, 0, 0}, dv[3] = {0, 0, 0}, l1 = 0.0;
            if (Em != 0.0) {
                const double h[3] = {g.dx, g.dy, g.dz};
                double gr[3][3];  // gr[a][c] = d u_c / d x_a at the centroid
#pragma unroll
                for (int ax = 0; ax < 3; ++ax)
#pragma unroll
                    for (int c = 0; c < 3; ++c)  // 1/(4h) hoisted: a division is ~10 FP64 instructions
                        gr[ax][c] = xh[3 * (1 << ax) + c] * (0.25 / h[ax]);
                const double ex = gr[0][0], ey = gr[1][1], ez = gr[2][2];
                const double gxy = gr[1][0] + gr[0][1], gyz = gr[2][1] + gr[1][2], gxz = gr[2][0] + gr[0][2];
                sv[0] = Em * (a.lam1 * ex + a.lam2 * (ey + ez));
                sv[1] = Em * (a.lam1 * ey + a.lam2 * (ex + ez));
                sv[2] = Em * (a.lam1 * ez + a.lam2 * (ex + ey));
                sv[3] = Em * a.mu * gxy;
                sv[4] = Em * a.mu * gyz;
                sv[5] = Em * a.mu * gxz;
                vmv = sqrt(0.5 * ((sv[0] - sv[1]) * (sv[0] - sv[1]) + (sv[1] - sv[2]) * (sv[1] - sv[2]) +
                                  (sv[2] - sv[0]) * (sv[2] - sv[0])) +
                           3.0 * (sv[3] * sv[3] + sv[4] * sv[4] + sv[5] * sv[5]));
                if constexpr (EIG) {
                    double S[3][3] = {{sv[0], sv[3], sv[5]}, {sv[3], sv[1], sv[4]}, {sv[5], sv[4], sv[2]}};
                    eig3_jacobi(S, pr, dv);
                    l1 = fabs(pr[0]) + fabs(pr[1]) + fabs(pr[2]);
                }
            }
            if (a.vm) a.vm[e] = vmv;
            if (a.voigt)
                for (int d = 0; d < 6; ++d) a.voigt[6 * e + d] = sv[d];
            if (a.principal)
                for (int d = 0; d < 3; ++d) a.principal[3 * e + d] = pr[d];
            if (a.dir)
                for (int d = 0; d < 3; ++d) a.dir[3 * e + d] = dv[d];
            if (a.l1) a.l1[e] = l1;
        }
    }
    if (a.vm_max_partials) {
        const double m = block_max<NT>(vmv, red);
        if (threadIdx.x == 0) a.vm_max_partials[blockIdx.x] = m;
    }
}

__global__ __launch_bounds__(NT) void k_max_reduce(const double* __restrict__ in, int nin, double* out) {
    __shared__ double red[NT / 64];
    double m = -INFINITY;
    for (int i = threadIdx.x; i < nin; i += NT) m = fmax(m, in[i]);
    m = block_max<NT>(m, red);
    if (threadIdx.x == 0) *out = m;
}

__global__ __launch_bounds__(NT) void k_simp_modulus(long ne, const double* __restrict__ x, double penal, double E0,
                                                     double Emin, double* __restrict__ E) {
    for (long e = gtid(); e < ne; e += gstride()) E[e] = Emin + pow(x[e], penal) * (E0 - Emin);
}

__device__ __forceinline__ double oc_candidate(double x, double dc, double dv, double lmid, double move,
                                               double xmin) {
    const double be = -dc / (dv * lmid);
    const double cand = x * sqrt(be > 0.0 ? be : 0.0);
    const double lo = fmax(xmin, x - move), hi = fmin(1.0, x + move);
    return fmin(fmax(cand, lo), hi);
}

// [l1, l2] = [0, 1e9] (top88), or with lguess > 0 the guess [lguess/16,
// 16 lguess] for the 8-point search, whose first pass widens the bracket to
// 0 or 1e9 on the side where the volume does not cross inside it (st->oc_widen)
__global__ void k_oc_init(PcgState* st, double lguess) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const bool warm = lguess > 0.0;
        st->scal[0] = warm ? lguess / 16.0 : 0.0;  // l1
        st->scal[1] = warm ? lguess * 16.0 : 1e9;  // l2
        st->scal[2] = 0.0;  // lmid
        st->scal[3] = 0.0;  // change
        st->iter = 0;
        st->done = 0;
        st->oc_widen = warm ? 1 : 0;
        for (int i = 0; i < 4; ++i) st->ticket[i] = 0;
    }
}

// one bisection step of top88's OC (lines 55-64)
__global__ __launch_bounds__(NT) void k_oc_eval(long ne, const double* __restrict__ x, const double* __restrict__ dc,
                                                const double* __restrict__ dv, double volfrac, double move,
                                                double xmin, double rel_tol, PcgState* st, double* partials) {
    __shared__ double red[NT / 64];
    if (st->done) return;
    const double lmid = 0.5 * (st->scal[0] + st->scal[1]);
    double s = 0.0;
    for (long e = gtid(); e < ne; e += gstride()) s += oc_candidate(x[e], dc[e], dv[e], lmid, move, xmin);
    const double v = block_sum<NT>(s, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = v;
}

__global__ __launch_bounds__(NT) void k_fin_oc_eval(PcgState* st, const double* partials, unsigned nb, long ne,
                                                    double volfrac, double rel_tol) {
    __shared__ double red[NT / 64];
    if (st->done) return;
    const double t = reduce_partials<NT>(partials, 1, 0, nb, red);
    if (threadIdx.x == 0) {
        const double lmid = 0.5 * (st->scal[0] + st->scal[1]);
        if (t > volfrac * (double)ne) st->scal[0] = lmid;
        else st->scal[1] = lmid;
        st->scal[2] = lmid;
        st->iter += 1;
        const double l1 = st->scal[0], l2 = st->scal[1];
        if ((l2 - l1) / (l1 + l2) <= rel_tol || st->iter >= 200) st->done = 1;
    }
}

// The device-resident loop's search for the same multiplier: per pass the
// volume at kOcK equally spaced interior points of [l1, l2] (one read of x,
// dc, dv; the candidate is oc_candidate's, evaluated as x sqrt(-dc/dv) times
// 1/sqrt(lambda): one square root per element, not per point), and the
// bracket shrinks to the interval where the volume crosses the target,
// (kOcK + 1)x per pass against 2x per bisection step; same stopping rule.
constexpr int kOcK = 8;
__device__ __forceinline__ double oc_point(double l1, double l2, int j) {
    return l1 + (l2 - l1) * (double)j / (double)(kOcK + 1);
}
__global__ __launch_bounds__(NT) void k_oc_eval_k(long ne, const double* __restrict__ x,
                                                  const double* __restrict__ dc, const double* __restrict__ dv,
                                                  double move, double xmin, const PcgState* st, double* partials) {
    __shared__ double red[NT / 64];
    if (st->done) return;
    const double l1 = st->scal[0], l2 = st->scal[1];
    double isl[kOcK], sum[kOcK];  // 1/sqrt(lambda_j) (lambda_j > 0: l2 > l1 >= 0)
#pragma unroll
    for (int j = 0; j < kOcK; ++j) {
        isl[j] = 1.0 / sqrt(oc_point(l1, l2, j + 1));
        sum[j] = 0.0;
    }
    for (long e = gtid(); e < ne; e += gstride()) {
        const double xe = x[e], be = -dc[e] / dv[e];
        const double b = xe * sqrt(be > 0.0 ? be : 0.0);
        const double lo = fmax(xmin, xe - move), hi = fmin(1.0, xe + move);
#pragma unroll
        for (int j = 0; j < kOcK; ++j) sum[j] += fmin(fmax(b * isl[j], lo), hi);
    }
#pragma unroll
    for (int j = 0; j < kOcK; ++j) {  // partials of candidate j at [j nb + block]
        const double v = block_sum<NT>(sum[j], red);
        if (threadIdx.x == 0) partials[(long)j * gridDim.x + blockIdx.x] = v;
    }
}

__global__ __launch_bounds__(NT) void k_fin_oc_eval_k(PcgState* st, const double* partials, unsigned nb, long ne,
                                                      double volfrac, double rel_tol) {
    __shared__ double red[NT / 64];
    if (st->done) return;
    double t[kOcK];
#pragma unroll
    for (int j = 0; j < kOcK; ++j) t[j] = reduce_partials<NT>(partials + (long)j * nb, 1, 0, nb, red);
    if (threadIdx.x == 0) {
        const double l1 = st->scal[0], l2 = st->scal[1], target = volfrac * (double)ne;
        int j = 0;  // the first candidate whose volume is within the target
        while (j < kOcK && t[j] > target) ++j;
        double lo = oc_point(l1, l2, j), hi = j == kOcK ? l2 : oc_point(l1, l2, j + 1);
        if (st->oc_widen) {  // first pass from a guessed bracket: the crossing may lie outside it
            if (j == 0) lo = 0.0;
            if (j == kOcK) hi = 1e9;
            st->oc_widen = 0;
        }
        st->scal[0] = lo;
        st->scal[1] = hi;
        st->scal[2] = 0.5 * (lo + hi);
        st->iter += 1;
        if ((hi - lo) / (hi + lo) <= rel_tol || st->iter >= 200) st->done = 1;
    }
}

__global__ __launch_bounds__(NT) void k_oc_apply(long ne, const double* __restrict__ x,
                                                 const double* __restrict__ dc, const double* __restrict__ dv,
                                                 double move, double xmin, double* __restrict__ xnew, PcgState* st,
                                                 double* partials) {
    __shared__ double red[NT / 64];
    const double lmid = st->scal[2];
    double ch = 0.0;
    for (long e = gtid(); e < ne; e += gstride()) {
        const double v = oc_candidate(x[e], dc[e], dv[e], lmid, move, xmin);
        ch = fmax(ch, fabs(v - x[e]));
        xnew[e] = v;
    }
    const double v = block_max<NT>(ch, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = v;
}

__global__ __launch_bounds__(NT) void k_fin_max(double* out, const double* partials, unsigned nb) {
    __shared__ double red[NT / 64];
    const double m = reduce_partials_max<NT>(partials, nb, red);
    if (threadIdx.x == 0) *out = fmax(m, 0.0);
}

// ---- flexible PCG (variable preconditioner z = P(r)) -----------------------
// Flexible CG over the two-direction preconditioner of the learned solve
// (truncated FCG): z = a u + b y, then p = z - sum_j c_j P_j and
// Ap = Az - sum_j c_j AP_j over the last m stored directions (A p from the
// stored images, no operator application)
// Fused forms of the learned flexible CG's vector work (one pass each):
// k_fcg_dots: the iteration's dot products, out[0..4] = u'Au, u'Ay, y'Ay, u'r,
// y'r (W = [u | y], AW = [Au | Ay]) and out[5 + 2j], out[6 + 2j] = Au'P_j,
// Ay'P_j over the cf.m stored directions, as per-block partials
// [block][kFcgDots] that k_fin_fcg_dots reduces (one block per dot).
__global__ __launch_bounds__(NT) void k_fcg_dots(long n, const double* __restrict__ W, const double* __restrict__ AW,
                                                 const double* __restrict__ r, FcgCoef cf,
                                                 const double* __restrict__ P, double* __restrict__ partials) {
    __shared__ double red[NT / 64];
    double s[kFcgDots];
#pragma unroll
    for (int d = 0; d < kFcgDots; ++d) s[d] = 0.0;
    const double* Pj[kFcgKeep - 1];
#pragma unroll
    for (int j = 0; j < kFcgKeep - 1; ++j) Pj[j] = P + (long)cf.slot[j < cf.m ? j : 0] * n;
    for (long i = gtid(); i < n; i += gstride()) {
        const double u = W[i], y = W[n + i], au = AW[i], ay = AW[n + i], ri = r[i];
        s[0] += u * au;
        s[1] += u * ay;
        s[2] += y * ay;
        s[3] += u * ri;
        s[4] += y * ri;
#pragma unroll
        for (int j = 0; j < kFcgKeep - 1; ++j)
            if (j < cf.m) {
                const double pj = Pj[j][i];
                s[5 + 2 * j] += au * pj;
                s[6 + 2 * j] += ay * pj;
            }
    }
#pragma unroll
    for (int d = 0; d < kFcgDots; ++d) {
        const double v = block_sum<NT>(s[d], red);
        if (threadIdx.x == 0) partials[(long)blockIdx.x * kFcgDots + d] = v;
    }
}
__global__ __launch_bounds__(NT) void k_fin_fcg_dots(const double* partials, unsigned nb, int ndot, double* out) {
    __shared__ double red[NT / 64];
    const int d = blockIdx.x;
    if (d >= ndot) return;
    const double t = reduce_partials<NT>(partials, kFcgDots, d, nb, red);
    if (threadIdx.x == 0) out[d] = t;
}

// k_fcg_direction2: the new direction p and Ap, with p'Ap and p'r; its
// finaliser leaves scal[0] = p'Ap, scal[1] = p'r, scal[2] = alpha = p'r / p'Ap
// (0 and scal[3] = 1 when either is not positive: breakdown, the step does
// nothing) for k_fcg_step2, which then forms r'r (scal[4])
__global__ __launch_bounds__(NT) void k_fcg_direction2(long n, double a, double b, const double* __restrict__ W,
                                                       const double* __restrict__ AW, FcgCoef cf,
                                                       const double* __restrict__ P, const double* __restrict__ AP,
                                                       const double* __restrict__ r, double* __restrict__ p,
                                                       double* __restrict__ Ap, double* __restrict__ partials) {
    __shared__ double red[NT / 64];
    double pap = 0.0, pr = 0.0;
    for (long i = gtid(); i < n; i += gstride()) {
        double pi = a * W[i] + b * W[n + i];
        double api = a * AW[i] + b * AW[n + i];
        for (int j = 0; j < cf.m; ++j) {
            pi -= cf.c[j] * P[(long)cf.slot[j] * n + i];
            api -= cf.c[j] * AP[(long)cf.slot[j] * n + i];
        }
        p[i] = pi;
        Ap[i] = api;
        pap += pi * api;
        pr += pi * r[i];
    }
    const double v0 = block_sum<NT>(pap, red);
    const double v1 = block_sum<NT>(pr, red);
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = v0;
        partials[2 * blockIdx.x + 1] = v1;
    }
}
__global__ __launch_bounds__(NT) void k_fin_fcg_alpha(const double* partials, unsigned nb, double* scal) {
    __shared__ double red[NT / 64];
    double v[2];
    reduce_partials_n<NT, 2>(partials, 2, nb, red, v);
    if (threadIdx.x == 0) {
        const bool ok = v[0] > 0.0 && v[1] > 0.0;  // also false for NaN
        scal[0] = v[0];
        scal[1] = v[1];
        scal[2] = ok ? v[1] / v[0] : 0.0;
        scal[3] = ok ? 0.0 : 1.0;
    }
}
// x += alpha p ; r = free ? r - alpha Ap : 0 with alpha = scal[2]; r'r partials
__global__ __launch_bounds__(NT) void k_fcg_step2(long n, const double* __restrict__ scal, double* __restrict__ x,
                                                  const double* __restrict__ p, double* __restrict__ r,
                                                  const double* __restrict__ Ap, const double* __restrict__ minv,
                                                  double* __restrict__ partials) {
    __shared__ double red[NT / 64];
    const double alpha = scal[2];
    double rr = 0.0;
    for (long i = gtid(); i < n; i += gstride()) {
        x[i] += alpha * p[i];
        const double ri = minv[i] != 0.0 ? r[i] - alpha * Ap[i] : 0.0;
        r[i] = ri;
        rr += ri * ri;
    }
    const double v = block_sum<NT>(rr, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = v;
}
__global__ __launch_bounds__(NT) void k_fin_fcg_rr(const double* partials, unsigned nb, double* scal) {
    __shared__ double red[NT / 64];
    const double t = reduce_partials<NT>(partials, 1, 0, nb, red);
    if (threadIdx.x == 0) scal[4] = t;
}

// e = a - b (masked by minv != 0)
__global__ __launch_bounds__(NT) void k_masked_diff(long n, const double* __restrict__ a, const double* __restrict__ b,
                                                    const double* __restrict__ minv, double* __restrict__ e) {
    for (long i = gtid(); i < n; i += gstride()) e[i] = minv[i] != 0.0 ? a[i] - b[i] : 0.0;
}

__global__ __launch_bounds__(NT) void k_mul(long n, const double* __restrict__ a, const double* __restrict__ b,
                                            double* __restrict__ out) {
    for (long i = gtid(); i < n; i += gstride()) out[i] = a[i] * b[i];
}

// Best-iterate snapshot (GPUSolver.jl:492-499): xbest = x when the state's
// residual is below 1e-2 and below the last snapshot's.  The state is only
// read here, so every block takes the same decision; k_snapshot_fin records
// the snapshot's residual afterwards.
__global__ __launch_bounds__(NT) void k_snapshot(long n, const double* __restrict__ x, double* __restrict__ xbest,
                                                 const PcgState* st) {
    const double rel = st->rel;
    if (!(rel < 1e-2 && rel < st->snap_rel)) return;
    typedef Vec16<double>::type VT;
    const long nv = n / 2;
    for (long v = gtid(); v < nv; v += gstride()) ((VT*)xbest)[v] = ((const VT*)x)[v];
    for (long i = 2 * nv + gtid(); i < n; i += gstride()) xbest[i] = x[i];
}
__global__ void k_snapshot_fin(PcgState* st) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const double rel = st->rel;
    if (rel < 1e-2 && rel < st->snap_rel) st->snap_rel = rel;
}
// end of a per-iteration-snapshot solve: a best iterate flagged by the last
// finaliser had no further update kernel to copy it
__global__ __launch_bounds__(NT) void k_snapshot_pending(long n, const double* __restrict__ x,
                                                         double* __restrict__ xbest, const PcgState* st) {
    if (!st->snap_pending) return;
    for (long i = gtid(); i < n; i += gstride()) xbest[i] = x[i];
}
__global__ void k_snapshot_clear(PcgState* st) {
    if (blockIdx.x == 0 && threadIdx.x == 0) st->snap_pending = 0;
}

// Element selection of a solve (GPUSolver.jl:128-129, Solver.jl:60-67):
// per block, the count of rho >= thr (active) and of rho > thr_keep (kept
// by the void pruning)
__global__ __launch_bounds__(NT) void k_count_active(long ne, const double* __restrict__ rho, double thr,
                                                     double thr_keep, double* partials) {
    __shared__ double red[NT / 64];
    double na = 0.0, nk = 0.0;
    for (long e = gtid(); e < ne; e += gstride()) {
        const double r = rho[e];
        na += r >= thr ? 1.0 : 0.0;
        nk += r > thr_keep ? 1.0 : 0.0;
    }
    const double v0 = block_sum<NT>(na, red);
    const double v1 = block_sum<NT>(nk, red);
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x + 0] = v0;
        partials[2 * blockIdx.x + 1] = v1;
    }
}
__global__ __launch_bounds__(NT) void k_fin_sum2(const double* partials, unsigned nb, double* out) {
    __shared__ double red[NT / 64];
    double sv[2];
    reduce_partials_n<NT, 2>(partials, 2, nb, red, sv);
    if (threadIdx.x == 0) {
        out[0] = sv[0];
        out[1] = sv[1];
    }
}
// moduli E rho of the elements in the operator: rho >= thr and, with the
// pruning, rho > thr_keep (the linear factor of GPUSolver.jl:147); 0 removes
// the element, and the DOFs it alone carried leave the system (zero diagonal)
__global__ __launch_bounds__(NT) void k_solve_moduli(long ne, const double* __restrict__ rho, double E, double thr,
                                                     double thr_keep, int prune, double* __restrict__ Eout) {
    for (long e = gtid(); e < ne; e += gstride()) {
        const double r = rho[e];
        const bool on = r >= thr && (!prune || r > thr_keep);
        Eout[e] = on ? E * r : 0.0;
    }
}
// per block, the largest |diag| over free DOFs (GPUSolver.jl:315-316)
__global__ __launch_bounds__(NT) void k_free_absmax(long n, const double* __restrict__ diag,
                                                    const uint8_t* __restrict__ fixed, double* partials) {
    __shared__ double red[NT / 64];
    double m = 0.0;
    for (long i = gtid(); i < n; i += gstride())
        if (fixed == nullptr || fixed[i] == 0) m = fmax(m, fabs(diag[i]));
    const double v = block_max<NT>(m, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = v;
}

inline unsigned blocks_for(long n) { return (unsigned)((n + NT - 1) / NT); }

}  // namespace

// ------------------------------------------------------------- launchers
void launch_pcg_reset(PcgState* st, double tol, double shift, int max_iter, int check_every, hipStream_t s,
                      int nrhs, int snap_mode) {
    hipLaunchKernelGGL(k_pcg_reset, dim3(1), dim3(64), 0, s, st, tol, shift, max_iter, check_every, nrhs, snap_mode);
}
template <typename T>
void launch_pcg_init(long n, const T* b, T* x, const T* q, const T* minv, T* r, T* p, double shift, PcgState* st,
                     double* partials, hipStream_t s, int nrhs) {
    hipLaunchKernelGGL((k_pcg_init<T>), dim3(kVecBlocks, nrhs), dim3(NT), 0, s, n, b, x, q, minv, r, p, shift, st,
                       partials);
    hipLaunchKernelGGL(k_fin_init, dim3(nrhs), dim3(NT), 0, s, st, partials, (unsigned)kVecBlocks);
}
template <typename T>
void launch_pcg_update(long n, T* x, const T* p, T* r, const T* q, const T* minv, PcgState* st, double* partials,
                       hipStream_t s, int nrhs, bool set_beta, const double* pq_partials, int npq, T* xbest) {
    hipLaunchKernelGGL((k_pcg_update<T>), dim3(kVecBlocks, nrhs), dim3(NT), 0, s, n, x, p, r, q, minv, st, partials,
                       pq_partials, (unsigned)(pq_partials ? npq : 0), xbest);
    // without set_beta the explicit preconditioner's finaliser reads these
    // partials (r'r at index 2b+1) after its own r'z and finishes the iteration
    if (set_beta)
        hipLaunchKernelGGL(k_fin_update, dim3(nrhs), dim3(NT), 0, s, st, partials, (unsigned)kVecBlocks, 1);
}
template <typename T>
void launch_pcg_update_free(long n, T* x, const T* p, T* r, const T* q, PcgState* st, const double* pq_partials,
                            int npq, hipStream_t s, T* xbest, int nrhs) {
    hipLaunchKernelGGL((k_pcg_update_free<T>), dim3(kVecBlocks, nrhs), dim3(NT), 0, s, n, x, p, r, q, st, pq_partials,
                       (unsigned)(pq_partials ? npq : 0), xbest);
}
template <typename T>
void launch_pcg_update_r(long n, T* r, const T* q, PcgState* st, const double* pq_partials, int npq, hipStream_t s,
                         int nrhs) {
    hipLaunchKernelGGL((k_pcg_update_r<T>), dim3(kVecBlocks, nrhs), dim3(NT), 0, s, n, r, q, st, pq_partials,
                       (unsigned)(pq_partials ? npq : 0));
}
template <typename T>
void launch_pcg_pupdate(long n, const T* r, const T* minv, T* p, const PcgState* st, hipStream_t s, int nrhs) {
    hipLaunchKernelGGL((k_pcg_pupdate<T>), dim3(kVecBlocks, nrhs), dim3(NT), 0, s, n, r, minv, p, st);
}
void launch_copy16(long n, const void* src, void* dst, hipStream_t s) {
    hipLaunchKernelGGL(k_copy16, dim3(4 * kVecBlocks), dim3(NT), 0, s, n, (const uint4*)src, (uint4*)dst);
}
void launch_dot(long n, const double* a, const double* b, double* out, unsigned* ticket, double* partials,
                hipStream_t s) {
    (void)ticket;
    hipLaunchKernelGGL(k_dot, dim3(kVecBlocks), dim3(NT), 0, s, n, a, b, partials);
    hipLaunchKernelGGL(k_fin_dot, dim3(1), dim3(NT), 0, s, out, partials, (unsigned)kVecBlocks);
}
void launch_fill(long n, double v, double* a, hipStream_t s) {
    hipLaunchKernelGGL(k_fill, dim3(kVecBlocks), dim3(NT), 0, s, n, v, a);
}
template <typename T>
void launch_rz_park(long n, const T* r, const T* z, PcgState* st, double* partials, hipStream_t s, int nrhs) {
    hipLaunchKernelGGL((k_rz_part<T>), dim3(kVecBlocks, nrhs), dim3(NT), 0, s, n, r, z, partials);
    hipLaunchKernelGGL(k_fin_rz_park, dim3(nrhs), dim3(NT), 0, s, st, partials, (unsigned)kVecBlocks);
}
template void launch_rz_park<float>(long, const float*, const float*, PcgState*, double*, hipStream_t, int);
template void launch_rz_park<double>(long, const double*, const double*, PcgState*, double*, hipStream_t, int);
void launch_fill32(long n, float v, float* a, hipStream_t s) {
    hipLaunchKernelGGL(k_fill32, dim3(kVecBlocks), dim3(NT), 0, s, n, v, a);
}
template <typename Tin, typename Tout>
void launch_convert(long n, const Tin* a, Tout* b, hipStream_t s) {
    hipLaunchKernelGGL((k_convert<Tin, Tout>), dim3(kVecBlocks), dim3(NT), 0, s, n, a, b);
}
void launch_true_residual(long n, const double* b, const double* q, const double* x, const double* minv64,
                          double shift, double* r64, float* r32, PcgState* st, double* partials, int mode,
                          hipStream_t s, int nrhs) {
    hipLaunchKernelGGL(k_true_residual, dim3(kVecBlocks, nrhs), dim3(NT), 0, s, n, b, q, x, minv64, shift, r64, r32, st,
                       partials, mode);
    hipLaunchKernelGGL(k_fin_true, dim3(nrhs), dim3(NT), 0, s, st, partials, (unsigned)kVecBlocks, mode);
}
template <typename T>
void launch_fix_p(long n, T* p, const T* pb, const PcgState* st, hipStream_t s, int nrhs) {
    hipLaunchKernelGGL((k_fix_p<T>), dim3(kVecBlocks, nrhs), dim3(NT), 0, s, n, p, pb, st);
}
template void launch_fix_p<float>(long, float*, const float*, const PcgState*, hipStream_t, int);
template void launch_fix_p<double>(long, double*, const double*, const PcgState*, hipStream_t, int);
void launch_ir_resume(PcgState* st, double tol_final, double delta, int first, hipStream_t s, int nrhs) {
    hipLaunchKernelGGL(k_ir_resume, dim3(1), dim3(64), 0, s, st, tol_final, delta, first, nrhs);
}
template <typename T>
void launch_mask(long n, const T* minv, T* x, hipStream_t s, int nrhs) {
    hipLaunchKernelGGL((k_mask<T>), dim3(kVecBlocks, nrhs), dim3(NT), 0, s, n, minv, x);
}
template void launch_mask<double>(long, const double*, double*, hipStream_t, int);
template void launch_mask<float>(long, const float*, float*, hipStream_t, int);
void launch_accumulate(long n, double* x64, float* x32, hipStream_t s, int nrhs) {
    hipLaunchKernelGGL(k_accumulate, dim3(kVecBlocks, nrhs), dim3(NT), 0, s, n, x64, x32);
}

void launch_snapshot_pending(long n, const double* x, double* xbest, PcgState* st, hipStream_t s) {
    hipLaunchKernelGGL(k_snapshot_pending, dim3(kVecBlocks), dim3(NT), 0, s, n, x, xbest, (const PcgState*)st);
    hipLaunchKernelGGL(k_snapshot_clear, dim3(1), dim3(64), 0, s, st);
}
void launch_snapshot(long n, const double* x, double* xbest, PcgState* st, hipStream_t s) {
    hipLaunchKernelGGL(k_snapshot, dim3(kVecBlocks), dim3(NT), 0, s, n, x, xbest, (const PcgState*)st);
    hipLaunchKernelGGL(k_snapshot_fin, dim3(1), dim3(64), 0, s, st);
}
void launch_count_active(long ne, const double* rho, double thr, double thr_keep, double* partials, double* out,
                         hipStream_t s) {
    hipLaunchKernelGGL(k_count_active, dim3(kVecBlocks), dim3(NT), 0, s, ne, rho, thr, thr_keep, partials);
    hipLaunchKernelGGL(k_fin_sum2, dim3(1), dim3(NT), 0, s, (const double*)partials, (unsigned)kVecBlocks, out);
}
void launch_solve_moduli(long ne, const double* rho, double E, double thr, double thr_keep, int prune, double* Eout,
                         hipStream_t s) {
    hipLaunchKernelGGL(k_solve_moduli, dim3(kVecBlocks), dim3(NT), 0, s, ne, rho, E, thr, thr_keep, prune, Eout);
}
void launch_free_absmax(long n, const double* diag, const uint8_t* fixed, double* partials, double* out,
                        hipStream_t s) {
    hipLaunchKernelGGL(k_free_absmax, dim3(kVecBlocks), dim3(NT), 0, s, n, diag, fixed, partials);
    launch_max_reduce(kVecBlocks, partials, kVecBlocks, out, s);
}

void launch_helm_rhs(const GridArgs& g, const double* in_e, double vol8, double* b, hipStream_t s) {
    hipLaunchKernelGGL(k_helm_rhs, dim3(blocks_for(g.nn)), dim3(NT), 0, s, g, in_e, vol8, b);
}
void launch_helm_minv(const GridArgs& g, const HelmArgs& A, double* minv, hipStream_t s) {
    hipLaunchKernelGGL(k_helm_minv, dim3(blocks_for(g.nn)), dim3(NT), 0, s, g, A, minv);
}
// the 1-D factors of A's box (HelmSF) when they reproduce all 64 entries of
// A.Ae to 1e-13 of the largest: entry (p, q) = R^2 (Kx My Mz + Mx Ky Mz +
// Mx My Kz) + Mx My Mz with the per-axis factor of "same coordinate" (K: k,
// M: 2m) or "different" (K: -k, M: m)
static bool helm_sf_params(const HelmArgs& A, HelmSF* c) {
    if (!(A.h[0] > 0.0 && A.h[1] > 0.0 && A.h[2] > 0.0)) return false;
    c->kx = 1.0 / A.h[0];
    c->ky = 1.0 / A.h[1];
    c->kz = 1.0 / A.h[2];
    c->mx = A.h[0] / 6.0;
    c->my = A.h[1] / 6.0;
    c->mz = A.h[2] / 6.0;
    c->R2 = A.R2;
    double amax = 0.0;
    for (int t = 0; t < 64; ++t) amax = std::max(amax, std::fabs(A.Ae[t]));
    for (int p = 0; p < 8; ++p)
        for (int q = 0; q < 8; ++q) {
            const bool sx = (p & 1) == (q & 1), sy = ((p >> 1) & 1) == ((q >> 1) & 1), sz = (p >> 2) == (q >> 2);
            const double Kx = sx ? c->kx : -c->kx, Ky = sy ? c->ky : -c->ky, Kz = sz ? c->kz : -c->kz;
            const double Mx = sx ? 2.0 * c->mx : c->mx, My = sy ? 2.0 * c->my : c->my, Mz = sz ? 2.0 * c->mz : c->mz;
            const double v = c->R2 * (Kx * My * Mz + Mx * Ky * Mz + Mx * My * Kz) + Mx * My * Mz;
            if (std::fabs(A.Ae[8 * kBinToRef[p] + kBinToRef[q]] - v) > 1e-13 * amax) return false;
        }
    return true;
}

bool helm_sf_available(const HelmArgs& A) {
    HelmSF c;
    return helm_sf_params(A, &c);
}
int helm_matvec_blocks(const GridArgs& g) {
    int a, b;
    return std::max(helm_tiles(g, &a, &b), helm_sf_tiles(g, &a, &b));
}
int launch_helm_matvec_fused(const GridArgs& g, const HelmArgs& A, const double* p_old, const double* r,
                             const double* minv, double* p_out, double* q, PcgState* st, double* partials,
                             hipStream_t s) {
    HelmSF c;
    if (!helm_sf_params(A, &c)) return -1;
    int ntx, nty;
    const int nb = helm_sf_tiles(g, &ntx, &nty);
    hipLaunchKernelGGL((k_helm_sf<true>), dim3(nb), dim3(kHsfX * kHsfY), 0, s, g, c, p_old, r, minv, p_out, q,
                       (int)MV_PCG, st, partials, ntx, nty);
    return nb;
}
int launch_helm_matvec(const GridArgs& g, const HelmArgs& A, const double* x, double* y, int mode, PcgState* st,
                       double* partials, hipStream_t s) {
    int ntx, nty;
    HelmSF c;
    if (helm_sf_params(A, &c)) {
        const int nb = helm_sf_tiles(g, &ntx, &nty);
        hipLaunchKernelGGL((k_helm_sf<false>), dim3(nb), dim3(kHsfX * kHsfY), 0, s, g, c, x, nullptr, nullptr,
                           nullptr, y, mode, st, partials, ntx, nty);
        return nb;
    }
    const int nb = helm_tiles(g, &ntx, &nty);
    // the eight box values, taken from node 0's row and checked against every entry
    HelmBox hb{};
    for (int sc = 0; sc < 8; ++sc) {
        const int b = ((sc & 1) ? 0 : 1) + 2 * ((sc & 2) ? 0 : 1) + 4 * ((sc & 4) ? 0 : 1);
        hb.c[sc] = A.Ae[8 * kBinToRef[0] + kBinToRef[b]];
    }
    double amax = 0.0;
    for (int t = 0; t < 64; ++t) amax = std::max(amax, std::fabs(A.Ae[t]));
    bool box = true;
    for (int p = 0; p < 8 && box; ++p)
        for (int q = 0; q < 8 && box; ++q) {
            const int sc = ((p & 1) == (q & 1)) + 2 * (((p >> 1) & 1) == ((q >> 1) & 1)) + 4 * ((p >> 2) == (q >> 2));
            box = std::fabs(A.Ae[8 * kBinToRef[p] + kBinToRef[q]] - hb.c[sc]) <= 1e-14 * amax;
        }
    if (box)
        hipLaunchKernelGGL((k_helm_matvec<true>), dim3(nb), dim3(NT), 0, s, g, A, hb, x, y, mode, st, partials, ntx,
                           nty);
    else
        hipLaunchKernelGGL((k_helm_matvec<false>), dim3(nb), dim3(NT), 0, s, g, A, hb, x, y, mode, st, partials,
                           ntx, nty);
    return nb;
}
void launch_fin_alpha(PcgState* st, const double* partials, int nblocks, int nrhs, hipStream_t s) {
    hipLaunchKernelGGL(k_fin_alpha, dim3(nrhs), dim3(NT), 0, s, st, partials, (unsigned)nblocks);
}
void launch_helm_extract(const GridArgs& g, const double* xn, double* out_e, hipStream_t s) {
    hipLaunchKernelGGL(k_helm_extract, dim3(blocks_for(g.ne)), dim3(NT), 0, s, g, xn, out_e);
}
int element_post_blocks(const GridArgs& g) { return (int)blocks_for(g.ne); }
void launch_element_post(const GridArgs& g, const ParityK<double>& K, const ElemArgs& a, hipStream_t s) {
    if (a.principal || a.l1 || a.dir)
        hipLaunchKernelGGL(k_element_post<true>, dim3(blocks_for(g.ne)), dim3(NT), 0, s, g, K, a);
    else
        hipLaunchKernelGGL(k_element_post<false>, dim3(blocks_for(g.ne)), dim3(NT), 0, s, g, K, a);
}
void launch_max_reduce(long n, const double* partials_in, int nin, double* out, hipStream_t s) {
    (void)n;
    hipLaunchKernelGGL(k_max_reduce, dim3(1), dim3(NT), 0, s, partials_in, nin, out);
}
void launch_simp_modulus(long ne, const double* xphys, double penal, double E0, double Emin, double* E,
                         hipStream_t s) {
    hipLaunchKernelGGL(k_simp_modulus, dim3(kVecBlocks), dim3(NT), 0, s, ne, xphys, penal, E0, Emin, E);
}
void launch_fcg_dots(long n, const double* W, const double* AW, const double* r, const FcgCoef& cf, const double* P,
                     double* partials, double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_fcg_dots, dim3(kVecBlocks), dim3(NT), 0, s, n, W, AW, r, cf, P, partials);
    hipLaunchKernelGGL(k_fin_fcg_dots, dim3(5 + 2 * cf.m), dim3(NT), 0, s, partials, (unsigned)kVecBlocks,
                       5 + 2 * cf.m, out);
}
void launch_fcg_direction2(long n, double a, double b, const double* W, const double* AW, const FcgCoef& cf,
                           const double* P, const double* AP, const double* r, double* p, double* Ap, double* partials,
                           double* scal, hipStream_t s) {
    hipLaunchKernelGGL(k_fcg_direction2, dim3(kVecBlocks), dim3(NT), 0, s, n, a, b, W, AW, cf, P, AP, r, p, Ap,
                       partials);
    hipLaunchKernelGGL(k_fin_fcg_alpha, dim3(1), dim3(NT), 0, s, partials, (unsigned)kVecBlocks, scal);
}
void launch_fcg_step2(long n, const double* scal_alpha, double* x, const double* p, double* r, const double* Ap,
                      const double* minv, double* partials, double* scal, hipStream_t s) {
    hipLaunchKernelGGL(k_fcg_step2, dim3(kVecBlocks), dim3(NT), 0, s, n, scal_alpha, x, p, r, Ap, minv, partials);
    hipLaunchKernelGGL(k_fin_fcg_rr, dim3(1), dim3(NT), 0, s, partials, (unsigned)kVecBlocks, scal);
}
void launch_mul(long n, const double* a, const double* b, double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_mul, dim3(kVecBlocks), dim3(NT), 0, s, n, a, b, out);
}
void launch_masked_diff(long n, const double* a, const double* b, const double* minv, double* e, hipStream_t s) {
    hipLaunchKernelGGL(k_masked_diff, dim3(kVecBlocks), dim3(NT), 0, s, n, a, b, minv, e);
}
void launch_oc_init(PcgState* st, hipStream_t s, double lguess) {
    hipLaunchKernelGGL(k_oc_init, dim3(1), dim3(64), 0, s, st, lguess);
}
void launch_oc_eval_k(long ne, const double* x, const double* dc, const double* dv, double volfrac, double move,
                      double xmin, double rel_tol, PcgState* st, double* partials, hipStream_t s) {
    hipLaunchKernelGGL(k_oc_eval_k, dim3(kVecBlocks), dim3(NT), 0, s, ne, x, dc, dv, move, xmin, st, partials);
    hipLaunchKernelGGL(k_fin_oc_eval_k, dim3(1), dim3(NT), 0, s, st, partials, (unsigned)kVecBlocks, ne, volfrac,
                       rel_tol);
}
void launch_oc_eval(long ne, const double* x, const double* dc, const double* dv, double volfrac, double move,
                    double xmin, double rel_tol, PcgState* st, double* partials, hipStream_t s) {
    hipLaunchKernelGGL(k_oc_eval, dim3(kVecBlocks), dim3(NT), 0, s, ne, x, dc, dv, volfrac, move, xmin, rel_tol, st,
                       partials);
    hipLaunchKernelGGL(k_fin_oc_eval, dim3(1), dim3(NT), 0, s, st, partials, (unsigned)kVecBlocks, ne, volfrac, rel_tol);
}
void launch_oc_apply(long ne, const double* x, const double* dc, const double* dv, double move, double xmin,
                     double* xnew, PcgState* st, double* partials, hipStream_t s) {
    hipLaunchKernelGGL(k_oc_apply, dim3(kVecBlocks), dim3(NT), 0, s, ne, x, dc, dv, move, xmin, xnew, st, partials);
    hipLaunchKernelGGL(k_fin_max, dim3(1), dim3(NT), 0, s, &st->scal[3], partials, (unsigned)kVecBlocks);
}

template void launch_pcg_init<double>(long, const double*, double*, const double*, const double*, double*, double*,
                                      double, PcgState*, double*, hipStream_t, int);
template void launch_pcg_init<float>(long, const float*, float*, const float*, const float*, float*, float*, double,
                                     PcgState*, double*, hipStream_t, int);
template void launch_pcg_update<double>(long, double*, const double*, double*, const double*, const double*,
                                        PcgState*, double*, hipStream_t, int, bool, const double*, int, double*);
template void launch_pcg_update<float>(long, float*, const float*, float*, const float*, const float*, PcgState*,
                                       double*, hipStream_t, int, bool, const double*, int, float*);
template void launch_pcg_update_free<double>(long, double*, const double*, double*, const double*, PcgState*,
                                             const double*, int, hipStream_t, double*, int);
template void launch_pcg_update_free<float>(long, float*, const float*, float*, const float*, PcgState*,
                                            const double*, int, hipStream_t, float*, int);
template void launch_pcg_update_r<float>(long, float*, const float*, PcgState*, const double*, int, hipStream_t,
                                         int);
template void launch_pcg_pupdate<double>(long, const double*, const double*, double*, const PcgState*, hipStream_t,
                                         int);
template void launch_pcg_pupdate<float>(long, const float*, const float*, float*, const PcgState*, hipStream_t, int);
template void launch_convert<double, float>(long, const double*, float*, hipStream_t);
template void launch_convert<float, double>(long, const float*, double*, hipStream_t);

}  // namespace hx
