// hx_elastic.hip -- matrix-free elasticity operator y = sum_e E_e Ke x_e on a
// structured hex grid, owner-computes (no atomics, deterministic).
//
// Replaces the reference's element-parallel kernel with 24 atomic adds per
// element (GPUSolver.jl:37-79) by a z-marching tile: a 256-thread workgroup
// owns a (TX-1)x(TY-1) column of nodes over a chunk of z planes.  Per element
// layer it stages one node plane of x in LDS (ring of two planes), every
// thread applies one element (parity basis, 234 flops) and parks its 24
// outputs in LDS, then every owned node sums the 4 elements below-and-above
// it in a fixed order and carries the upper half to the next layer.  Each
// node value of y is written exactly once; x and E are read ~1.2x (halo).
#include "hx_launch.h"

#include <type_traits>

namespace hx {

namespace {

template <typename T>
struct DenseK {
    const T* k;  // 24x24 binary-order, row-major, unit E
};

typedef float floatx4 __attribute__((ext_vector_type(4)));

// Element-operator modes of the element-local z-marching kernel (the parity
// mode there is the reference-shaped form kept for comparison; production
// parity runs in k_matvec_sf below)
enum { EL_PARITY = 0, EL_DENSE = 1, EL_MFMA = 2 };

template <typename T, int TX, int TY, int MODE, bool FUSE, int PD>
__global__ __launch_bounds__(TX* TY) void k_matvec_zmarch(GridArgs g, ParityK<T> K, DenseK<T> Kd, MvArgs<T> a,
                                                          int kchunk, int ntx, int nty, int ntiles) {
    constexpr int NT = TX * TY;
    constexpr int PX = TX + 1, PY = TY + 1;
    constexpr int PLANE = 3 * PX * PY;
    constexpr int NL = (PLANE + NT - 1) / NT;
    static_assert(MODE != EL_MFMA || (TX == 16 && TY == 16), "MFMA layout assumes a 16x16 element tile");
    __shared__ T xs[2][3][PY][PX];
    __shared__ T ys[24][TY][TX];
    __shared__ T es[TY][TX];  // E of the element layer being processed
    __shared__ double red[NT / 64];

    // multi-RHS: blockIdx.y selects the column (own x, y, state, partials)
    a.x += (long)blockIdx.y * a.cs;
    a.y += (long)blockIdx.y * a.cs;
    if constexpr (FUSE) {
        a.r += (long)blockIdx.y * a.cs;
        a.p_out += (long)blockIdx.y * a.cs;
    }
    if (a.mode == MV_PCG) {
        a.st += blockIdx.y;
        a.partials += (long)blockIdx.y * gridDim.x;
    }
    if (a.mode == MV_PCG && a.st->done) return;
    const double beta = FUSE ? a.st->beta : 0.0;

    // XCD-aware tile order: the 8 XCDs take blocks round-robin, so give each
    // XCD a contiguous range of tiles (neighbouring tiles share halo planes).
    const unsigned nb = gridDim.x;  // multiple of 8
    const unsigned tile = (blockIdx.x & 7u) * (nb >> 3) + (blockIdx.x >> 3);
    const bool active = tile < (unsigned)ntiles;

    const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    double pq = 0.0;

    if (active) {
        const int cx = tile % ntx, cy = (tile / ntx) % nty, cz = tile / (ntx * nty);
        const int i0 = cx * (TX - 1), j0 = cy * (TY - 1);
        const int k0 = cz * kchunk, k1 = min(k0 + kchunk, g.NZ);
        const int ie = i0 - 1 + tx, je = j0 - 1 + ty;
        const bool ecol = ie >= 0 && ie < g.nx && je >= 0 && je < g.ny;
        const bool nthr = threadIdx.x < (TX - 1) * (TY - 1);
        const int ox = threadIdx.x % (TX - 1), oy = threadIdx.x / (TX - 1);
        const int ni = i0 + ox, nj = j0 + oy;
        const bool nown = nthr && ni < g.NX && nj < g.NY;
        const long plane_stride = (long)g.NX * g.NY;

        // Plane loads are issued unconditionally from clamped addresses and
        // masked when parked in LDS: a load under a divergent branch makes the
        // compiler wait for it inside the branch, which serialises the
        // prefetch of plane k+2 with the element work of plane k.
        // Plane k + 1 + PD is requested while element layer k is processed: a
        // ring of PD register slots keeps PD planes in flight per block (the
        // slot of a plane is a compile-time index, see the unrolled loop).
        const int ie_c = min(max(ie, 0), g.nx - 1), je_c = min(max(je, 0), g.ny - 1);
        T rx[PD][NL], rr[PD][FUSE ? NL : 1], rm[PD][FUSE ? NL : 1], re[PD];
        unsigned valid[PD];
        auto plane_load = [&](int k, int q) {
            const bool kin = k >= 0 && k < g.NZ;
            const long kc = min(max(k, 0), g.NZ - 1);
            valid[q] = 0;
#pragma unroll
            for (int m = 0; m < NL; ++m) {
                const int idx = min(threadIdx.x + m * NT, (unsigned)PLANE - 1);
                const int row = idx / (3 * PX), col3 = idx - row * 3 * PX, cx3 = col3 / 3;
                const int ii = i0 - 1 + cx3, jj = j0 - 1 + row;
                const bool ok = kin && (int)(threadIdx.x + m * NT) < PLANE && ii >= 0 && ii < g.NX && jj >= 0 &&
                                jj < g.NY;
                const int iic = min(max(ii, 0), g.NX - 1), jjc = min(max(jj, 0), g.NY - 1);
                const long gi = 3 * (kc * plane_stride + (long)jjc * g.NX + iic) + (col3 - 3 * cx3);
                rx[q][m] = a.x[gi];
                if constexpr (FUSE) {
                    rr[q][m] = a.r[gi];
                    rm[q][m] = a.minv[gi];
                }
                valid[q] |= (unsigned)ok << m;
            }
            // E of element layer k-1 travels with plane k, so one wait covers both
            re[q] = a.E[(long)min(max(k - 1, 0), g.nz - 1) * g.nx * g.ny + (long)je_c * g.nx + ie_c];
        };
        auto plane_store = [&](int k, int q) {
            const int sl = k & 1;
#pragma unroll
            for (int m = 0; m < NL; ++m) {
                const int idx = threadIdx.x + m * NT;
                T v;
                if constexpr (FUSE) v = T((double)rm[q][m] * (double)rr[q][m] + beta * (double)rx[q][m]);
                else v = rx[q][m];
                v = ((valid[q] >> m) & 1u) ? v : T(0);
                if (idx < PLANE) {
                    const int row = idx / (3 * PX), col3 = idx - row * 3 * PX;
                    xs[sl][col3 % 3][row][col3 / 3] = v;
                }
            }
            es[ty][tx] = (ecol && k - 1 >= 0 && k - 1 < g.nz) ? re[q] : T(0);
        };

        // prologue: plane k0-1 straight to LDS, planes k0 .. k0-1+PD into the
        // ring (plane k0-1+j lives in slot j % PD)
        plane_load(k0 - 1, 0);
        plane_store(k0 - 1, 0);
#pragma unroll
        for (int j = 1; j <= PD; ++j)
            if (k0 - 1 + j <= k1) plane_load(k0 - 1 + j, j % PD);
        T carry[3] = {T(0), T(0), T(0)};

        // MFMA operand B = Ke^T (binary order), two 16-column blocks, 6 k-steps
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        const int li = lane & 15, lk = lane >> 4;
        T kb0[MODE == EL_MFMA ? 6 : 1], kb1[MODE == EL_MFMA ? 6 : 1];
        if constexpr (MODE == EL_MFMA) {
#pragma unroll
            for (int s = 0; s < 6; ++s) {
                kb0[s] = Kd.k[24 * li + 4 * s + lk];
                kb1[s] = (li < 8) ? Kd.k[24 * (16 + li) + 4 * s + lk] : T(0);
            }
        }

        for (int kb = k0 - 1; kb < k1; kb += PD) {
#pragma unroll
        for (int u = 0; u < PD; ++u) {
            const int ke = kb + u;  // element layer; plane ke+1 sits in slot (u+1) % PD
            if (ke >= k1) break;
            plane_store(ke + 1, (u + 1) % PD);  // also parks E of layer ke in es
            __syncthreads();
            if (ke + 1 + PD <= k1) plane_load(ke + 1 + PD, (u + 1) % PD);
            const T ee = es[ty][tx];

            if constexpr (MODE == EL_MFMA) {
                // ---- element phase on MFMA: each wave multiplies the gathered
                // [16 elements x 24 DOFs] rows of 4 tile rows by Ke^T (24x24)
#pragma unroll
                for (int gq = 0; gq < 4; ++gq) {
                    const int ety = 4 * wave + gq;
                    floatx4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                    for (int s = 0; s < 6; ++s) {
                        const int k = 4 * s + lk, L = k / 3, c = k - 3 * L;
                        const float av = xs[(ke + (L >> 2)) & 1][c][ety + ((L >> 1) & 1)][li + (L & 1)];
                        acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(av, kb0[s], acc0, 0, 0, 0);
                        acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(av, kb1[s], acc1, 0, 0, 0);
                    }
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int etx = 4 * lk + v;
                        const float e = es[ety][etx];
                        ys[li][ety][etx] = e * acc0[v];
                        if (li < 8) ys[16 + li][ety][etx] = e * acc1[v];
                    }
                }
            } else {
            // ---- element phase: one element per thread
            T y[24];
            if (ee != T(0)) {
                T u[24];
#pragma unroll
                for (int L = 0; L < 8; ++L) {
                    const int bx = L & 1, by = (L >> 1) & 1, bz = L >> 2;
                    const int sl = (ke + bz) & 1;
#pragma unroll
                    for (int c = 0; c < 3; ++c) u[3 * L + c] = xs[sl][c][ty + by][tx + bx];
                }
                if constexpr (MODE == EL_PARITY) {
                    element_apply_parity(u, y, K, ee);
                } else {
#pragma unroll
                    for (int r = 0; r < 24; ++r) {
                        T s = T(0);
#pragma unroll
                        for (int c = 0; c < 24; ++c) s += Kd.k[24 * r + c] * u[c];
                        y[r] = ee * s;
                    }
                }
            } else {
#pragma unroll
                for (int r = 0; r < 24; ++r) y[r] = T(0);
            }
#pragma unroll
            for (int r = 0; r < 24; ++r) ys[r][ty][tx] = y[r];
            }
            __syncthreads();

            // ---- node phase: gather the 4 elements around each owned node
            if (nthr) {
                T bot[3] = {T(0), T(0), T(0)}, top[3] = {T(0), T(0), T(0)};
#pragma unroll
                for (int dy = 0; dy < 2; ++dy)
#pragma unroll
                    for (int dx = 0; dx < 2; ++dx) {
                        const int L = (1 - dx) + 2 * (1 - dy);
#pragma unroll
                        for (int c = 0; c < 3; ++c) {
                            bot[c] += ys[3 * L + c][oy + dy][ox + dx];
                            top[c] += ys[3 * (L + 4) + c][oy + dy][ox + dx];
                        }
                    }
                if (ke >= k0 && nown) {
                    const int sl = ke & 1;
                    const long base = 3 * ((long)ke * plane_stride + (long)nj * g.NX + ni);
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        const T xv = xs[sl][c][oy + 1][ox + 1];
                        const T yv = carry[c] + bot[c] + a.shift * xv;
                        a.y[base + c] = yv;
                        if constexpr (FUSE) a.p_out[base + c] = xv;
                        pq += (double)xv * (double)yv;
                    }
                }
#pragma unroll
                for (int c = 0; c < 3; ++c) carry[c] = top[c];
            }
            __syncthreads();
        }
        }
    }

    if (a.mode == MV_PCG) {  // per-block p'q; k_fin_alpha reduces them
        const double v = block_sum<NT>(pq, red);
        if (threadIdx.x == 0) a.partials[blockIdx.x] = v;
    }
}


// ---------------------------------------------------------------------------
// Sum-factorised z-marching operator in the parity basis (the production
// contraction).  The butterfly that takes an element's 8 node values to its
// parity modes factorises over x, y and z.  Its x and y stages act on one
// node plane only: for the face (xy square of an element column) in plane k,
//   F_k[s], s = sx + 2 sy  (4 modes x 3 components),
// is computed once per plane and shared by the element layers below and
// above; the element's modes are xh[s] = F_lo[s] + F_hi[s], xh[s+4] =
// F_hi[s] - F_lo[s].  The inverse runs backwards: z-inverse per element, the
// two elements sharing a face add their face modes (a register carry along
// z, the owner-computes sum over z), then the y- and x-inverse move the face
// totals to the nodes.  A node receives from 4 faces: one LDS exchange with
// the face below in y and one DPP lane shift (row_shr:1) from the face to
// the left in x -- the 16-element tile row is one 16-lane DPP row.
// Per element and layer: ~180 flops, 6 LDS reads, 6+6 LDS exchange words and
// 2 barriers, against 234 flops, 27 reads, 24+24 exchange words and 3
// barriers for the element-local form.
// lane l <- lane l-1: within 16-lane DPP rows (row_shr:1) for 16-wide tile
// rows, across the wave (wave_shr:1) for 32-wide ones; lane 0 of a row / the
// wave receives 0 (bound_ctrl).  A tile row's first thread is never an owned
// node, so what crosses a row boundary is discarded.
template <int TX>
__device__ __forceinline__ int lane_from_left_i(int v) {
    static_assert(TX == 16 || TX == 32, "tile rows of 16 or 32 lanes");
    return __builtin_amdgcn_update_dpp(0, v, TX == 16 ? 0x111 : 0x138, 0xF, 0xF, true);
}
template <int TX>
__device__ __forceinline__ float lane_from_left(float v) {
    return __int_as_float(lane_from_left_i<TX>(__float_as_int(v)));
}
template <int TX>
__device__ __forceinline__ double lane_from_left(double v) {
    const int lo = lane_from_left_i<TX>(__double2loint(v));
    const int hi = lane_from_left_i<TX>(__double2hiint(v));
    return __hiloint2double(hi, lo);
}

// FUSE: 0 plain operator; 1 fused p = D^-1 r + beta p_old; 2 fused p = z + beta p_old
// with z handed in a.r (an explicit preconditioner's output; a.minv unused).
// The two smoothing forms of a V-cycle level, S = a.minv the smoother's
// diagonal (0 where the DOF is not solved for):
// 3 (pre-smoothing from zero) x = S r and the residual y = r - K x where
//   S != 0 (0 elsewhere); a.x unused;
// 4 (post-smoothing) x = S r + mask(P e), the coarse correction e (a.ec, node
//   grid cNX x cNY x cNZ) interpolated trilinearly where S != 0, then
//   y = x + S (r - K x), with the per-block partial sums of r'y over the owned
//   nodes when a.partials is set (a.partials[b]), and of r'r over the owned
//   DOFs with S != 0 (a.partials[gridDim.x + b]).  At its start a workgroup interpolates e in
//   z onto every fine plane of its chunk over the tile's coarse footprint, in
//   dynamic LDS ((kchunk + 2) planes, smoothing_post_kchunk); the march then
//   reads the four xy corners of each staged node.
// 6 (residual of a given iterate) y = r - K x where S != 0 (0 elsewhere), x = a.x;
// 7 (post-smoothing from a given iterate) x = a.x + mask(P e), then y as in 4.
// Forms 3, 4, 6 and 7 return at once when a.st is set and its solve is done.
// (the FP32 post-smoothing form with one plane in flight is held to 128 VGPRs:
// four waves per SIMD instead of three; the compiler spills two of them, 12 B
// of scratch per lane)
template <typename T, int TX, int TY, int FUSE, int PD>
__global__ __launch_bounds__(TX * TY, (FUSE == 4 && sizeof(T) == 4 && PD == 1) ? 4 : 1) void k_matvec_sf(
    GridArgs g, ParityK<T> K, MvArgs<T> a, int kchunk, int ntx, int nty, int ntiles) {
    constexpr int NT = TX * TY;
    constexpr int PX = TX + 1, PY = TY + 1;
    constexpr int NEXTRA = 3 * (PY + TX);  // plane values beyond the thread grid: column TX and row TY
    static_assert(NEXTRA <= NT, "one extra plane value per thread");
    static_assert(PD == 1 || PD == 2, "planes in flight");
    constexpr int OROW = 3 * (TX - 1);       // values of one owned node row
    constexpr int OVALS = (TY - 1) * OROW;   // owned values of one plane
    __shared__ T xs[3][PY][PX];  // one node plane (x, or the fused p), component-major
    __shared__ T ex[6][NT];      // y-exchange: row ty+1 half of each face's y-inverse
    constexpr bool POUT = FUSE == 1 || FUSE == 2;  // forms that write the staged vector to p_out
    constexpr bool SMOOTH = FUSE >= 3;
    constexpr bool PRO = FUSE == 4 || FUSE == 7;  // fused prolongation of a coarse correction
    constexpr bool XB = FUSE == 6 || FUSE == 7;   // smoothing forms from a given iterate a.x
    // coarse footprint of the tile's node columns (i0-1 .. i0-1+TX, j0-1 .. j0-1+TY)
    constexpr int CPX = TX / 2 + 2, CPY = TY / 2 + 2, CVALS = 3 * CPX * CPY;
    // form 4: the z-interpolated coarse correction, fine plane f of the chunk at [f CVALS + (c CPY + ly) CPX + lx]
    extern __shared__ __align__(16) unsigned char sf_dyn[];
    [[maybe_unused]] T* const czl = reinterpret_cast<T*>(sf_dyn);
    __shared__ T so[POUT ? 2 : 1][OVALS];  // results of the owned nodes, node-major, for row-contiguous stores
    __shared__ double red[NT / 64];

    if constexpr (!SMOOTH || XB) a.x += (long)blockIdx.y * a.cs;  // forms 3 and 4 read no x
    a.y += (long)blockIdx.y * a.cs;
    if constexpr (FUSE) a.r += (long)blockIdx.y * a.cs;
    if constexpr (POUT) a.p_out += (long)blockIdx.y * a.cs;
    if (a.mode == MV_PCG) {
        a.st += blockIdx.y;
        a.partials += (long)blockIdx.y * gridDim.x;
    }
    if (a.mode == MV_PCG && a.st->done) return;
    if constexpr (SMOOTH) {  // column blockIdx.y: its state, coarse correction and r'y, r'r partials
        if (a.st != nullptr) a.st += blockIdx.y;
        if constexpr (PRO) {
            a.ec += (long)blockIdx.y * a.ccs;
            if (a.partials != nullptr) a.partials += 2L * gridDim.x * blockIdx.y;
        }
        if (a.st != nullptr && a.st->done) return;
    }
    const T beta = (FUSE == 1 || FUSE == 2) ? T(a.st->beta) : T(0);

    const unsigned nb = gridDim.x;  // multiple of 8: XCD-contiguous tile ranges
    const unsigned tile = (blockIdx.x & 7u) * (nb >> 3) + (blockIdx.x >> 3);
    const bool active = tile < (unsigned)ntiles;
    const int tx = threadIdx.x & (TX - 1), ty = threadIdx.x / TX;
    double pq = 0.0;
    [[maybe_unused]] double rr = 0.0;  // form 4: r'r over the solved owned DOFs

    if (active) {
        const int cx = tile % ntx, cy = (tile / ntx) % nty, cz = tile / (ntx * nty);
        const int i0 = cx * (TX - 1), j0 = cy * (TY - 1);
        const int k0 = cz * kchunk, k1 = min(k0 + kchunk, g.NZ);
        // thread (tx,ty): face / element column (gi, gj), node (gi, gj) owned if tx,ty >= 1
        const int gi = i0 - 1 + tx, gj = j0 - 1 + ty;
        const bool own = tx >= 1 && ty >= 1 && gi < g.NX && gj < g.NY;
        const bool ecol = gi >= 0 && gi < g.nx && gj >= 0 && gj < g.ny;
        const long pstride = 3L * g.NX * g.NY;
        const int off_own = 3 * (min(max(gj, 0), g.NY - 1) * g.NX + min(max(gi, 0), g.NX - 1));
        const int eoff = min(max(gj, 0), g.ny - 1) * g.nx + min(max(gi, 0), g.nx - 1);
        // extra value of thread t < NEXTRA: component t % 3 of node t / 3 of
        // (TX, 0..TY) then (0..TX-1, TY); the other threads load a clamped copy
        const int xe = min((int)threadIdx.x, NEXTRA - 1), xn = xe / 3, xc = xe - 3 * xn;
        const int xtx = xn < PY ? TX : xn - PY, xty = xn < PY ? xn : TY;
        const bool xstore = (int)threadIdx.x < NEXTRA;
        const int off_x =
            3 * (min(max(j0 - 1 + xty, 0), g.NY - 1) * g.NX + min(max(i0 - 1 + xtx, 0), g.NX - 1)) + xc;

        // form 4: the tile's coarse footprint starts at coarse node (cbx, cby);
        // the chunk stages fine planes fbase .. fbase + nf - 1 (clamped to the grid)
        [[maybe_unused]] const int cbx = (i0 - 1) < 0 ? -1 : (i0 - 1) >> 1;
        [[maybe_unused]] const int cby = (j0 - 1) < 0 ? -1 : (j0 - 1) >> 1;
        [[maybe_unused]] const int fbase = max(k0 - 1, 0), nf = min(k1, g.NZ - 1) - fbase + 1;
        // fill: fine plane 2K holds coarse plane K, 2K+1 the mean of K and K+1.
        // A thread walks coarse columns (one value of the xy footprint) and
        // loads each coarse plane of the column once, kFillGroup fine planes
        // at a time (all loads of a group issued before any is waited for,
        // clamped addresses); the fine planes' parity against the coarse ones
        // is fixed per chunk, so the register index of every tap is static.
        auto coarse_fill = [&]() __attribute__((always_inline)) {
            constexpr int kFillGroup = 16, kFillK = kFillGroup / 2 + 1;
            const long pl = 3L * a.cNY * a.cNX;
            for (int ii = threadIdx.x; ii < CVALS; ii += NT) {
                const int ly = ii / (3 * CPX), rem = ii - ly * 3 * CPX, lx = rem / 3, c = rem - 3 * lx;
                const int I = min(max(cbx + lx, 0), a.cNX - 1), J = min(max(cby + ly, 0), a.cNY - 1);
                const T* col = a.ec + 3 * ((long)J * a.cNX + I) + c;
                T* dst = czl + (c * CPY + ly) * CPX + lx;
                for (int fs = 0; fs < nf; fs += kFillGroup) {
                    const int Kc = (fbase + fs) >> 1;
                    T v[kFillK];
#pragma unroll
                    for (int q = 0; q < kFillK; ++q) v[q] = col[min(Kc + q, a.cNZ - 1) * pl];
                    auto emit = [&](auto parity) __attribute__((always_inline)) {
                        constexpr int pp = decltype(parity)::value;
#pragma unroll
                        for (int f = 0; f < kFillGroup; ++f) {
                            const int q = (f + pp) >> 1;
                            if (fs + f < nf)
                                dst[(fs + f) * CVALS] = ((f + pp) & 1) ? T(0.5) * (v[q] + v[q + 1]) : v[q];
                        }
                    };
                    if (fbase & 1) emit(std::integral_constant<int, 1>{});
                    else emit(std::integral_constant<int, 0>{});
                }
            }
        };
        // trilinear interpolation of the correction: per thread, the LDS offset of
        // the lower xy corner and the four xy weights of its node and of its extra
        // value's node (fixed over the march), applied to the staged plane's
        // z-interpolated footprint
        // corner offset and the node's x / y parities packed in one int: o << 2 | py << 1 | px
        auto corners = [&](int fi, int fj) __attribute__((always_inline)) {
            const int lx = (fi < 0 ? -1 : fi >> 1) - cbx, ly = (fj < 0 ? -1 : fj >> 1) - cby;
            return ((ly * CPX + lx) << 2) | ((fj & 1) << 1) | (fi & 1);
        };
        [[maybe_unused]] const int cn_own = corners(gi, gj), cn_x = corners(i0 - 1 + xtx, j0 - 1 + xty);
        [[maybe_unused]] auto coarse_at = [&](int k, int c, int mf) __attribute__((always_inline)) -> T {
            const int f = min(max(mf, 0), g.NZ - 1) - fbase;
            const T* P = &czl[f * CVALS + c * CPY * CPX + (k >> 2)];
            const T wx = (k & 1) ? T(0.5) : T(0), wy = (k & 2) ? T(0.5) : T(0);
            const T lo = P[0] + wx * (P[1] - P[0]), hi = P[CPX] + wx * (P[CPX + 1] - P[CPX]);
            return lo + wy * (hi - lo);
        };
        // PD register sets of a prefetched plane: x (and r, D^-1 when fused)
        // at this thread's node and extra value, plus E of element layer
        // k-1.  Addresses are clamped, so every load is unconditional (a load
        // under a branch makes the compiler's wait counts conservative);
        // out-of-range node values only ever meet E = 0.
        struct Slot {
            T x[3], r[FUSE ? 3 : 1], m[(FUSE == 1 || SMOOTH) ? 3 : 1], qx, qr, qm, e;
            int k;
        };
        Slot sl[PD];
        auto plane_load = [&](Slot& q, int k) __attribute__((always_inline)) {
            q.k = k;
            const long po = (long)min(max(k, 0), g.NZ - 1) * pstride;
            if constexpr (FUSE != 3 && FUSE != 4) {
                const T* X = a.x + po;
#pragma unroll
                for (int c = 0; c < 3; ++c) q.x[c] = X[off_own + c];
                q.qx = X[off_x];
            }
            if constexpr (FUSE != 0) {
                const T* R = a.r + po;
#pragma unroll
                for (int c = 0; c < 3; ++c) q.r[c] = R[off_own + c];
                q.qr = R[off_x];
            }
            if constexpr (FUSE == 1 || SMOOTH) {
                const T* M = a.minv + po;
#pragma unroll
                for (int c = 0; c < 3; ++c) q.m[c] = M[off_own + c];
                q.qm = M[off_x];
            }
            q.e = a.E[(long)min(max(k - 1, 0), g.nz - 1) * g.nx * g.ny + eoff];
        };
        // park a prefetched plane in LDS; sets this thread's node value (and,
        // smoothing forms, its r and S) in xnew, rnew, mnew and the modulus of
        // its element in layer q.k - 1.  The node values of two planes live
        // in the two register sets of Xn, Rn, Mn, whose roles alternate with
        // the layer parity (no copies between layers).
        T ee;
        T Xn[2][3], Rn[2][SMOOTH ? 3 : 1], Mn[2][SMOOTH ? 3 : 1];
        auto plane_store = [&](const Slot& q, T* xnew, T* rnew, T* mnew) __attribute__((always_inline)) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                T v;
                if constexpr (FUSE == 1) v = q.m[c] * q.r[c] + beta * q.x[c];  // p = D^-1 r + beta p
                else if constexpr (FUSE == 2) v = q.r[c] + beta * q.x[c];        // p = z + beta p
                else if constexpr (FUSE == 3) v = q.m[c] * q.r[c];              // x = S r
                else if constexpr (FUSE == 4)                                     // x = S r + mask(P e)
                    v = q.m[c] * q.r[c] + (q.m[c] != T(0) ? coarse_at(cn_own, c, q.k) : T(0));
                else if constexpr (FUSE == 7)                                     // x = x0 + mask(P e)
                    v = q.x[c] + (q.m[c] != T(0) ? coarse_at(cn_own, c, q.k) : T(0));
                else v = q.x[c];
                xs[c][ty][tx] = v;
                xnew[c] = v;
                if constexpr (SMOOTH) {
                    rnew[c] = q.r[c];
                    mnew[c] = q.m[c];
                }
            }
            T v;
            if constexpr (FUSE == 1) v = q.qm * q.qr + beta * q.qx;
            else if constexpr (FUSE == 2) v = q.qr + beta * q.qx;
            else if constexpr (FUSE == 3) v = q.qm * q.qr;
            else if constexpr (FUSE == 4) {
                v = q.qm * q.qr;
                if (xstore && q.qm != T(0)) v += coarse_at(cn_x, xc, q.k);
            } else if constexpr (FUSE == 7) {
                v = q.qx;
                if (xstore && q.qm != T(0)) v += coarse_at(cn_x, xc, q.k);
            }
            else v = q.qx;
            if (xstore) xs[xc][xty][xtx] = v;
            ee = (ecol && q.k - 1 >= 0 && q.k - 1 < g.nz) ? q.e : T(0);
            asm volatile("" : "+v"(ee));  // settle ee here, before the next plane is requested
        };
        auto face_modes = [&](T* F) __attribute__((always_inline)) {  // F[3 s + c]
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const T x00 = xs[c][ty][tx], x10 = xs[c][ty][tx + 1];
                const T x01 = xs[c][ty + 1][tx], x11 = xs[c][ty + 1][tx + 1];
                const T A0 = x00 + x10, B0 = x10 - x00, A1 = x01 + x11, B1 = x11 - x01;
                F[c] = A0 + A1;
                F[3 + c] = B0 + B1;
                F[6 + c] = A1 - A0;
                F[9 + c] = B1 - B0;
            }
        };

        T F0[12], F1[12], carry[12];
        // Node results of plane ko are parked in LDS (so) at the end of a
        // layer and written out after the next layer's first barrier, ahead of
        // its prefetch, as row-contiguous runs of 3*(TX-1) values: a lane
        // writes consecutive words instead of one component of one node.
        int ko = -1;
        auto flush = [&]() __attribute__((always_inline)) {
            if (ko < k0) return;
            T* Y = a.y + (long)ko * pstride;
            T* P = POUT ? a.p_out + (long)ko * pstride : nullptr;
#pragma unroll
            for (int m = 0; m < (OVALS + NT - 1) / NT; ++m) {
                const int idx = threadIdx.x + m * NT;
                const int row = idx / OROW, col = idx - row * OROW;
                if (idx < OVALS && i0 + col / 3 < g.NX && j0 + row < g.NY) {
                    const int o = 3 * ((j0 + row) * g.NX + i0) + col;
                    Y[o] = so[0][idx];
                    if constexpr (POUT) P[o] = so[POUT ? 1 : 0][idx];
                }
            }
        };
        // element layer ke: Flo holds the face modes of plane ke, Fhi receives
        // plane ke+1's, which sits in slot q; plane ke+1+PD is requested into
        // q.  Register set `cur` holds plane ke's node values, `nxt` receives
        // plane ke+1's.
        auto step = [&](int ke, const T* Flo, T* Fhi, Slot& q, int cur) __attribute__((always_inline)) {
            const int nxt = cur ^ 1;
            const T* xo = Xn[cur];
            [[maybe_unused]] const T* ro = Rn[cur];
            [[maybe_unused]] const T* mo = Mn[cur];
            plane_store(q, Xn[nxt], Rn[nxt], Mn[nxt]);  // plane ke+1, E of layer ke
            __syncthreads();
            flush();
            asm volatile("" ::: "memory");
            plane_load(q, min(ke + 1 + PD, k1));  // unconditional (past k1: re-reads plane k1, a cache hit)
            asm volatile("" ::: "memory");  // keep the prefetch ahead of this layer's arithmetic
            face_modes(Fhi);

            // element (gi, gj, ke): modes, contraction, z-inverse; E scales the
            // face totals (12 products and 12 fused ones) rather than the
            // modes (21 products)
            T xh[24];
#pragma unroll
            for (int i = 0; i < 12; ++i) {
                xh[i] = Flo[i] + Fhi[i];
                xh[12 + i] = Fhi[i] - Flo[i];
            }
            T yh[24];
#pragma unroll
            for (int i = 0; i < 24; ++i) yh[i] = T(0);
#pragma unroll
            for (int n = 0; n < kParityNnz; ++n)
                yh[kParityPattern[n].r] += K.v[kParitySym.v[n]] * xh[kParityPattern[n].c];
            // face totals of plane ke: top half of layer ke-1 (carry) + bottom half of layer ke
            T H[12];
#pragma unroll
            for (int i = 0; i < 12; ++i) {
                H[i] = fma(ee, yh[i] - yh[12 + i], carry[i]);
                carry[i] = ee * (yh[i] + yh[12 + i]);
            }
            // y-inverse: row ty (A0,B0) stays, row ty+1 (A1,B1) goes to the face above in y
            T eA[3], eB[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                eA[c] = H[c] - H[6 + c];
                eB[c] = H[3 + c] - H[9 + c];
                ex[c][threadIdx.x] = H[c] + H[6 + c];
                ex[3 + c][threadIdx.x] = H[3 + c] + H[9 + c];
            }
            __syncthreads();
            const int src = ty >= 1 ? threadIdx.x - TX : threadIdx.x;
            T node[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const T A = eA[c] + ex[c][src], B = eB[c] + ex[3 + c][src];
                node[c] = (A - B) + lane_from_left<TX>(A + B);  // x-inverse: own half + left edge's half
            }
            ko = ke;
            const int so_i = (ty - 1) * OROW + 3 * (tx - 1);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const T yv = SMOOTH ? node[c] : node[c] + a.shift * xo[c];  // the level operators carry no shift
                T out = yv;
                if constexpr (FUSE == 3 || FUSE == 6) out = mo[c] != T(0) ? ro[c] - yv : T(0);  // residual
                else if constexpr (FUSE == 4 || FUSE == 7) out = xo[c] + mo[c] * (ro[c] - yv);  // x + S (r - K x)
                if (tx >= 1 && ty >= 1) {
                    so[0][so_i + c] = out;
                    if constexpr (POUT) so[POUT ? 1 : 0][so_i + c] = xo[c];
                }
                if constexpr (FUSE == 4 || FUSE == 7) {
                    if (ke >= k0 && own) {
                        pq += (double)ro[c] * (double)out;
                        if (mo[c] != T(0)) rr += (double)ro[c] * (double)ro[c];
                    }
                } else if constexpr (FUSE != 3 && FUSE != 6) {  // the residual forms reduce nothing
                    if (ke >= k0 && own) pq += (double)xo[c] * (double)yv;
                }
            }
        };

        plane_load(sl[0], k0 - 1);
        if constexpr (PRO) {  // the chunk's z-interpolated correction, before the first staging
            coarse_fill();
            __syncthreads();
        }
        plane_store(sl[0], Xn[0], Rn[0], Mn[0]);  // plane k0-1: the first layer's `cur`
        __syncthreads();
        face_modes(F0);
        // plane k0+j sits in slot j % PD
#pragma unroll
        for (int j = 0; j < PD; ++j) plane_load(sl[j], min(k0 + j, k1));
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 12; ++i) carry[i] = T(0);
        // two layers per trip (face-mode ping-pong); with PD = 2 the slot
        // index then follows the layer parity as well
        for (int ke = k0 - 1; ke < k1; ke += 2) {
            step(ke, F0, F1, sl[0], 0);
            if (ke + 1 < k1) step(ke + 1, F1, F0, sl[PD - 1], 1);
        }
        __syncthreads();
        flush();
    }

    if (FUSE != 3 && FUSE != 6 && (a.mode == MV_PCG || (PRO && a.partials != nullptr))) {  // p'q, or r'y
        const double v = block_sum<NT>(pq, red);
        if (threadIdx.x == 0) a.partials[blockIdx.x] = v;
        if constexpr (PRO) {  // and r'r after all blocks' r'y
            const double w = block_sum<NT>(rr, red);
            if (threadIdx.x == 0) a.partials[gridDim.x + blockIdx.x] = w;
        }
    }
}

// Reference-simple variant: one thread per node gathers its 8 elements with
// the dense Ke (x from global memory / cache).  Used to cross-check.
__global__ __launch_bounds__(256) void k_matvec_gather(GridArgs g, const double* __restrict__ Kb,
                                                       MvArgs<double> a) {
    __shared__ double red[4];
    a.x += (long)blockIdx.y * a.cs;
    a.y += (long)blockIdx.y * a.cs;
    if (a.mode == MV_PCG) {
        a.st += blockIdx.y;
        a.partials += (long)blockIdx.y * gridDim.x;
    }
    if (a.mode == MV_PCG && a.st->done) return;
    const long n = (long)blockIdx.x * 256 + threadIdx.x;
    double pq = 0.0;
    if (n < g.nn) {
        const int i = (int)(n % g.NX), j = (int)((n / g.NX) % g.NY), k = (int)(n / ((long)g.NX * g.NY));
        double acc[3] = {0, 0, 0};
        for (int dz = 0; dz < 2; ++dz)
            for (int dy = 0; dy < 2; ++dy)
                for (int dx = 0; dx < 2; ++dx) {
                    const int ie = i - 1 + dx, je = j - 1 + dy, ke = k - 1 + dz;
                    if (ie < 0 || je < 0 || ke < 0 || ie >= g.nx || je >= g.ny || ke >= g.nz) continue;
                    const double E = a.E[(long)ke * g.nx * g.ny + (long)je * g.nx + ie];
                    if (E == 0.0) continue;
                    const int Ln = (1 - dx) + 2 * (1 - dy) + 4 * (1 - dz);  // this node inside the element
                    for (int L = 0; L < 8; ++L) {
                        const long m = (long)(ie + (L & 1)) + (long)g.NX * ((je + ((L >> 1) & 1)) +
                                                                             (long)g.NY * (ke + (L >> 2)));
                        for (int d = 0; d < 3; ++d) {
                            const double xv = a.x[3 * m + d];
                            for (int c = 0; c < 3; ++c) acc[c] += E * Kb[24 * (3 * Ln + c) + 3 * L + d] * xv;
                        }
                    }
                }
        for (int c = 0; c < 3; ++c) {
            const double xv = a.x[3 * n + c];
            const double yv = acc[c] + a.shift * xv;
            a.y[3 * n + c] = yv;
            pq += xv * yv;
        }
    }
    if (a.mode == MV_PCG) {
        const double v = block_sum<256>(pq, red);
        if (threadIdx.x == 0) a.partials[blockIdx.x] = v;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_elastic_minv(GridArgs g, double kd0, double kd1, double kd2,
                                                      const T* __restrict__ E, const uint8_t* __restrict__ fixed,
                                                      double shift, T* __restrict__ minv) {
    const long n = (long)blockIdx.x * 256 + threadIdx.x;
    if (n >= g.nn) return;
    const int i = (int)(n % g.NX), j = (int)((n / g.NX) % g.NY), k = (int)(n / ((long)g.NX * g.NY));
    double S = 0.0;  // out-of-grid elements add the zero word (same sum, every load in flight)
#pragma unroll
    for (int dz = 0; dz < 2; ++dz)
#pragma unroll
        for (int dy = 0; dy < 2; ++dy)
#pragma unroll
            for (int dx = 0; dx < 2; ++dx) {
                const int ie = i - 1 + dx, je = j - 1 + dy, ke = k - 1 + dz;
                const bool in = ie >= 0 && je >= 0 && ke >= 0 && ie < g.nx && je < g.ny && ke < g.nz;
                S += (double)*(in ? E + ((long)ke * g.nx * g.ny + (long)je * g.nx + ie)
                                  : reinterpret_cast<const T*>(kZeroWord));
            }
    const double kd[3] = {kd0, kd1, kd2};
    uint8_t fx[3] = {0, 0, 0};
    if (fixed != nullptr)
#pragma unroll
        for (int c = 0; c < 3; ++c) fx[c] = fixed[3 * n + c];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const bool fr = fx[c] == 0 && S > 0.0;
        minv[3 * n + c] = fr ? T(1.0 / (kd[c] * S + shift)) : T(0);
    }
}

__global__ __launch_bounds__(256) void k_elastic_diag(GridArgs g, double kd0, double kd1, double kd2,
                                                      const double* __restrict__ E, double* __restrict__ diag) {
    const long n = (long)blockIdx.x * 256 + threadIdx.x;
    if (n >= g.nn) return;
    const int i = (int)(n % g.NX), j = (int)((n / g.NX) % g.NY), k = (int)(n / ((long)g.NX * g.NY));
    double S = 0.0;
    for (int dz = 0; dz < 2; ++dz)
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) {
                const int ie = i - 1 + dx, je = j - 1 + dy, ke = k - 1 + dz;
                if (ie < 0 || je < 0 || ke < 0 || ie >= g.nx || je >= g.ny || ke >= g.nz) continue;
                S += E[(long)ke * g.nx * g.ny + (long)je * g.nx + ie];
            }
    diag[3 * n + 0] = kd0 * S;
    diag[3 * n + 1] = kd1 * S;
    diag[3 * n + 2] = kd2 * S;
}

inline int round8(int n) { return (n + 7) & ~7; }

// values of a TX x TY tile's footprint on one coarse plane (3 components)
template <int TX, int TY>
constexpr int post_footprint() {
    return 3 * (TX / 2 + 2) * (TY / 2 + 2);
}
constexpr long kPostLdsBytes = 64 * 1024;  // dynamic LDS of the post-smoothing form

template <int TX, int TY>
int zmarch_tiles(const GridArgs& g, int kchunk, int* ntx, int* nty) {
    *ntx = (g.NX + TX - 2) / (TX - 1);
    *nty = (g.NY + TY - 2) / (TY - 1);
    const int nz = (g.NZ + kchunk - 1) / kchunk;
    return *ntx * *nty * nz;
}

}  // namespace

int smoothing_post_kchunk(int tile, int kchunk, int elem_bytes) {
    const int cv = tile == 1 ? post_footprint<16, 32>() : tile == 2 ? post_footprint<32, 8>() : post_footprint<16, 16>();
    while (kchunk > 1 && (long)(kchunk + 2) * cv * elem_bytes > kPostLdsBytes) kchunk /= 2;
    return kchunk;
}

int matvec_parity_blocks(const GridArgs& g, int tile, int kchunk) {
    int a, b;
    return round8(tile == 1   ? zmarch_tiles<16, 32>(g, kchunk, &a, &b)
                  : tile == 2 ? zmarch_tiles<32, 8>(g, kchunk, &a, &b)
                              : zmarch_tiles<16, 16>(g, kchunk, &a, &b));
}
int matvec_dense_blocks(const GridArgs& g, int kchunk) {
    int a, b;
    return round8(zmarch_tiles<16, 16>(g, kchunk, &a, &b));
}
int matvec_gather_blocks(const GridArgs& g) { return (int)((g.nn + 255) / 256); }

namespace {

template <typename T, int TX, int TY, int MODE>
int launch_zmarch(const GridArgs& g, DenseK<T> Kd, const MvArgs<T>& a, const MvShape& sh, hipStream_t s, int nrhs) {
    int ntx, nty;
    const int nt = zmarch_tiles<TX, TY>(g, sh.kchunk, &ntx, &nty);
    const int nb = round8(nt);
    const ParityK<T> none{};
    if (a.p_out != nullptr)
        hipLaunchKernelGGL((k_matvec_zmarch<T, TX, TY, MODE, true, 1>), dim3(nb, nrhs), dim3(TX * TY), 0, s, g, none,
                           Kd, a, sh.kchunk, ntx, nty, nt);
    else
        hipLaunchKernelGGL((k_matvec_zmarch<T, TX, TY, MODE, false, 1>), dim3(nb, nrhs), dim3(TX * TY), 0, s, g, none,
                           Kd, a, sh.kchunk, ntx, nty, nt);
    return nb;
}

template <typename T, int TX, int TY, int PD>
int launch_sf(const GridArgs& g, const ParityK<T>& K, const MvArgs<T>& a, const MvShape& sh_in, hipStream_t s,
              int nrhs) {
    MvShape sh = sh_in;
    size_t dyn = 0;
    if (a.form == 4 || a.form == 7) {  // the z-interpolated coarse correction of the chunk, (kchunk + 2) footprints
        sh.kchunk = smoothing_post_kchunk(sh.tile, sh.kchunk, (int)sizeof(T));
        dyn = (size_t)(sh.kchunk + 2) * post_footprint<TX, TY>() * sizeof(T);
        static bool lds_set = false;  // first use: the start of a solve, outside graph capture
        if (!lds_set) {
            (void)hipFuncSetAttribute((const void*)k_matvec_sf<T, TX, TY, 4, PD>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)kPostLdsBytes);
            (void)hipFuncSetAttribute((const void*)k_matvec_sf<T, TX, TY, 7, PD>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)kPostLdsBytes);
            lds_set = true;
        }
    }
    int ntx, nty;
    const int nt = zmarch_tiles<TX, TY>(g, sh.kchunk, &ntx, &nty);
    const int nb = round8(nt);
    if (a.form == 3)
        hipLaunchKernelGGL((k_matvec_sf<T, TX, TY, 3, PD>), dim3(nb, nrhs), dim3(TX * TY), 0, s, g, K, a, sh.kchunk,
                           ntx, nty, nt);
    else if (a.form == 4)
        hipLaunchKernelGGL((k_matvec_sf<T, TX, TY, 4, PD>), dim3(nb, nrhs), dim3(TX * TY), dyn, s, g, K, a, sh.kchunk,
                           ntx, nty, nt);
    else if (a.form == 6)
        hipLaunchKernelGGL((k_matvec_sf<T, TX, TY, 6, PD>), dim3(nb, nrhs), dim3(TX * TY), 0, s, g, K, a, sh.kchunk,
                           ntx, nty, nt);
    else if (a.form == 7)
        hipLaunchKernelGGL((k_matvec_sf<T, TX, TY, 7, PD>), dim3(nb, nrhs), dim3(TX * TY), dyn, s, g, K, a, sh.kchunk,
                           ntx, nty, nt);
    else if (a.p_out != nullptr && a.minv == nullptr)
        hipLaunchKernelGGL((k_matvec_sf<T, TX, TY, 2, PD>), dim3(nb, nrhs), dim3(TX * TY), 0, s, g, K, a, sh.kchunk,
                           ntx, nty, nt);
    else if (a.p_out != nullptr)
        hipLaunchKernelGGL((k_matvec_sf<T, TX, TY, 1, PD>), dim3(nb, nrhs), dim3(TX * TY), 0, s, g, K, a, sh.kchunk,
                           ntx, nty, nt);
    else
        hipLaunchKernelGGL((k_matvec_sf<T, TX, TY, 0, PD>), dim3(nb, nrhs), dim3(TX * TY), 0, s, g, K, a, sh.kchunk,
                           ntx, nty, nt);
    return nb;
}

}  // namespace

template <typename T>
int launch_matvec_parity(const GridArgs& g, const ParityK<T>& K, const MvArgs<T>& a, const MvShape& sh,
                         hipStream_t s, int nrhs) {
    int nb;
    if (sh.pdepth == 2)
        nb = sh.tile == 1   ? launch_sf<T, 16, 32, 2>(g, K, a, sh, s, nrhs)
             : sh.tile == 2 ? launch_sf<T, 32, 8, 2>(g, K, a, sh, s, nrhs)
                            : launch_sf<T, 16, 16, 2>(g, K, a, sh, s, nrhs);
    else
        nb = sh.tile == 1   ? launch_sf<T, 16, 32, 1>(g, K, a, sh, s, nrhs)
             : sh.tile == 2 ? launch_sf<T, 32, 8, 1>(g, K, a, sh, s, nrhs)
                            : launch_sf<T, 16, 16, 1>(g, K, a, sh, s, nrhs);
    if (a.mode == MV_PCG && !a.defer_alpha) launch_fin_alpha(a.st, a.partials, nb, nrhs, s);
    return nb;
}

template <typename T>
int launch_matvec_dense(const GridArgs& g, const T* Kbin, const MvArgs<T>& a, const MvShape& sh, hipStream_t s,
                        int nrhs) {
    const int nb = launch_zmarch<T, 16, 16, EL_DENSE>(g, DenseK<T>{Kbin}, a, sh, s, nrhs);
    if (a.mode == MV_PCG && !a.defer_alpha) launch_fin_alpha(a.st, a.partials, nb, nrhs, s);
    return nb;
}

int launch_matvec_mfma(const GridArgs& g, const float* Kbin, const MvArgs<float>& a, const MvShape& sh,
                       hipStream_t s, int nrhs) {
    const int nb = launch_zmarch<float, 16, 16, EL_MFMA>(g, DenseK<float>{Kbin}, a, sh, s, nrhs);
    if (a.mode == MV_PCG && !a.defer_alpha) launch_fin_alpha(a.st, a.partials, nb, nrhs, s);
    return nb;
}

int launch_matvec_gather(const GridArgs& g, const double* Kbin, const MvArgs<double>& a, hipStream_t s,
                         int nrhs) {
    const int nb = matvec_gather_blocks(g);
    hipLaunchKernelGGL(k_matvec_gather, dim3(nb, nrhs), dim3(256), 0, s, g, Kbin, a);
    if (a.mode == MV_PCG && !a.defer_alpha) launch_fin_alpha(a.st, a.partials, nb, nrhs, s);
    return nb;
}

template <typename T>
void launch_elastic_minv(const GridArgs& g, const double kd[3], const T* E, const uint8_t* fixed, double shift,
                         T* minv, hipStream_t s) {
    hipLaunchKernelGGL((k_elastic_minv<T>), dim3((unsigned)((g.nn + 255) / 256)), dim3(256), 0, s, g, kd[0], kd[1],
                       kd[2], E, fixed, shift, minv);
}

void launch_elastic_diag(const GridArgs& g, const double kd[3], const double* E, double* diag, hipStream_t s) {
    hipLaunchKernelGGL(k_elastic_diag, dim3((unsigned)((g.nn + 255) / 256)), dim3(256), 0, s, g, kd[0], kd[1], kd[2],
                       E, diag);
}

template int launch_matvec_parity<double>(const GridArgs&, const ParityK<double>&, const MvArgs<double>&, const MvShape&,
                                           hipStream_t, int);
template int launch_matvec_parity<float>(const GridArgs&, const ParityK<float>&, const MvArgs<float>&, const MvShape&,
                                          hipStream_t, int);
template int launch_matvec_dense<double>(const GridArgs&, const double*, const MvArgs<double>&, const MvShape&, hipStream_t,
                                          int);
template int launch_matvec_dense<float>(const GridArgs&, const float*, const MvArgs<float>&, const MvShape&, hipStream_t,
                                         int);
template void launch_elastic_minv<double>(const GridArgs&, const double*, const double*, const uint8_t*, double,
                                          double*, hipStream_t);
template void launch_elastic_minv<float>(const GridArgs&, const double*, const float*, const uint8_t*, double,
                                         float*, hipStream_t);

}  // namespace hx
