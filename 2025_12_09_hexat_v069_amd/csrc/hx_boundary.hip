// hx_boundary.hip -- boundary conditions and loads mapped onto the structured
// grid on the device (Boundary.get_nodes_by_location, get_bc_indicator,
// find_nearest_node, apply_external_forces!, Boundary.jl:56-291), and the
// density remap of a regrid (MeshRefiner.refine_mesh_and_fields,
// MeshRefiner.jl:131-150).
//
// A location is three components, each "every node" or an FP32 coordinate
// with its FP32 tolerance (resolved on the host from the YAML grammar exactly
// as the reference resolves them); a node (i,j,k) sits at the FP32 coordinates
// Float32(i)*dx, ... of Mesh.generate_mesh and matches when every component
// does: |coord - value| <= tol.
#include "hx_launch.h"

namespace hx {

namespace {

constexpr int BT = 256;

__device__ __forceinline__ void node_xyz(const GridArgs& g, float dx, float dy, float dz, long n, int& i, int& j,
                                         int& k, float& x, float& y, float& z) {
    i = (int)(n % g.NX);
    j = (int)((n / g.NX) % g.NY);
    k = (int)(n / ((long)g.NX * g.NY));
    x = (float)i * dx;
    y = (float)j * dy;
    z = (float)k * dz;
}

__device__ __forceinline__ bool matches(const LocSpec& s, float x, float y, float z) {
    const float c[3] = {x, y, z};
    bool m = true;
#pragma unroll
    for (int d = 0; d < 3; ++d) m = m && (s.all[d] || fabsf(c[d] - s.val[d]) <= s.tol[d]);
    return m;
}

__global__ void __launch_bounds__(BT) k_bc_mark(GridArgs g, float dx, float dy, float dz, LocSpec s, int dof_bits,
                                                uint8_t* __restrict__ fixed) {
    const long n = (long)blockIdx.x * BT + threadIdx.x;
    if (n >= g.nn) return;
    int i, j, k;
    float x, y, z;
    node_xyz(g, dx, dy, dz, n, i, j, k, x, y, z);
    if (!matches(s, x, y, z)) return;
#pragma unroll
    for (int d = 0; d < 3; ++d)
        if (dof_bits & (1 << d)) fixed[3 * n + d] = 1;
}

// per block: matched nodes (count) and the nearest-node keys (min FP32
// squared distance)
__global__ void __launch_bounds__(BT) k_loc_count(GridArgs g, float dx, float dy, float dz, LocSpec s,
                                                  unsigned* __restrict__ cnt) {
    __shared__ unsigned ws[BT / 64];
    const long n = (long)blockIdx.x * BT + threadIdx.x;
    unsigned m = 0;
    if (n < g.nn) {
        int i, j, k;
        float x, y, z;
        node_xyz(g, dx, dy, dz, n, i, j, k, x, y, z);
        m = matches(s, x, y, z) ? 1u : 0u;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m += __shfl_xor(m, o, 64);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) cnt[blockIdx.x] = ws[0] + ws[1] + ws[2] + ws[3];
}

// Find_nearest_node's squared distance, in FP32 in the reference's order
// (Boundary.jl:201-205), contraction off
__device__ __forceinline__ float dist2(float x, float y, float z, const float* t) {
#pragma clang fp contract(off)
    const float a = x - t[0], b = y - t[1], c = z - t[2];
    return a * a + b * b + c * c;
}

__global__ void __launch_bounds__(BT) k_near_min(GridArgs g, float dx, float dy, float dz, float tx, float ty,
                                                 float tz, float* __restrict__ part) {
    __shared__ float ws[BT / 64];
    const long n = (long)blockIdx.x * BT + threadIdx.x;
    const float t[3] = {tx, ty, tz};
    float m = INFINITY;
    if (n < g.nn) {
        int i, j, k;
        float x, y, z;
        node_xyz(g, dx, dy, dz, n, i, j, k, x, y, z);
        m = dist2(x, y, z, t);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fminf(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = fminf(fminf(ws[0], ws[1]), fminf(ws[2], ws[3]));
}

// among the nodes within 1e-9 of the least squared distance, the largest
// (x, y, z) -- on the grid, the largest (i, j, k) -- as a 63-bit key
__global__ void __launch_bounds__(BT) k_near_key(GridArgs g, float dx, float dy, float dz, float tx, float ty,
                                                 float tz, const float* __restrict__ dmin,
                                                 unsigned long long* __restrict__ part) {
    __shared__ unsigned long long ws[BT / 64];
    const long n = (long)blockIdx.x * BT + threadIdx.x;
    const float t[3] = {tx, ty, tz};
    unsigned long long key = 0;
    if (n < g.nn) {
        int i, j, k;
        float x, y, z;
        node_xyz(g, dx, dy, dz, n, i, j, k, x, y, z);
        const float d = dist2(x, y, z, t);
        if (fabs((double)d - (double)*dmin) <= 1e-9)
            key = (1ull << 63) | ((unsigned long long)i << 42) | ((unsigned long long)j << 21) | (unsigned long long)k;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long v = __shfl_xor(key, o, 64);
        key = v > key ? v : key;
    }
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = key;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long r = ws[0];
        for (int w = 1; w < BT / 64; ++w) r = ws[w] > r ? ws[w] : r;
        part[blockIdx.x] = r;
    }
}

// one-block finalisers: sum of counts; min of distances; max of keys
__global__ void __launch_bounds__(1024) k_fin_u32(const unsigned* __restrict__ p, long nb, unsigned* out) {
    __shared__ unsigned long long ws[16];
    unsigned long long s = 0;
    for (long b = threadIdx.x; b < nb; b += 1024) s += p[b];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int w = 0; w < 16; ++w) t += ws[w];
        *out = (unsigned)t;
    }
}
__global__ void __launch_bounds__(1024) k_fin_fmin(const float* __restrict__ p, long nb, float* out) {
    __shared__ float ws[16];
    float m = INFINITY;
    for (long b = threadIdx.x; b < nb; b += 1024) m = fminf(m, p[b]);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fminf(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        float t = ws[0];
        for (int w = 1; w < 16; ++w) t = fminf(t, ws[w]);
        *out = t;
    }
}
__global__ void __launch_bounds__(1024) k_fin_kmax(const unsigned long long* __restrict__ p, long nb,
                                                   unsigned long long* out) {
    __shared__ unsigned long long ws[16];
    unsigned long long m = 0;
    for (long b = threadIdx.x; b < nb; b += 1024) m = p[b] > m ? p[b] : m;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long v = __shfl_xor(m, o, 64);
        m = v > m ? v : m;
    }
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = ws[0];
        for (int w = 1; w < 16; ++w) t = ws[w] > t ? ws[w] : t;
        *out = t;
    }
}

// F[3n + d] += scale * f[d] on the matched nodes, scale = 1 / count with a
// location (the total split evenly, Boundary.jl:281), else 1
__global__ void __launch_bounds__(BT) k_force_add(GridArgs g, float dx, float dy, float dz, LocSpec s,
                                                  const unsigned* __restrict__ count, double f0, double f1, double f2,
                                                  double* __restrict__ F) {
    const long n = (long)blockIdx.x * BT + threadIdx.x;
    if (n >= g.nn) return;
    int i, j, k;
    float x, y, z;
    node_xyz(g, dx, dy, dz, n, i, j, k, x, y, z);
    if (!matches(s, x, y, z)) return;
    const double sc = 1.0 / (double)*count;
    F[3 * n + 0] += sc * f0;
    F[3 * n + 1] += sc * f1;
    F[3 * n + 2] += sc * f2;
}

// the nearest node (key's i, j, k) receives the whole force
__global__ void k_force_point(GridArgs g, const unsigned long long* __restrict__ key, double f0, double f1, double f2,
                              double* __restrict__ F) {
    if (blockIdx.x || threadIdx.x) return;
    const unsigned long long v = *key;
    if (!(v >> 63)) return;
    const long i = (long)((v >> 42) & 0x1FFFFF), j = (long)((v >> 21) & 0x1FFFFF), k = (long)(v & 0x1FFFFF);
    const long n = i + (long)g.NX * (j + (long)g.NY * k);
    F[3 * n + 0] += f0;
    F[3 * n + 1] += f1;
    F[3 * n + 2] += f2;
}

// Regrid (MeshRefiner.jl:131-150): each new element takes the fields of the
// old element its centroid falls in: c = (i + 1/2 as Float32) * new_h,
// old index clamp(floor(c / old_h), 0, old_n - 1), FP32 as the reference.
__global__ void __launch_bounds__(BT) k_regrid(int onx, int ony, int onz, int nnx, int nny, int nnz, float ndx,
                                               float ndy, float ndz, float odx, float ody, float odz, int nf,
                                               const float* __restrict__ oldf, float* __restrict__ newf) {
#pragma clang fp contract(off)
    const long ne_new = (long)nnx * nny * nnz, ne_old = (long)onx * ony * onz;
    const long e = (long)blockIdx.x * BT + threadIdx.x;
    if (e >= ne_new) return;
    const int ix = (int)(e % nnx), iy = (int)((e / nnx) % nny), iz = (int)(e / ((long)nnx * nny));
    const float cx = ((float)(ix + 1) - 0.5f) * ndx, cy = ((float)(iy + 1) - 0.5f) * ndy,
                cz = ((float)(iz + 1) - 0.5f) * ndz;
    const int ox = min(max((int)floorf(cx / odx), 0), onx - 1);
    const int oy = min(max((int)floorf(cy / ody), 0), ony - 1);
    const int oz = min(max((int)floorf(cz / odz), 0), onz - 1);
    const long o = ox + (long)onx * (oy + (long)ony * oz);
    for (int f = 0; f < nf; ++f) newf[f * ne_new + e] = oldf[f * ne_old + o];
}

}  // namespace

long boundary_blocks(const GridArgs& g) { return (g.nn + BT - 1) / BT; }

void launch_bc_mark(const GridArgs& g, const LocSpec& s, int dof_bits, uint8_t* fixed, hipStream_t st) {
    hipLaunchKernelGGL(k_bc_mark, dim3((unsigned)boundary_blocks(g)), dim3(BT), 0, st, g, (float)g.dx, (float)g.dy,
                       (float)g.dz, s, dof_bits, fixed);
}

void launch_loc_count(const GridArgs& g, const LocSpec& s, unsigned* part, unsigned* count, hipStream_t st) {
    const long nb = boundary_blocks(g);
    hipLaunchKernelGGL(k_loc_count, dim3((unsigned)nb), dim3(BT), 0, st, g, (float)g.dx, (float)g.dy, (float)g.dz, s,
                       part);
    hipLaunchKernelGGL(k_fin_u32, dim3(1), dim3(1024), 0, st, (const unsigned*)part, nb, count);
}

void launch_force_add(const GridArgs& g, const LocSpec& s, const unsigned* count, const double f[3], double* F,
                      hipStream_t st) {
    hipLaunchKernelGGL(k_force_add, dim3((unsigned)boundary_blocks(g)), dim3(BT), 0, st, g, (float)g.dx, (float)g.dy,
                       (float)g.dz, s, count, f[0], f[1], f[2], F);
}

void launch_nearest(const GridArgs& g, const float t[3], float* dpart, float* dmin, unsigned long long* kpart,
                    unsigned long long* key, hipStream_t st) {
    const long nb = boundary_blocks(g);
    const float dx = (float)g.dx, dy = (float)g.dy, dz = (float)g.dz;
    hipLaunchKernelGGL(k_near_min, dim3((unsigned)nb), dim3(BT), 0, st, g, dx, dy, dz, t[0], t[1], t[2], dpart);
    hipLaunchKernelGGL(k_fin_fmin, dim3(1), dim3(1024), 0, st, (const float*)dpart, nb, dmin);
    hipLaunchKernelGGL(k_near_key, dim3((unsigned)nb), dim3(BT), 0, st, g, dx, dy, dz, t[0], t[1], t[2],
                       (const float*)dmin, kpart);
    hipLaunchKernelGGL(k_fin_kmax, dim3(1), dim3(1024), 0, st, (const unsigned long long*)kpart, nb, key);
}

void launch_force_point(const GridArgs& g, const unsigned long long* key, const double f[3], double* F,
                        hipStream_t st) {
    hipLaunchKernelGGL(k_force_point, dim3(1), dim3(64), 0, st, g, key, f[0], f[1], f[2], F);
}

void launch_regrid(int onx, int ony, int onz, int nnx, int nny, int nnz, const float nh[3], const float oh[3], int nf,
                   const float* oldf, float* newf, hipStream_t st) {
    const long ne = (long)nnx * nny * nnz;
    hipLaunchKernelGGL(k_regrid, dim3((unsigned)((ne + BT - 1) / BT)), dim3(BT), 0, st, onx, ony, onz, nnx, nny, nnz,
                       nh[0], nh[1], nh[2], oh[0], oh[1], oh[2], nf, oldf, newf);
}

}  // namespace hx
