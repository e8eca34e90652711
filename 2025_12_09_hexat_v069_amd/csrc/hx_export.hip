// hx_export.hip -- payload of the viewer/checkpoint file: the elements whose
// density reaches a threshold, kept in element order, with their centroids,
// densities and signed L1 stress (Postprocessing.export_binary_for_web,
// Postprocessing.jl:159-195).
//
// An ordered stream compaction in three launches over tiles of 2048 elements:
// count the kept elements of each tile, scan the tile counts in one
// workgroup, then write each kept element at its tile's offset plus its rank
// inside the tile (wave ballots, popcounts and a 4-entry LDS prefix per
// 256-element step).  The centroid is computed from (i,j,k), not read: the
// structured grid has no node array on the device.  Bytes per element: 4 for
// the density test (read twice, the second time mostly from L2) plus 16 read
// (l1, 3 principal values) and 20 written for each kept element.
#include <mutex>

#include "hx_launch.h"

namespace hx {

namespace {

constexpr int WT = 256;             // threads per workgroup
constexpr int WSTEPS = 8;           // 256-element steps per tile
constexpr int WTILE = WT * WSTEPS;  // elements per tile

__device__ __forceinline__ bool kept(const float* density, long e, long ne, float thr) {
    return e < ne && density[e] >= thr;  // NaN is never kept, as findall(x -> x >= t)
}

__global__ void __launch_bounds__(WT) k_web_count(const float* __restrict__ density, long ne, float thr,
                                                  unsigned* __restrict__ tile_count) {
    __shared__ unsigned wsum[WT / 64];
    const long base = (long)blockIdx.x * WTILE;
    unsigned n = 0;
#pragma unroll
    for (int s = 0; s < WSTEPS; ++s) {
        const long e = base + s * WT + threadIdx.x;
        n += __popcll(__ballot(kept(density, e, ne, thr)));
    }
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = n;
    __syncthreads();
    if (threadIdx.x == 0) tile_count[blockIdx.x] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

// Exclusive scan of ntiles counts in one workgroup of 1024 threads;
// offs[ntiles] receives the total, and so does *total_out when given.
__global__ void __launch_bounds__(1024) k_web_scan(const unsigned* __restrict__ tile_count, int ntiles,
                                                   unsigned* __restrict__ offs, unsigned* __restrict__ total_out) {
    __shared__ unsigned part[1024];
    unsigned carry = 0;
    for (int c0 = 0; c0 < ntiles; c0 += 1024) {
        const int t = c0 + (int)threadIdx.x;
        const unsigned v = t < ntiles ? tile_count[t] : 0u;
        part[threadIdx.x] = v;
        __syncthreads();
        // Hillis-Steele inclusive scan over the chunk
        for (int o = 1; o < 1024; o <<= 1) {
            const unsigned a = threadIdx.x >= (unsigned)o ? part[threadIdx.x - o] : 0u;
            __syncthreads();
            part[threadIdx.x] += a;
            __syncthreads();
        }
        if (t < ntiles) offs[t] = carry + part[threadIdx.x] - v;
        carry += part[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        offs[ntiles] = carry;
        if (total_out) *total_out = carry;
    }
}

// Centroid of element (i,j,k) as MeshUtilities.element_centroid
// (MeshUtilities.jl:18-32) computes it from Mesh.generate_mesh's node
// coordinates (Mesh.jl:36-38, Float32((i-1))*dx): the eight corner values
// summed in FP32 in node order n1..n8, then scaled by 0.125f.  Contraction
// into FMAs is off so that every sum rounds as the reference's does.
__device__ __forceinline__ float corner_sum(float a, float b, const int* pat) {
#pragma clang fp contract(off)
    float s = 0.0f;
#pragma unroll
    for (int n = 0; n < 8; ++n) s = s + (pat[n] ? b : a);
    return s * 0.125f;
}

__device__ __forceinline__ float fsign(float v) { return v > 0.0f ? 1.0f : (v < 0.0f ? -1.0f : v); }

__global__ void __launch_bounds__(WT) k_web_write(const float* __restrict__ density, const float* __restrict__ l1,
                                                  const float* __restrict__ principal, long ne, float thr, int nx,
                                                  int ny, float dx, float dy, float dz,
                                                  const unsigned* __restrict__ offs, float* __restrict__ cent,
                                                  float* __restrict__ dens_out, float* __restrict__ sl1_out) {
#pragma clang fp contract(off)
    __shared__ unsigned wcnt[WT / 64];
    // corner patterns of the node order (-,-,-),(+,-,-),(+,+,-),(-,+,-),(-,-,+),(+,-,+),(+,+,+),(-,+,+)
    const int px[8] = {0, 1, 1, 0, 0, 1, 1, 0};
    const int py[8] = {0, 0, 1, 1, 0, 0, 1, 1};
    const int pz[8] = {0, 0, 0, 0, 1, 1, 1, 1};
    const long base = (long)blockIdx.x * WTILE;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
    unsigned run = offs[blockIdx.x];
    for (int s = 0; s < WSTEPS; ++s) {
        const long e = base + s * WT + threadIdx.x;
        const bool k = kept(density, e, ne, thr);
        const unsigned long long m = __ballot(k);
        if (lane == 0) wcnt[wave] = (unsigned)__popcll(m);
        __syncthreads();
        unsigned before = 0, total = 0;
#pragma unroll
        for (int w = 0; w < WT / 64; ++w) {
            before += w < wave ? wcnt[w] : 0u;
            total += wcnt[w];
        }
        if (k) {
            const unsigned pos = run + before + (unsigned)__popcll(m & below);
            const int i = (int)(e % nx);
            const int j = (int)((e / nx) % ny);
            const int kk = (int)(e / ((long)nx * ny));
            cent[3L * pos + 0] = corner_sum((float)i * dx, (float)(i + 1) * dx, px);
            cent[3L * pos + 1] = corner_sum((float)j * dy, (float)(j + 1) * dy, py);
            cent[3L * pos + 2] = corner_sum((float)kk * dz, (float)(kk + 1) * dz, pz);
            dens_out[pos] = density[e];
            // sign of the principal stress of largest magnitude, first one on ties
            // (Postprocessing.jl:187-195)
            const float s1 = principal[3 * e], s2 = principal[3 * e + 1], s3 = principal[3 * e + 2];
            float amax = fabsf(s1), sg = fsign(s1);
            if (fabsf(s2) > amax) { amax = fabsf(s2); sg = fsign(s2); }
            if (fabsf(s3) > amax) { sg = fsign(s3); }
            sl1_out[pos] = l1[e] * sg;
        }
        run += total;
        __syncthreads();  // wcnt is rewritten by the next step
    }
}

// ---- legacy VTK (ExportVTK.export_solution, ExportVTK.jl:20-172) ------------
// Every value the file holds is an FP32 (or Int32) word in big-endian byte
// order (`hton`); the kernels write those words, the host writes the ASCII
// section headers around them.

__device__ __forceinline__ unsigned be32(float v) { return __builtin_bswap32(__float_as_uint(v)); }
__device__ __forceinline__ unsigned be32i(int v) { return __builtin_bswap32((unsigned)v); }

// sanitize_data (ExportVTK.jl:35-42): NaN and +-Inf -> 0; then, if the largest
// magnitude exceeds 1e10, a clamp to +-1e10 -- which changes nothing when it
// does not, so the clamp is applied to every value
__device__ __forceinline__ float sanitize(float v) {
    if (!(v == v) || isinf(v)) return 0.0f;
    return fminf(fmaxf(v, -1.0e10f), 1.0e10f);
}

// Point data, one thread per node: coordinates Float32(i)*dx (Mesh.jl:36-38),
// displacement and external force as FP32, the boundary-condition indicator
// (1 where fixed), and the displacement magnitude sqrt((d1^2 + d2^2) + d3^2)
// in FP32 (ExportVTK.jl:58-85).  u, f, fixed: 3 nn (nullable: zeros).
__global__ void __launch_bounds__(WT) k_vtk_points(GridArgs g, float dx, float dy, float dz,
                                                   const double* __restrict__ u, const double* __restrict__ f,
                                                   const uint8_t* __restrict__ fixed, unsigned* __restrict__ pts,
                                                   unsigned* __restrict__ disp, unsigned* __restrict__ frc,
                                                   unsigned* __restrict__ bcs, unsigned* __restrict__ dmag) {
#pragma clang fp contract(off)
    const long n = (long)blockIdx.x * WT + threadIdx.x;
    if (n >= g.nn) return;
    const int i = (int)(n % g.NX), j = (int)((n / g.NX) % g.NY), k = (int)(n / ((long)g.NX * g.NY));
    pts[3 * n + 0] = be32((float)i * dx);
    pts[3 * n + 1] = be32((float)j * dy);
    pts[3 * n + 2] = be32((float)k * dz);
    float d[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        d[c] = u ? sanitize((float)u[3 * n + c]) : 0.0f;
        disp[3 * n + c] = be32(d[c]);
        frc[3 * n + c] = be32(f ? sanitize((float)f[3 * n + c]) : 0.0f);
        bcs[3 * n + c] = be32((fixed && fixed[3 * n + c]) ? 1.0f : 0.0f);
    }
    const float s2 = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2];
    dmag[n] = be32(sqrtf(s2));
}

// Cells of the kept elements (density >= threshold, element order; the
// ranks as k_web_write's): 8 and the 0-based node ids in the reference's node
// order, and the cell scalars L1 (sanitized), von Mises and density.
__global__ void __launch_bounds__(WT) k_vtk_cells(const float* __restrict__ density, const float* __restrict__ l1,
                                                  const float* __restrict__ vm, long ne, float thr, int nx, int ny,
                                                  const unsigned* __restrict__ offs, unsigned* __restrict__ cells,
                                                  unsigned* __restrict__ l1_out, unsigned* __restrict__ vm_out,
                                                  unsigned* __restrict__ dens_out) {
    __shared__ unsigned wcnt[WT / 64];
    const long base = (long)blockIdx.x * WTILE;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
    const long NX = nx + 1, NY = ny + 1;
    unsigned run = offs[blockIdx.x];
    for (int s = 0; s < WSTEPS; ++s) {
        const long e = base + s * WT + threadIdx.x;
        const bool k = kept(density, e, ne, thr);
        const unsigned long long m = __ballot(k);
        if (lane == 0) wcnt[wave] = (unsigned)__popcll(m);
        __syncthreads();
        unsigned before = 0, total = 0;
#pragma unroll
        for (int w = 0; w < WT / 64; ++w) {
            before += w < wave ? wcnt[w] : 0u;
            total += wcnt[w];
        }
        if (k) {
            const unsigned pos = run + before + (unsigned)__popcll(m & below);
            const long i = e % nx, j = (e / nx) % ny, kk = e / ((long)nx * ny);
            const long n0 = i + NX * (j + NY * kk);
            const long nodes[8] = {n0, n0 + 1, n0 + 1 + NX, n0 + NX,
                                   n0 + NX * NY, n0 + 1 + NX * NY, n0 + 1 + NX + NX * NY, n0 + NX + NX * NY};
            unsigned* c = cells + 9L * pos;
            c[0] = be32i(8);
#pragma unroll
            for (int a = 0; a < 8; ++a) c[1 + a] = be32i((int)nodes[a]);
            l1_out[pos] = be32(sanitize(l1[e]));
            vm_out[pos] = be32(vm[e]);
            dens_out[pos] = be32(density[e]);
        }
        run += total;
        __syncthreads();
    }
}

}  // namespace

void launch_vtk_pack(const GridArgs& g, const double* u, const double* f, const uint8_t* fixed, const float* density,
                     const float* l1, const float* vm, float thr, unsigned* tile_count, unsigned* offs,
                     unsigned* total, unsigned* pts, unsigned* disp, unsigned* frc, unsigned* bcs, unsigned* dmag,
                     unsigned* cells, unsigned* l1_out, unsigned* vm_out, unsigned* dens_out, hipStream_t s) {
    hipLaunchKernelGGL(k_vtk_points, dim3((unsigned)((g.nn + WT - 1) / WT)), dim3(WT), 0, s, g, (float)g.dx,
                       (float)g.dy, (float)g.dz, u, f, fixed, pts, disp, frc, bcs, dmag);
    const int nt = web_tiles(g.ne);
    hipLaunchKernelGGL(k_web_count, dim3(nt), dim3(WT), 0, s, density, g.ne, thr, tile_count);
    hipLaunchKernelGGL(k_web_scan, dim3(1), dim3(1024), 0, s, tile_count, nt, offs, total);
    hipLaunchKernelGGL(k_vtk_cells, dim3(nt), dim3(WT), 0, s, density, l1, vm, g.ne, thr, g.nx, g.ny, offs, cells,
                       l1_out, vm_out, dens_out);
}

// Tile counts and offsets of the packing entry point that allocates nothing
// (hx_web_pack_device): kWebScratchTiles counts, then kWebScratchTiles + 1
// offsets.  Each GPU holds its own copy until the process ends; its address
// comes from hipGetSymbolAddress (web_scratch_address), never from the
// array's name in host code.  Every entry of a call's range is written by
// that call before it is read, so nothing is cleared between calls; calls
// that overlap on one GPU race on it.
static __device__ __align__(256) unsigned scratch[2 * kWebScratchTiles + 1];

int web_scratch_address(int device, unsigned** out) {
    static std::mutex mu;
    static unsigned* ptr[kWebScratchDevices] = {};
    if (device < 0 || device >= kWebScratchDevices) return 1;
    std::lock_guard<std::mutex> lk(mu);
    if (!ptr[device]) {
        int cur = -1;
        if (hipGetDevice(&cur) != hipSuccess || cur != device) return 2;  // the device must be current
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return 3;
        ptr[device] = (unsigned*)p;
    }
    *out = ptr[device];
    return 0;
}

int web_tiles(long ne) { return (int)((ne + WTILE - 1) / WTILE); }

void launch_web_pack(const GridArgs& g, const float* density, const float* l1, const float* principal, float thr,
                     unsigned* tile_count, unsigned* offs, float* cent, float* dens_out, float* sl1_out,
                     hipStream_t s, unsigned* total_out) {
    const int nt = web_tiles(g.ne);
    if (nt == 0) return;
    hipLaunchKernelGGL(k_web_count, dim3(nt), dim3(WT), 0, s, density, g.ne, thr, tile_count);
    hipLaunchKernelGGL(k_web_scan, dim3(1), dim3(1024), 0, s, tile_count, nt, offs, total_out);
    hipLaunchKernelGGL(k_web_write, dim3(nt), dim3(WT), 0, s, density, l1, principal, g.ne, thr, g.nx, g.ny,
                       (float)g.dx, (float)g.dy, (float)g.dz, offs, cent, dens_out, sl1_out);
}

}  // namespace hx
