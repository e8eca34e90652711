// hx_stl.hip -- isosurface of the density for the STL export
// (Postprocessing.export_smooth_watertight_stl, Postprocessing.jl:53-148).
//
// 1. Nodal density: the mean of the densities of the elements around each
//    node (get_smooth_nodal_densities, :19-37), FP32.
// 2. Sampling grid: the nodal field refined `s` times per element by
//    trilinear interpolation (trilinear_interpolate, :39-47, in the
//    reference's operation order) with one layer of zeros around it
//    (:62-92), so every isosurface closes.
// 3. Marching cubes over the sampling grid: one pass counts each cell's
//    triangles, one workgroup scans the counts, one pass writes them in cell
//    order.  The case table is generated here rather than transcribed: for
//    each of the 256 corner patterns the face crossings are linked into
//    closed loops -- on a face with four crossings the two segments cut off
//    the corners at or above the level, the same rule for both cells that
//    share the face, so the surface has no cracks -- and each loop becomes a
//    fan of triangles turned to face away from the corners at or above the
//    level.  A vertex on a cell edge is interpolated from the edge's lower
//    end, so the two to four cells that share an edge compute the same bits.
//    The reference calls MarchingCubes.jl (a third-party table); the
//    triangulation therefore differs from it in the choice of triangles, not
//    in the surface's position.
#include <algorithm>
#include <mutex>
#include <vector>

#include "hx_launch.h"

namespace hx {

namespace {

constexpr int ST = 256;

// corner c = bx + 2 by + 4 bz; edge e joins kEdge[e][0] < kEdge[e][1]
constexpr int kEdge[12][2] = {{0, 1}, {2, 3}, {4, 5}, {6, 7}, {0, 2}, {1, 3},
                              {4, 6}, {5, 7}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};

int edge_of(int a, int b) {
    if (a > b) std::swap(a, b);
    for (int e = 0; e < 12; ++e)
        if (kEdge[e][0] == a && kEdge[e][1] == b) return e;
    return -1;
}

// triangles of every case: up to 5 per case, 3 edge ids each, -1 terminated
struct McTable {
    int8_t tri[256][16];
    uint8_t ntri[256];
};

McTable build_table() {
    McTable t{};
    // faces as corner cycles
    const int faces[6][4] = {{0, 2, 6, 4}, {1, 3, 7, 5}, {0, 1, 5, 4}, {2, 3, 7, 6}, {0, 1, 3, 2}, {4, 5, 7, 6}};
    auto cpos = [](int c, double* p) {
        p[0] = c & 1;
        p[1] = (c >> 1) & 1;
        p[2] = (c >> 2) & 1;
    };
    for (int cs = 0; cs < 256; ++cs) {
        auto in = [&](int c) { return (cs >> c) & 1; };
        // segment graph over crossed edges: two neighbours per crossed edge
        int nb[12][2];
        int deg[12] = {};
        for (int e = 0; e < 12; ++e) nb[e][0] = nb[e][1] = -1;
        auto link = [&](int a, int b) {
            nb[a][deg[a]++] = b;
            nb[b][deg[b]++] = a;
        };
        for (int f = 0; f < 6; ++f) {
            int ex[4];
            int nx = 0;
            for (int k = 0; k < 4; ++k) {
                const int a = faces[f][k], b = faces[f][(k + 1) & 3];
                if (in(a) != in(b)) ex[nx++] = k;  // edge k of the cycle: corners k, k+1
            }
            if (nx == 2) {
                link(edge_of(faces[f][ex[0]], faces[f][(ex[0] + 1) & 3]),
                     edge_of(faces[f][ex[1]], faces[f][(ex[1] + 1) & 3]));
            } else if (nx == 4) {  // every cycle edge crossed: cut off each inside corner
                for (int k = 0; k < 4; ++k) {
                    const int c = faces[f][k];
                    if (!in(c)) continue;
                    const int prev = faces[f][(k + 3) & 3], next = faces[f][(k + 1) & 3];
                    link(edge_of(prev, c), edge_of(c, next));
                }
            }
        }
        // walk the loops
        bool used[12] = {};
        int nt = 0;
        for (int e0 = 0; e0 < 12; ++e0) {
            if (deg[e0] != 2 || used[e0]) continue;
            std::vector<int> loop;
            int prev = -1, cur = e0;
            while (!used[cur]) {
                used[cur] = true;
                loop.push_back(cur);
                const int nxt = nb[cur][0] != prev ? nb[cur][0] : nb[cur][1];
                prev = cur;
                cur = nxt;
            }
            // orientation: the loop's normal (Newell, edge midpoints) points away
            // from the inside ends of its edges
            double P[12][3], nrm[3] = {0, 0, 0}, cen[3] = {0, 0, 0}, ins[3] = {0, 0, 0};
            const int m = (int)loop.size();
            for (int k = 0; k < m; ++k) {
                double a[3], b[3];
                cpos(kEdge[loop[k]][0], a);
                cpos(kEdge[loop[k]][1], b);
                for (int d = 0; d < 3; ++d) P[k][d] = 0.5 * (a[d] + b[d]);
                const int ci = in(kEdge[loop[k]][0]) ? kEdge[loop[k]][0] : kEdge[loop[k]][1];
                double q[3];
                cpos(ci, q);
                for (int d = 0; d < 3; ++d) {
                    cen[d] += P[k][d] / m;
                    ins[d] += q[d] / m;
                }
            }
            for (int k = 0; k < m; ++k) {
                const double* u = P[k];
                const double* v = P[(k + 1) % m];
                nrm[0] += (u[1] - v[1]) * (u[2] + v[2]);
                nrm[1] += (u[2] - v[2]) * (u[0] + v[0]);
                nrm[2] += (u[0] - v[0]) * (u[1] + v[1]);
            }
            const double dotv = nrm[0] * (cen[0] - ins[0]) + nrm[1] * (cen[1] - ins[1]) + nrm[2] * (cen[2] - ins[2]);
            if (dotv < 0) std::reverse(loop.begin(), loop.end());
            for (int k = 1; k + 1 < m; ++k) {
                t.tri[cs][3 * nt + 0] = (int8_t)loop[0];
                t.tri[cs][3 * nt + 1] = (int8_t)loop[k];
                t.tri[cs][3 * nt + 2] = (int8_t)loop[k + 1];
                ++nt;
            }
        }
        t.ntri[cs] = (uint8_t)nt;
        for (int k = 3 * nt; k < 16; ++k) t.tri[cs][k] = -1;
    }
    return t;
}

// 1. nodal mean of the element densities (FP32 sum in element order of the
// reference's loop -- elements k-major, then j, then i -- over the up to 8
// elements around the node, divided by their count)
__global__ void __launch_bounds__(ST) k_stl_nodal(GridArgs g, const float* __restrict__ rho, float* __restrict__ nod) {
#pragma clang fp contract(off)
    const long n = (long)blockIdx.x * ST + threadIdx.x;
    if (n >= g.nn) return;
    const int i = (int)(n % g.NX), j = (int)((n / g.NX) % g.NY), k = (int)(n / ((long)g.NX * g.NY));
    float s = 0.0f;
    int cnt = 0;
    for (int dk = -1; dk <= 0; ++dk)
        for (int dj = -1; dj <= 0; ++dj)
            for (int di = -1; di <= 0; ++di) {
                const int ie = i + di, je = j + dj, ke = k + dk;
                if (ie < 0 || je < 0 || ke < 0 || ie >= g.nx || je >= g.ny || ke >= g.nz) continue;
                s = s + rho[ie + (long)g.nx * (je + (long)g.ny * ke)];
                ++cnt;
            }
    nod[n] = cnt > 0 ? s / (float)cnt : 0.0f;
}

// 2. sampling grid, (s nx + 3) x (s ny + 3) x (s nz + 3) points with one layer of zeros
__global__ void __launch_bounds__(ST) k_stl_fine(GridArgs g, int sub, const float* __restrict__ nod,
                                                 float* __restrict__ fine, int FX, int FY, int FZ) {
#pragma clang fp contract(off)
    const long p = (long)blockIdx.x * ST + threadIdx.x;
    if (p >= (long)FX * FY * FZ) return;
    const int a = (int)(p % FX), b = (int)((p / FX) % FY), c = (int)(p / ((long)FX * FY));
    if (a == 0 || b == 0 || c == 0 || a == FX - 1 || b == FY - 1 || c == FZ - 1) {
        fine[p] = 0.0f;
        return;
    }
    const int ix = a - 1, iy = b - 1, iz = c - 1;
    int cx = ix / sub, cy = iy / sub, cz = iz / sub;
    if (cx >= g.nx) cx = g.nx - 1;
    if (cy >= g.ny) cy = g.ny - 1;
    if (cz >= g.nz) cz = g.nz - 1;
    const float xd = (float)(ix - cx * sub) / (float)sub, yd = (float)(iy - cy * sub) / (float)sub,
                zd = (float)(iz - cz * sub) / (float)sub;
    auto N = [&](int di, int dj, int dk) {
        return nod[(cx + di) + (long)g.NX * ((cy + dj) + (long)g.NY * (cz + dk))];
    };
    // vals 1..8 = (i,j,k), (i+1,j,k), (i+1,j+1,k), (i,j+1,k), then k+1 (Postprocessing.jl:84-85)
    const float v1 = N(0, 0, 0), v2 = N(1, 0, 0), v3 = N(1, 1, 0), v4 = N(0, 1, 0);
    const float v5 = N(0, 0, 1), v6 = N(1, 0, 1), v7 = N(1, 1, 1), v8 = N(0, 1, 1);
    const float c00 = v1 * (1.0f - xd) + v2 * xd;
    const float c01 = v4 * (1.0f - xd) + v3 * xd;
    const float c10 = v5 * (1.0f - xd) + v6 * xd;
    const float c11 = v8 * (1.0f - xd) + v7 * xd;
    const float c0 = c00 * (1.0f - yd) + c01 * yd;
    const float c1 = c10 * (1.0f - yd) + c11 * yd;
    fine[p] = c0 * (1.0f - zd) + c1 * zd;
}

// kEdge[e][end] on the device: x edges 0-3, y edges 4-7, z edges 8-11
__device__ __forceinline__ int kEdgeDev(int e, int end) {
    int c0;
    if (e < 4) return 2 * e + end;
    if (e < 8) {
        const int k = e - 4;
        c0 = (k & 1) + 4 * (k >> 1);
        return c0 + 2 * end;
    }
    c0 = e - 8;
    return c0 + 4 * end;
}

__device__ __forceinline__ int cell_case(const float* __restrict__ fine, long base, int FX, long FXY, float iso) {
    int cs = 0;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const long q = base + (c & 1) + ((c >> 1) & 1) * (long)FX + ((c >> 2) & 1) * FXY;
        cs |= (fine[q] >= iso ? 1 : 0) << c;
    }
    return cs;
}

// 3a. triangles per cell, summed per workgroup of 256 cells
__global__ void __launch_bounds__(ST) k_mc_count(const float* __restrict__ fine, int FX, int FY, int FZ, float iso,
                                                 const uint8_t* __restrict__ ntri, unsigned* __restrict__ blk) {
    __shared__ unsigned ws[ST / 64];
    const long ncell = (long)(FX - 1) * (FY - 1) * (FZ - 1);
    const long cell = (long)blockIdx.x * ST + threadIdx.x;
    unsigned n = 0;
    if (cell < ncell) {
        const int a = (int)(cell % (FX - 1)), b = (int)((cell / (FX - 1)) % (FY - 1)),
                  c = (int)(cell / ((long)(FX - 1) * (FY - 1)));
        const long FXY = (long)FX * FY;
        n = ntri[cell_case(fine, a + (long)FX * b + FXY * c, FX, FXY, iso)];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) n += __shfl_xor(n, o, 64);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = n;
    __syncthreads();
    if (threadIdx.x == 0) blk[blockIdx.x] = ws[0] + ws[1] + ws[2] + ws[3];
}

// exclusive scan of nb block counts in one workgroup; offs[nb] = total
__global__ void __launch_bounds__(1024) k_mc_scan(const unsigned* __restrict__ blk, int nb, unsigned* __restrict__ offs) {
    __shared__ unsigned part[1024];
    unsigned carry = 0;
    for (int c0 = 0; c0 < nb; c0 += 1024) {
        const int t = c0 + (int)threadIdx.x;
        const unsigned v = t < nb ? blk[t] : 0u;
        part[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < 1024; o <<= 1) {
            const unsigned a = threadIdx.x >= (unsigned)o ? part[threadIdx.x - o] : 0u;
            __syncthreads();
            part[threadIdx.x] += a;
            __syncthreads();
        }
        if (t < nb) offs[t] = carry + part[threadIdx.x] - v;
        carry += part[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) offs[nb] = carry;
}

// 3b. write each cell's triangles at its block offset + its rank in the block:
// 12 floats per triangle (unit normal or zeros, then the three vertices)
__global__ void __launch_bounds__(ST) k_mc_write(const float* __restrict__ fine, int FX, int FY, int FZ, float iso,
                                                 float x0, float y0, float z0, float hx, float hy, float hz,
                                                 const int8_t* __restrict__ tri, const uint8_t* __restrict__ ntri,
                                                 const unsigned* __restrict__ offs, unsigned cap,
                                                 float* __restrict__ out) {
    __shared__ unsigned ws[ST / 64];
    const long ncell = (long)(FX - 1) * (FY - 1) * (FZ - 1);
    const long cell = (long)blockIdx.x * ST + threadIdx.x;
    int cs = 0, nt = 0, a = 0, b = 0, c = 0;
    const long FXY = (long)FX * FY;
    if (cell < ncell) {
        a = (int)(cell % (FX - 1));
        b = (int)((cell / (FX - 1)) % (FY - 1));
        c = (int)(cell / ((long)(FX - 1) * (FY - 1)));
        cs = cell_case(fine, a + (long)FX * b + FXY * c, FX, FXY, iso);
        nt = ntri[cs];
    }
    // rank of this cell's first triangle in the block: inclusive wave scan, then the waves before
    unsigned incl = (unsigned)nt;
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned v = __shfl_up(incl, o, 64);
        if (lane >= o) incl += v;
    }
    if (lane == 63) ws[threadIdx.x >> 6] = incl;
    __syncthreads();
    unsigned before = incl - (unsigned)nt;
    for (int w = 0; w < (int)(threadIdx.x >> 6); ++w) before += ws[w];
    if (nt == 0) return;
    const unsigned first = offs[blockIdx.x] + before;
    const long base = a + (long)FX * b + FXY * c;
    for (int t = 0; t < nt; ++t) {
        if (first + t >= cap) return;
        float V[3][3];
        for (int v = 0; v < 3; ++v) {
            const int e = tri[cs * 16 + 3 * t + v];
            const int c0 = kEdgeDev(e, 0), c1 = kEdgeDev(e, 1);  // c0 < c1: c0 is the edge's lower end
            const long q0 = base + (c0 & 1) + ((c0 >> 1) & 1) * (long)FX + ((c0 >> 2) & 1) * FXY;
            const long q1 = base + (c1 & 1) + ((c1 >> 1) & 1) * (long)FX + ((c1 >> 2) & 1) * FXY;
            const float f0 = fine[q0], f1 = fine[q1];
            const float s = (iso - f0) / (f1 - f0);
            const int ga = a + (c0 & 1), gb = b + ((c0 >> 1) & 1), gc = c + ((c0 >> 2) & 1);
            float p[3] = {x0 + (float)ga * hx, y0 + (float)gb * hy, z0 + (float)gc * hz};
            const int axis = (c1 ^ c0) == 1 ? 0 : ((c1 ^ c0) == 2 ? 1 : 2);
            const float h = axis == 0 ? hx : (axis == 1 ? hy : hz);
            p[axis] = p[axis] + s * h;
            V[v][0] = p[0];
            V[v][1] = p[1];
            V[v][2] = p[2];
        }
        const float e1x = V[1][0] - V[0][0], e1y = V[1][1] - V[0][1], e1z = V[1][2] - V[0][2];
        const float e2x = V[2][0] - V[0][0], e2y = V[2][1] - V[0][1], e2z = V[2][2] - V[0][2];
        float nx = e1y * e2z - e1z * e2y, ny = e1z * e2x - e1x * e2z, nz = e1x * e2y - e1y * e2x;
        const float mag = sqrtf(nx * nx + ny * ny + nz * nz);
        if (mag > 1e-12f) {
            nx /= mag;
            ny /= mag;
            nz /= mag;
        } else {
            nx = ny = nz = 0.0f;
        }
        float* o = out + 12L * (first + t);
        o[0] = nx;
        o[1] = ny;
        o[2] = nz;
        for (int v = 0; v < 3; ++v)
            for (int d = 0; d < 3; ++d) o[3 + 3 * v + d] = V[v][d];
    }
}

}  // namespace

const void* mc_table_host(int part) {
    static McTable t;
    static std::once_flag once;
    std::call_once(once, [] { t = build_table(); });
    return part == 0 ? (const void*)t.tri : (const void*)t.ntri;
}

void launch_stl_field(const GridArgs& g, int sub, const float* rho, float* nod, float* fine, hipStream_t s) {
    const int FX = sub * g.nx + 3, FY = sub * g.ny + 3, FZ = sub * g.nz + 3;
    hipLaunchKernelGGL(k_stl_nodal, dim3((unsigned)((g.nn + ST - 1) / ST)), dim3(ST), 0, s, g, rho, nod);
    const long np = (long)FX * FY * FZ;
    hipLaunchKernelGGL(k_stl_fine, dim3((unsigned)((np + ST - 1) / ST)), dim3(ST), 0, s, g, sub, (const float*)nod,
                       fine, FX, FY, FZ);
}

long mc_blocks(int FX, int FY, int FZ) { return ((long)(FX - 1) * (FY - 1) * (FZ - 1) + ST - 1) / ST; }

void launch_mc_count(const float* fine, int FX, int FY, int FZ, float iso, const uint8_t* ntri, unsigned* blk,
                     unsigned* offs, hipStream_t s) {
    const long nb = mc_blocks(FX, FY, FZ);
    hipLaunchKernelGGL(k_mc_count, dim3((unsigned)nb), dim3(ST), 0, s, fine, FX, FY, FZ, iso, ntri, blk);
    hipLaunchKernelGGL(k_mc_scan, dim3(1), dim3(1024), 0, s, (const unsigned*)blk, (int)nb, offs);
}

void launch_mc_write(const float* fine, int FX, int FY, int FZ, float iso, float x0, float y0, float z0, float hx,
                     float hy, float hz, const int8_t* tri, const uint8_t* ntri, const unsigned* offs, unsigned cap,
                     float* out, hipStream_t s) {
    const long nb = mc_blocks(FX, FY, FZ);
    hipLaunchKernelGGL(k_mc_write, dim3((unsigned)nb), dim3(ST), 0, s, fine, FX, FY, FZ, iso, x0, y0, z0, hx, hy, hz,
                       tri, ntri, offs, cap, out);
}

}  // namespace hx
