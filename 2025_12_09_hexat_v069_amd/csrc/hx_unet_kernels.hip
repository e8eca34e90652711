// hx_unet_kernels.hip -- kernels of the learned preconditioner: a 3-D U-Net
// on the node grid, BF16 activations and weights, FP32 accumulation.
//
// Activations are channels-last, [pixel][C] with C a multiple of 8, pixel =
// i + nx*(j + ny*k) of that level's node grid.  A 3x3x3 convolution is an
// implicit GEMM: rows = output pixels, K = (tap, input channel), columns =
// output channels.  Each wave owns 16 output pixels and all output channels;
// one v_mfma_f32_16x16x32_bf16 consumes four 8-channel groups (one per lane
// quarter), so K is walked as a flat list of (tap, channel-group) pairs.
#include <hip/hip_runtime.h>

#include "hx_unet.h"

namespace hx {

namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ float bf2f(uint16_t v) { return __uint_as_float((uint32_t)v << 16); }
__device__ __forceinline__ uint16_t f2bf(float f) {  // round to nearest even
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

struct U16x8 {
    uint4 v;
};

__device__ __forceinline__ bf16x8 as_bf16x8(const uint4& v) {
    union {
        uint4 u;
        bf16x8 b;
    } c;
    c.u = v;
    return c.b;
}

// tap t = (dx+1) + 3*(dy+1) + 9*(dz+1)
__device__ __forceinline__ void tap_off(int t, int& dx, int& dy, int& dz) {
    dx = t % 3 - 1;
    dy = (t / 3) % 3 - 1;
    dz = t / 9 - 1;
}

constexpr int kWaves = 4;

__global__ __launch_bounds__(256) void k_conv3d(ConvArgs a) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lq = lane >> 4;
    const long npix_out = (long)a.vout.nx * a.vout.ny * a.vout.nz;
    const long p0 = ((long)blockIdx.x * kWaves + wave) * 16;
    if (p0 >= npix_out) return;
    const long p = p0 + li;  // this lane's A-row pixel
    const bool prow = p < npix_out;
    int ox = 0, oy = 0, oz = 0;
    if (prow) {
        ox = (int)(p % a.vout.nx);
        oy = (int)((p / a.vout.nx) % a.vout.ny);
        oz = (int)(p / ((long)a.vout.nx * a.vout.ny));
    }
    const int cin = a.c0 + a.c1;
    const int ng = cin / 8;
    const int total = 27 * ng;
    const int nob = a.cout / 16;
    floatx4 acc[kMaxOutBlocks];
#pragma unroll
    for (int b = 0; b < kMaxOutBlocks; ++b) acc[b] = floatx4{0.f, 0.f, 0.f, 0.f};

    for (int g0 = 0; g0 < total; g0 += 4) {
        const int gi = g0 + lq;
        uint4 av = make_uint4(0, 0, 0, 0);
        int t = 0, cg = 0;
        if (gi < total) {
            t = gi / ng;
            cg = gi - t * ng;
            if (prow) {
                int dx, dy, dz;
                tap_off(t, dx, dy, dz);
                int sx, sy, sz;
                bool ok;
                if (!a.transposed) {  // gather: src = stride*out + off
                    sx = a.stride * ox + dx;
                    sy = a.stride * oy + dy;
                    sz = a.stride * oz + dz;
                    ok = true;
                } else {  // adjoint: out = stride*src + off  ->  src = (out - off)/stride
                    const int ux = ox - dx, uy = oy - dy, uz = oz - dz;
                    ok = ux >= 0 && uy >= 0 && uz >= 0 && ux % a.stride == 0 && uy % a.stride == 0 &&
                         uz % a.stride == 0;
                    sx = ux / a.stride;
                    sy = uy / a.stride;
                    sz = uz / a.stride;
                }
                ok = ok && sx >= 0 && sy >= 0 && sz >= 0 && sx < a.vin.nx && sy < a.vin.ny && sz < a.vin.nz;
                if (ok) {
                    const int ch = 8 * cg;
                    if (ch < a.c0) {
                        long sp;
                        if (a.up0) {  // in0 lives on the coarse grid (nearest upsample)
                            sp = (long)(sx >> 1) + (long)a.vc0.nx * ((sy >> 1) + (long)a.vc0.ny * (sz >> 1));
                        } else {
                            sp = (long)sx + (long)a.vin.nx * (sy + (long)a.vin.ny * sz);
                        }
                        av = *reinterpret_cast<const uint4*>(a.in0 + sp * a.c0 + ch);
                    } else {
                        const long sp = (long)sx + (long)a.vin.nx * (sy + (long)a.vin.ny * sz);
                        av = *reinterpret_cast<const uint4*>(a.in1 + sp * a.c1 + (ch - a.c0));
                    }
                }
            }
        }
        const bf16x8 A = as_bf16x8(av);
#pragma unroll
        for (int b = 0; b < kMaxOutBlocks; ++b) {
            if (b < nob) {
                uint4 bv = make_uint4(0, 0, 0, 0);
                if (gi < total) {
                    const int j = 16 * b + li;  // output channel
                    bv = *reinterpret_cast<const uint4*>(a.w + ((long)j * 27 + t) * cin + 8 * cg);
                }
                acc[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, as_bf16x8(bv), acc[b], 0, 0, 0);
            }
        }
    }
    // epilogue: lane holds D[row = 4*lq + v][col = li] of each output block
#pragma unroll
    for (int b = 0; b < kMaxOutBlocks; ++b) {
        if (b >= nob) continue;
        const int j = 16 * b + li;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const long q = p0 + 4 * lq + v;
            if (q >= npix_out) continue;
            float y = acc[b][v];
            if (a.bias) y += a.bias[j];
            if (a.relu) y = fmaxf(y, 0.f);
            if (a.mask && bf2f(a.mask[q * a.cout + j]) <= 0.f) y = 0.f;
            if (a.out_f) {
                if (a.accumulate) a.out_f[q * a.cout + j] += y;
                else a.out_f[q * a.cout + j] = y;
            }
            if (a.out_bf) a.out_bf[q * a.cout + j] = f2bf(a.accumulate && a.out_f ? a.out_f[q * a.cout + j] : y);
        }
    }
}

// dW[j][t][c] += sum_p dY[p][j] * X(src(p,t))[c]   (FP32 VALU, LDS staged),
// db[j] += sum_p dY[p][j].  One block per (tap, pixel chunk).
__global__ __launch_bounds__(256) void k_conv3d_wgrad(WGradArgs a) {
    extern __shared__ float sm[];
    const int cin = a.c0 + a.c1, cout = a.cout;
    const int t = blockIdx.y;
    const long npix = (long)a.vout.nx * a.vout.ny * a.vout.nz;
    const long pbeg = (long)blockIdx.x * a.chunk, pend = min(npix, pbeg + a.chunk);
    constexpr int TP = 32;  // pixels staged per step
    float* sdy = sm;                  // [TP][cout]
    float* sx = sm + TP * cout;       // [TP][cin]
    const int npairs = cout * cin;
    constexpr int MAXP = 24;          // pairs per thread (cout*cin <= 256*24)
    float accw[MAXP];
    float accb = 0.f;
#pragma unroll
    for (int k = 0; k < MAXP; ++k) accw[k] = 0.f;
    int dx, dy, dz;
    tap_off(t, dx, dy, dz);
    for (long pb = pbeg; pb < pend; pb += TP) {
        __syncthreads();
        for (int idx = threadIdx.x; idx < TP * cout; idx += 256) {
            const int pp = idx / cout, j = idx - pp * cout;
            const long p = pb + pp;
            sdy[idx] = (p < pend) ? bf2f(a.dy[p * cout + j]) : 0.f;
        }
        for (int idx = threadIdx.x; idx < TP * cin; idx += 256) {
            const int pp = idx / cin, c = idx - pp * cin;
            const long p = pb + pp;
            float v = 0.f;
            if (p < pend) {
                const int ox = (int)(p % a.vout.nx), oy = (int)((p / a.vout.nx) % a.vout.ny),
                          oz = (int)(p / ((long)a.vout.nx * a.vout.ny));
                const int sx = a.stride * ox + dx, sy = a.stride * oy + dy, sz = a.stride * oz + dz;
                if (sx >= 0 && sy >= 0 && sz >= 0 && sx < a.vin.nx && sy < a.vin.ny && sz < a.vin.nz) {
                    if (c < a.c0) {
                        const long sp = a.up0 ? (long)(sx >> 1) + (long)a.vc0.nx * ((sy >> 1) + (long)a.vc0.ny * (sz >> 1))
                                              : (long)sx + (long)a.vin.nx * (sy + (long)a.vin.ny * sz);
                        v = bf2f(a.in0[sp * a.c0 + c]);
                    } else {
                        const long sp = (long)sx + (long)a.vin.nx * (sy + (long)a.vin.ny * sz);
                        v = bf2f(a.in1[sp * a.c1 + (c - a.c0)]);
                    }
                }
            }
            sx[idx] = v;
        }
        __syncthreads();
        const int np = (int)min((long)TP, pend - pb);
        for (int pp = 0; pp < np; ++pp) {
#pragma unroll
            for (int k = 0; k < MAXP; ++k) {
                const int pr = threadIdx.x + 256 * k;
                if (pr < npairs) {
                    const int j = pr / cin, c = pr - j * cin;
                    accw[k] += sdy[pp * cout + j] * sx[pp * cin + c];
                }
            }
            if (t == 13 && threadIdx.x < cout) accb += sdy[pp * cout + threadIdx.x];
        }
    }
#pragma unroll
    for (int k = 0; k < MAXP; ++k) {
        const int pr = threadIdx.x + 256 * k;
        if (pr < npairs) {
            const int j = pr / cin, c = pr - j * cin;
            atomicAdd(&a.dw[((long)j * 27 + t) * cin + c], accw[k]);
        }
    }
    if (t == 13 && threadIdx.x < cout && a.db) atomicAdd(&a.db[threadIdx.x], accb);
}

// coarse[P][c] += sum over the fine pixels that map to P (adjoint of the
// nearest upsample)
__global__ __launch_bounds__(256) void k_upsample_adjoint(const float* __restrict__ fine, int c, Vol vf, Vol vc,
                                                          float* __restrict__ coarse) {
    const long n = (long)vc.nx * vc.ny * vc.nz * c;
    for (long idx = (long)blockIdx.x * 256 + threadIdx.x; idx < n; idx += (long)gridDim.x * 256) {
        const long P = idx / c;
        const int ch = (int)(idx - P * c);
        const int X = (int)(P % vc.nx), Y = (int)((P / vc.nx) % vc.ny), Z = (int)(P / ((long)vc.nx * vc.ny));
        float s = 0.f;
        for (int dz = 0; dz < 2; ++dz)
            for (int dy = 0; dy < 2; ++dy)
                for (int dx = 0; dx < 2; ++dx) {
                    const int x = 2 * X + dx, y = 2 * Y + dy, z = 2 * Z + dz;
                    if (x < vf.nx && y < vf.ny && z < vf.nz)
                        s += fine[((long)x + (long)vf.nx * (y + (long)vf.ny * z)) * c + ch];
                }
        coarse[idx] += s;
    }
}

__global__ __launch_bounds__(256) void k_f32_to_bf16(long n, const float* __restrict__ a, uint16_t* __restrict__ b,
                                                     const uint16_t* __restrict__ mask) {
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        float v = a[i];
        if (mask && bf2f(mask[i]) <= 0.f) v = 0.f;
        b[i] = f2bf(v);
    }
}

// W[j][t][c] -> Wt[c][t][j] (weights of the adjoint convolution)
__global__ __launch_bounds__(256) void k_weight_transpose(const uint16_t* __restrict__ w, int cout, int cin,
                                                          uint16_t* __restrict__ wt) {
    const long n = (long)cout * 27 * cin;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        const int j = (int)(i / (27L * cin));
        const int t = (int)((i / cin) % 27);
        const int c = (int)(i % cin);
        wt[((long)c * 27 + t) * cout + j] = w[i];
    }
}

// Adam step on FP32 master weights, refresh the BF16 copy
__global__ __launch_bounds__(256) void k_adam(long n, float* __restrict__ w, float* __restrict__ g,
                                              float* __restrict__ m, float* __restrict__ v,
                                              uint16_t* __restrict__ wbf, float lr, float b1, float b2, float eps,
                                              float c1, float c2, float clip) {
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        float gi = g[i] * clip;
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        w[i] = wi;
        if (wbf) wbf[i] = f2bf(wi);
        g[i] = 0.f;
    }
}

__global__ __launch_bounds__(256) void k_fillf(long n, float v, float* a) {
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) a[i] = v;
}

// splitmix64 finaliser (Steele, Lea & Flood 2014): a counter-based hash that
// gives every lattice value its own stream without state
__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
// standard normal value of lattice key q (Box-Muller on two hashed uniforms in (0, 1))
__device__ __forceinline__ double lattice_normal(unsigned long long seed, unsigned long long q) {
    const unsigned long long h1 = mix64(seed ^ mix64(q)), h2 = mix64(h1 ^ 0xD1B54A32D192ED03ull);
    const double u1 = ((double)(h1 >> 11) + 0.5) * 0x1p-53, u2 = (double)(h2 >> 11) * 0x1p-53;
    return sqrt(-2.0 * log(u1)) * cospi(2.0 * u2);
}

// Manufactured displacement for training: on every level l = 1..levels the
// standard normal values of the lattice of spacing 2^l (one per lattice point
// and component) interpolated trilinearly to the nodes, weighted 2^(l-levels)
// (smoother parts larger), summed; 0 where minv = 0 (not solved for).
__global__ __launch_bounds__(256) void k_smooth_field(GridArgs g, unsigned long long seed, int levels,
                                                      const double* __restrict__ minv, double* __restrict__ e) {
    const long n = (long)blockIdx.x * 256 + threadIdx.x;
    if (n >= g.nn) return;
    const int i = (int)(n % g.NX), j = (int)((n / g.NX) % g.NY), k = (int)(n / ((long)g.NX * g.NY));
    double v[3] = {0.0, 0.0, 0.0};
    for (int l = 1; l <= levels; ++l) {
        const int h = 1 << l;
        const long LX = (g.NX + h - 1) / h + 1, LY = (g.NY + h - 1) / h + 1;
        const int I = i >> l, J = j >> l, K = k >> l;
        const double fx = (double)(i & (h - 1)) / h, fy = (double)(j & (h - 1)) / h, fz = (double)(k & (h - 1)) / h;
        const double w = ldexp(1.0, l - levels);
        for (int c8 = 0; c8 < 8; ++c8) {
            const int ox = c8 & 1, oy = (c8 >> 1) & 1, oz = c8 >> 2;
            const double wt = w * (ox ? fx : 1.0 - fx) * (oy ? fy : 1.0 - fy) * (oz ? fz : 1.0 - fz);
            const unsigned long long pt =
                ((unsigned long long)l << 56) ^ (unsigned long long)(((long)(K + oz) * LY + (J + oy)) * LX + (I + ox));
#pragma unroll
            for (int c = 0; c < 3; ++c) v[c] += wt * lattice_normal(seed, 3ull * pt + c);
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) e[3 * n + c] = minv[3 * n + c] != 0.0 ? v[c] : 0.0;
}

// Network input from a residual: ch 0..2 = r/scale at the node, ch 3 = the
// node's mean element modulus on a log scale, ch 4..7 = 0.
__global__ __launch_bounds__(256) void k_pack_input(GridArgs g, const double* __restrict__ r, const double* __restrict__ E,
                                                    double inv_scale, const double* __restrict__ rr_dev,
                                                    uint16_t* __restrict__ out) {
    const long n = (long)blockIdx.x * 256 + threadIdx.x;
    if (n >= g.nn) return;
    if (rr_dev) inv_scale = 1.0 / sqrt(*rr_dev / (3.0 * (double)g.nn));  // 1 / rms(r) from r'r on the device
    const int i = (int)(n % g.NX), j = (int)((n / g.NX) % g.NY), k = (int)(n / ((long)g.NX * g.NY));
    double s = 0.0;
    int cnt = 0;
    for (int dz = 0; dz < 2; ++dz)
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) {
                const int ie = i - 1 + dx, je = j - 1 + dy, ke = k - 1 + dz;
                if (ie < 0 || je < 0 || ke < 0 || ie >= g.nx || je >= g.ny || ke >= g.nz) continue;
                s += E[(long)ke * g.nx * g.ny + (long)je * g.nx + ie];
                ++cnt;
            }
    const double em = cnt ? s / cnt : 0.0;
    const float dens = em > 0.0 ? (float)(log10(em) / 9.0 + 1.0) : 0.f;  // E in [1e-9, 1] -> [0, 1]
    uint16_t* o = out + n * 8;
    for (int c = 0; c < 3; ++c) o[c] = f2bf((float)(r[3 * n + c] * inv_scale));
    o[3] = f2bf(dens);
    for (int c = 4; c < 8; ++c) o[c] = 0;
}

// z = minv r + scale * net(r)  (fixed DOFs stay 0)
__global__ __launch_bounds__(256) void k_unpack_output(long nn, const float* __restrict__ y, int cy,
                                                       const double* __restrict__ r, const double* __restrict__ minv,
                                                       double scale, double* __restrict__ z) {
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < 3 * nn; i += (long)gridDim.x * 256) {
        const long n = i / 3;
        const int c = (int)(i - 3 * n);
        const double m = minv[i];
        z[i] = m != 0.0 ? m * r[i] + scale * (double)y[n * cy + c] : 0.0;
    }
}

// training loss: L = sum |y - t|^2 / sum |t|^2 over free DOFs; dL/dy into g
__global__ __launch_bounds__(256) void k_loss_grad(long nn, const float* __restrict__ y, int cy,
                                                   const double* __restrict__ tgt, const double* __restrict__ minv,
                                                   double tscale, double inv_tnorm2, float* __restrict__ g,
                                                   double* __restrict__ loss_acc) {
    __shared__ double red[4];
    double s = 0.0;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < 3 * nn; i += (long)gridDim.x * 256) {
        const long n = i / 3;
        const int c = (int)(i - 3 * n);
        const double d = minv[i] != 0.0 ? (double)y[n * cy + c] - tgt[i] * tscale : 0.0;
        g[n * cy + c] = (float)(2.0 * d * inv_tnorm2);
        s += d * d;
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(loss_acc, (red[0] + red[1] + red[2] + red[3]) * inv_tnorm2);
}

inline unsigned nblk(long n) { return (unsigned)std::min<long>((n + 255) / 256, 65535L * 4); }

}  // namespace

void launch_conv3d(const ConvArgs& a, hipStream_t s) {
    const long npix = (long)a.vout.nx * a.vout.ny * a.vout.nz;
    const unsigned nb = (unsigned)((npix + 16 * kWaves - 1) / (16 * kWaves));
    hipLaunchKernelGGL(k_conv3d, dim3(nb), dim3(256), 0, s, a);
}

void launch_conv3d_wgrad(const WGradArgs& a, hipStream_t s) {
    const long npix = (long)a.vout.nx * a.vout.ny * a.vout.nz;
    const unsigned nb = (unsigned)((npix + a.chunk - 1) / a.chunk);
    const size_t shm = 32 * (size_t)(a.cout + a.c0 + a.c1) * sizeof(float);
    hipLaunchKernelGGL(k_conv3d_wgrad, dim3(nb, 27), dim3(256), shm, s, a);
}

void launch_upsample_adjoint(const float* fine, int c, Vol vf, Vol vc, float* coarse, hipStream_t s) {
    hipLaunchKernelGGL(k_upsample_adjoint, dim3(nblk((long)vc.nx * vc.ny * vc.nz * c)), dim3(256), 0, s, fine, c, vf,
                       vc, coarse);
}
void launch_f32_to_bf16(long n, const float* a, uint16_t* b, const uint16_t* mask, hipStream_t s) {
    hipLaunchKernelGGL(k_f32_to_bf16, dim3(nblk(n)), dim3(256), 0, s, n, a, b, mask);
}
void launch_weight_transpose(const uint16_t* w, int cout, int cin, uint16_t* wt, hipStream_t s) {
    hipLaunchKernelGGL(k_weight_transpose, dim3(nblk((long)cout * 27 * cin)), dim3(256), 0, s, w, cout, cin, wt);
}
void launch_adam(long n, float* w, float* g, float* m, float* v, uint16_t* wbf, float lr, float b1, float b2,
                 float eps, float c1, float c2, float clip, hipStream_t s) {
    hipLaunchKernelGGL(k_adam, dim3(nblk(n)), dim3(256), 0, s, n, w, g, m, v, wbf, lr, b1, b2, eps, c1, c2, clip);
}
void launch_fillf(long n, float v, float* a, hipStream_t s) {
    hipLaunchKernelGGL(k_fillf, dim3(nblk(n)), dim3(256), 0, s, n, v, a);
}
void launch_smooth_field(const GridArgs& g, unsigned long long seed, int levels, const double* minv, double* e,
                         hipStream_t s) {
    hipLaunchKernelGGL(k_smooth_field, dim3((unsigned)((g.nn + 255) / 256)), dim3(256), 0, s, g, seed, levels, minv,
                       e);
}
void launch_pack_input(const GridArgs& g, const double* r, const double* E, double inv_scale, uint16_t* out,
                       hipStream_t s, const double* rr_dev) {
    hipLaunchKernelGGL(k_pack_input, dim3((unsigned)((g.nn + 255) / 256)), dim3(256), 0, s, g, r, E, inv_scale,
                       rr_dev, out);
}
void launch_unpack_output(long nn, const float* y, int cy, const double* r, const double* minv, double scale,
                          double* z, hipStream_t s) {
    hipLaunchKernelGGL(k_unpack_output, dim3(nblk(3 * nn)), dim3(256), 0, s, nn, y, cy, r, minv, scale, z);
}
void launch_loss_grad(long nn, const float* y, int cy, const double* tgt, const double* minv, double tscale,
                      double inv_tnorm2, float* g, double* loss_acc, hipStream_t s) {
    hipLaunchKernelGGL(k_loss_grad, dim3(nblk(3 * nn)), dim3(256), 0, s, nn, y, cy, tgt, minv, tscale, inv_tnorm2, g,
                       loss_acc);
}

}  // namespace hx
