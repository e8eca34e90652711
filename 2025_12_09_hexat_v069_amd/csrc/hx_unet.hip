// hx_unet.hip -- the learned preconditioner: network definition, forward and
// backward passes, Adam training on (density, residual, error) triples that
// the library's own Jacobi-PCG generates, and the flexible-PCG solve that
// uses it.
//
// Network (node grid G0 of the problem, levels G1..G3 each half the size):
//   in  [G0][8]  : residual r/|r|_rms (3 ch), log-modulus density (1 ch), 0
//   L0  stride-2 conv  8->16  ReLU   G0->G1   (learned restriction)
//   L1  conv 16->16 ReLU (skip s1)   G1
//   L2  stride-2 conv 16->32 ReLU    G1->G2
//   L3  conv 32->32 ReLU (skip s2)   G2
//   L4  stride-2 conv 32->64 ReLU    G2->G3
//   L5  conv 64->64 ReLU             G3
//   L6  conv [up(b3) 64 | s2 32] -> 32 ReLU   G2
//   L7  conv [up(u2) 32 | s1 16] -> 16 ReLU   G1
//   L8  conv 16->16 (first 3 channels used)    G1
//   out trilinear prolongation G1->G0;  z = D^-1 r + |r|_rms * out
// The three levels carry 16/32/64 channels with 3x3x3 kernels; the finest
// level works at half the node resolution because the Jacobi term already
// treats the highest frequencies, which keeps one application at ~1/8 of the
// cost of a full-resolution first level.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "hx_launch.h"
#include "hx_unet.h"

namespace hx {

namespace {

uint16_t host_f2bf(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

template <typename T>
struct Buf {
    T* p = nullptr;
    size_t n = 0;
    bool ensure(size_t m) {
        if (m <= n && p) return true;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        if (hipMalloc((void**)&p, std::max<size_t>(m, 1) * sizeof(T)) != hipSuccess) return false;
        n = m;
        return true;
    }
    ~Buf() {
        if (p) (void)hipFree(p);
    }
};

__host__ __device__ Vol half(Vol v) { return Vol{(v.nx + 1) / 2, (v.ny + 1) / 2, (v.nz + 1) / 2}; }
__host__ __device__ long npix(Vol v) { return (long)v.nx * v.ny * v.nz; }

struct Layer {
    int cin0, cin1, cout, stride, relu;
    int up0;            // in0 comes from the next coarser level
    long woff, boff;    // offsets into the parameter vector
};

// prolongation G1 -> G0 (trilinear on the node lattice) and its adjoint
__global__ void k_prolong(const float* __restrict__ yc, int cy, Vol vc, Vol vf, float* __restrict__ yf) {
    const long n = npix(vf);
    for (long p = (long)blockIdx.x * 256 + threadIdx.x; p < n; p += (long)gridDim.x * 256) {
        const int i = (int)(p % vf.nx), j = (int)((p / vf.nx) % vf.ny), k = (int)(p / ((long)vf.nx * vf.ny));
        int I[3][2];
        float W[3][2];
        const int ii[3] = {i, j, k}, nc[3] = {vc.nx, vc.ny, vc.nz};
        for (int d = 0; d < 3; ++d) {
            const int a = ii[d] >> 1;
            if ((ii[d] & 1) == 0 || a + 1 >= nc[d]) {
                I[d][0] = a; W[d][0] = 1.f; I[d][1] = a; W[d][1] = 0.f;
            } else {
                I[d][0] = a; W[d][0] = 0.5f; I[d][1] = a + 1; W[d][1] = 0.5f;
            }
        }
        float s[3] = {0.f, 0.f, 0.f};
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b)
                for (int c = 0; c < 2; ++c) {
                    const float w = W[0][a] * W[1][b] * W[2][c];
                    if (w == 0.f) continue;
                    const long q = (long)I[0][a] + (long)vc.nx * (I[1][b] + (long)vc.ny * I[2][c]);
                    for (int ch = 0; ch < 3; ++ch) s[ch] += w * yc[q * cy + ch];
                }
        for (int ch = 0; ch < 3; ++ch) yf[p * 3 + ch] = s[ch];
    }
}

__device__ float prolong_w(int i, int I, int nc) {  // weight of coarse I in fine i (1-D)
    const int a = i >> 1;
    if ((i & 1) == 0 || a + 1 >= nc) return (I == a) ? 1.f : 0.f;
    return (I == a || I == a + 1) ? 0.5f : 0.f;
}

__global__ void k_prolong_adjoint(const float* __restrict__ gf, Vol vf, Vol vc, int cy, float* __restrict__ gc) {
    const long n = npix(vc);
    for (long P = (long)blockIdx.x * 256 + threadIdx.x; P < n; P += (long)gridDim.x * 256) {
        const int I = (int)(P % vc.nx), J = (int)((P / vc.nx) % vc.ny), K = (int)(P / ((long)vc.nx * vc.ny));
        float s[3] = {0.f, 0.f, 0.f};
        for (int k = 2 * K - 1; k <= 2 * K + 1; ++k) {
            if (k < 0 || k >= vf.nz) continue;
            const float wk = prolong_w(k, K, vc.nz);
            if (wk == 0.f) continue;
            for (int j = 2 * J - 1; j <= 2 * J + 1; ++j) {
                if (j < 0 || j >= vf.ny) continue;
                const float wj = prolong_w(j, J, vc.ny);
                if (wj == 0.f) continue;
                for (int i = 2 * I - 1; i <= 2 * I + 1; ++i) {
                    if (i < 0 || i >= vf.nx) continue;
                    const float wi = prolong_w(i, I, vc.nx);
                    if (wi == 0.f) continue;
                    const long q = (long)i + (long)vf.nx * (j + (long)vf.ny * k);
                    for (int ch = 0; ch < 3; ++ch) s[ch] += wi * wj * wk * gf[q * 3 + ch];
                }
            }
        }
        for (int ch = 0; ch < cy; ++ch) gc[P * cy + ch] = ch < 3 ? s[ch] : 0.f;
    }
}

// y (fine network output, FP32) -> FP64 direction on free DOFs, times scale
__global__ void k_direction(long nn, const float* __restrict__ yf, const double* __restrict__ minv, double scale,
                            double* __restrict__ y) {
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < 3 * nn; i += (long)gridDim.x * 256)
        y[i] = minv[i] != 0.0 ? scale * (double)yf[i] : 0.0;
}

// dL/dy = -2 b (r - a Au - b Ay) / e'r on free DOFs (envelope of the 2x2
// energy minimisation over span{D^-1 r, y})
__global__ void k_subspace_grad(long nn, const double* __restrict__ r, const double* __restrict__ Au,
                                const double* __restrict__ Ay, const double* __restrict__ minv, double a, double b,
                                double scale, float* __restrict__ gf) {
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < 3 * nn; i += (long)gridDim.x * 256)
        gf[i] = minv[i] != 0.0 ? (float)(scale * (r[i] - a * Au[i] - b * Ay[i])) : 0.f;
}

unsigned gb(long n) { return (unsigned)std::min<long>((n + 255) / 256, 8192); }

}  // namespace

// ---------------------------------------------------------------- network
struct UNet {
    GridArgs g{};
    Vol v0{}, v1{}, v2{}, v3{};
    std::vector<Layer> L;
    long nparam = 0;
    bool ready = false;     // weights initialised
    bool trained = false;
    Buf<float> w32, grad, m, v;   // master weights / Adam moments
    Buf<uint16_t> wbf, wt;        // BF16 weights and transposed copy
    // activations (bf16) and gradients (f32)
    Buf<uint16_t> x0, a1, s1, a2, s2, a3, b3, u2, u1, dyb;
    Buf<float> y1, yf, gy1, gyf, da1, ds1, da2, ds2, da3, db3, du2, du1, tmp;
    Buf<double> sc;  // scratch scalars
    int adam_t = 0;

    long wsize(int l) const { return (long)L[l].cout * 27 * (L[l].cin0 + L[l].cin1); }

    void define() {
        L.clear();
        auto add = [&](int c0, int c1, int co, int st, int relu, int up) {
            Layer l{c0, c1, co, st, relu, up, nparam, 0};
            nparam += (long)co * 27 * (c0 + c1);
            l.boff = nparam;
            nparam += co;
            L.push_back(l);
        };
        add(8, 0, 16, 2, 1, 0);   // L0
        add(16, 0, 16, 1, 1, 0);  // L1
        add(16, 0, 32, 2, 1, 0);  // L2
        add(32, 0, 32, 1, 1, 0);  // L3
        add(32, 0, 64, 2, 1, 0);  // L4
        add(64, 0, 64, 1, 1, 0);  // L5
        add(64, 32, 32, 1, 1, 1); // L6
        add(32, 16, 16, 1, 1, 1); // L7
        add(16, 0, 16, 1, 0, 0);  // L8
    }

    bool alloc_params() {
        return w32.ensure(nparam) && grad.ensure(nparam) && m.ensure(nparam) && v.ensure(nparam) &&
               wbf.ensure(nparam) && wt.ensure(nparam) && sc.ensure(8);
    }

    bool bind(const GridArgs& gg) {
        g = gg;
        v0 = Vol{g.NX, g.NY, g.NZ};
        v1 = half(v0);
        v2 = half(v1);
        v3 = half(v2);
        const long p0 = npix(v0), p1 = npix(v1), p2 = npix(v2), p3 = npix(v3);
        return x0.ensure(p0 * 8) && a1.ensure(p1 * 16) && s1.ensure(p1 * 16) && a2.ensure(p2 * 32) &&
               s2.ensure(p2 * 32) && a3.ensure(p3 * 64) && b3.ensure(p3 * 64) && u2.ensure(p2 * 32) &&
               u1.ensure(p1 * 16) && y1.ensure(p1 * 16) && yf.ensure(p0 * 3) && gy1.ensure(p1 * 16) &&
               gyf.ensure(p0 * 3) && da1.ensure(p1 * 16) && ds1.ensure(p1 * 16) && da2.ensure(p2 * 32) &&
               ds2.ensure(p2 * 32) && da3.ensure(p3 * 64) && db3.ensure(p3 * 64) && du2.ensure(p2 * 32) &&
               du1.ensure(p1 * 16) && dyb.ensure(std::max({p1 * 16, p2 * 32, p3 * 64})) &&
               tmp.ensure(std::max({p1 * 64, p2 * 64, p3 * 64}));
    }

    void init_weights(unsigned seed, hipStream_t s) {
        std::vector<float> h(nparam, 0.f);
        std::mt19937 rng(seed);
        for (size_t l = 0; l < L.size(); ++l) {
            const int fan_in = 27 * (L[l].cin0 + L[l].cin1);
            const float sd = (l + 1 == L.size()) ? 0.1f * std::sqrt(1.f / fan_in) : std::sqrt(2.f / fan_in);
            std::normal_distribution<float> nd(0.f, sd);
            for (long i = 0; i < wsize(l); ++i) h[L[l].woff + i] = nd(rng);
        }
        std::vector<uint16_t> hb(nparam);
        for (long i = 0; i < nparam; ++i) hb[i] = host_f2bf(h[i]);
        (void)hipMemcpyAsync(w32.p, h.data(), nparam * 4, hipMemcpyHostToDevice, s);
        (void)hipMemcpyAsync(wbf.p, hb.data(), nparam * 2, hipMemcpyHostToDevice, s);
        launch_fillf(nparam, 0.f, grad.p, s);
        launch_fillf(nparam, 0.f, m.p, s);
        launch_fillf(nparam, 0.f, v.p, s);
        (void)hipStreamSynchronize(s);
        adam_t = 0;
        ready = true;
    }

    ConvArgs conv(int l, const uint16_t* in0, Vol vc0, const uint16_t* in1, Vol vin, Vol vout, uint16_t* out_bf,
                  float* out_f) const {
        ConvArgs a{};
        a.in0 = in0;
        a.c0 = L[l].cin0;
        a.up0 = L[l].up0;
        a.vc0 = vc0;
        a.in1 = in1;
        a.c1 = L[l].cin1;
        a.vin = vin;
        a.vout = vout;
        a.stride = L[l].stride;
        a.transposed = 0;
        a.w = wbf.p + L[l].woff;
        a.bias = w32.p + L[l].boff;
        a.cout = L[l].cout;
        a.relu = L[l].relu;
        a.out_bf = out_bf;
        a.out_f = out_f;
        return a;
    }

    // forward from the packed input x0; leaves the fine 3-channel output in yf
    void forward(hipStream_t s) {
        launch_conv3d(conv(0, x0.p, v0, nullptr, v0, v1, a1.p, nullptr), s);
        launch_conv3d(conv(1, a1.p, v1, nullptr, v1, v1, s1.p, nullptr), s);
        launch_conv3d(conv(2, s1.p, v1, nullptr, v1, v2, a2.p, nullptr), s);
        launch_conv3d(conv(3, a2.p, v2, nullptr, v2, v2, s2.p, nullptr), s);
        launch_conv3d(conv(4, s2.p, v2, nullptr, v2, v3, a3.p, nullptr), s);
        launch_conv3d(conv(5, a3.p, v3, nullptr, v3, v3, b3.p, nullptr), s);
        launch_conv3d(conv(6, b3.p, v3, s2.p, v2, v2, u2.p, nullptr), s);
        launch_conv3d(conv(7, u2.p, v2, s1.p, v1, v1, u1.p, nullptr), s);
        launch_conv3d(conv(8, u1.p, v1, nullptr, v1, v1, nullptr, y1.p), s);
        hipLaunchKernelGGL(k_prolong, dim3(gb(npix(v0))), dim3(256), 0, s, y1.p, 16, v1, v0, yf.p);
    }

    // adjoint convolution of layer l: output channels [c_lo, c_lo+c_n) of its
    // input, at the layer's input grid, written/accumulated into out (f32)
    void adjoint(int l, const uint16_t* dy, Vol vout_fwd, Vol vin_fwd, int c_lo, int c_n, float* out, int acc,
                 hipStream_t s) {
        const int cin = L[l].cin0 + L[l].cin1;
        ConvArgs a{};
        a.in0 = dy;
        a.c0 = L[l].cout;
        a.up0 = 0;
        a.vc0 = vout_fwd;
        a.in1 = nullptr;
        a.c1 = 0;
        a.vin = vout_fwd;
        a.vout = vin_fwd;
        a.stride = L[l].stride;
        a.transposed = 1;
        a.w = wt.p + L[l].woff + (long)c_lo * 27 * L[l].cout;
        a.bias = nullptr;
        a.cout = c_n;
        a.relu = 0;
        a.out_f = out;
        a.accumulate = acc;
        (void)cin;
        launch_conv3d(a, s);
    }

    void wgrad(int l, const uint16_t* dy, const uint16_t* in0, Vol vc0, const uint16_t* in1, Vol vin, Vol vout,
               hipStream_t s) {
        WGradArgs a{};
        a.dy = dy;
        a.cout = L[l].cout;
        a.in0 = in0;
        a.c0 = L[l].cin0;
        a.up0 = L[l].up0;
        a.vc0 = vc0;
        a.in1 = in1;
        a.c1 = L[l].cin1;
        a.vin = vin;
        a.vout = vout;
        a.stride = L[l].stride;
        a.dw = grad.p + L[l].woff;
        a.db = grad.p + L[l].boff;
        a.chunk = 2048;
        launch_conv3d_wgrad(a, s);
    }

    // backward from dL/dyf (gyf); accumulates parameter gradients in grad
    void backward(hipStream_t s) {
        for (size_t l = 0; l < L.size(); ++l)
            launch_weight_transpose(wbf.p + L[l].woff, L[l].cout, L[l].cin0 + L[l].cin1, wt.p + L[l].woff, s);
        const long p1 = npix(v1), p2 = npix(v2), p3 = npix(v3);
        hipLaunchKernelGGL(k_prolong_adjoint, dim3(gb(p1)), dim3(256), 0, s, gyf.p, v0, v1, 16, gy1.p);
        // L8 (no ReLU)
        launch_f32_to_bf16(p1 * 16, gy1.p, dyb.p, nullptr, s);
        wgrad(8, dyb.p, u1.p, v1, nullptr, v1, v1, s);
        adjoint(8, dyb.p, v1, v1, 0, 16, du1.p, 0, s);
        // L7: in = [up(u2) | s1]
        launch_f32_to_bf16(p1 * 16, du1.p, dyb.p, u1.p, s);  // through L7's ReLU
        wgrad(7, dyb.p, u2.p, v2, s1.p, v1, v1, s);
        adjoint(7, dyb.p, v1, v1, 0, 32, tmp.p, 0, s);
        launch_fillf(p2 * 32, 0.f, du2.p, s);
        launch_upsample_adjoint(tmp.p, 32, v1, v2, du2.p, s);
        adjoint(7, dyb.p, v1, v1, 32, 16, ds1.p, 0, s);
        // L6: in = [up(b3) | s2]
        launch_f32_to_bf16(p2 * 32, du2.p, dyb.p, u2.p, s);
        wgrad(6, dyb.p, b3.p, v3, s2.p, v2, v2, s);
        adjoint(6, dyb.p, v2, v2, 0, 64, tmp.p, 0, s);
        launch_fillf(p3 * 64, 0.f, db3.p, s);
        launch_upsample_adjoint(tmp.p, 64, v2, v3, db3.p, s);
        adjoint(6, dyb.p, v2, v2, 64, 32, ds2.p, 0, s);
        // L5
        launch_f32_to_bf16(p3 * 64, db3.p, dyb.p, b3.p, s);
        wgrad(5, dyb.p, a3.p, v3, nullptr, v3, v3, s);
        adjoint(5, dyb.p, v3, v3, 0, 64, da3.p, 0, s);
        // L4 (stride 2, G2 -> G3)
        launch_f32_to_bf16(p3 * 64, da3.p, dyb.p, a3.p, s);
        wgrad(4, dyb.p, s2.p, v2, nullptr, v2, v3, s);
        adjoint(4, dyb.p, v3, v2, 0, 32, ds2.p, 1, s);
        // L3
        launch_f32_to_bf16(p2 * 32, ds2.p, dyb.p, s2.p, s);
        wgrad(3, dyb.p, a2.p, v2, nullptr, v2, v2, s);
        adjoint(3, dyb.p, v2, v2, 0, 32, da2.p, 0, s);
        // L2 (stride 2, G1 -> G2)
        launch_f32_to_bf16(p2 * 32, da2.p, dyb.p, a2.p, s);
        wgrad(2, dyb.p, s1.p, v1, nullptr, v1, v2, s);
        adjoint(2, dyb.p, v2, v1, 0, 16, ds1.p, 1, s);
        // L1
        launch_f32_to_bf16(p1 * 16, ds1.p, dyb.p, s1.p, s);
        wgrad(1, dyb.p, a1.p, v1, nullptr, v1, v1, s);
        adjoint(1, dyb.p, v1, v1, 0, 16, da1.p, 0, s);
        // L0 (stride 2, G0 -> G1)
        launch_f32_to_bf16(p1 * 16, da1.p, dyb.p, a1.p, s);
        wgrad(0, dyb.p, x0.p, v0, nullptr, v0, v1, s);
    }

    void adam(float lr, float clip, hipStream_t s) {
        ++adam_t;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.f - std::pow(b1, (float)adam_t), c2 = 1.f - std::pow(b2, (float)adam_t);
        launch_adam(nparam, w32.p, grad.p, m.p, v.p, wbf.p, lr, b1, b2, 1e-8f, c1, c2, clip, s);
    }
};

// -------------------------------------------------------------- public glue
UNet* unet_create() {
    UNet* u = new UNet();
    u->define();
    return u;
}
void unet_destroy(UNet* u) { delete u; }
long unet_nparam(const UNet* u) { return u->nparam; }
bool unet_ready(const UNet* u) { return u && u->ready; }

bool unet_bind(UNet* u, const GridArgs& g, unsigned seed, hipStream_t s) {
    if (!u->alloc_params()) return false;
    if (!u->bind(g)) return false;
    if (!u->ready) u->init_weights(seed, s);
    return true;
}

// Training pieces for the subspace objective: the caller forms the network's
// direction y, applies the operator and solves the 2x2 energy problem over
// span{D^-1 r, y}, then hands the output gradient back.
void unet_forward_sample(UNet* u, const double* r, const double* E, double inv_scale, hipStream_t s,
                         const double* rr_dev) {
    launch_pack_input(u->g, r, E, inv_scale, u->x0.p, s, rr_dev);
    u->forward(s);
}
// y = scale * net output on free DOFs
void unet_direction(UNet* u, const double* minv, double scale, double* y, hipStream_t s) {
    hipLaunchKernelGGL(k_direction, dim3(gb(3 * u->g.nn)), dim3(256), 0, s, u->g.nn, u->yf.p, minv, scale, y);
}
// output gradient: scale * (r - a Au - b Ay) on free DOFs
void unet_subspace_grad(UNet* u, const double* r, const double* Au, const double* Ay, const double* minv, double a,
                        double b, double scale, hipStream_t s) {
    hipLaunchKernelGGL(k_subspace_grad, dim3(gb(3 * u->g.nn)), dim3(256), 0, s, u->g.nn, r, Au, Ay, minv, a, b, scale,
                       u->gyf.p);
}
void unet_backward_step(UNet* u, float lr, hipStream_t s) {
    u->backward(s);
    u->adam(lr, 1.0f, s);
    u->trained = true;
}

int unet_weights_bf16(UNet* u, uint16_t* out, hipStream_t s) {
    if (out) {
        (void)hipMemcpyAsync(out, u->wbf.p, u->nparam * sizeof(uint16_t), hipMemcpyDeviceToHost, s);
        (void)hipStreamSynchronize(s);
    }
    return (int)u->nparam;
}

// ---- test hooks --------------------------------------------------------
void unet_init(UNet* u, unsigned seed, hipStream_t s) { u->init_weights(seed, s); }
void unet_params_f32(UNet* u, float* out, hipStream_t s) {
    (void)hipMemcpyAsync(out, u->w32.p, u->nparam * sizeof(float), hipMemcpyDeviceToHost, s);
    (void)hipStreamSynchronize(s);
}
void unet_forward_packed(UNet* u, const uint16_t* x0_host, float* y_host, hipStream_t s) {
    const long p0 = npix(u->v0);
    (void)hipMemcpyAsync(u->x0.p, x0_host, p0 * 8 * sizeof(uint16_t), hipMemcpyHostToDevice, s);
    u->forward(s);
    (void)hipMemcpyAsync(y_host, u->yf.p, p0 * 3 * sizeof(float), hipMemcpyDeviceToHost, s);
    (void)hipStreamSynchronize(s);
}
void unet_backward_host(UNet* u, const float* gy_host, float* grad_host, hipStream_t s) {
    const long p0 = npix(u->v0);
    launch_fillf(u->nparam, 0.f, u->grad.p, s);
    (void)hipMemcpyAsync(u->gyf.p, gy_host, p0 * 3 * sizeof(float), hipMemcpyHostToDevice, s);
    u->backward(s);
    (void)hipMemcpyAsync(grad_host, u->grad.p, u->nparam * sizeof(float), hipMemcpyDeviceToHost, s);
    (void)hipStreamSynchronize(s);
    launch_fillf(u->nparam, 0.f, u->grad.p, s);
}

}  // namespace hx
