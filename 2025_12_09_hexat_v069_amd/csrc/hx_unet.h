// hx_unet.h -- learned preconditioner (3-D U-Net, BF16/MFMA) declarations.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hx_internal.h"

namespace hx {

struct Vol {
    int nx, ny, nz;
};

constexpr int kMaxOutBlocks = 6;  // output channels per conv <= 96

// y = conv3x3x3(concat(up?(in0), in1)) (+bias, ReLU, mask)
struct ConvArgs {
    const uint16_t* in0;  // [pix][c0] bf16
    int c0;
    int up0;              // in0 is on the coarse grid vc0 (nearest upsample to vin)
    Vol vc0;
    const uint16_t* in1;  // [pix][c1] bf16 (nullable when c1 == 0)
    int c1;
    Vol vin;              // grid of the (fine) input
    Vol vout;             // grid of the output
    int stride;           // 1 or 2
    int transposed;       // 0: gather src = stride*out + off; 1: adjoint (out = stride*src + off)
    const uint16_t* w;    // [cout][27][c0+c1] bf16
    const float* bias;    // [cout] (nullable)
    int cout;             // multiple of 16, <= 16*kMaxOutBlocks
    int relu;
    const uint16_t* mask; // [pix][cout] bf16 (nullable): zero where mask <= 0
    uint16_t* out_bf;     // [pix][cout] bf16 (nullable)
    float* out_f;         // [pix][cout] f32 (nullable)
    int accumulate;       // out_f += y
};

// dW[j][t][c] += sum_p dy[p][j] x(src(p,t))[c]; db[j] += sum_p dy[p][j]
struct WGradArgs {
    const uint16_t* dy;   // [pix_out][cout]
    int cout;
    const uint16_t* in0;
    int c0;
    int up0;
    Vol vc0;
    const uint16_t* in1;
    int c1;
    Vol vin, vout;
    int stride;
    float* dw;            // [cout][27][cin] f32
    float* db;            // [cout] f32 (nullable)
    int chunk;            // output pixels per block
};

void launch_conv3d(const ConvArgs& a, hipStream_t s);
void launch_conv3d_wgrad(const WGradArgs& a, hipStream_t s);
void launch_upsample_adjoint(const float* fine, int c, Vol vf, Vol vc, float* coarse, hipStream_t s);
void launch_f32_to_bf16(long n, const float* a, uint16_t* b, const uint16_t* mask, hipStream_t s);
void launch_weight_transpose(const uint16_t* w, int cout, int cin, uint16_t* wt, hipStream_t s);
void launch_adam(long n, float* w, float* g, float* m, float* v, uint16_t* wbf, float lr, float b1, float b2,
                 float eps, float c1, float c2, float clip, hipStream_t s);
void launch_fillf(long n, float v, float* a, hipStream_t s);
// rr_dev non-null: the scale is 1 / rms(r) from r'r held on the device there (inv_scale unused)
void launch_pack_input(const GridArgs& g, const double* r, const double* E, double inv_scale, uint16_t* out,
                       hipStream_t s, const double* rr_dev = nullptr);
// e = sum_{l=1..levels} 2^(l-levels) * trilinear interpolation of hashed standard
// normal values on the lattice of spacing 2^l; 0 where minv = 0
void launch_smooth_field(const GridArgs& g, unsigned long long seed, int levels, const double* minv, double* e,
                         hipStream_t s);
void launch_unpack_output(long nn, const float* y, int cy, const double* r, const double* minv, double scale,
                          double* z, hipStream_t s);
void launch_loss_grad(long nn, const float* y, int cy, const double* tgt, const double* minv, double tscale,
                      double inv_tnorm2, float* g, double* loss_acc, hipStream_t s);

}  // namespace hx
