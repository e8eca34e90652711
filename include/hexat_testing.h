/*
 * hexat_testing.h -- test hooks of libhexat.so for the learned
 * preconditioner's network.  They expose one forward and one backward pass
 * of the U-Net on caller buffers so the tests can compare them with a CPU
 * reference.  No reference counterpart; not used by the solver path.
 */
#ifndef HEXAT_TESTING_H
#define HEXAT_TESTING_H

#include <stdint.h>

#include "hexat.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Device memory for tests of the entry points that take device pointers
 * (hx_web_pack_device): allocate, free, and copy host <-> device
 * (to_device 1: src host, dst device; 0: src device, dst host). */
int hx_test_device_alloc(hx_ctx* ctx, long bytes, void** ptr);
int hx_test_device_free(hx_ctx* ctx, void* ptr);
int hx_test_memcpy(hx_ctx* ctx, void* dst, const void* src, long bytes, int to_device);

/* Plain device copy of `bytes` bytes (a multiple of 16) between two buffers
 * the call allocates, `reps` launches of one copy kernel; average launch time
 * in ms.  The reference kernel for scaling hardware byte counters. */
int hx_test_copy(hx_ctx* ctx, long bytes, int reps, double* avg_ms);

/* One application of the multilevel preconditioner (HX_PC_MULTILEVEL) for
 * moduli Ee and the DOF mask of hx_set_fixed: z = M^-1 r, in FP64
 * (precision HX_FP64) or with FP32 vectors (HX_FP32_IR); *levels receives
 * the number of coarse levels.  No reference counterpart. */
int hx_test_ml_apply(hx_ctx* ctx, const double* Ee, const double* r, double* z, int precision, int* levels);
/* The same for the V-cycle form (HX_PC_VCYCLE); *vlevels receives the number
 * of smoothed levels (every level above the coarsest).  No reference
 * counterpart. */
int hx_test_vc_apply(hx_ctx* ctx, const double* Ee, const double* r, double* z, int precision, int* levels,
                     int* vlevels);

/* The SIMP loop's OC multiplier search (8 points per pass, hx_to_step's) on
 * caller arrays (Ne each): from [0, 1e9], or with lguess > 0 from
 * [lguess/16, 16 lguess], which its first pass widens to 0 or 1e9 on the side
 * where the volume does not cross inside; x_new, the multiplier and the
 * number of passes.  No reference counterpart (top88's OC, lines 55-64). */
int hx_test_oc_search(hx_ctx* ctx, const double* x, const double* dc, const double* dv, double volfrac, double move,
                      double xmin, double rel_tol, double lguess, double* xnew, double* lmid, int* passes);

/* The manufactured displacement the learned preconditioner's training uses
 * (hx_learn_stats::manufactured) on the current grid, every DOF solved: on
 * levels l = 1..levels the standard normal values of the lattice of spacing
 * 2^l (splitmix64 of the seed and the lattice point, Box-Muller),
 * interpolated trilinearly and weighted 2^(l - levels), summed; ndof values.
 * No reference counterpart. */
int hx_test_smooth_field(hx_ctx* ctx, unsigned long long seed, int levels, double* e);

/* number of network parameters (conv weights then bias, layer by layer) */
int hx_test_unet_nparam(hx_ctx* ctx, long* nparam);
/* (re)initialise the network weights with a seed; copy the FP32 master
 * parameters out (nullable) */
int hx_test_unet_init(hx_ctx* ctx, int seed, float* params_f32);
/* forward pass on the current grid: x0 is the packed input, nNodes x 8 BF16
 * (node-major); y is the fine output, nNodes x 3 FP32 */
int hx_test_unet_forward(hx_ctx* ctx, const uint16_t* x0, float* y);
/* backward pass of the last forward for dL/dy = gy (nNodes x 3); returns the
 * parameter gradient (nparam FP32) */
int hx_test_unet_backward(hx_ctx* ctx, const float* gy, float* grad);

#ifdef __cplusplus
}
#endif
#endif
