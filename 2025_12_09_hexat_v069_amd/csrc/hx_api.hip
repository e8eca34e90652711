// hx_api.hip -- C ABI of libhexat.so: context, device workspace, solve
// drivers (FP64 Jacobi-PCG, FP32 PCG with FP64 reliable updates, Helmholtz
// filter), element post-processing, OC and the device-resident SIMP loop.
//
// The drivers keep every vector and scalar of a solve in HBM.  The host
// launches iterations in batches and reads the 128-byte PcgState back once
// per batch to decide whether to stop; kernels of an already finished solve
// return at once, so a batch may overshoot by a few empty launches only.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include <random>

#include "../../include/hexat.h"
#include "../../include/hexat_testing.h"
#include "hx_launch.h"
#include "hx_unet.h"

namespace hx {
void element_stiffness(double dx, double dy, double dz, double nu, double E, double* Ke);
void scalar_matrices(double dx, double dy, double dz, int quirk, double* Ks, double* Ms);
struct UNet;
UNet* unet_create();
void unet_destroy(UNet* u);
long unet_nparam(const UNet* u);
bool unet_ready(const UNet* u);
bool unet_bind(UNet* u, const GridArgs& g, unsigned seed, hipStream_t s);
void unet_forward_sample(UNet* u, const double* r, const double* E, double inv_scale, hipStream_t s,
                         const double* rr_dev = nullptr);
void unet_direction(UNet* u, const double* minv, double scale, double* y, hipStream_t s);
void unet_subspace_grad(UNet* u, const double* r, const double* Au, const double* Ay, const double* minv, double a,
                        double b, double scale, hipStream_t s);
void unet_backward_step(UNet* u, float lr, hipStream_t s);
int unet_weights_bf16(UNet* u, uint16_t* out, hipStream_t s);
void unet_init(UNet* u, unsigned seed, hipStream_t s);
void unet_params_f32(UNet* u, float* out, hipStream_t s);
void unet_forward_packed(UNet* u, const uint16_t* x0_host, float* y_host, hipStream_t s);
void unet_backward_host(UNet* u, const float* gy_host, float* grad_host, hipStream_t s);
}  // namespace hx

using namespace hx;

namespace {

// Generation of the device buffers: advanced by every (re)allocation and
// release, so a captured graph whose kernel arguments may name a freed buffer
// is never replayed (run_batch drops the cache when it has moved).
unsigned long g_buf_gen = 0;

// Device buffer that frees itself: owned by one object at a time (moves,
// no copies), so the timing and tuning helpers' local buffers and the
// context's members are all released.
template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) {
        o.p = nullptr;
        o.n = 0;
    }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            release();
            p = o.p;
            n = o.n;
            o.p = nullptr;
            o.n = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    hipError_t ensure(size_t m) {
        if (m <= n && p) return hipSuccess;
        release();
        ++g_buf_gen;
        hipError_t e = hipMalloc((void**)&p, std::max<size_t>(m, 1) * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            n = 0;
            return e;
        }
        n = m;
        return hipSuccess;
    }
    void release() {
        if (p) {
            (void)hipFree(p);
            ++g_buf_gen;
        }
        p = nullptr;
        n = 0;
    }
};

struct TimerPair {
    hipEvent_t a, b;
    int cls;
};

constexpr double kIrDelta = 0.01;  // reliable update when ||r|| drops by this factor (HEXAT_IR_DELTA)
constexpr int kMaxRhs = 8;         // columns of a block solve

}  // namespace

struct hx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    bool has_grid = false;
    GridArgs g{};
    double nu = 0.3;
    double Ke[576];
    double Kbin[576];
    ParityK<double> kp64{};
    ParityK<float> kp32{};
    double kd[3] = {0, 0, 0};
    bool has_fixed = false;
    int mv_variant = HX_MV_PARITY;
    int mv32_variant = HX_MV_PARITY;
    MvShape shape64, shape32;  // launch shapes of the FP64 / FP32 operator
    // launch shapes of the V-cycle's pre- / post-smoothing operators (kchunk 0:
    // unset, the operator's shape of the solve's precision)
    MvShape shape_pre{0, 0, 1}, shape_post{0, 0, 1};

    DevBuf<uint8_t> fixed;
    DevBuf<double> E, f, x, r, p, q, minv, Kbin_d, p2;
    DevBuf<float> Kbin32_d, E32, x32, r32, p32, q32, minv32, p32b;
    DevBuf<double> partials;
    PcgState* st = nullptr;       // kMaxRhs device states
    PcgState* st_host = nullptr;  // pinned mirror
    // block (multi-RHS) buffers, nrhs * ndof each
    DevBuf<double> b_f, b_x, b_r, b_p, b_q, b_p2;
    DevBuf<float> b_x32, b_r32, b_p32, b_q32, b_p32b;
    DevBuf<double> h_b, h_x, h_r, h_p, h_q, h_minv, h_in, h_out, h_p2;
    DevBuf<double> e_pr, e_vm, e_sv, e_l1, e_dir, e_ce, e_dc, e_aux, small;
    DevBuf<float> w_rho, w_l1, w_pr, w_cent, w_rho_out, w_sl1;  // hx_web_pack
    DevBuf<unsigned> w_cnt, w_offs;
    unsigned* web_scratch = nullptr;  // device address of the packing's static scratch (hx_web_pack_device)
    DevBuf<unsigned> v_pts, v_cells;  // hx_vtk_pack words
    DevBuf<float> v_l1, v_vm;
    DevBuf<double> v_u, v_f;
    DevBuf<uint8_t> v_fixed;
    DevBuf<float> s_nod, s_fine, s_tris;  // hx_stl_pack
    DevBuf<unsigned> b_part, b_cnt;       // boundary mapping
    DevBuf<double> b_F;
    DevBuf<float> b_dpart, b_dmin;
    DevBuf<unsigned long long> b_kpart, b_key;
    DevBuf<float> r_old, r_new;           // regrid
    DevBuf<unsigned> s_blk, s_offs;
    DevBuf<int8_t> s_tri;
    DevBuf<uint8_t> s_ntri;
    // SIMP session
    bool to_ready = false;
    hx_to_params tp{};
    int to_iter = 0;
    double oc_lambda = 0.0;  // the last TO iteration's OC multiplier (0: none yet)
    DevBuf<double> t_x, t_xphys, t_xnew, t_dc, t_dcf, t_dv, t_vm, t_ones;
    // learned preconditioner
    UNet* unet = nullptr;
    bool unet_trained = false;
    DevBuf<double> l_z, l_zold, l_P, l_AP, l_dots, l_ustar, l_samples, l_rho, l_man, l_part;
    int l_nsamples = 0;
    long l_sample_stride = 0;
    // multilevel preconditioner: level moduli, inverse diagonals, work vectors
    int ml_max_levels = kMlMaxLevels;
    bool ml_exact = true;  // solve the first level with at most kMlExactMax DOFs exactly (and stop there)
    DevBuf<double> ml_ainv;
    DevBuf<float> ml_ainv32;
    DevBuf<double> ml_E[kMlMaxLevels + 1], ml_r[kMlMaxLevels + 1], ml_y[kMlMaxLevels + 1],
        ml_minv[kMlMaxLevels + 1], ml_z;
    DevBuf<float> ml_r32[kMlMaxLevels + 1], ml_y32[kMlMaxLevels + 1], ml_minv32[kMlMaxLevels + 1], ml_z32;
    MlHierarchy<double> ml64;
    MlHierarchy<float> ml32;
    // V-cycle form: smoother weight, per smoothed level the smoother diagonal,
    // residual and operator moduli
    double vc_omega = 1.95;
    int vc_min_blocks = 512;  // coarse levels' launch width in workgroups (two per compute unit)
    // FP32 form: reliable update each time a column's recursive residual falls by this factor
    double ir_delta = kIrDelta;
    int fcg_keep = kFcgKeep;  // search directions the learned flexible CG keeps (the newest keep-1 orthogonalise)
    DevBuf<double> vc_s[kMlMaxLevels + 1], vc_res[kMlMaxLevels + 1], vc_E[kMlMaxLevels + 1];
    DevBuf<float> vc_s32[kMlMaxLevels + 1], vc_res32[kMlMaxLevels + 1], vc_E32[kMlMaxLevels + 1];
    DevBuf<double> vc_cstage;  // the one-workgroup pass's level data, contiguous (MlHierarchy::cstage)
    DevBuf<float> vc_cstage32;
    // captured iteration batches (hipGraph), keyed by kernel arguments
    bool graphs_on = true;
    std::vector<std::pair<std::string, hipGraphExec_t>> graphs;
    unsigned long graphs_gen = 0;  // g_buf_gen when the cached graphs were captured
    // best-iterate snapshot of the single-column solves, the best accepted
    // iterate of hx_solve_system's attempts, its densities
    DevBuf<double> x_best, x_glob, s_rho;
    // what the last single-column solve saw: best recursive residual, the
    // recursive residual it stopped at, and whether the snapshot was returned
    double last_best_rel = 0.0, last_rec_rel = 0.0;
    bool last_restored = false;
    // timing
    bool timing = false;
    std::vector<TimerPair> pending;
    std::vector<hipEvent_t> free_events;
    double acc_ms[HX_K_COUNT] = {0};
    long cnt[HX_K_COUNT] = {0};
};

namespace {

int fail(hx_ctx* c, int code, const std::string& m) {
    if (c) c->err = m;
    return code;
}

#define HXC(call)                                                                                     \
    do {                                                                                              \
        hipError_t e_ = (call);                                                                       \
        if (e_ != hipSuccess) return fail(c, HX_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define HXA(buf, n)                                                                                   \
    do {                                                                                              \
        hipError_t e_ = (buf).ensure(n);                                                              \
        if (e_ != hipSuccess) return fail(c, HX_ERR_OOM, std::string("allocating " #buf ": ") + hipGetErrorString(e_)); \
    } while (0)

#define HXK()                                                                                         \
    do {                                                                                              \
        hipError_t e_ = hipGetLastError();                                                            \
        if (e_ != hipSuccess) return fail(c, HX_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e_)); \
    } while (0)

hipEvent_t get_event(hx_ctx* c) {
    if (!c->free_events.empty()) {
        hipEvent_t e = c->free_events.back();
        c->free_events.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

// Brackets one launch with events when timing is on.
struct TimeScope {
    hx_ctx* c;
    int cls;
    hipEvent_t a = nullptr;
    TimeScope(hx_ctx* c_, int cls_) : c(c_), cls(cls_) {
        if (c->timing) {
            a = get_event(c);
            (void)hipEventRecord(a, c->stream);
        }
    }
    ~TimeScope() {
        if (a) {
            hipEvent_t b = get_event(c);
            (void)hipEventRecord(b, c->stream);
            c->pending.push_back({a, b, cls});
        }
    }
};

void flush_timing(hx_ctx* c) {
    for (auto& t : c->pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) {
            c->acc_ms[t.cls] += ms;
            c->cnt[t.cls] += 1;
        }
        c->free_events.push_back(t.a);
        c->free_events.push_back(t.b);
    }
    c->pending.clear();
}

// Destroys the cached graphs once the stream has drained, so no exec is
// destroyed while a launch of it may still be running.
void clear_graphs(hx_ctx* c) {
    if (!c->graphs.empty() && c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto& g : c->graphs) (void)hipGraphExecDestroy(g.second);
    c->graphs.clear();
    c->graphs_gen = g_buf_gen;
}

// Launch one batch of iterations.  Without timing, the batch is captured
// once into a hipGraph (kernel nodes only) and replayed; the key holds every
// kernel argument that can change between calls.
template <typename F>
int run_batch(hx_ctx* c, const std::string& key, F&& launch_batch) {
    if (c->timing || !c->graphs_on) {
        launch_batch();
        HXK();
        return HX_OK;
    }
    if (c->graphs_gen != g_buf_gen) clear_graphs(c);  // a buffer moved since the capture
    hipGraphExec_t exec = nullptr;
    for (auto& g : c->graphs)
        if (g.first == key) exec = g.second;
    if (!exec) {
        hipGraph_t graph = nullptr;
        HXC(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        launch_batch();
        const hipError_t le = hipGetLastError();
        HXC(hipStreamEndCapture(c->stream, &graph));
        if (le != hipSuccess) {
            (void)hipGraphDestroy(graph);
            return fail(c, HX_ERR_HIP, std::string("capture: ") + hipGetErrorString(le));
        }
        const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ie != hipSuccess) return fail(c, HX_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(ie));
        if (c->graphs.size() > 64) clear_graphs(c);
        c->graphs.emplace_back(key, exec);
    }
    HXC(hipGraphLaunch(exec, c->stream));
    return HX_OK;
}

std::string batch_key(hx_ctx* c, const char* kind, int nr, int batch, int variant, double shift, const void* a,
                      const void* b, const void* d, const void* e, const void* f) {
    char buf[320];
    std::snprintf(buf, sizeof buf, "%s|%d|%d|%d|%a|%p|%p|%p|%p|%p|%d|%d|%d|%d|%d|%d|%d|%d|%d|%d|%d|%d", kind, nr,
                  batch, variant, shift, a, b, d, e, f, c->shape64.tile, c->shape64.kchunk, c->shape64.pdepth,
                  c->shape32.tile, c->shape32.kchunk, c->shape32.pdepth, c->shape_pre.tile, c->shape_pre.kchunk,
                  c->shape_pre.pdepth, c->shape_post.tile, c->shape_post.kchunk, c->shape_post.pdepth);
    return buf;
}

int sync_state(hx_ctx* c, int nrhs = 1) {
    HXC(hipMemcpyAsync(c->st_host, c->st, nrhs * sizeof(PcgState), hipMemcpyDeviceToHost, c->stream));
    HXC(hipStreamSynchronize(c->stream));
    flush_timing(c);
    return HX_OK;
}

int status_to_code(int s) {
    switch (s) {
        case 0: return HX_OK;
        case 3: return HX_ERR_NOT_CONVERGED;
        case 4: return HX_ERR_BREAKDOWN;
        case 5: return HX_ERR_DIVERGED;
        case 10: return HX_OK;
        default: return HX_ERR_STATE;
    }
}

long ndof(const hx_ctx* c) { return 3 * c->g.nn; }

// y = K(E) x (+ shift x); MV_PCG mode also reduces p'q into the state.
long ndof(const hx_ctx* c);

// The operator's p'q partials live after the vector kernels' and the
// V-cycle's (per column 2 per block of kVecBlocks, then 2 per fine
// post-smoothing workgroup, up to 4096 of them), so a PCG update that reduces
// them itself (defer_alpha) never overwrites them while other blocks still
// read.
constexpr long kOpPartials = 2L * kMaxRhs * (kVecBlocks + 4096);

// returns the operator's workgroups per column (its p'q partials) in *npq
int mv64(hx_ctx* c, const double* x, double* y, const double* E, double shift, int mode, int variant,
         int nrhs = 1, const double* fr = nullptr, const double* fminv = nullptr, double* fp_out = nullptr,
         bool defer_alpha = false, int* npq = nullptr) {
    MvArgs<double> a{x, y, E, shift, mode, c->st, c->partials.p + kOpPartials, ndof(c), fr, fminv, fp_out};
    a.defer_alpha = defer_alpha ? 1 : 0;
    int v = variant == HX_MV_AUTO ? c->mv_variant : variant;
    if (fp_out && v == HX_MV_GATHER) v = HX_MV_PARITY;  // the gather kernel has no fused p-update
    TimeScope ts(c, HX_K_MATVEC);
    int nb;
    if (v == HX_MV_DENSE) nb = launch_matvec_dense<double>(c->g, c->Kbin_d.p, a, c->shape64, c->stream, nrhs);
    else if (v == HX_MV_GATHER) nb = launch_matvec_gather(c->g, c->Kbin_d.p, a, c->stream, nrhs);
    else nb = launch_matvec_parity<double>(c->g, c->kp64, a, c->shape64, c->stream, nrhs);
    if (npq) *npq = nb;
    return HX_OK;
}

int mv32(hx_ctx* c, const float* x, float* y, const float* E, double shift, int mode, int variant, int nrhs = 1,
         const float* fr = nullptr, const float* fminv = nullptr, float* fp_out = nullptr, bool defer_alpha = false,
         int* npq = nullptr) {
    MvArgs<float> a{x, y, E, (float)shift, mode, c->st, c->partials.p + kOpPartials, ndof(c), fr, fminv, fp_out};
    a.defer_alpha = defer_alpha ? 1 : 0;
    TimeScope ts(c, HX_K_MATVEC32);
    int nb;
    if (variant == HX_MV_DENSE || variant == HX_MV_MFMA) {
        MvShape sh = c->shape32;  // the dense and MFMA contractions use the 16x16 tile
        sh.tile = 0;
        if (variant == HX_MV_DENSE) nb = launch_matvec_dense<float>(c->g, c->Kbin32_d.p, a, sh, c->stream, nrhs);
        else nb = launch_matvec_mfma(c->g, c->Kbin32_d.p, a, sh, c->stream, nrhs);
    } else {
        nb = launch_matvec_parity<float>(c->g, c->kp32, a, c->shape32, c->stream, nrhs);
    }
    if (npq) *npq = nb;
    return HX_OK;
}

int ensure_grid(hx_ctx* c) {
    if (!c->has_grid) return fail(c, HX_ERR_STATE, "no grid set (call hx_grid_set first)");
    return HX_OK;
}

// Vectors of one solve: nrhs columns of ndof values each.
struct SolveVecs {
    int nrhs;
    double *f, *x, *r, *p, *q;
    float *x32, *r32, *p32, *q32;
    double* p2;   // second buffer of the double-buffered search direction
    float* p32b;
    // single column: best-iterate snapshot (GPUSolver.jl:492-499), returned
    // in x when the solve fails; nullptr: none (block solves)
    double* xbest = nullptr;
};

// A solve that did not converge returns its best iterate (GPUSolver.jl:603-606):
// the snapshot, which starts as the initial iterate (:441) and is replaced
// at check points by any better iterate with a residual below 1e-2.
bool restore_best(hx_ctx* c, const SolveVecs& v, long n) {
    const int st = c->st_host[0].status;
    if (!v.xbest || v.nrhs != 1 || st == 0 || st == 10) return false;
    (void)hipMemcpyAsync(v.x, v.xbest, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream);
    return true;
}

bool all_done(const hx_ctx* c, int nrhs) {
    for (int k = 0; k < nrhs; ++k)
        if (!c->st_host[k].done) return false;
    return true;
}

void fill_stats(const hx_ctx* c, int nrhs, int refinements, hx_pcg_stats* s) {
    if (!s) return;
    for (int k = 0; k < nrhs; ++k) {
        const PcgState& h = c->st_host[k];
        s[k].iterations = h.iter;
        s[k].status = h.status == 10 ? HX_ZERO_RHS : status_to_code(h.status);
        s[k].rel_res = h.status == 10 ? 0.0 : h.rel;
        s[k].refinements = refinements;
        s[k].fallbacks = 0;
    }
}

int worst_code(const hx_ctx* c, int nrhs) {
    int rc = HX_OK;
    for (int k = 0; k < nrhs; ++k) {
        const int cd = status_to_code(c->st_host[k].status);
        if (cd != HX_OK) rc = cd;
    }
    return rc;
}

// Levels of the multilevel preconditioner for the current moduli: coarse
// moduli by 2x2x2 averaging, coarse inverse diagonals, work vectors; for the
// V-cycle form also the smoothed levels' smoother diagonals, operator moduli
// and vectors.  The fine level's moduli (E, E32) and inverse diagonal (minv,
// minv32) are the solve's own.
// the hierarchy of the current moduli; its vectors hold nrhs columns
int ml_setup(hx_ctx* c, bool f32, bool vcycle = false, int nrhs = 1) {
    ml_prepare_kernels();
    int L = ml_levels_for(c->g, std::min(c->ml_max_levels, kMlMaxLevels));
    bool exact = false;
    if (c->ml_exact) {  // the first level small enough to solve exactly ends the hierarchy
        GridArgs gl = c->g;
        for (int l = 1; l <= L; ++l) {
            gl = ml_coarse_grid(gl);
            if (3 * gl.nn <= kMlExactMax) {
                L = l;
                exact = true;
                break;
            }
        }
    }
    const long n = ndof(c);
    MlHierarchy<double>& h = c->ml64;
    MlHierarchy<float>& h32 = c->ml32;
    h.levels = h32.levels = L;
    h.min_blocks = h32.min_blocks = c->vc_min_blocks;
    h.ainv = nullptr;
    h32.ainv = nullptr;
    h.g[0] = h32.g[0] = c->g;
    h.minv[0] = c->minv.p;
    h32.minv[0] = c->minv32.p;
    if (f32) HXA(c->ml_z32, n * nrhs);
    else HXA(c->ml_z, n * nrhs);
    GridArgs gl = c->g;
    const double* Ef = c->E.p;
    for (int l = 1; l <= L; ++l) {
        const GridArgs gc = ml_coarse_grid(gl);
        const long nc = 3 * gc.nn;
        h.g[l] = h32.g[l] = gc;  // both: the grids do not depend on the precision
        HXA(c->ml_E[l], gc.ne);
        launch_ml_coarsen_E(gl, Ef, gc, c->ml_E[l].p, nullptr, c->stream);
        const double kdl[3] = {c->kd[0] * (1 << l), c->kd[1] * (1 << l), c->kd[2] * (1 << l)};
        const uint8_t* fx = c->has_fixed ? c->fixed.p : nullptr;
        if (f32) {
            HXA(c->ml_r32[l], nc * nrhs);
            HXA(c->ml_y32[l], nc * nrhs);
            HXA(c->ml_minv32[l], nc);
            launch_ml_minv<float>(gc, kdl, c->ml_E[l].p, fx, 1 << l, c->g, c->ml_minv32[l].p, c->stream);
            h32.r[l] = c->ml_r32[l].p;
            h32.y[l] = c->ml_y32[l].p;
            h32.minv[l] = c->ml_minv32[l].p;
            h32.w[l] = 1.0;
        } else {
            HXA(c->ml_r[l], nc * nrhs);
            HXA(c->ml_y[l], nc * nrhs);
            HXA(c->ml_minv[l], nc);
            launch_ml_minv<double>(gc, kdl, c->ml_E[l].p, fx, 1 << l, c->g, c->ml_minv[l].p, c->stream);
            h.r[l] = c->ml_r[l].p;
            h.y[l] = c->ml_y[l].p;
            h.minv[l] = c->ml_minv[l].p;
            h.w[l] = 1.0;
        }
        gl = gc;
        Ef = c->ml_E[l].p;
    }
    if (exact) {
        const long N = 3 * gl.nn;
        const uint8_t* fx = c->has_fixed ? c->fixed.p : nullptr;
        if (f32) {
            HXA(c->ml_ainv32, N * N);
            launch_ml_exact_setup<float>(gl, Ef, c->Kbin_d.p, (double)(1 << L), fx, 1 << L, c->g, c->ml_ainv32.p,
                                         c->stream);
            h32.ainv = c->ml_ainv32.p;
        } else {
            HXA(c->ml_ainv, N * N);
            launch_ml_exact_setup<double>(gl, Ef, c->Kbin_d.p, (double)(1 << L), fx, 1 << L, c->g, c->ml_ainv.p,
                                          c->stream);
            h.ainv = c->ml_ainv.p;
        }
    }
    if (vcycle) {
        const int nv = f32 ? vc_levels(h32) : vc_levels(h);
        const uint8_t* fx = c->has_fixed ? c->fixed.p : nullptr;
        for (int l = 0; l < nv; ++l) {
            const GridArgs& gl = f32 ? h32.g[l] : h.g[l];
            const long nl = 3 * gl.nn;
            const double* El = l == 0 ? c->E.p : c->ml_E[l].p;
            const double hl = (double)(1 << l);
            if (f32) {
                HXA(c->vc_s32[l], nl);
                HXA(c->vc_res32[l], nl * nrhs);
                launch_vc_smoother<float>(gl, El, hl, c->Kbin_d.p, fx, 1 << l, c->g, c->vc_omega, c->vc_s32[l].p,
                                          c->stream);
                if (l > 0) {
                    HXA(c->vc_E32[l], gl.ne);
                    launch_vc_scale_E<float>(gl.ne, El, hl, c->vc_E32[l].p, c->stream);
                }
                h32.sinv[l] = c->vc_s32[l].p;
                h32.vres[l] = c->vc_res32[l].p;
                h32.Eop[l] = l == 0 ? c->E32.p : c->vc_E32[l].p;
            } else {
                HXA(c->vc_s[l], nl);
                HXA(c->vc_res[l], nl * nrhs);
                launch_vc_smoother<double>(gl, El, hl, c->Kbin_d.p, fx, 1 << l, c->g, c->vc_omega, c->vc_s[l].p,
                                           c->stream);
                if (l > 0) {
                    HXA(c->vc_E[l], gl.ne);
                    launch_vc_scale_E<double>(gl.ne, El, hl, c->vc_E[l].p, c->stream);
                }
                h.sinv[l] = c->vc_s[l].p;
                h.vres[l] = c->vc_res[l].p;
                h.Eop[l] = l == 0 ? c->E.p : c->vc_E[l].p;
            }
        }
        int rc;
        if (f32) {
            HXA(c->vc_cstage32, std::max(1L, vc_coarse_stage_values(h32)));
            rc = launch_vc_coarse_stage(h32, c->vc_cstage32.p, c->stream);
        } else {
            HXA(c->vc_cstage, std::max(1L, vc_coarse_stage_values(h)));
            rc = launch_vc_coarse_stage(h, c->vc_cstage.p, c->stream);
        }
        if (rc) return fail(c, rc, "staging the V-cycle's coarse level data");
    }
    HXK();
    return HX_OK;
}

// launch shape of a V-cycle smoothing form (post: form 4, else form 3)
MvShape vc_shape(const hx_ctx* c, bool f32, bool post) {
    const MvShape& s = post ? c->shape_post : c->shape_pre;
    return s.kchunk > 0 ? s : (f32 ? c->shape32 : c->shape64);
}

// The V-cycle PCG updates x and r without the mask (launch_pcg_update_free)
// and takes r'r from the fine post-smoothing operator, given a coarse level
bool vc_free_update(const hx_ctx* c, bool vcycle, bool f32) {
    return vcycle && (f32 ? c->ml32.levels : c->ml64.levels) >= 1;
}

// z = M^-1 r of the explicit preconditioner into ml_z / ml_z32 (fin_mode as
// launch_ml_apply's); nrhs columns (the V-cycle only)
int pc_apply64(hx_ctx* c, bool vcycle, const double* r, int fin_mode, int nrhs = 1) {
    if (vcycle) {
        if (launch_vc_apply<double>(c->ml64, c->kp64, vc_shape(c, false, false), vc_shape(c, false, true), r,
                                    c->ml_z.p, c->st, c->partials.p, kOpPartials, fin_mode, c->stream,
                                    vc_free_update(c, true, false), nrhs) != HX_OK)
            return fail(c, HX_ERR_BAD_ARG, "grid too large for the V-cycle's partial sums, or columns without a "
                                           "coarse level");
    } else {
        launch_ml_apply<double>(c->ml64, r, c->minv.p, c->ml_z.p, c->st, c->partials.p, fin_mode, c->stream);
    }
    return HX_OK;
}
// xup, pup: the V-cycle's one-workgroup pass also does xup += alpha pup
int pc_apply32(hx_ctx* c, bool vcycle, const float* r, int fin_mode, int nrhs = 1, float* xup = nullptr,
               const float* pup = nullptr) {
    if (vcycle) {
        if (launch_vc_apply<float>(c->ml32, c->kp32, vc_shape(c, true, false), vc_shape(c, true, true), r,
                                   c->ml_z32.p, c->st, c->partials.p, kOpPartials, fin_mode, c->stream,
                                   vc_free_update(c, true, true), nrhs, xup, pup) != HX_OK)
            return fail(c, HX_ERR_BAD_ARG, "grid too large for the V-cycle's partial sums, or columns without a "
                                           "coarse level");
    } else {
        launch_ml_apply<float>(c->ml32, r, c->minv32.p, c->ml_z32.p, c->st, c->partials.p, fin_mode, c->stream);
    }
    return HX_OK;
}

// ---------------------------------------------------------------- FP64 PCG
int pcg_fp64(hx_ctx* c, const hx_pcg_opts& o, hx_pcg_stats* s, const SolveVecs& v) {
    const long n = ndof(c);
    const int nr = v.nrhs;
    int batch = o.batch > 0 ? o.batch : (n * nr < 3000000 ? 64 : 16);
    batch = batch < 2 ? 2 : (batch & ~1);  // even: the double-buffered p returns to buffer 0
    launch_elastic_minv<double>(c->g, c->kd, c->E.p, c->has_fixed ? c->fixed.p : nullptr, o.shift, c->minv.p,
                                c->stream);
    // multilevel: explicit z = M^-1 r after each update; the operator forms
    // p = z + beta p_old (parity contraction, which has that fused form)
    const bool vc = o.precond == HX_PC_VCYCLE;
    const bool ml = o.precond == HX_PC_MULTILEVEL || vc;
    int rc = ml ? ml_setup(c, false, vc, nr) : HX_OK;
    if (rc) return rc;
    const bool free_upd = vc_free_update(c, vc, false);
    launch_pcg_reset(c->st, o.tol, o.shift, o.max_iter, 1, c->stream, nr, v.xbest ? 1 : 0);
    if (o.warm_start) launch_mask<double>(n, c->minv.p, v.x, c->stream, nr);
    else launch_fill(n * nr, 0.0, v.x, c->stream);
    if (v.xbest) HXC(hipMemcpyAsync(v.xbest, v.x, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    mv64(c, v.x, v.q, c->E.p, 0.0, MV_PLAIN, o.matvec, nr);
    launch_pcg_init<double>(n, v.f, v.x, v.q, c->minv.p, v.r, v.p, o.shift, c->st, c->partials.p, c->stream, nr);
    if (ml && (rc = pc_apply64(c, vc, v.r, 1, nr))) return rc;
    HXK();
    if ((rc = sync_state(c, nr))) return rc;
    const int mvv = ml ? HX_MV_PARITY : o.matvec;
    const std::string key = batch_key(c, vc ? (free_upd ? "f64vcf" : "f64vc") : ml ? "f64ml" : "f64", nr, batch,
                                      mvv == HX_MV_AUTO ? c->mv_variant : mvv, o.shift, v.x, v.p, v.q, v.r, v.p2);
    // p is double-buffered (v.p, v.p2): each matvec forms p_new = D^-1 r + beta p_old
    // from the plane it stages anyway and writes it beside q; batches are even.
    while (!all_done(c, nr)) {
        rc = run_batch(c, key, [&] {
            for (int b = 0; b < batch; ++b) {
                double* pin = (b & 1) ? v.p2 : v.p;
                double* pout = (b & 1) ? v.p : v.p2;
                int npq = 0;  // alpha: reduced by the update from the operator's partials
                if (ml)
                    mv64(c, pin, v.q, c->E.p, o.shift, MV_PCG, HX_MV_PARITY, nr, c->ml_z.p, nullptr, pout, true, &npq);
                else mv64(c, pin, v.q, c->E.p, o.shift, MV_PCG, o.matvec, nr, v.r, c->minv.p, pout, true, &npq);
                {
                    TimeScope ts(c, HX_K_UPDATE);
                    if (free_upd)
                        launch_pcg_update_free<double>(n, v.x, pout, v.r, v.q, c->st, c->partials.p + kOpPartials,
                                                       npq, c->stream, v.xbest, nr);
                    else
                        launch_pcg_update<double>(n, v.x, pout, v.r, v.q, c->minv.p, c->st, c->partials.p,
                                                  c->stream, nr, !ml, c->partials.p + kOpPartials, npq, v.xbest);
                }
                if (ml) {
                    TimeScope ts(c, HX_K_PRECOND);
                    (void)pc_apply64(c, vc, v.r, 0, nr);  // its one failure is caught at the start call
                }
            }
        });
        if (rc) return rc;
        if ((rc = sync_state(c, nr))) return rc;
    }
    if (v.xbest) {  // a best iterate the last finaliser flagged
        launch_snapshot_pending(n, v.x, v.xbest, c->st, c->stream);
        HXK();
    }
    std::vector<PcgState> fin(c->st_host, c->st_host + nr);
    c->last_best_rel = fin[0].best_rel;
    c->last_rec_rel = fin[0].rel;
    c->last_restored = restore_best(c, v, n);
    // true residual of the returned iterate (FP64)
    mv64(c, v.x, v.q, c->E.p, 0.0, MV_PLAIN, o.matvec, nr);
    launch_true_residual(n, v.f, v.q, v.x, c->minv.p, o.shift, nullptr, nullptr, c->st, c->partials.p, 1, c->stream,
                         nr);
    HXK();
    if ((rc = sync_state(c, nr))) return rc;
    for (int k = 0; k < nr; ++k) {  // keep the solve's status/iterations, report the true residual
        const double rel = c->st_host[k].rel;
        c->st_host[k] = fin[k];
        c->st_host[k].rel = rel;
    }
    fill_stats(c, nr, 0, s);
    return worst_code(c, nr);
}

// ------------------------------------------- FP32 PCG + FP64 reliable updates
int ensure_fp32(hx_ctx* c) {
    const long n = ndof(c);
    HXA(c->E32, c->g.ne);
    HXA(c->x32, n);
    HXA(c->r32, n);
    HXA(c->p32, n);
    HXA(c->q32, n);
    HXA(c->minv32, n);
    HXA(c->p32b, n);
    return HX_OK;
}

// FP32 recurrences (element contraction, vectors) with FP64 dot products and
// alpha/beta; every time a column's recursive residual falls by kIrDelta, its
// FP32 correction is folded into the FP64 solution and the residual is
// replaced by the true FP64 residual b - K x (reliable update).  Converged
// when the true residual is below tol.
int pcg_fp32ir(hx_ctx* c, const hx_pcg_opts& o, hx_pcg_stats* s, const SolveVecs& v) {
    const long n = ndof(c);
    const int nr = v.nrhs;
    int rc = ensure_fp32(c);
    if (rc) return rc;
    int batch = o.batch > 0 ? o.batch : (n * nr < 3000000 ? 64 : 16);
    batch = batch < 2 ? 2 : (batch & ~1);
    const int v32 = (o.matvec == HX_MV_DENSE || o.matvec == HX_MV_MFMA) ? o.matvec
                    : (o.matvec == HX_MV_AUTO ? c->mv32_variant : HX_MV_PARITY);
    launch_convert<double, float>(c->g.ne, c->E.p, c->E32.p, c->stream);
    launch_elastic_minv<double>(c->g, c->kd, c->E.p, c->has_fixed ? c->fixed.p : nullptr, o.shift, c->minv.p,
                                c->stream);
    launch_convert<double, float>(n, c->minv.p, c->minv32.p, c->stream);
    const bool vc = o.precond == HX_PC_VCYCLE;
    const bool ml = o.precond == HX_PC_MULTILEVEL || vc;  // explicit z = M^-1 r (see pcg_fp64)
    if (ml && (rc = ml_setup(c, true, vc, nr))) return rc;
    const bool free_upd = vc_free_update(c, vc, true);
    launch_pcg_reset(c->st, o.tol, o.shift, o.max_iter, 1, c->stream, nr, v.xbest ? 2 : 0);
    if (o.warm_start) launch_mask<double>(n, c->minv.p, v.x, c->stream, nr);
    else launch_fill(n * nr, 0.0, v.x, c->stream);
    if (v.xbest) HXC(hipMemcpyAsync(v.xbest, v.x, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    launch_fill32(n * nr, 0.0f, v.x32, c->stream);
    launch_fill32(n * nr, 0.0f, v.p32, c->stream);
    mv64(c, v.x, v.q, c->E.p, 0.0, MV_PLAIN, HX_MV_PARITY, nr);
    launch_true_residual(n, v.f, v.q, v.x, c->minv.p, o.shift, nullptr, v.r32, c->st, c->partials.p, 3, c->stream,
                         nr);
    if (ml && (rc = pc_apply32(c, vc, v.r32, 2, nr))) return rc;
    launch_ir_resume(c->st, o.tol, c->ir_delta, 1, c->stream, nr);
    HXK();
    if ((rc = sync_state(c, nr))) return rc;
    int refinements = 0;
    // with a one-workgroup pass in the V-cycle, its spare workgroups take x += alpha p
    // and the update kernel does r -= alpha q alone
    const bool x_in_pass = free_upd && vc_coarse_pass(c->ml32);
    const std::string key = batch_key(c, vc ? (x_in_pass ? "f32vcx" : free_upd ? "f32vcf" : "f32vc") : ml ? "f32ml"
                                                                                                       : "f32",
                                      nr, batch, ml ? HX_MV_PARITY : v32, o.shift, v.x32, v.p32, v.q32, v.r32,
                                      v.p32b);
    while (!all_done(c, nr)) {
        while (!all_done(c, nr)) {
            rc = run_batch(c, key, [&] {
                for (int b = 0; b < batch; ++b) {
                    float* pin = (b & 1) ? v.p32b : v.p32;
                    float* pout = (b & 1) ? v.p32 : v.p32b;
                    int npq = 0;  // alpha: reduced by the update from the operator's partials
                    if (ml)
                        mv32(c, pin, v.q32, c->E32.p, o.shift, MV_PCG, HX_MV_PARITY, nr, c->ml_z32.p, nullptr, pout,
                             true, &npq);
                    else
                        mv32(c, pin, v.q32, c->E32.p, o.shift, MV_PCG, v32, nr, v.r32, c->minv32.p, pout, true, &npq);
                    {
                        TimeScope ts(c, HX_K_UPDATE);
                        if (x_in_pass)
                            launch_pcg_update_r<float>(n, v.r32, v.q32, c->st, c->partials.p + kOpPartials, npq,
                                                       c->stream, nr);
                        else if (free_upd)
                            launch_pcg_update_free<float>(n, v.x32, pout, v.r32, v.q32, c->st,
                                                          c->partials.p + kOpPartials, npq, c->stream, nullptr,
                                                          nr);
                        else
                            launch_pcg_update<float>(n, v.x32, pout, v.r32, v.q32, c->minv32.p, c->st,
                                                     c->partials.p, c->stream, nr, !ml, c->partials.p + kOpPartials,
                                                     npq);
                    }
                    if (ml) {
                        TimeScope ts(c, HX_K_PRECOND);
                        (void)pc_apply32(c, vc, v.r32, 0, nr, x_in_pass ? v.x32 : nullptr,
                                         x_in_pass ? pout : nullptr);  // its one failure is caught at the start call
                    }
                }
            });
            if (rc) return rc;
            if ((rc = sync_state(c, nr))) return rc;
        }
        // reliable update of every column: fold the FP32 correction, replace r
        launch_fix_p<float>(n, v.p32, v.p32b, c->st, c->stream, nr);
        launch_accumulate(n, v.x, v.x32, c->stream, nr);
        mv64(c, v.x, v.q, c->E.p, 0.0, MV_PLAIN, HX_MV_PARITY, nr);
        launch_true_residual(n, v.f, v.q, v.x, c->minv.p, o.shift, nullptr, v.r32, c->st, c->partials.p, 2,
                             c->stream, nr);
        // check point: x is the FP64 iterate and rel its true residual, unless
        // the phase failed (its folded correction is then not an iterate)
        if (v.xbest && c->st_host[0].status != 4 && c->st_host[0].status != 5)
            launch_snapshot(n, v.x, v.xbest, c->st, c->stream);
        // the phase's last preconditioner application left z = M^-1 r of the
        // recursive residual, which the true one matches to FP32 rounding:
        // the resumed phase keeps it and takes r'z with the true r
        if (ml) launch_rz_park<float>(n, v.r32, c->ml_z32.p, c->st, c->partials.p, c->stream, nr);
        launch_ir_resume(c->st, o.tol, c->ir_delta, 0, c->stream, nr);
        HXK();
        ++refinements;
        if ((rc = sync_state(c, nr))) return rc;
    }
    c->last_best_rel = c->st_host[0].best_rel;
    c->last_rec_rel = c->st_host[0].rel;
    c->last_restored = false;
    if (restore_best(c, v, n)) {  // report the true residual of the snapshot
        c->last_restored = true;
        const int status = c->st_host[0].status, iters = c->st_host[0].iter;
        mv64(c, v.x, v.q, c->E.p, 0.0, MV_PLAIN, HX_MV_PARITY, 1);
        launch_true_residual(n, v.f, v.q, v.x, c->minv.p, o.shift, nullptr, nullptr, c->st, c->partials.p, 0,
                             c->stream, 1);
        HXK();
        if ((rc = sync_state(c, 1))) return rc;
        c->st_host[0].status = status;
        c->st_host[0].iter = iters;
    }
    fill_stats(c, nr, refinements, s);
    return worst_code(c, nr);
}

// ------------------------------------------ flexible PCG, learned preconditioner
// z = D^-1 r + |r|_rms net(r, E); Polak-Ribiere beta.  If it does not reach
// tol within max_iter (or breaks down), Jacobi-PCG takes over from its iterate.
// Flexible CG with the learned preconditioner (GPUSolver.jl's Jacobi-PCG loop
// with the preconditioner swapped).  Each iteration forms u = D^-1 r and the
// network's direction y = net(r/|r|_rms, E), applies the operator to both in
// one two-column launch and takes z = a u + b y with (a, b) minimising the
// energy error over span{u, y} -- which needs no knowledge of the error since
// A e = r:  [u'Au u'Ay; u'Ay y'Ay] (a, b) = (u'r, y'r).  The search direction
// is z A-orthogonalised against the last keep-1 ones (truncated flexible CG;
// keep = 8, HEXAT_FCG_KEEP: 4 kept took 1.6x the iterations, 12 and 16 no
// fewer), whose images under A are kept, so A p costs no operator
// application.  The vector work is three fused passes (the dot products; the
// direction with p'Ap and p'r; the step with r'r), alpha stays on the device,
// and the host synchronises once per iteration.  Jacobi-PCG takes over from
// the current iterate when the flexible iteration stalls or breaks down, and
// the true residual judges convergence.
// stall: no 1% drop of the best residual for max(kLearnedStall, it/2)
// iterations (CG residuals plateau for ~100 iterations on healthy runs).  The
// first kLearnedRestarts stalls restart the flexible iteration from the true
// residual with an empty direction ring (truncated flexible CG loses the
// A-orthogonality it relies on over thousands of iterations); the next one
// hands over to Jacobi-PCG.  HEXAT_LEARNED_STALL, read at each solve, fixes the
// window instead (iterations, >= 1), so a short window forces the restarts.
constexpr int kLearnedStall = 200;
constexpr int kLearnedRestarts = 2;

extern "C++" {
// dot products of n-vectors, one sync for the lot
int dots_host(hx_ctx* c, long n, std::initializer_list<std::pair<const double*, const double*>> pr, double* out) {
    int k = 0;
    for (const auto& q : pr) launch_dot(n, q.first, q.second, c->l_dots.p + k++, nullptr, c->partials.p, c->stream);
    HXC(hipMemcpyAsync(out, c->l_dots.p, k * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HXC(hipStreamSynchronize(c->stream));
    return HX_OK;
}
}  // extern "C++"

int pcg_learned(hx_ctx* c, const hx_pcg_opts& o, hx_pcg_stats* s) {
    if (!c->unet || !c->unet_trained) return fail(c, HX_ERR_STATE, "learned preconditioner not trained");
    if (!unet_bind(c->unet, c->g, 1234u, c->stream)) return fail(c, HX_ERR_OOM, "U-Net activations");
    const long n = ndof(c);
    HXA(c->l_z, 2 * n);     // W = [u | y]
    HXA(c->l_zold, 2 * n);  // A W
    HXA(c->l_dots, kFcgDots + 8);
    hipEvent_t ev0 = get_event(c), ev1 = get_event(c);
    (void)hipEventRecord(ev0, c->stream);
    launch_elastic_minv<double>(c->g, c->kd, c->E.p, c->has_fixed ? c->fixed.p : nullptr, o.shift, c->minv.p,
                                c->stream);
    launch_pcg_reset(c->st, o.tol, o.shift, o.max_iter, 1, c->stream, 1);
    if (o.warm_start) launch_mask<double>(n, c->minv.p, c->x.p, c->stream, 1);
    else launch_fill(n, 0.0, c->x.p, c->stream);
    mv64(c, c->x.p, c->q.p, c->E.p, 0.0, MV_PLAIN, HX_MV_AUTO, 1);
    launch_pcg_init<double>(n, c->f.p, c->x.p, c->q.p, c->minv.p, c->r.p, c->p.p, o.shift, c->st, c->partials.p,
                            c->stream, 1);
    HXK();
    int rc = sync_state(c, 1);
    if (rc) return rc;
    const double normb = c->st_host->normb;
    double rr = c->st_host->rr, rel = c->st_host->rel, best = rel;
    int it = 0, best_it = 0, status = c->st_host->status == 10 ? 10 : -1, restarts = 0;
    bool stalled = false;
    double* W = c->l_z.p;
    double* AW = c->l_zold.p;
    HXA(c->l_P, (long)kFcgKeep * n);   // stored search directions
    HXA(c->l_AP, (long)kFcgKeep * n);  // and their images under A
    HXA(c->l_part, (long)kVecBlocks * kFcgDots);
    double* dots = c->l_dots.p;              // [0, kFcgDots): the iteration's dot products
    double* scal = c->l_dots.p + kFcgDots;   // p'Ap, p'r, alpha, breakdown flag, r'r of the last step
    double pAp_j[kFcgKeep];
    // ring of `keep` slots (<= kFcgKeep); the newest keep-1 directions orthogonalise
    const int keep = c->fcg_keep;
    int nstored = 0, next = 0, last = -1, restart_it = 0, window = 0;
    if (const char* ev = std::getenv("HEXAT_LEARNED_STALL")) window = std::max(0, std::atoi(ev));
    // Per iteration one host synchronisation: the dot products of this
    // iteration arrive with the last step's p'Ap, breakdown flag and r'r, so
    // its convergence is judged one iteration late (the last network
    // evaluation of a converged solve is discarded)
    HXC(hipMemcpyAsync(scal + 4, &rr, sizeof(double), hipMemcpyHostToDevice, c->stream));  // r'r of the start
    while (status < 0) {
        launch_mul(n, c->minv.p, c->r.p, W, c->stream);                       // u
        // the network sees r / rms(r) with r'r of the last step (on the device);
        // y keeps the network's scale (the energy-optimal pair (a, b) absorbs it)
        unet_forward_sample(c->unet, c->r.p, c->E.p, 0.0, c->stream, scal + 4);
        unet_direction(c->unet, c->minv.p, 1.0, W + n, c->stream);            // y
        mv64(c, W, AW, c->E.p, o.shift, MV_PLAIN, HX_MV_AUTO, 2);             // [Au | Ay]
        FcgCoef cf{};
        cf.m = std::min(nstored, keep - 1);
        for (int j = 0; j < cf.m; ++j) cf.slot[j] = (next - 1 - j + 2 * keep) % keep;
        launch_fcg_dots(n, W, AW, c->r.p, cf, c->l_P.p, c->l_part.p, dots, c->stream);
        double h[kFcgDots + 5];
        HXC(hipMemcpyAsync(h, dots, (kFcgDots + 5) * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HXC(hipStreamSynchronize(c->stream));
        if (last >= 0) {  // the last step: its direction's p'Ap, breakdown, residual
            if (h[kFcgDots + 3] != 0.0) {
                status = 4;
                break;
            }
            pAp_j[last] = h[kFcgDots + 0];
            rr = h[kFcgDots + 4];
            rel = std::sqrt(rr) / normb;
            ++it;
            if (!(rel == rel) || rel > 1e3) {
                status = 5;
                break;
            }
            if (rel < 0.99 * best) best = rel, best_it = it;
        }
        if (rel < o.tol) {
            status = 0;
            break;
        }
        if (it >= o.max_iter) {
            status = 3;
            break;
        }
        if (it - best_it >= (window > 0 ? window : std::max(kLearnedStall, (it - restart_it) / 2))) {
            if (restarts == kLearnedRestarts) {
                stalled = true;
                break;
            }
            // restart: r = b - A x in FP64, an empty ring, the stall window and best from here
            ++restarts;
            mv64(c, c->x.p, c->q.p, c->E.p, 0.0, MV_PLAIN, HX_MV_AUTO, 1);
            launch_pcg_init<double>(n, c->f.p, c->x.p, c->q.p, c->minv.p, c->r.p, c->p.p, o.shift, c->st,
                                    c->partials.p, c->stream, 1);
            HXK();
            if ((rc = sync_state(c, 1))) return rc;
            rr = c->st_host->rr;
            rel = std::sqrt(rr) / normb;
            best = rel;
            best_it = restart_it = it;
            nstored = 0;
            next = 0;
            last = -1;
            HXC(hipMemcpyAsync(scal + 4, &rr, sizeof(double), hipMemcpyHostToDevice, c->stream));
            continue;
        }
        const double uAu = h[0], uAy = h[1], yAy = h[2], ur = h[3], yr = h[4];
        const double det = uAu * yAy - uAy * uAy;
        double a = ur / uAu, b = 0.0;  // Jacobi alone when y adds nothing usable
        if (det > 1e-12 * uAu * yAy) a = (ur * yAy - yr * uAy) / det, b = (yr * uAu - ur * uAy) / det;
        for (int j = 0; j < cf.m; ++j)  // c_j = (A z)' P_j / P_j' A P_j
            cf.c[j] = (a * h[5 + 2 * j] + b * h[6 + 2 * j]) / pAp_j[cf.slot[j]];
        double* pn = c->l_P.p + (long)next * n;
        double* apn = c->l_AP.p + (long)next * n;
        launch_fcg_direction2(n, a, b, W, AW, cf, c->l_P.p, c->l_AP.p, c->r.p, pn, apn, c->l_part.p, scal,
                              c->stream);
        launch_fcg_step2(n, scal, c->x.p, pn, c->r.p, apn, c->minv.p, c->l_part.p, scal, c->stream);
        HXK();
        last = next;
        next = (next + 1) % keep;
        nstored = std::min(nstored + 1, keep);
    }
    (void)hipEventRecord(ev1, c->stream);
    (void)hipEventSynchronize(ev1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ev0, ev1);
    c->free_events.push_back(ev0);
    c->free_events.push_back(ev1);
    if (status == 10) {  // zero right-hand side
        fill_stats(c, 1, 0, s);
        if (s) s->fallbacks = 0;
        return worst_code(c, 1);
    }
    if (status == 0) {  // the true residual judges convergence
        mv64(c, c->x.p, c->q.p, c->E.p, 0.0, MV_PLAIN, HX_MV_AUTO, 1);
        launch_true_residual(n, c->f.p, c->q.p, c->x.p, c->minv.p, o.shift, nullptr, nullptr, c->st, c->partials.p, 1,
                             c->stream, 1);
        HXK();
        if ((rc = sync_state(c, 1))) return rc;
        if (c->st_host->rel <= 1.5 * o.tol) {
            if (s) {
                s->iterations = it;
                s->status = HX_OK;
                s->rel_res = c->st_host->rel;
                s->seconds = ms * 1e-3;
                s->refinements = restarts;
                s->fallbacks = 0;
            }
            return HX_OK;
        }
    }
    // Jacobi-PCG from the current iterate
    (void)stalled;
    hx_pcg_opts jo = o;
    jo.precond = HX_PC_JACOBI;
    jo.warm_start = 1;
    jo.max_iter = std::max(o.max_iter - it, o.max_iter / 2);
    rc = pcg_fp64(c, jo, s, SolveVecs{1, c->f.p, c->x.p, c->r.p, c->p.p, c->q.p, nullptr, nullptr, nullptr, nullptr,
                                      c->p2.p, nullptr});
    if (rc == HX_ERR_HIP || rc == HX_ERR_OOM) return rc;
    if (s) {
        s->iterations += it;
        s->seconds += ms * 1e-3;
        s->refinements = restarts;
        s->fallbacks = 1;
    }
    return rc;
}

SolveVecs single_vecs(hx_ctx* c) {
    return SolveVecs{1, c->f.p, c->x.p, c->r.p, c->p.p, c->q.p, c->x32.p, c->r32.p, c->p32.p, c->q32.p,
                     c->p2.p, c->p32b.p, c->x_best.p};
}

// ------------------------------------------------------------- Helmholtz
int helm_solve(hx_ctx* c, double R, const hx_filter_opts& o, const double* in_dev, double* out_dev,
               hx_filter_stats* s) {
    const long nn = c->g.nn;
    HXA(c->h_b, nn);
    HXA(c->h_x, nn);
    HXA(c->h_r, nn);
    HXA(c->h_p, nn);
    HXA(c->h_q, nn);
    HXA(c->h_minv, nn);
    double Ks[64], Ms[64];
    scalar_matrices(c->g.dx, c->g.dy, c->g.dz, o.quirk, Ks, Ms);
    HelmArgs A;
    for (int t = 0; t < 64; ++t) A.Ae[t] = R * R * Ks[t] + Ms[t];  // GPUHelmholtz.jl:128-129
    A.h[0] = c->g.dx;  // the box scalar_matrices integrated over (the quirk's is 1.5 dz tall)
    A.h[1] = c->g.dy;
    A.h[2] = o.quirk ? 1.5 * c->g.dz : c->g.dz;
    A.R2 = R * R;
    const double vol8 = c->g.dx * c->g.dy * c->g.dz / 8.0;
    launch_helm_rhs(c->g, in_dev, vol8, c->h_b.p, c->stream);
    launch_helm_minv(c->g, A, c->h_minv.p, c->stream);
    launch_pcg_reset(c->st, o.tol, 0.0, o.max_iter > 0 ? o.max_iter : 200, o.check_every, c->stream);
    launch_fill(nn, 0.0, c->h_x.p, c->stream);
    launch_fill(nn, 0.0, c->h_q.p, c->stream);
    launch_pcg_init<double>(nn, c->h_b.p, c->h_x.p, c->h_q.p, c->h_minv.p, c->h_r.p, c->h_p.p, 0.0, c->st,
                            c->partials.p, c->stream);
    HXK();
    int rc = sync_state(c);
    if (rc) return rc;
    if (c->st_host->status == 10) {  // GPUHelmholtz.jl:179-182: zero field returned as is
        HXC(hipMemcpyAsync(out_dev, in_dev, c->g.ne * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        if (s) {
            s->iterations = 0;
            s->status = HX_OK;
            s->rel_res = 0.0;
        }
        return HX_OK;
    }
    // batches of 16 iterations, replayed as one graph per operator and
    // buffer set: the operator leaves its p'q partials apart from the
    // update's, which forms alpha from them (no finaliser between the two)
    const int batch = 16;
    unsigned long long ah = 1469598103934665603ull;  // FNV-1a of the element matrix
    for (int t = 0; t < 64; ++t) {
        unsigned long long bits;
        std::memcpy(&bits, &A.Ae[t], sizeof bits);
        for (int q = 0; q < 8; ++q) ah = (ah ^ ((bits >> (8 * q)) & 0xff)) * 1099511628211ull;
    }
    // the sum-factorised operator forms p = minv r + beta p_old itself from a
    // second p buffer (batches even: p returns to h_p); otherwise the gather
    // form and a separate p-update
    const bool fused = helm_sf_available(A);
    if (fused) HXA(c->h_p2, nn);
    char key[224];
    std::snprintf(key, sizeof key, "helm %d %d %d %d %llx %p %p %p %p %p %p", c->g.nx, c->g.ny, c->g.nz, fused ? 1 : 0,
                  ah, (void*)c->h_x.p, (void*)c->h_r.p, (void*)c->h_p.p, (void*)c->h_q.p, (void*)c->h_minv.p,
                  fused ? (void*)c->h_p2.p : nullptr);
    double* const op_partials = c->partials.p + kOpPartials;
    while (!c->st_host->done) {
        rc = run_batch(c, key, [&] {
            for (int b = 0; b < batch; ++b) {
                int npq;
                double* pin = (fused && (b & 1)) ? c->h_p2.p : c->h_p.p;
                double* pout = (fused && !(b & 1)) ? c->h_p2.p : c->h_p.p;
                {
                    TimeScope ts(c, HX_K_HELM);
                    npq = fused ? launch_helm_matvec_fused(c->g, A, pin, c->h_r.p, c->h_minv.p, pout, c->h_q.p, c->st,
                                                           op_partials, c->stream)
                                : launch_helm_matvec(c->g, A, c->h_p.p, c->h_q.p, MV_PCG, c->st, op_partials,
                                                     c->stream);
                }
                launch_pcg_update<double>(nn, c->h_x.p, pout, c->h_r.p, c->h_q.p, c->h_minv.p, c->st, c->partials.p,
                                          c->stream, 1, true, op_partials, npq);
                if (!fused) launch_pcg_pupdate<double>(nn, c->h_r.p, c->h_minv.p, c->h_p.p, c->st, c->stream);
            }
        });
        if (rc) return rc;
        if ((rc = sync_state(c))) return rc;
    }
    launch_helm_extract(c->g, c->h_x.p, out_dev, c->stream);
    HXK();
    if ((rc = sync_state(c))) return rc;
    const int status = c->st_host->status;
    if (s) {
        s->iterations = c->st_host->iter;
        s->status = status == 3 ? HX_OK : status_to_code(status);  // the reference stops at max_iter silently
        s->rel_res = c->st_host->rel;
    }
    return (status == 4 || status == 5) ? status_to_code(status) : HX_OK;
}

int elem_post(hx_ctx* c, const ElemArgs& a) {
    ParityK<double> k = c->kp64;
    TimeScope ts(c, HX_K_ELEM);
    launch_element_post(c->g, k, a, c->stream);
    return HX_OK;
}

void material_consts(double nu, double* l1, double* l2, double* mu) {
    const double f = 1.0 / ((1.0 + nu) * (1.0 - 2.0 * nu));  // Element.jl:45-62 with E = 1
    *l1 = (1.0 - nu) * f;
    *l2 = nu * f;
    *mu = 0.5 * (1.0 - 2.0 * nu) * f;
}

// OC update with the multiplier found by bisection (top88, hx_oc_update) or,
// with kary, by the 8-point search (9x bracket reduction per pass, the same
// stopping rule; the device-resident loop)
// lguess > 0 (8-point search only): start from a bracket around this multiplier
int oc_run(hx_ctx* c, long ne, const double* x, const double* dc, const double* dv, double volfrac, double move,
           double xmin, double rel_tol, double* xnew, bool kary = false, double lguess = 0.0) {
    launch_oc_init(c->st, c->stream, kary ? lguess : 0.0);
    int rc;
    // chunks of passes replayed as one graph per field set and parameters
    // (a finished search returns at once from the remaining passes)
    char key[256];
    std::snprintf(key, sizeof key, "oc %d %ld %p %p %p %p %.17g %.17g %.17g %.17g", kary ? 1 : 0, ne, (const void*)x,
                  (const void*)dc, (const void*)dv, (void*)c->partials.p, volfrac, move, xmin, rel_tol);
    for (int chunk = 0; chunk < 13; ++chunk) {
        rc = run_batch(c, key, [&] {
            for (int k = 0; k < (kary ? 8 : 16); ++k)
                if (kary)
                    launch_oc_eval_k(ne, x, dc, dv, volfrac, move, xmin, rel_tol, c->st, c->partials.p, c->stream);
                else launch_oc_eval(ne, x, dc, dv, volfrac, move, xmin, rel_tol, c->st, c->partials.p, c->stream);
        });
        if (rc) return rc;
        if ((rc = sync_state(c))) return rc;
        if (c->st_host->done) break;
    }
    launch_oc_apply(ne, x, dc, dv, move, xmin, xnew, c->st, c->partials.p, c->stream);
    HXK();
    return sync_state(c);
}

int upload(hx_ctx* c, void* dst, const void* src, size_t bytes) {
    HXC(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    return HX_OK;
}
int download(hx_ctx* c, void* dst, const void* src, size_t bytes) {
    HXC(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    HXC(hipStreamSynchronize(c->stream));
    return HX_OK;
}

}  // namespace

// =================================================================== C ABI
extern "C" {

int hx_version(void) { return 11; }

int hx_device_count(int* count) {
    if (!count) return HX_ERR_BAD_ARG;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
    return HX_OK;
}

int hx_ctx_create(int device, hx_ctx** out) {
    if (!out) return HX_ERR_BAD_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return HX_ERR_NO_DEVICE;
    if (device < 0 || device >= n) return HX_ERR_BAD_ARG;
    if (hipSetDevice(device) != hipSuccess) return HX_ERR_HIP;
    hx_ctx* c = new hx_ctx();
    c->device = device;
    if (const char* ev = std::getenv("HEXAT_NO_GRAPH")) c->graphs_on = ev[0] == '0';
    // cap on the multilevel preconditioner's coarse levels (default: every 2:1 coarsening)
    if (const char* ev = std::getenv("HEXAT_ML_LEVELS")) c->ml_max_levels = std::max(0, std::atoi(ev));
    if (const char* ev = std::getenv("HEXAT_FCG_KEEP")) {  // learned flexible CG: directions kept, 2..8
        const int k = std::atoi(ev);
        if (k >= 2 && k <= kFcgKeep) c->fcg_keep = k;
    }
    if (const char* ev = std::getenv("HEXAT_ML_EXACT")) c->ml_exact = ev[0] != '0';
    if (const char* ev = std::getenv("HEXAT_IR_DELTA")) {  // reliable-update factor of the FP32 form, 0 < d < 1
        const double d = std::atof(ev);
        if (d > 0.0 && d < 1.0) c->ir_delta = d;
    }
    if (const char* ev = std::getenv("HEXAT_VC_OMEGA")) {  // V-cycle smoother weight, 0 < omega < 2
        const double w = std::atof(ev);
        if (w > 0.0 && w < 2.0) c->vc_omega = w;
    }
    if (const char* ev = std::getenv("HEXAT_VC_MINBLOCKS")) {  // coarse V-cycle levels' launch width
        const int b = std::atoi(ev);
        if (b >= 1) c->vc_min_blocks = b;
    }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc((void**)&c->st, kMaxRhs * sizeof(PcgState)) != hipSuccess ||
        hipHostMalloc((void**)&c->st_host, kMaxRhs * sizeof(PcgState), hipHostMallocDefault) != hipSuccess) {
        hx_ctx_destroy(c);
        return HX_ERR_HIP;
    }
    std::memset(c->st_host, 0, kMaxRhs * sizeof(PcgState));
    if (web_scratch_address(device, &c->web_scratch) != 0) {  // device current (hipSetDevice above)
        hx_ctx_destroy(c);
        return HX_ERR_HIP;
    }
    *out = c;
    return HX_OK;
}

void hx_ctx_destroy(hx_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    flush_timing(c);
    clear_graphs(c);
    if (c->unet) unet_destroy(c->unet);
    for (auto e : c->free_events) (void)hipEventDestroy(e);
    DevBuf<double>* d[] = {&c->E, &c->f, &c->x, &c->r, &c->p, &c->q, &c->minv, &c->Kbin_d, &c->partials,
                           &c->h_b, &c->h_x, &c->h_r, &c->h_p, &c->h_q, &c->h_minv, &c->h_in, &c->h_out, &c->h_p2,
                           &c->e_pr, &c->e_vm, &c->e_sv, &c->e_l1, &c->e_dir, &c->e_ce, &c->e_dc, &c->e_aux,
                           &c->small, &c->t_x, &c->t_xphys, &c->t_xnew, &c->t_dc, &c->t_dcf, &c->t_dv, &c->t_vm,
                           &c->t_ones, &c->b_f, &c->b_x, &c->b_r, &c->b_p, &c->b_q, &c->b_p2, &c->p2, &c->l_z, &c->l_zold,
                           &c->l_ustar, &c->l_samples, &c->l_rho, &c->x_best, &c->x_glob, &c->s_rho};
    for (auto b : d) b->release();
    DevBuf<float>* fl[] = {&c->Kbin32_d, &c->E32, &c->x32, &c->r32, &c->p32, &c->q32, &c->minv32,
                           &c->b_x32, &c->b_r32, &c->b_p32, &c->b_q32, &c->p32b, &c->b_p32b,
                           &c->w_rho, &c->w_l1, &c->w_pr, &c->w_cent, &c->w_rho_out, &c->w_sl1};
    for (auto b : fl) b->release();
    c->w_cnt.release();
    c->w_offs.release();
    c->fixed.release();
    if (c->st) (void)hipFree(c->st);
    if (c->st_host) (void)hipHostFree(c->st_host);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* hx_last_error(const hx_ctx* c) { return c ? c->err.c_str() : "null context"; }

int hx_element_stiffness(double dx, double dy, double dz, double nu, double E, double* Ke) {
    if (!Ke || !(dx > 0) || !(dy > 0) || !(dz > 0) || !(nu > -1.0 && nu < 0.5)) return HX_ERR_BAD_ARG;
    element_stiffness(dx, dy, dz, nu, E, Ke);
    return HX_OK;
}

int hx_scalar_matrices(double dx, double dy, double dz, int quirk, double* Ks, double* Ms) {
    if (!Ks || !Ms || !(dx > 0) || !(dy > 0) || !(dz > 0)) return HX_ERR_BAD_ARG;
    scalar_matrices(dx, dy, dz, quirk, Ks, Ms);
    return HX_OK;
}

int hx_grid_set(hx_ctx* c, int nx, int ny, int nz, double dx, double dy, double dz, double nu) {
    if (!c) return HX_ERR_BAD_ARG;
    if (nx < 1 || ny < 1 || nz < 1 || !(dx > 0) || !(dy > 0) || !(dz > 0) || !(nu > -1.0 && nu < 0.5))
        return fail(c, HX_ERR_BAD_ARG, "invalid grid");
    const long NX = nx + 1, NY = ny + 1, NZ = nz + 1;
    if (3 * NX * NY * NZ > (long)INT32_MAX) return fail(c, HX_ERR_BAD_ARG, "grid too large");
    (void)hipSetDevice(c->device);
    clear_graphs(c);
    GridArgs g{nx, ny, nz, (int)NX, (int)NY, (int)NZ, NX * NY * NZ, (long)nx * ny * nz, dx, dy, dz};
    c->g = g;
    c->nu = nu;
    element_stiffness(dx, dy, dz, nu, 1.0, c->Ke);
    // binary local order and parity-basis image
    for (int a = 0; a < 8; ++a)
        for (int b = 0; b < 8; ++b)
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                    c->Kbin[24 * (3 * kRefToBin[a] + i) + 3 * kRefToBin[b] + j] = c->Ke[24 * (3 * a + i) + 3 * b + j];
    auto V = [](int L, int s) {  // prod over axes in s of xi(L)
        double v = 1.0;
        for (int ax = 0; ax < 3; ++ax)
            if (s & (1 << ax)) v *= ((L >> ax) & 1) ? 1.0 : -1.0;
        return v;
    };
    double Kp[24][24];
    double kmax = 0.0;
    for (int s = 0; s < 8; ++s)
        for (int ci = 0; ci < 3; ++ci)
            for (int t = 0; t < 8; ++t)
                for (int di = 0; di < 3; ++di) {
                    double acc = 0.0;
                    for (int L = 0; L < 8; ++L)
                        for (int M = 0; M < 8; ++M)
                            acc += V(L, s) * V(M, t) * c->Kbin[24 * (3 * L + ci) + 3 * M + di];
                    Kp[3 * s + ci][3 * t + di] = acc / 64.0;
                    kmax = std::max(kmax, std::fabs(acc / 64.0));
                }
    for (int i = 0; i < 24; ++i)  // exact symmetry, so mirrored entries may share one value
        for (int j = 0; j < i; ++j) Kp[i][j] = Kp[j][i] = 0.5 * (Kp[i][j] + Kp[j][i]);
    bool inpat[24][24] = {};
    for (int t = 0; t < kParityNnz; ++t) {
        inpat[kParityPattern[t].r][kParityPattern[t].c] = true;
        c->kp64.v[t] = Kp[kParityPattern[t].r][kParityPattern[t].c];
        c->kp32.v[t] = (float)c->kp64.v[t];
    }
    for (int i = 0; i < 24; ++i)
        for (int j = 0; j < 24; ++j)
            if (!inpat[i][j] && std::fabs(Kp[i][j]) > 1e-12 * kmax)
                return fail(c, HX_ERR_STATE, "element stiffness outside the parity pattern");
    for (int d = 0; d < 3; ++d) c->kd[d] = c->Ke[24 * d + d];
    for (int a = 1; a < 8; ++a)
        for (int d = 0; d < 3; ++d)
            if (std::fabs(c->Ke[24 * (3 * a + d) + 3 * a + d] - c->kd[d]) > 1e-12 * std::fabs(c->kd[d]))
                return fail(c, HX_ERR_STATE, "unequal element diagonal");
    const long n = 3 * g.nn;
    HXA(c->E, g.ne);
    HXA(c->f, n);
    HXA(c->x, n);
    HXA(c->r, n);
    HXA(c->p, n);
    HXA(c->q, n);
    HXA(c->minv, n);
    HXA(c->p2, n);
    HXA(c->x_best, n);
    HXA(c->Kbin_d, 576);
    HXA(c->Kbin32_d, 576);
    const long np = std::max<long>({3L * kVecBlocks + 8, (long)matvec_parity_blocks(g, 0, 4) + 8,
                                    (long)matvec_parity_blocks(g, 1, 4) + 8, (long)matvec_parity_blocks(g, 2, 4) + 8,
                                    (long)matvec_gather_blocks(g) + 8,
                                    (long)helm_matvec_blocks(g) + 8, (long)element_post_blocks(g) + 8});
    HXA(c->partials, kOpPartials + np * kMaxRhs);
    HXA(c->small, 64);
    float kb32[576];
    for (int t = 0; t < 576; ++t) kb32[t] = (float)c->Kbin[t];
    HXC(hipMemcpy(c->Kbin_d.p, c->Kbin, 576 * sizeof(double), hipMemcpyHostToDevice));
    HXC(hipMemcpy(c->Kbin32_d.p, kb32, 576 * sizeof(float), hipMemcpyHostToDevice));
    // default: everything free
    HXA(c->fixed, n);
    HXC(hipMemsetAsync(c->fixed.p, 0, n, c->stream));
    launch_fill(n, 0.0, c->x.p, c->stream);
    launch_fill(n, 0.0, c->p.p, c->stream);
    HXC(hipStreamSynchronize(c->stream));
    c->has_fixed = false;
    c->has_grid = true;
    c->to_ready = false;
    return HX_OK;
}

int hx_set_fixed(hx_ctx* c, const uint8_t* fixed) {
    if (!c) return HX_ERR_BAD_ARG;
    if (int rc = ensure_grid(c)) return rc;
    const long n = ndof(c);
    if (!fixed) {
        HXC(hipMemsetAsync(c->fixed.p, 0, n, c->stream));
        c->has_fixed = false;
    } else {
        HXC(hipMemcpyAsync(c->fixed.p, fixed, n, hipMemcpyHostToDevice, c->stream));
        c->has_fixed = true;
    }
    HXC(hipStreamSynchronize(c->stream));
    return HX_OK;
}

int hx_matvec(hx_ctx* c, const double* Ee, const double* x, double* y, int variant) {
    if (!c || !Ee || !x || !y) return HX_ERR_BAD_ARG;
    if (int rc = ensure_grid(c)) return rc;
    if (variant < HX_MV_AUTO || variant > HX_MV_GATHER) return fail(c, HX_ERR_BAD_ARG, "matvec variant");
    const long n = ndof(c);
    int rc;
    if ((rc = upload(c, c->E.p, Ee, c->g.ne * sizeof(double))) || (rc = upload(c, c->p.p, x, n * sizeof(double))))
        return rc;
    mv64(c, c->p.p, c->q.p, c->E.p, 0.0, MV_PLAIN, variant);
    HXK();
    return download(c, y, c->q.p, n * sizeof(double));
}

int hx_diagonal(hx_ctx* c, const double* Ee, double* diag) {
    if (!c || !Ee || !diag) return HX_ERR_BAD_ARG;
    if (int rc = ensure_grid(c)) return rc;
    int rc = upload(c, c->E.p, Ee, c->g.ne * sizeof(double));
    if (rc) return rc;
    launch_elastic_diag(c->g, c->kd, c->E.p, c->q.p, c->stream);
    HXK();
    return download(c, diag, c->q.p, ndof(c) * sizeof(double));
}

static int run_pcg(hx_ctx* c, const hx_pcg_opts& o, hx_pcg_stats* s, const SolveVecs& v) {
    if (o.precond == HX_PC_LEARNED && v.nrhs == 1) return pcg_learned(c, o, s);
    return (o.precision == HX_FP32_IR) ? pcg_fp32ir(c, o, s, v) : pcg_fp64(c, o, s, v);
}

static int pcg_dispatch(hx_ctx* c, const hx_pcg_opts& o, hx_pcg_stats* s) {
    if (o.precision == HX_FP32_IR)
        if (int rc = ensure_fp32(c)) return rc;
    hipEvent_t a = get_event(c), b = get_event(c);
    (void)hipEventRecord(a, c->stream);
    const int rc = run_pcg(c, o, s, single_vecs(c));
    (void)hipEventRecord(b, c->stream);
    (void)hipEventSynchronize(b);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, a, b);
    c->free_events.push_back(a);
    c->free_events.push_back(b);
    if (s) s->seconds = ms * 1e-3;
    return rc;
}

static int check_opts(hx_ctx* c, const hx_pcg_opts* o) {
    if (!o) return fail(c, HX_ERR_BAD_ARG, "null options");
    if (!(o->tol > 0) || o->max_iter < 1 || !(o->shift >= 0)) return fail(c, HX_ERR_BAD_ARG, "invalid PCG options");
    if (o->precision != HX_FP64 && o->precision != HX_FP32_IR) return fail(c, HX_ERR_BAD_ARG, "precision");
    if (o->precond != HX_PC_JACOBI && o->precond != HX_PC_LEARNED && o->precond != HX_PC_MULTILEVEL &&
        o->precond != HX_PC_VCYCLE)
        return fail(c, HX_ERR_BAD_ARG, "preconditioner");
    if (o->precond == HX_PC_LEARNED && !c->unet_trained)
        return fail(c, HX_ERR_STATE, "learned preconditioner not trained (hx_learned_train)");
    if (o->matvec < HX_MV_AUTO || o->matvec > HX_MV_MFMA) return fail(c, HX_ERR_BAD_ARG, "matvec variant");
    if (o->matvec == HX_MV_MFMA && o->precision != HX_FP32_IR)
        return fail(c, HX_ERR_BAD_ARG, "the MFMA contraction is FP32 (precision form only)");
    return HX_OK;
}

// GPUSolver.jl:132-134: a solve needs at least one element with stiffness
bool any_active(const double* Ee, long ne) {
    for (long e = 0; e < ne; ++e)
        if (Ee[e] != 0.0) return true;
    return false;
}

int hx_pcg_solve(hx_ctx* c, const double* Ee, const double* f, double* u, const hx_pcg_opts* o, hx_pcg_stats* s) {
    if (!c || !Ee || !f || !u) return HX_ERR_BAD_ARG;
    if (int rc = ensure_grid(c)) return rc;
    if (int rc = check_opts(c, o)) return rc;
    if (!any_active(Ee, c->g.ne)) return fail(c, HX_ERR_NO_ACTIVE, "no active elements (every modulus is 0)");
    const long n = ndof(c);
    int rc;
    if ((rc = upload(c, c->E.p, Ee, c->g.ne * sizeof(double))) || (rc = upload(c, c->f.p, f, n * sizeof(double))))
        return rc;
    if (o->warm_start && (rc = upload(c, c->x.p, u, n * sizeof(double)))) return rc;
    const int prc = pcg_dispatch(c, *o, s);
    if ((rc = download(c, u, c->x.p, n * sizeof(double)))) return rc;
    return prc;
}

int hx_solve_system(hx_ctx* c, const double* density, const double* f, double* u, const hx_solve_opts* o,
                    hx_solve_stats* s) {
    if (!c || !density || !f || !u || !o) return HX_ERR_BAD_ARG;
    if (int rc = ensure_grid(c)) return rc;
    if (!(o->tol > 0) || o->max_iter < 1 || !(o->shift_factor >= 0) || o->max_attempts < 1 ||
        !(o->min_stiffness_threshold >= 0))
        return fail(c, HX_ERR_BAD_ARG, "invalid solve options");
    if (o->precision != HX_FP64 && o->precision != HX_FP32_IR) return fail(c, HX_ERR_BAD_ARG, "precision");
    if (o->precond != HX_PC_JACOBI && o->precond != HX_PC_MULTILEVEL && o->precond != HX_PC_VCYCLE)
        return fail(c, HX_ERR_BAD_ARG, "preconditioner (jacobi, multilevel or vcycle)");
    const long ne = c->g.ne, n = ndof(c);
    hx_solve_stats st{};
    HXA(c->s_rho, ne);
    HXA(c->x_glob, n);
    int rc;
    if ((rc = upload(c, c->s_rho.p, density, ne * sizeof(double))) || (rc = upload(c, c->f.p, f, n * sizeof(double))))
        return rc;
    if (o->warm_start && (rc = upload(c, c->x.p, u, n * sizeof(double)))) return rc;
    if (o->precision == HX_FP32_IR && (rc = ensure_fp32(c))) return rc;
    hipEvent_t ev0 = get_event(c), ev1 = get_event(c);
    struct Back {  // events back to the pool on every return
        hx_ctx* c;
        hipEvent_t a, b;
        ~Back() {
            c->free_events.push_back(a);
            c->free_events.push_back(b);
        }
    } back{c, ev0, ev1};
    (void)hipEventRecord(ev0, c->stream);
    // element selection (Solver.jl:61-65: the pruning threshold is an FP32 product)
    const double thr = o->min_stiffness_threshold;
    const double thr_keep = (double)((float)thr * 1.01f);
    launch_count_active(ne, c->s_rho.p, thr, thr_keep, c->partials.p, c->small.p, c->stream);
    HXK();
    double counts[2];
    if ((rc = download(c, counts, c->small.p, 2 * sizeof(double)))) return rc;
    st.n_active = (long)counts[0];
    const long n_keep = (long)counts[1];
    st.pruned = (o->prune_voids && (double)n_keep < 0.99 * (double)ne) ? 1 : 0;
    st.n_kept = st.pruned ? n_keep : st.n_active;
    if (st.n_kept == 0) {
        if (s) *s = st;
        return fail(c, HX_ERR_NO_ACTIVE, "no active elements found (threshold " + std::to_string(thr) + ")");
    }
    launch_solve_moduli(ne, c->s_rho.p, o->E, thr, thr_keep, st.pruned, c->E.p, c->stream);
    // shift = factor * the largest free diagonal entry (GPUSolver.jl:315-316, 391)
    launch_elastic_diag(c->g, c->kd, c->E.p, c->q.p, c->stream);
    launch_free_absmax(n, c->q.p, c->has_fixed ? c->fixed.p : nullptr, c->partials.p, c->small.p + 2, c->stream);
    HXK();
    double max_diag = 0.0;
    if ((rc = download(c, &max_diag, c->small.p + 2, sizeof(double)))) return rc;
    double factor = o->shift_factor;
    bool success = false, have_glob = false;
    double glob_rel = INFINITY;
    int last_code = HX_OK;
    for (int attempt = 1; attempt <= o->max_attempts; ++attempt) {
        const double shift = factor * max_diag;
        hx_pcg_opts po{o->tol, o->max_iter, shift, o->precision, o->precond, (o->warm_start && attempt == 1) ? 1 : 0,
                       HX_MV_AUTO, 0};
        hx_pcg_stats ps{};
        const int prc = run_pcg(c, po, &ps, single_vecs(c));
        if (prc == HX_ERR_HIP || prc == HX_ERR_OOM || prc == HX_ERR_STATE || prc == HX_ERR_BAD_ARG) return prc;
        st.attempts = attempt;
        st.iterations += ps.iterations;
        st.shift = shift;
        st.rel_res = c->last_rec_rel;
        st.best_rel = c->last_best_rel;
        st.true_rel = ps.rel_res;
        if (ps.status == HX_ZERO_RHS) {  // zeros, 0.0, "None"
            launch_fill(n, 0.0, c->x.p, c->stream);
            st.outcome = HX_SOLVE_ZERO_RHS;
            st.rel_res = st.true_rel = st.best_rel = 0.0;
            success = true;
            break;
        }
        if (prc == HX_OK) {
            st.outcome = HX_SOLVE_CONVERGED;
            success = true;
            break;
        }
        last_code = prc;
        const double best = c->last_best_rel;
        const bool diverged = prc == HX_ERR_BREAKDOWN || prc == HX_ERR_DIVERGED;
        const bool failed_badly = diverged || !(best <= 0.1);
        if (!failed_badly && best < 5e-2) {  // acceptable stagnation (GPUSolver.jl:549-557): x is the best iterate
            st.outcome = HX_SOLVE_ACCEPTED;
            success = true;
            break;
        }
        if (failed_badly) {
            factor *= 10.0;
        } else {
            if (best < glob_rel) {
                glob_rel = best;
                HXC(hipMemcpyAsync(c->x_glob.p, c->x.p, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
                have_glob = true;
            }
            factor *= 5.0;
        }
    }
    int code = HX_OK;
    if (!success) {  // "All solver attempts failed. Returning best stagnated result." (GPUSolver.jl:603-606)
        st.outcome = HX_SOLVE_FAILED;
        code = HX_ERR_NOT_CONVERGED;
        if (have_glob) HXC(hipMemcpyAsync(c->x.p, c->x_glob.p, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        else launch_fill(n, 0.0, c->x.p, c->stream);
        mv64(c, c->x.p, c->q.p, c->E.p, 0.0, MV_PLAIN, HX_MV_AUTO);
        launch_true_residual(n, c->f.p, c->q.p, c->x.p, c->minv.p, st.shift, nullptr, nullptr, c->st, c->partials.p, 0,
                             c->stream);
        HXK();
        if ((rc = sync_state(c))) return rc;
        st.true_rel = c->st_host[0].rel;
        (void)last_code;
    }
    (void)hipEventRecord(ev1, c->stream);
    if ((rc = download(c, u, c->x.p, n * sizeof(double)))) return rc;
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ev0, ev1);
    st.seconds = ms * 1e-3;
    st.status = code;
    if (s) *s = st;
    if (code != HX_OK) return fail(c, code, "all solver attempts failed; best stagnated result returned");
    return HX_OK;
}

int hx_block_pcg_solve(hx_ctx* c, int nrhs, const double* Ee, const double* F, double* U, const hx_pcg_opts* o,
                       hx_pcg_stats* s) {
    if (!c || !Ee || !F || !U) return HX_ERR_BAD_ARG;
    if (int rc = ensure_grid(c)) return rc;
    if (nrhs < 1 || nrhs > kMaxRhs) return fail(c, HX_ERR_BAD_ARG, "nrhs must be 1..8");
    if (o && (o->precond == HX_PC_MULTILEVEL || o->precond == HX_PC_LEARNED) && nrhs > 1)
        return fail(c, HX_ERR_BAD_ARG, "the additive multilevel and learned preconditioners take one right-hand side");
    if (int rc = check_opts(c, o)) return rc;
    if (o->precond == HX_PC_VCYCLE && nrhs > 1 && ml_levels_for(c->g, std::min(c->ml_max_levels, kMlMaxLevels)) < 1)
        return fail(c, HX_ERR_BAD_ARG, "a V-cycle block solve needs a coarse level (even element counts)");
    const long n = ndof(c), nt = n * nrhs;
    HXA(c->b_f, nt);
    HXA(c->b_x, nt);
    HXA(c->b_r, nt);
    HXA(c->b_p, nt);
    HXA(c->b_q, nt);
    HXA(c->b_p2, nt);
    if (o->precision == HX_FP32_IR) {
        if (int rc = ensure_fp32(c)) return rc;
        HXA(c->b_x32, nt);
        HXA(c->b_r32, nt);
        HXA(c->b_p32, nt);
        HXA(c->b_q32, nt);
        HXA(c->b_p32b, nt);
    }
    int rc;
    if ((rc = upload(c, c->E.p, Ee, c->g.ne * sizeof(double))) || (rc = upload(c, c->b_f.p, F, nt * sizeof(double))))
        return rc;
    if (o->warm_start && (rc = upload(c, c->b_x.p, U, nt * sizeof(double)))) return rc;
    SolveVecs v{nrhs, c->b_f.p, c->b_x.p, c->b_r.p, c->b_p.p, c->b_q.p, c->b_x32.p, c->b_r32.p, c->b_p32.p,
                c->b_q32.p, c->b_p2.p, c->b_p32b.p};
    hipEvent_t a = get_event(c), b = get_event(c);
    (void)hipEventRecord(a, c->stream);
    const int prc = run_pcg(c, *o, s, v);
    (void)hipEventRecord(b, c->stream);
    (void)hipEventSynchronize(b);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, a, b);
    c->free_events.push_back(a);
    c->free_events.push_back(b);
    if (s)
        for (int k = 0; k < nrhs; ++k) s[k].seconds = ms * 1e-3;
    if ((rc = download(c, U, c->b_x.p, nt * sizeof(double)))) return rc;
    return prc;
}

int hx_helmholtz_filter(hx_ctx* c, double R, const hx_filter_opts* o, const double* in_e, double* out_e,
                        hx_filter_stats* s) {
    if (!c || !o || !in_e || !out_e || !(R >= 0) || !(o->tol > 0)) return HX_ERR_BAD_ARG;
    if (int rc = ensure_grid(c)) return rc;
    const long ne = c->g.ne;
    HXA(c->h_in, ne);
    HXA(c->h_out, ne);
    int rc;
    if ((rc = upload(c, c->h_in.p, in_e, ne * sizeof(double)))) return rc;
    hipEvent_t a = get_event(c), b = get_event(c);
    (void)hipEventRecord(a, c->stream);
    const int hrc = helm_solve(c, R, *o, c->h_in.p, c->h_out.p, s);
    (void)hipEventRecord(b, c->stream);
    (void)hipEventSynchronize(b);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, a, b);
    c->free_events.push_back(a);
    c->free_events.push_back(b);
    if (s) s->seconds = ms * 1e-3;
    if ((rc = download(c, out_e, c->h_out.p, ne * sizeof(double)))) return rc;
    return hrc;
}

int hx_element_stress(hx_ctx* c, const double* Ee, const double* u, double* principal, double* vm, double* voigt,
                      double* l1, double* dir) {
    if (!c || !Ee || !u) return HX_ERR_BAD_ARG;
    if (int rc = ensure_grid(c)) return rc;
    const long ne = c->g.ne;
    HXA(c->e_pr, 3 * ne);
    HXA(c->e_vm, ne);
    HXA(c->e_sv, 6 * ne);
    HXA(c->e_l1, ne);
    HXA(c->e_dir, 3 * ne);
    int rc;
    if ((rc = upload(c, c->E.p, Ee, ne * sizeof(double))) || (rc = upload(c, c->x.p, u, ndof(c) * sizeof(double))))
        return rc;
    ElemArgs a{};
    a.u = c->x.p;
    a.Emod = c->E.p;
    material_consts(c->nu, &a.lam1, &a.lam2, &a.mu);
    a.principal = principal ? c->e_pr.p : nullptr;
    a.vm = vm ? c->e_vm.p : nullptr;
    a.voigt = voigt ? c->e_sv.p : nullptr;
    a.l1 = l1 ? c->e_l1.p : nullptr;
    a.dir = dir ? c->e_dir.p : nullptr;
    elem_post(c, a);
    HXK();
    if (principal && (rc = download(c, principal, c->e_pr.p, 3 * ne * sizeof(double)))) return rc;
    if (vm && (rc = download(c, vm, c->e_vm.p, ne * sizeof(double)))) return rc;
    if (voigt && (rc = download(c, voigt, c->e_sv.p, 6 * ne * sizeof(double)))) return rc;
    if (l1 && (rc = download(c, l1, c->e_l1.p, ne * sizeof(double)))) return rc;
    if (dir && (rc = download(c, dir, c->e_dir.p, 3 * ne * sizeof(double)))) return rc;
    HXC(hipStreamSynchronize(c->stream));
    return HX_OK;
}

int hx_web_pack(hx_ctx* c, const float* density, const float* l1, const float* principal, float threshold,
                float* centroids, float* density_out, float* signed_l1, uint32_t* count) {
    if (!c || !density || !l1 || !principal || !centroids || !density_out || !signed_l1 || !count)
        return HX_ERR_BAD_ARG;
    if (int rc = ensure_grid(c)) return rc;
    const long ne = c->g.ne;
    const int nt = web_tiles(ne);
    HXA(c->w_rho, ne);
    HXA(c->w_l1, ne);
    HXA(c->w_pr, 3 * ne);
    HXA(c->w_cent, 3 * ne);
    HXA(c->w_rho_out, ne);
    HXA(c->w_sl1, ne);
    HXA(c->w_cnt, nt);
    HXA(c->w_offs, nt + 1);
    int rc;
    if ((rc = upload(c, c->w_rho.p, density, ne * sizeof(float))) || (rc = upload(c, c->w_l1.p, l1, ne * sizeof(float))) ||
        (rc = upload(c, c->w_pr.p, principal, 3 * ne * sizeof(float))))
        return rc;
    launch_web_pack(c->g, c->w_rho.p, c->w_l1.p, c->w_pr.p, threshold, c->w_cnt.p, c->w_offs.p, c->w_cent.p,
                    c->w_rho_out.p, c->w_sl1.p, c->stream);
    HXK();
    unsigned n = 0;
    if ((rc = download(c, &n, c->w_offs.p + nt, sizeof(unsigned)))) return rc;
    if ((long)n > ne) return fail(c, HX_ERR_HIP, "hx_web_pack: kept count exceeds the element count");
    *count = n;
    if (n == 0) return HX_OK;
    if ((rc = download(c, centroids, c->w_cent.p, 3L * n * sizeof(float))) ||
        (rc = download(c, density_out, c->w_rho_out.p, (long)n * sizeof(float))) ||
        (rc = download(c, signed_l1, c->w_sl1.p, (long)n * sizeof(float))))
        return rc;
    return HX_OK;
}

int hx_web_pack_device(hx_ctx* c, const float* density, const float* l1, const float* principal, float threshold,
                       float* centroids, float* density_out, float* signed_l1, uint32_t* count) {
    if (!c || !density || !l1 || !principal || !centroids || !density_out || !signed_l1 || !count)
        return HX_ERR_BAD_ARG;
    if (int rc = ensure_grid(c)) return rc;
    const int nt = web_tiles(c->g.ne);
    if (nt > kWebScratchTiles)
        return fail(c, HX_ERR_BAD_ARG, "hx_web_pack_device: at most " + std::to_string((long)kWebScratchTiles * 2048) +
                                           " elements (the scratch's tiles)");
    if (!c->web_scratch) return fail(c, HX_ERR_STATE, "hx_web_pack_device: no scratch address on this device");
    HXC(hipSetDevice(c->device));
    launch_web_pack(c->g, density, l1, principal, threshold, c->web_scratch, c->web_scratch + kWebScratchTiles,
                    centroids, density_out, signed_l1, c->stream, count);
    HXK();
    HXC(hipStreamSynchronize(c->stream));
    return HX_OK;
}

int hx_vtk_pack(hx_ctx* c, const double* u, const double* f, const uint8_t* fixed, const float* density,
                const float* l1, const float* vm, float threshold, uint32_t* points, uint32_t* cells,
                uint32_t* cell_scalars, uint32_t* count) {
    if (!c || !density || !points || !cells || !cell_scalars || !count) return HX_ERR_BAD_ARG;
    if (int rc = ensure_grid(c)) return rc;
    const long ne = c->g.ne, nn = c->g.nn, n = ndof(c);
    const int nt = web_tiles(ne);
    HXA(c->v_pts, 13 * nn);
    HXA(c->v_cells, 12 * ne);
    HXA(c->w_rho, ne);
    HXA(c->v_l1, ne);
    HXA(c->v_vm, ne);
    HXA(c->w_cnt, nt);
    HXA(c->w_offs, nt + 1);
    int rc;
    if ((rc = upload(c, c->w_rho.p, density, ne * sizeof(float)))) return rc;
    if (l1) { if ((rc = upload(c, c->v_l1.p, l1, ne * sizeof(float)))) return rc; }
    else launch_fill32(ne, 0.0f, c->v_l1.p, c->stream);
    if (vm) { if ((rc = upload(c, c->v_vm.p, vm, ne * sizeof(float)))) return rc; }
    else launch_fill32(ne, 0.0f, c->v_vm.p, c->stream);
    if (u) {
        HXA(c->v_u, n);
        if ((rc = upload(c, c->v_u.p, u, n * sizeof(double)))) return rc;
    }
    if (f) {
        HXA(c->v_f, n);
        if ((rc = upload(c, c->v_f.p, f, n * sizeof(double)))) return rc;
    }
    if (fixed) {
        HXA(c->v_fixed, n);
        if ((rc = upload(c, c->v_fixed.p, fixed, n))) return rc;
    }
    unsigned* P = c->v_pts.p;
    unsigned* Cw = c->v_cells.p;
    launch_vtk_pack(c->g, u ? c->v_u.p : nullptr, f ? c->v_f.p : nullptr, fixed ? c->v_fixed.p : nullptr, c->w_rho.p,
                    c->v_l1.p, c->v_vm.p, threshold, c->w_cnt.p, c->w_offs.p, c->w_offs.p + nt, P, P + 3 * nn,
                    P + 6 * nn, P + 9 * nn, P + 12 * nn, Cw, Cw + 9 * ne, Cw + 10 * ne, Cw + 11 * ne, c->stream);
    HXK();
    unsigned kept = 0;
    if ((rc = download(c, &kept, c->w_offs.p + nt, sizeof(unsigned)))) return rc;
    if ((long)kept > ne) return fail(c, HX_ERR_HIP, "hx_vtk_pack: kept count exceeds the element count");
    *count = kept;
    if ((rc = download(c, points, P, 13 * nn * sizeof(unsigned)))) return rc;
    if (kept) {
        if ((rc = download(c, cells, Cw, 9L * kept * sizeof(unsigned)))) return rc;
        for (int q = 0; q < 3; ++q)
            if ((rc = download(c, cell_scalars + (long)q * kept, Cw + (9 + q) * ne, (long)kept * sizeof(unsigned))))
                return rc;
    }
    return HX_OK;
}

int hx_stl_pack(hx_ctx* c, const float* density, int subdivision, float threshold, float* triangles,
                long max_triangles, long* count, float* field) {
    if (!c || !density || !count || subdivision < 1 || max_triangles < 0 || (max_triangles > 0 && !triangles))
        return HX_ERR_BAD_ARG;
    if (int rc = ensure_grid(c)) return rc;
    const GridArgs& g = c->g;
    const long ne = g.ne, nn = g.nn;
    const int FX = subdivision * g.nx + 3, FY = subdivision * g.ny + 3, FZ = subdivision * g.nz + 3;
    const long np = (long)FX * FY * FZ, nb = mc_blocks(FX, FY, FZ);
    if (nb > 64L * 1024 * 1024) return fail(c, HX_ERR_BAD_ARG, "hx_stl_pack: sampling grid too large");
    HXA(c->w_rho, ne);
    HXA(c->s_nod, nn);
    HXA(c->s_fine, np);
    HXA(c->s_blk, nb);
    HXA(c->s_offs, nb + 1);
    int rc;
    if (!c->s_tri.p) {
        HXA(c->s_tri, 256 * 16);
        HXA(c->s_ntri, 256);
        if ((rc = upload(c, c->s_tri.p, mc_table_host(0), 256 * 16)) || (rc = upload(c, c->s_ntri.p, mc_table_host(1), 256)))
            return rc;
    }
    if ((rc = upload(c, c->w_rho.p, density, ne * sizeof(float)))) return rc;
    launch_stl_field(g, subdivision, c->w_rho.p, c->s_nod.p, c->s_fine.p, c->stream);
    launch_mc_count(c->s_fine.p, FX, FY, FZ, threshold, c->s_ntri.p, c->s_blk.p, c->s_offs.p, c->stream);
    HXK();
    unsigned total = 0;
    if ((rc = download(c, &total, c->s_offs.p + nb, sizeof(unsigned)))) return rc;
    *count = total;
    if (field && (rc = download(c, field, c->s_fine.p, np * sizeof(float)))) return rc;
    const long nw = std::min<long>(total, max_triangles);
    if (nw > 0) {
        HXA(c->s_tris, 12 * nw);
        const float sx = (float)g.dx / (float)subdivision, sy = (float)g.dy / (float)subdivision,
                    sz = (float)g.dz / (float)subdivision;
        launch_mc_write(c->s_fine.p, FX, FY, FZ, threshold, -sx, -sy, -sz, sx, sy, sz, c->s_tri.p, c->s_ntri.p,
                        c->s_offs.p, (unsigned)nw, c->s_tris.p, c->stream);
        HXK();
        if ((rc = download(c, triangles, c->s_tris.p, 12 * nw * sizeof(float)))) return rc;
    }
    return HX_OK;
}

namespace {
// the FP32 location of Boundary.interpret_location_component (Boundary.jl:79-115)
// on this grid: tolerance 1e-4 * max(1, |cmax - cmin|) with cmin = 0 and cmax
// the last node's coordinate Float32(n) * h
LocSpec loc_spec(const hx_ctx* c, const hx_location& l) {
    LocSpec s{};
    const int n[3] = {c->g.nx, c->g.ny, c->g.nz};
    const float h[3] = {(float)c->g.dx, (float)c->g.dy, (float)c->g.dz};
    for (int d = 0; d < 3; ++d) {
        s.all[d] = l.all[d] ? 1 : 0;
        s.val[d] = l.value[d];
        const float cmax = (float)n[d] * h[d];
        s.tol[d] = 1.0e-4f * std::max(1.0f, std::fabs(cmax - 0.0f));
    }
    return s;
}
}  // namespace

int hx_bc_indicator(hx_ctx* c, int nspec, const hx_location* locs, const int* dof_bits, uint8_t* fixed_out) {
    if (!c || nspec < 0 || (nspec > 0 && (!locs || !dof_bits))) return HX_ERR_BAD_ARG;
    if (int rc = ensure_grid(c)) return rc;
    const long n = ndof(c);
    HXC(hipMemsetAsync(c->fixed.p, 0, n, c->stream));
    for (int q = 0; q < nspec; ++q) {
        if (dof_bits[q] & ~7) return fail(c, HX_ERR_BAD_ARG, "Invalid DoF index (must be 1..3)");
        launch_bc_mark(c->g, loc_spec(c, locs[q]), dof_bits[q], c->fixed.p, c->stream);
    }
    HXK();
    c->has_fixed = true;
    if (fixed_out) return download(c, fixed_out, c->fixed.p, n);
    HXC(hipStreamSynchronize(c->stream));
    return HX_OK;
}

int hx_external_forces(hx_ctx* c, int nspec, const hx_location* locs, const double* force, const int* point_fallback,
                       double* F, uint32_t* matched) {
    if (!c || !F || nspec < 0 || (nspec > 0 && (!locs || !force))) return HX_ERR_BAD_ARG;
    if (int rc = ensure_grid(c)) return rc;
    const long n = ndof(c), nb = boundary_blocks(c->g);
    HXA(c->b_part, nb);
    HXA(c->b_cnt, nspec > 0 ? nspec : 1);
    HXA(c->b_dpart, nb);
    HXA(c->b_dmin, 1);
    HXA(c->b_kpart, nb);
    HXA(c->b_key, 1);
    HXA(c->b_F, n);
    int rc;
    if ((rc = upload(c, c->b_F.p, F, n * sizeof(double)))) return rc;
    std::vector<unsigned> cnt(nspec > 0 ? nspec : 1, 0);
    for (int q = 0; q < nspec; ++q) {  // the specs in order, as the reference adds them
        const LocSpec s = loc_spec(c, locs[q]);
        launch_loc_count(c->g, s, c->b_part.p, c->b_cnt.p + q, c->stream);
        HXK();
        if ((rc = download(c, &cnt[q], c->b_cnt.p + q, sizeof(unsigned)))) return rc;
        if (cnt[q] > 0) {
            launch_force_add(c->g, s, c->b_cnt.p + q, force + 3 * q, c->b_F.p, c->stream);
        } else if (point_fallback && point_fallback[q] && !s.all[0] && !s.all[1] && !s.all[2]) {
            // no node on the point: its nearest node takes the whole force (Boundary.jl:254-265)
            launch_nearest(c->g, s.val, c->b_dpart.p, c->b_dmin.p, c->b_kpart.p, c->b_key.p, c->stream);
            launch_force_point(c->g, c->b_key.p, force + 3 * q, c->b_F.p, c->stream);
            cnt[q] = 1;
        }
        HXK();
    }
    if (matched)
        for (int q = 0; q < nspec; ++q) matched[q] = cnt[q];
    return download(c, F, c->b_F.p, n * sizeof(double));
}

int hx_regrid(hx_ctx* c, int onx, int ony, int onz, int nnx, int nny, int nnz, float len_x, float len_y, float len_z,
              int nfields, const float* old_fields, float* new_fields) {
    if (!c || onx < 1 || ony < 1 || onz < 1 || nnx < 1 || nny < 1 || nnz < 1 || nfields < 1 || !old_fields ||
        !new_fields || !(len_x > 0) || !(len_y > 0) || !(len_z > 0))
        return HX_ERR_BAD_ARG;
    const long ono = (long)onx * ony * onz, nno = (long)nnx * nny * nnz;
    HXA(c->r_old, ono * nfields);
    HXA(c->r_new, nno * nfields);
    int rc;
    if ((rc = upload(c, c->r_old.p, old_fields, ono * nfields * sizeof(float)))) return rc;
    // spacings as the reference forms them: Float32 length / Int count (MeshRefiner.jl:122-129,
    // Helpers.calculate_element_distribution)
    const float nh[3] = {len_x / (float)nnx, len_y / (float)nny, len_z / (float)nnz};
    const float oh[3] = {len_x / (float)onx, len_y / (float)ony, len_z / (float)onz};
    launch_regrid(onx, ony, onz, nnx, nny, nnz, nh, oh, nfields, c->r_old.p, c->r_new.p, c->stream);
    HXK();
    return download(c, new_fields, c->r_new.p, nno * nfields * sizeof(float));
}

int hx_compliance_sens(hx_ctx* c, const double* xphys, double penal, double E0, double Emin, const double* u,
                       const double* f, double* ce, double* dc, double* compliance) {
    if (!c || !u || (dc && !xphys) || (compliance && !f)) return HX_ERR_BAD_ARG;
    if (int rc = ensure_grid(c)) return rc;
    const long ne = c->g.ne, n = ndof(c);
    HXA(c->e_ce, ne);
    HXA(c->e_dc, ne);
    HXA(c->e_aux, ne);
    int rc;
    if ((rc = upload(c, c->x.p, u, n * sizeof(double)))) return rc;
    if (xphys && (rc = upload(c, c->e_aux.p, xphys, ne * sizeof(double)))) return rc;
    if (f && (rc = upload(c, c->f.p, f, n * sizeof(double)))) return rc;
    ElemArgs a{};
    a.u = c->x.p;
    a.xphys = xphys ? c->e_aux.p : nullptr;
    a.penal = penal;
    a.E0 = E0;
    a.Emin = Emin;
    a.ce = c->e_ce.p;
    a.dc = dc ? c->e_dc.p : nullptr;
    elem_post(c, a);
    if (compliance) launch_dot(n, c->f.p, c->x.p, c->small.p, &c->st->ticket[0], c->partials.p, c->stream);
    HXK();
    if (ce && (rc = download(c, ce, c->e_ce.p, ne * sizeof(double)))) return rc;
    if (dc && (rc = download(c, dc, c->e_dc.p, ne * sizeof(double)))) return rc;
    if (compliance && (rc = download(c, compliance, c->small.p, sizeof(double)))) return rc;
    return HX_OK;
}

int hx_oc_update(hx_ctx* c, const double* x, const double* dc, const double* dv, double volfrac, double move,
                 double xmin, double rel_tol, double* xnew, double* lmid, int* steps, double* change) {
    if (!c || !x || !dc || !dv || !xnew || !(rel_tol > 0) || !(move > 0)) return HX_ERR_BAD_ARG;
    if (int rc = ensure_grid(c)) return rc;
    const long ne = c->g.ne;
    HXA(c->e_ce, ne);
    HXA(c->e_dc, ne);
    HXA(c->e_aux, ne);
    HXA(c->e_vm, ne);
    int rc;
    if ((rc = upload(c, c->e_aux.p, x, ne * sizeof(double))) || (rc = upload(c, c->e_dc.p, dc, ne * sizeof(double))) ||
        (rc = upload(c, c->e_ce.p, dv, ne * sizeof(double))))
        return rc;
    if ((rc = oc_run(c, ne, c->e_aux.p, c->e_dc.p, c->e_ce.p, volfrac, move, xmin, rel_tol, c->e_vm.p))) return rc;
    if (lmid) *lmid = c->st_host->scal[2];
    if (steps) *steps = c->st_host->iter;
    if (change) *change = c->st_host->scal[3];
    return download(c, xnew, c->e_vm.p, ne * sizeof(double));
}

// the device-resident loop's OC search (8-point, optionally from a guessed
// bracket around lguess) on caller arrays
int hx_test_oc_search(hx_ctx* c, const double* x, const double* dc, const double* dv, double volfrac, double move,
                      double xmin, double rel_tol, double lguess, double* xnew, double* lmid, int* passes) {
    if (!c || !x || !dc || !dv || !xnew || !(rel_tol > 0) || !(move > 0)) return HX_ERR_BAD_ARG;
    if (int rc = ensure_grid(c)) return rc;
    const long ne = c->g.ne;
    HXA(c->e_ce, ne);
    HXA(c->e_dc, ne);
    HXA(c->e_aux, ne);
    HXA(c->e_vm, ne);
    int rc;
    if ((rc = upload(c, c->e_aux.p, x, ne * sizeof(double))) || (rc = upload(c, c->e_dc.p, dc, ne * sizeof(double))) ||
        (rc = upload(c, c->e_ce.p, dv, ne * sizeof(double))))
        return rc;
    if ((rc = oc_run(c, ne, c->e_aux.p, c->e_dc.p, c->e_ce.p, volfrac, move, xmin, rel_tol, c->e_vm.p, true,
                     lguess)))
        return rc;
    if (lmid) *lmid = c->st_host->scal[2];
    if (passes) *passes = c->st_host->iter;
    return download(c, xnew, c->e_vm.p, ne * sizeof(double));
}

// the manufactured displacement of the learned preconditioner's training on
// the current grid, every DOF solved (no mask)
int hx_test_smooth_field(hx_ctx* c, unsigned long long seed, int levels, double* e) {
    if (!c || !e || levels < 1 || levels > 20) return HX_ERR_BAD_ARG;
    if (int rc = ensure_grid(c)) return rc;
    const long n = ndof(c);
    HXA(c->l_z, 2 * n);
    launch_fill(n, 1.0, c->l_z.p + n, c->stream);
    launch_smooth_field(c->g, seed, levels, c->l_z.p + n, c->l_z.p, c->stream);
    HXK();
    return download(c, e, c->l_z.p, n * sizeof(double));
}

// ------------------------------------------------------- SIMP/OC session
int hx_to_init(hx_ctx* c, const hx_to_params* p, const double* f, const double* x0) {
    if (!c || !p || !f || !x0) return HX_ERR_BAD_ARG;
    if (int rc = ensure_grid(c)) return rc;
    if (!(p->penal >= 1) || !(p->volfrac > 0 && p->volfrac < 1) || !(p->rmin >= 0) || !(p->pcg_tol > 0) ||
        p->pcg_max_iter < 1 || !(p->filter_tol > 0) || !(p->oc_rel_tol > 0) || !(p->move > 0))
        return fail(c, HX_ERR_BAD_ARG, "invalid TO parameters");
    if (p->precision != HX_FP64 && p->precision != HX_FP32_IR) return fail(c, HX_ERR_BAD_ARG, "precision");
    if (p->precond != HX_PC_JACOBI && p->precond != HX_PC_LEARNED && p->precond != HX_PC_MULTILEVEL &&
        p->precond != HX_PC_VCYCLE)
        return fail(c, HX_ERR_BAD_ARG, "preconditioner");
    if (p->precond == HX_PC_LEARNED && !c->unet_trained)
        return fail(c, HX_ERR_STATE, "learned preconditioner not trained (hx_learned_train)");
    const long ne = c->g.ne, n = ndof(c);
    HXA(c->t_x, ne);
    HXA(c->t_xphys, ne);
    HXA(c->t_xnew, ne);
    HXA(c->t_dc, ne);
    HXA(c->t_dcf, ne);
    HXA(c->t_dv, ne);
    HXA(c->t_vm, ne);
    HXA(c->t_ones, ne);
    int rc;
    if ((rc = upload(c, c->f.p, f, n * sizeof(double))) || (rc = upload(c, c->t_x.p, x0, ne * sizeof(double))))
        return rc;
    c->tp = *p;
    launch_fill(ne, 1.0, c->t_ones.p, c->stream);
    launch_fill(n, 0.0, c->x.p, c->stream);
    const double R = p->rmin / (2.0 * std::sqrt(3.0));  // Lazarov & Sigmund 2011
    hx_filter_opts fo{0, p->filter_tol, p->filter_max_iter > 0 ? p->filter_max_iter : 1000, 1};
    if ((rc = helm_solve(c, R, fo, c->t_x.p, c->t_xphys.p, nullptr))) return rc;
    if ((rc = helm_solve(c, R, fo, c->t_ones.p, c->t_dv.p, nullptr))) return rc;  // dv = H 1
    c->to_iter = 0;
    c->oc_lambda = 0.0;
    c->to_ready = true;
    return HX_OK;
}

int hx_to_step(hx_ctx* c, hx_to_stats* s) {
    if (!c) return HX_ERR_BAD_ARG;
    if (!c->to_ready) return fail(c, HX_ERR_STATE, "hx_to_init first");
    const hx_to_params& p = c->tp;
    const long ne = c->g.ne, n = ndof(c);
    struct Events {  // back to the pool on every return
        hx_ctx* c;
        hipEvent_t a, b;
        ~Events() {
            c->free_events.push_back(a);
            c->free_events.push_back(b);
        }
    } evs{c, get_event(c), get_event(c)};
    hipEvent_t ev0 = evs.a, ev1 = evs.b;
    (void)hipEventRecord(ev0, c->stream);
    int rc;
    // 1. SIMP moduli of the physical densities
    launch_simp_modulus(ne, c->t_xphys.p, p.penal, p.E0, p.Emin, c->E.p, c->stream);
    // 2. elasticity solve, warm-started from the previous displacement
    hx_pcg_opts po{p.pcg_tol, p.pcg_max_iter, 0.0, p.precision, p.precond, 1, HX_MV_AUTO, 0};
    hx_pcg_stats ps{};
    if (p.precision == HX_FP32_IR)
        if ((rc = ensure_fp32(c))) return rc;
    SolveVecs tv = single_vecs(c);
    tv.xbest = nullptr;  // no best-iterate copies: a failed solve ends the step anyway
    const int prc = run_pcg(c, po, &ps, tv);
    if (prc != HX_OK) {
        // the densities stay as they are: a failed solve (its last iterate is
        // left in the displacement) gives no sensitivities to update them with
        (void)hipStreamSynchronize(c->stream);
        if (s) {
            *s = hx_to_stats{};
            s->iter = c->to_iter;
            s->pcg_iterations = ps.iterations;
            s->pcg_rel_res = ps.rel_res;
            s->pcg_status = ps.status;
        }
        if (prc == HX_ERR_HIP || prc == HX_ERR_OOM) return prc;
        return fail(c, prc, "TO step: the elasticity solve failed (" + std::to_string(prc) +
                                "); densities not updated");
    }
    // 3. compliance, sensitivities, von Mises
    launch_dot(n, c->f.p, c->x.p, c->small.p, &c->st->ticket[0], c->partials.p, c->stream);
    ElemArgs a{};
    a.u = c->x.p;
    a.xphys = c->t_xphys.p;
    a.penal = p.penal;
    a.E0 = p.E0;
    a.Emin = p.Emin;
    a.dc = c->t_dc.p;
    if (p.with_stress) {
        a.Emod = c->E.p;
        a.vm = c->t_vm.p;
        a.vm_max_partials = c->e_aux.p;
        HXA(c->e_aux, std::max<long>(ne, element_post_blocks(c->g)));
        a.vm_max_partials = c->e_aux.p;
        material_consts(c->nu, &a.lam1, &a.lam2, &a.mu);
    }
    elem_post(c, a);
    if (p.with_stress) launch_max_reduce(ne, c->e_aux.p, element_post_blocks(c->g), c->small.p + 1, c->stream);
    HXK();
    // 4. sensitivity filter (the PDE filter is symmetric: dc_x = H dc_xphys)
    const double R = p.rmin / (2.0 * std::sqrt(3.0));
    hx_filter_opts fo{0, p.filter_tol, p.filter_max_iter > 0 ? p.filter_max_iter : 1000, 1};
    hx_filter_stats f1{}, f2{};
    if ((rc = helm_solve(c, R, fo, c->t_dc.p, c->t_dcf.p, &f1))) return rc;
    // 5. OC (volume of the mass-preserving filter equals that of x_new)
    // the multiplier moves little between TO iterations: the search starts around the last one
    if ((rc = oc_run(c, ne, c->t_x.p, c->t_dcf.p, c->t_dv.p, p.volfrac, p.move, p.xmin, p.oc_rel_tol, c->t_xnew.p,
                     true, c->oc_lambda)))
        return rc;
    c->oc_lambda = c->st_host->scal[2];
    const double change = c->st_host->scal[3];
    // 6. density filter
    if ((rc = helm_solve(c, R, fo, c->t_xnew.p, c->t_xphys.p, &f2))) return rc;
    std::swap(c->t_x, c->t_xnew);
    launch_dot(ne, c->t_xphys.p, c->t_ones.p, c->small.p + 2, &c->st->ticket[0], c->partials.p, c->stream);
    HXK();
    (void)hipEventRecord(ev1, c->stream);
    double sm[3] = {0, 0, 0};
    if ((rc = download(c, sm, c->small.p, 3 * sizeof(double)))) return rc;
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ev0, ev1);
    c->to_iter += 1;
    if (s) {
        s->iter = c->to_iter;
        s->compliance = sm[0];
        s->change = change;
        s->volume = sm[2] / (double)ne;
        s->pcg_iterations = ps.iterations;
        s->pcg_rel_res = ps.rel_res;
        s->pcg_status = ps.status;
        s->filter_iterations = f1.iterations + f2.iterations;
        s->max_vm = p.with_stress ? sm[1] : 0.0;
        s->seconds = ms * 1e-3;
    }
    return HX_OK;
}

int hx_to_read(hx_ctx* c, int field, double* out) {
    if (!c || !out) return HX_ERR_BAD_ARG;
    if (!c->to_ready) return fail(c, HX_ERR_STATE, "hx_to_init first");
    const long ne = c->g.ne;
    switch (field) {
        case HX_FIELD_X: return download(c, out, c->t_x.p, ne * sizeof(double));
        case HX_FIELD_XPHYS: return download(c, out, c->t_xphys.p, ne * sizeof(double));
        case HX_FIELD_U: return download(c, out, c->x.p, ndof(c) * sizeof(double));
        case HX_FIELD_VM: return download(c, out, c->t_vm.p, ne * sizeof(double));
        case HX_FIELD_DC: return download(c, out, c->t_dcf.p, ne * sizeof(double));
        default: return fail(c, HX_ERR_BAD_ARG, "unknown field");
    }
}

// Random training moduli: uniform densities smoothed by the PDE filter of a
// random radius in [1, 3], stretched to [0, 1], solid above a random
// threshold in [0.2, 0.7] and graded to 0.01 below it, SIMP p = 3.
static int random_moduli(hx_ctx* c, std::mt19937_64& rng, std::vector<double>& hrho, std::vector<double>& hE) {
    std::uniform_real_distribution<double> U(0.0, 1.0);
    const long ne = c->g.ne;
    int rc;
    for (long e = 0; e < ne; ++e) hrho[e] = U(rng);
    if ((rc = upload(c, c->l_rho.p, hrho.data(), ne * sizeof(double)))) return rc;
    hx_filter_opts fo{0, 1e-6, 500, 1};
    if ((rc = helm_solve(c, 1.0 + 2.0 * U(rng), fo, c->l_rho.p, c->l_rho.p, nullptr))) return rc;
    HXC(hipMemcpy(hrho.data(), c->l_rho.p, ne * sizeof(double), hipMemcpyDeviceToHost));
    double lo = 1e300, hi = -1e300;
    for (long e = 0; e < ne; ++e) lo = std::min(lo, hrho[e]), hi = std::max(hi, hrho[e]);
    const double thr = 0.2 + 0.5 * U(rng);
    for (long e = 0; e < ne; ++e) {
        double t = hi > lo ? (hrho[e] - lo) / (hi - lo) : 0.5;
        t = t > thr ? 1.0 : std::max(0.01, t / thr);
        hE[e] = 1e-9 + t * t * t;
    }
    return upload(c, c->E.p, hE.data(), ne * sizeof(double));
}

// Streaming training on manufactured solutions (hx_learn_stats::manufactured
// = m): per density field kManSolutions smooth random displacements e*, load
// b = A e* on the solved DOFs, Jacobi-PCG from zero, and at the first m of the
// iterations k = 0, 2, 8, 32 the pair (r_k = b - A x_k, e_k = e* - x_k) --
// A e_k = r_k exactly, so no reference solve is needed -- takes one Adam step
// on the subspace objective right away.
static int learned_train_manufactured(hx_ctx* c, hx_learn_stats* s) {
    constexpr int kManSolutions = 4;
    const int ks[] = {0, 2, 8, 32};
    const int nk = std::min(4, std::max(1, s->manufactured));
    const long n = ndof(c), ne = c->g.ne;
    HXA(c->l_man, 7 * n);
    HXA(c->l_rho, ne);
    HXA(c->l_dots, 8 + 2 * kFcgKeep);
    double* estar = c->l_man.p;
    double *ek = estar + n, *rk = ek + n, *u = rk + n, *Au = u + n, *y = Au + n, *Ay = y + n;
    std::vector<double> hE(ne), hrho(ne);
    int levels = 1;
    while ((2 << levels) <= std::max({c->g.nx, c->g.ny, c->g.nz})) ++levels;  // coarsest lattice ~ half the grid
    levels = std::max(1, levels - 1);
    const int per_field = kManSolutions * nk;
    const int nfield = std::max(1, (s->samples + per_field - 1) / per_field);
    int rc;
    double last = 0.0, lj = 0.0;
    long used = 0;
    for (int ep = 0; ep < s->epochs; ++ep) {
        std::mt19937_64 rng((uint64_t)s->seed * 7919u + 29u);  // the same stream every pass
        double acc = 0.0, accj = 0.0;
        long cnt = 0;
        for (int fi = 0; fi < nfield; ++fi) {
            if ((rc = random_moduli(c, rng, hrho, hE))) return rc;
            launch_elastic_minv<double>(c->g, c->kd, c->E.p, c->has_fixed ? c->fixed.p : nullptr, 0.0, c->minv.p,
                                        c->stream);
            for (int si = 0; si < kManSolutions; ++si) {
                const unsigned long long seed = rng();
                launch_smooth_field(c->g, seed, levels, c->minv.p, estar, c->stream);
                mv64(c, estar, c->f.p, c->E.p, 0.0, MV_PLAIN, HX_MV_AUTO, 1);
                launch_mask<double>(n, c->minv.p, c->f.p, c->stream, 1);  // b = A e* on the solved DOFs
                for (int kk = 0; kk < nk; ++kk) {
                    if (ks[kk] == 0) {
                        launch_fill(n, 0.0, c->x.p, c->stream);
                    } else {
                        hx_pcg_opts it{1e-30, ks[kk], 0.0, HX_FP64, HX_PC_JACOBI, 0, HX_MV_AUTO, 0};
                        rc = pcg_fp64(c, it, nullptr, single_vecs(c));
                        if (rc == HX_ERR_HIP || rc == HX_ERR_OOM) return rc;
                    }
                    mv64(c, c->x.p, c->q.p, c->E.p, 0.0, MV_PLAIN, HX_MV_AUTO, 1);
                    launch_true_residual(n, c->f.p, c->q.p, c->x.p, c->minv.p, 0.0, rk, nullptr, c->st, c->partials.p,
                                         1, c->stream, 1);
                    launch_masked_diff(n, estar, c->x.p, c->minv.p, ek, c->stream);
                    launch_mul(n, c->minv.p, rk, u, c->stream);  // u = D^-1 r
                    mv64(c, u, Au, c->E.p, 0.0, MV_PLAIN, HX_MV_AUTO, 1);
                    double h[4];
                    if ((rc = dots_host(c, n, {{rk, rk}, {ek, rk}, {u, Au}, {u, rk}}, h))) return rc;
                    const double rr = h[0], er = h[1], uAu = h[2], ur = h[3];
                    if (!(rr > 0.0) || !(er > 0.0) || !(uAu > 0.0)) continue;
                    unet_forward_sample(c->unet, rk, c->E.p, 1.0 / std::sqrt(rr / (double)n), c->stream);
                    unet_direction(c->unet, c->minv.p, 1.0, y, c->stream);
                    mv64(c, y, Ay, c->E.p, 0.0, MV_PLAIN, HX_MV_AUTO, 1);
                    double g[3];
                    if ((rc = dots_host(c, n, {{u, Ay}, {y, Ay}, {y, rk}}, g))) return rc;
                    const double uAy = g[0], yAy = g[1], yr = g[2];
                    const double jl = 1.0 - ur * ur / (uAu * er);
                    accj += jl;
                    ++cnt;
                    const double det = uAu * yAy - uAy * uAy;
                    if (!(det > 1e-12 * uAu * yAy)) {
                        acc += jl;
                        continue;
                    }
                    const double a = (ur * yAy - yr * uAy) / det, b = (yr * uAu - ur * uAy) / det;
                    acc += 1.0 - (a * ur + b * yr) / er;
                    unet_subspace_grad(c->unet, rk, Au, Ay, c->minv.p, a, b, -2.0 * b / er, c->stream);
                    unet_backward_step(c->unet, (float)s->lr, c->stream);
                }
            }
        }
        if (cnt == 0) return fail(c, HX_ERR_STATE, "no usable training samples");
        last = acc / (double)cnt;
        lj = accj / (double)cnt;
        used = cnt;
    }
    s->samples = (int)used;
    s->final_loss = last;
    s->jacobi_loss = lj;
    return HX_OK;
}

int hx_learned_train(hx_ctx* c, const double* Ee, const double* f, hx_learn_stats* s) {
    if (!c || !s) return HX_ERR_BAD_ARG;
    if (int rc = ensure_grid(c)) return rc;
    if (s->samples < 1 || s->epochs < 1 || !(s->lr > 0)) return fail(c, HX_ERR_BAD_ARG, "invalid training options");
    hipEvent_t ev0 = get_event(c), ev1 = get_event(c);
    (void)hipEventRecord(ev0, c->stream);
    if (!c->unet) c->unet = unet_create();
    if (!unet_bind(c->unet, c->g, (unsigned)s->seed, c->stream)) return fail(c, HX_ERR_OOM, "U-Net buffers");
    if (s->manufactured) {
        const int rc = learned_train_manufactured(c, s);
        (void)hipEventRecord(ev1, c->stream);
        (void)hipEventSynchronize(ev1);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, ev0, ev1);
        c->free_events.push_back(ev0);
        c->free_events.push_back(ev1);
        if (rc) return rc;
        c->unet_trained = true;
        s->seconds = ms * 1e-3;
        return HX_OK;
    }
    const long n = ndof(c), ne = c->g.ne;
    // residual/error pairs at these Jacobi-PCG iteration counts of each problem
    const int ks[] = {0, 2, 8, 32, 128, 512};
    const int nk = 6;
    const int nprob = std::max(1, (s->samples + nk - 1) / nk);
    const int nsamp = nprob * nk;
    const long stride = 4 * n + ne;  // r, e, minv, E, A D^-1 r
    HXA(c->l_samples, (size_t)nsamp * stride);
    HXA(c->l_ustar, n);
    HXA(c->l_rho, ne);
    std::mt19937_64 rng((uint64_t)s->seed * 7919u + 17u);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    std::vector<double> hE(ne), hf(n), hrho(ne);
    const int NX = c->g.NX, NY = c->g.NY, NZ = c->g.NZ;
    std::vector<uint8_t> hfixed(n, 0);
    if (c->has_fixed) HXC(hipMemcpy(hfixed.data(), c->fixed.p, n, hipMemcpyDeviceToHost));
    int rc;
    hx_pcg_opts solve{1e-12, 200000, 0.0, HX_FP64, HX_PC_JACOBI, 0, HX_MV_AUTO, 0};
    for (int pi = 0; pi < nprob; ++pi) {
        // problem: moduli and loads
        if (pi == 0 && Ee) {
            std::memcpy(hE.data(), Ee, ne * sizeof(double));
            if ((rc = upload(c, c->E.p, hE.data(), ne * sizeof(double)))) return rc;
        } else if ((rc = random_moduli(c, rng, hrho, hE))) {
            return rc;
        }
        if (pi == 0 && f) {
            std::memcpy(hf.data(), f, n * sizeof(double));
        } else {
            std::fill(hf.begin(), hf.end(), 0.0);
            const int nl = 1 + (int)(U(rng) * 4);
            for (int l = 0; l < nl; ++l) {  // unit loads at random free nodes of the x = max face
                for (int tries = 0; tries < 100; ++tries) {
                    const int j = (int)(U(rng) * NY), k = (int)(U(rng) * NZ);
                    const long nd = (long)(NX - 1) + (long)NX * (j + (long)NY * k);
                    if (hfixed[3 * nd] && hfixed[3 * nd + 1] && hfixed[3 * nd + 2]) continue;
                    double d[3] = {U(rng) - 0.5, U(rng) - 0.5, U(rng) - 0.5};
                    const double nrm = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) + 1e-12;
                    for (int q = 0; q < 3; ++q) hf[3 * nd + q] += d[q] / nrm;
                    break;
                }
            }
        }
        if ((rc = upload(c, c->f.p, hf.data(), n * sizeof(double)))) return rc;
        hx_pcg_stats ps{};
        rc = pcg_fp64(c, solve, &ps, single_vecs(c));
        if (rc == HX_ERR_HIP || rc == HX_ERR_OOM) return rc;
        HXC(hipMemcpyAsync(c->l_ustar.p, c->x.p, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        for (int kk = 0; kk < nk; ++kk) {
            double* base = c->l_samples.p + (long)(pi * nk + kk) * stride;
            double *sr = base, *se = base + n, *sm = base + 2 * n, *sE = base + 3 * n;
            if (ks[kk] == 0) {
                launch_fill(n, 0.0, c->x.p, c->stream);
            } else {
                hx_pcg_opts it{1e-30, ks[kk], 0.0, HX_FP64, HX_PC_JACOBI, 0, HX_MV_AUTO, 0};
                rc = pcg_fp64(c, it, nullptr, single_vecs(c));
                if (rc == HX_ERR_HIP || rc == HX_ERR_OOM) return rc;
            }
            mv64(c, c->x.p, c->q.p, c->E.p, 0.0, MV_PLAIN, HX_MV_AUTO, 1);
            launch_elastic_minv<double>(c->g, c->kd, c->E.p, c->has_fixed ? c->fixed.p : nullptr, 0.0, sm,
                                        c->stream);
            launch_true_residual(n, c->f.p, c->q.p, c->x.p, sm, 0.0, sr, nullptr, c->st, c->partials.p, 1,
                                 c->stream, 1);
            launch_masked_diff(n, c->l_ustar.p, c->x.p, sm, se, c->stream);
            HXC(hipMemcpyAsync(sE, c->E.p, ne * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
            launch_mul(n, sm, sr, c->q.p, c->stream);  // u = D^-1 r
            mv64(c, c->q.p, sE + ne, c->E.p, 0.0, MV_PLAIN, HX_MV_AUTO, 1);  // A u
            HXK();
        }
    }
    HXC(hipStreamSynchronize(c->stream));
    c->l_nsamples = nsamp;
    c->l_sample_stride = stride;
    // per-sample scalars: |r|^2, e'r (= |e|_A^2 since A e = r), and for
    // u = D^-1 r: u'Au, u'r
    DevBuf<double> nrm;
    HXA(nrm, 8);
    auto dots = [&](std::initializer_list<std::pair<const double*, const double*>> pr, double* out) -> int {
        int k = 0;
        for (const auto& q : pr) launch_dot(n, q.first, q.second, nrm.p + k++, nullptr, c->partials.p, c->stream);
        HXC(hipMemcpyAsync(out, nrm.p, k * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HXC(hipStreamSynchronize(c->stream));
        return HX_OK;
    };
    HXA(c->l_z, n);
    HXA(c->l_zold, n);
    std::vector<double> rr(nsamp), er(nsamp), uAu(nsamp), ur(nsamp);
    std::vector<int> order;
    double lj = 0.0;
    for (int i = 0; i < nsamp; ++i) {
        const double* base = c->l_samples.p + (long)i * stride;
        const double *sr = base, *se = base + n, *sm = base + 2 * n, *sAu = base + 3 * n + ne;
        launch_mul(n, sm, sr, c->l_z.p, c->stream);
        double h[4];
        if ((rc = dots({{sr, sr}, {se, sr}, {c->l_z.p, sAu}, {c->l_z.p, sr}}, h))) return rc;
        rr[i] = h[0], er[i] = h[1], uAu[i] = h[2], ur[i] = h[3];
        // keep samples whose residual is well above the reference solution's
        // (relative 1e-12): below ~1e-6 of the problem's first residual, e is
        // dominated by that solution's own error and A e = r no longer holds
        const double rr0 = rr[i - i % nk];
        if (rr[i] > 1e-12 * rr0 && er[i] > 0.0 && uAu[i] > 0.0) {
            order.push_back(i);
            lj += 1.0 - ur[i] * ur[i] / (uAu[i] * er[i]);
        }
    }
    if (order.empty()) return fail(c, HX_ERR_STATE, "no usable training samples");
    // Adam, batch size 1, seeded order, on the subspace objective
    //   L = 1 - max_{a,b} [2(a u'r + b y'r) - |a u + b y|_A^2] / e'r
    // (the energy left outside span{D^-1 r, y}; scale-invariant in y)
    double last = 0.0;
    for (int ep = 0; ep < s->epochs; ++ep) {
        std::shuffle(order.begin(), order.end(), rng);
        double acc = 0.0;
        for (int i : order) {
            double* base = c->l_samples.p + (long)i * stride;
            double *sr = base, *sm = base + 2 * n, *sE = base + 3 * n, *sAu = base + 3 * n + ne;
            const double inv = 1.0 / std::sqrt(rr[i] / (double)n);
            unet_forward_sample(c->unet, sr, sE, inv, c->stream);
            unet_direction(c->unet, sm, 1.0, c->l_z.p, c->stream);           // y
            mv64(c, c->l_z.p, c->l_zold.p, sE, 0.0, MV_PLAIN, HX_MV_AUTO, 1);  // A y
            double h[3];
            if ((rc = dots({{sAu, c->l_z.p}, {c->l_z.p, c->l_zold.p}, {c->l_z.p, sr}}, h))) return rc;
            const double uAy = h[0], yAy = h[1], yr = h[2];
            const double det = uAu[i] * yAy - uAy * uAy;
            if (!(det > 1e-12 * uAu[i] * yAy)) {  // y (almost) along D^-1 r: no usable gradient
                acc += 1.0 - ur[i] * ur[i] / (uAu[i] * er[i]);
                continue;
            }
            const double a = (ur[i] * yAy - yr * uAy) / det, b = (yr * uAu[i] - ur[i] * uAy) / det;
            acc += 1.0 - (a * ur[i] + b * yr) / er[i];
            unet_subspace_grad(c->unet, sr, sAu, c->l_zold.p, sm, a, b, -2.0 * b / er[i], c->stream);
            unet_backward_step(c->unet, (float)s->lr, c->stream);
        }
        last = acc / (double)order.size();
    }
    s->jacobi_loss = lj / (double)order.size();
    HXK();
    (void)hipEventRecord(ev1, c->stream);
    (void)hipEventSynchronize(ev1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ev0, ev1);
    c->free_events.push_back(ev0);
    c->free_events.push_back(ev1);
    c->unet_trained = true;
    s->samples = nsamp;
    s->final_loss = last;
    s->seconds = ms * 1e-3;
    return HX_OK;
}

int hx_learned_weights(hx_ctx* c, uint16_t* out, int* count) {
    if (!c || !count) return HX_ERR_BAD_ARG;
    if (!c->unet) return fail(c, HX_ERR_STATE, "no network (hx_learned_train first)");
    *count = unet_weights_bf16(c->unet, out, c->stream);
    return HX_OK;
}

int hx_test_unet_nparam(hx_ctx* c, long* nparam) {
    if (!c || !nparam) return HX_ERR_BAD_ARG;
    if (!c->unet) c->unet = unet_create();
    *nparam = unet_nparam(c->unet);
    return HX_OK;
}

int hx_test_unet_init(hx_ctx* c, int seed, float* params) {
    if (!c) return HX_ERR_BAD_ARG;
    if (int rc = ensure_grid(c)) return rc;
    if (!c->unet) c->unet = unet_create();
    if (!unet_bind(c->unet, c->g, (unsigned)seed, c->stream)) return fail(c, HX_ERR_OOM, "U-Net buffers");
    unet_init(c->unet, (unsigned)seed, c->stream);
    if (params) unet_params_f32(c->unet, params, c->stream);
    HXK();
    return HX_OK;
}

int hx_test_unet_forward(hx_ctx* c, const uint16_t* x0, float* y) {
    if (!c || !x0 || !y) return HX_ERR_BAD_ARG;
    if (!c->unet || !unet_ready(c->unet)) return fail(c, HX_ERR_STATE, "hx_test_unet_init first");
    if (!unet_bind(c->unet, c->g, 0u, c->stream)) return fail(c, HX_ERR_OOM, "U-Net buffers");
    unet_forward_packed(c->unet, x0, y, c->stream);
    HXK();
    return HX_OK;
}

int hx_test_unet_backward(hx_ctx* c, const float* gy, float* grad) {
    if (!c || !gy || !grad) return HX_ERR_BAD_ARG;
    if (!c->unet || !unet_ready(c->unet)) return fail(c, HX_ERR_STATE, "hx_test_unet_init first");
    unet_backward_host(c->unet, gy, grad, c->stream);
    HXK();
    return HX_OK;
}

// Time every launch shape of the operator as the PCG runs it (fused
// p-update, p'q partials, alpha finaliser) on the current grid with uniform
// moduli and a non-smooth vector, and keep the fastest per precision:
// FP64 parity {16x16, 16x32 tiles} x {8, 16, 32 z planes} x {1, 2 planes in
// flight} and dense; FP32 the same parity shapes, dense and MFMA.
extern "C++" {
template <typename T>
int time_smoothing_t(hx_ctx* c, int form, int reps, const MvShape* shapes, int nshapes, float* ms);
}

int hx_autotune(hx_ctx* c, int nrhs, hx_tune_result* out) {
    if (!c || nrhs < 1 || nrhs > kMaxRhs) return HX_ERR_BAD_ARG;
    if (int rc = ensure_grid(c)) return rc;
    const long n = ndof(c);
    if (int rc = ensure_fp32(c)) return rc;
    DevBuf<double> tx, ty, tr, tm, tp;
    DevBuf<float> fx, fy, fr, fm, fp;
    HXA(tx, n * nrhs);
    HXA(ty, n * nrhs);
    HXA(tr, n * nrhs);
    HXA(tm, n);
    HXA(tp, n * nrhs);
    HXA(fx, n * nrhs);
    HXA(fy, n * nrhs);
    HXA(fr, n * nrhs);
    HXA(fm, n);
    HXA(fp, n * nrhs);
    {
        std::vector<double> h(n);
        uint32_t z = 0x9e3779b9u;
        for (long i = 0; i < n; ++i) {
            z = z * 1664525u + 1013904223u;
            h[i] = (double)(z >> 8) * (1.0 / 16777216.0) - 0.5;
        }
        for (int j = 0; j < nrhs; ++j) {
            HXC(hipMemcpyAsync(tx.p + j * n, h.data(), n * sizeof(double), hipMemcpyHostToDevice, c->stream));
            HXC(hipMemcpyAsync(tr.p + j * n, h.data(), n * sizeof(double), hipMemcpyHostToDevice, c->stream));
        }
        HXC(hipStreamSynchronize(c->stream));
    }
    launch_fill(n, 1.0, tm.p, c->stream);
    launch_convert<double, float>(n * nrhs, tx.p, fx.p, c->stream);
    launch_convert<double, float>(n * nrhs, tr.p, fr.p, c->stream);
    launch_fill32(n, 1.0f, fm.p, c->stream);
    launch_fill(c->g.ne, 0.5, c->E.p, c->stream);
    launch_convert<double, float>(c->g.ne, c->E.p, c->E32.p, c->stream);
    const bool was_timing = c->timing;
    c->timing = false;
    hipEvent_t a = get_event(c), b = get_event(c);
    auto time_it = [&](auto&& launch) {
        launch_pcg_reset(c->st, 0.0, 0.0, 1 << 30, 1, c->stream, nrhs);
        for (int w = 0; w < 2; ++w) launch();
        (void)hipEventRecord(a, c->stream);
        const int reps = 5;
        for (int r = 0; r < reps; ++r) launch();
        (void)hipEventRecord(b, c->stream);
        (void)hipEventSynchronize(b);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, a, b);
        return (double)ms / reps;
    };
    auto run64 = [&](int v) {
        return time_it([&] { mv64(c, tx.p, ty.p, c->E.p, 0.0, MV_PCG, v, nrhs, tr.p, tm.p, tp.p); });
    };
    auto run32 = [&](int v) {
        return time_it([&] { mv32(c, fx.p, fy.p, c->E32.p, 0.0, MV_PCG, v, nrhs, fr.p, fm.p, fp.p); });
    };
    // chunks: 12 planes is the fastest FP32 form at 256x128x128 (fewer
    // rounds of workgroups over the chip than 8, shorter marches than 16)
    const int tiles[] = {0, 1, 2}, chunks[] = {8, 12, 16, 32}, depths[] = {1, 2};
    double best64 = 1e30, best32 = 1e30;
    MvShape b64 = c->shape64, b32 = c->shape32;
    int bv64 = HX_MV_PARITY, bv32 = HX_MV_PARITY;
    for (int t : tiles)
        for (int k : chunks)
            for (int d : depths) {
                c->shape64 = MvShape{t, k, d};
                c->shape32 = MvShape{t, k, d};
                const double m64 = run64(HX_MV_PARITY);
                if (m64 < best64) best64 = m64, b64 = c->shape64;
                const double m32 = run32(HX_MV_PARITY);
                if (m32 < best32) best32 = m32, b32 = c->shape32;
            }
    c->shape64 = b64;
    c->shape32 = b32;
    {  // the gather kernel has no fused p-update, so it is not a PCG candidate
        const double ms = run64(HX_MV_DENSE);
        if (ms < best64) best64 = ms, bv64 = HX_MV_DENSE;
    }
    double t32[3] = {best32, 0, 0};
    const int vv[3] = {HX_MV_PARITY, HX_MV_DENSE, HX_MV_MFMA};
    for (int i = 1; i < 3; ++i) {
        t32[i] = run32(vv[i]);
        if (t32[i] < best32) best32 = t32[i], bv32 = vv[i];
    }
    c->free_events.push_back(a);
    c->free_events.push_back(b);
    c->mv_variant = bv64;
    c->mv32_variant = bv32;
    // the V-cycle's fine-level smoothing operators (FP32, the production form)
    double best_pre = 0.0, best_post = 0.0;
    c->shape_pre = c->shape_post = MvShape{0, 0, 1};
    if (ml_levels_for(c->g, 1) >= 1) {
        std::vector<MvShape> cands;
        for (int t : tiles)
            for (int k : {8, 12, 16})
                for (int d : depths) cands.push_back(MvShape{t, k, d});
        std::vector<float> tp(cands.size()), tq(cands.size());
        const int reps = 5;
        int rs = time_smoothing_t<float>(c, 3, reps, cands.data(), (int)cands.size(), tp.data());
        if (rs == HX_OK) rs = time_smoothing_t<float>(c, 4, reps, cands.data(), (int)cands.size(), tq.data());
        for (size_t i = 0; rs == HX_OK && i < cands.size(); ++i) {
            const double mp = tp[i] / reps, mq = tq[i] / reps;
            if (mp > 0.0 && (best_pre == 0.0 || mp < best_pre)) best_pre = mp, c->shape_pre = cands[i];
            if (mq > 0.0 && (best_post == 0.0 || mq < best_post)) best_post = mq, c->shape_post = cands[i];
        }
    }
    c->timing = was_timing;
    clear_graphs(c);
    launch_pcg_reset(c->st, 0.0, 0.0, 0, 1, c->stream, nrhs);
    HXK();
    HXC(hipStreamSynchronize(c->stream));
    if (out) {
        out->tile = c->shape64.tile;
        out->kchunk = c->shape64.kchunk;
        out->variant64 = c->mv_variant;
        out->variant32 = c->mv32_variant;
        out->ms64 = best64;
        out->ms32 = best32;
        out->ms32_parity = t32[0];
        out->ms32_dense = t32[1];
        out->ms32_mfma = t32[2];
        out->tile32 = c->shape32.tile;
        out->kchunk32 = c->shape32.kchunk;
        out->pdepth = c->shape64.pdepth;
        out->pdepth32 = c->shape32.pdepth;
        out->tile_pre = c->shape_pre.tile;
        out->kchunk_pre = c->shape_pre.kchunk;
        out->pdepth_pre = c->shape_pre.pdepth;
        out->tile_post = c->shape_post.tile;
        out->kchunk_post = c->shape_post.kchunk;
        out->pdepth_post = c->shape_post.pdepth;
        out->ms_pre = best_pre;
        out->ms_post = best_post;
    }
    return HX_OK;
}

// Operator timing: `reps` back-to-back launches between two HIP events, each
// on its own set of vectors (kSets sets, > 2x the 256 MB last-level cache
// together at production sizes), so a launch does not find its inputs left
// in cache by the previous one -- as in the PCG, where the vector update
// streams everything between two operator applications.
extern "C++" {
template <typename T>
int time_operator_t(hx_ctx* c, int nrhs, int reps, float* ms, bool zform) {
    constexpr int kSets = 4;
    const long n = ndof(c) * nrhs;
    DevBuf<T> x[kSets], y[kSets], r[kSets], m[kSets], p[kSets];
    for (int k = 0; k < kSets; ++k) {
        HXA(x[k], n);
        HXA(y[k], n);
        HXA(r[k], n);
        HXA(m[k], ndof(c));
        HXA(p[k], n);
        if constexpr (sizeof(T) == 8) {
            launch_fill(n, 0.5, x[k].p, c->stream);
            launch_fill(n, 0.25, r[k].p, c->stream);
            launch_fill(ndof(c), 1.0, m[k].p, c->stream);
        } else {
            launch_fill32(n, 0.5f, x[k].p, c->stream);
            launch_fill32(n, 0.25f, r[k].p, c->stream);
            launch_fill32(ndof(c), 1.0f, m[k].p, c->stream);
        }
    }
    // the PCG-mode kernel without its p'q partial write: same loads and
    // stores per node (p_old, r, D^-1, E in; q, p out), fused p-update on;
    // zform: the explicit-preconditioner form (p_old, z, E in), parity contraction
    auto go = [&](int k) {
        const T* mk = zform ? nullptr : m[k].p;
        if constexpr (sizeof(T) == 8)
            mv64(c, x[k].p, y[k].p, c->E.p, 0.0, MV_PLAIN, zform ? HX_MV_PARITY : HX_MV_AUTO, nrhs, r[k].p, mk, p[k].p);
        else
            mv32(c, x[k].p, y[k].p, c->E32.p, 0.0, MV_PLAIN, zform ? HX_MV_PARITY : c->mv32_variant, nrhs, r[k].p, mk,
                 p[k].p);
    };
    hipEvent_t a = get_event(c), b = get_event(c);
    for (int k = 0; k < kSets; ++k) go(k);
    (void)hipEventRecord(a, c->stream);
    for (int i = 0; i < reps; ++i) go(i % kSets);
    (void)hipEventRecord(b, c->stream);
    HXC(hipEventSynchronize(b));
    (void)hipEventElapsedTime(ms, a, b);
    c->free_events.push_back(a);
    c->free_events.push_back(b);
    return HX_OK;
}

// back-to-back launches of the fine level's V-cycle smoothing operator (form
// 3 or 4) on kSets vector sets, at each of nshapes launch shapes (ms[i]: the
// total of `reps` launches at shapes[i], 0 for a shape whose partial sums
// would not fit); form 4 reads a coarse correction on the first coarsening's
// node grid and writes its r'z partials, as in the solve
template <typename T>
int time_smoothing_t(hx_ctx* c, int form, int reps, const MvShape* shapes, int nshapes, float* ms) {
    constexpr int kSets = 4;
    const long n = ndof(c);
    const GridArgs gc = ml_coarse_grid(c->g);
    DevBuf<T> r[kSets], s[kSets], y[kSets], e[kSets];
    for (int k = 0; k < kSets; ++k) {
        HXA(r[k], n);
        HXA(s[k], n);
        HXA(y[k], n);
        HXA(e[k], 3 * gc.nn);
        if constexpr (sizeof(T) == 8) {
            launch_fill(n, 0.25, r[k].p, c->stream);
            launch_fill(n, 0.5, s[k].p, c->stream);
            launch_fill(3 * gc.nn, 0.125, e[k].p, c->stream);
        } else {
            launch_fill32(n, 0.25f, r[k].p, c->stream);
            launch_fill32(n, 0.5f, s[k].p, c->stream);
            launch_fill32(3 * gc.nn, 0.125f, e[k].p, c->stream);
        }
    }
    hipEvent_t a = get_event(c), b = get_event(c);
    for (int i = 0; i < nshapes; ++i) {
        const MvShape sh = shapes[i];
        ms[i] = 0.f;
        const int kc = form == 4 ? smoothing_post_kchunk(sh.tile, sh.kchunk, (int)sizeof(T)) : sh.kchunk;
        if (2L * kVecBlocks + matvec_parity_blocks(c->g, sh.tile, kc) > kOpPartials) continue;
        auto go = [&](int k) {
            MvArgs<T> a{nullptr, y[k].p, nullptr, T(0), MV_PLAIN, nullptr,
                        form == 4 ? c->partials.p + 2 * kVecBlocks : nullptr, n, r[k].p, s[k].p, nullptr};
            a.form = form;
            a.ec = e[k].p;
            a.cNX = gc.NX;
            a.cNY = gc.NY;
            a.cNZ = gc.NZ;
            if constexpr (sizeof(T) == 8) {
                a.E = c->E.p;
                launch_matvec_parity<double>(c->g, c->kp64, a, sh, c->stream, 1);
            } else {
                a.E = c->E32.p;
                launch_matvec_parity<float>(c->g, c->kp32, a, sh, c->stream, 1);
            }
        };
        for (int k = 0; k < kSets; ++k) go(k);
        (void)hipEventRecord(a, c->stream);
        for (int it = 0; it < reps; ++it) go(it % kSets);
        (void)hipEventRecord(b, c->stream);
        HXC(hipEventSynchronize(b));
        (void)hipEventElapsedTime(&ms[i], a, b);
    }
    c->free_events.push_back(a);
    c->free_events.push_back(b);
    return HX_OK;
}
}  // extern "C++"

int hx_time_smoothing(hx_ctx* c, int precision, int form, int reps, double* avg_ms) {
    if (!c || reps < 1 || !avg_ms || (form != 3 && form != 4)) return HX_ERR_BAD_ARG;
    if (precision != HX_FP64 && precision != HX_FP32_IR) return HX_ERR_BAD_ARG;
    if (int rc = ensure_grid(c)) return rc;
    if (ml_levels_for(c->g, 1) < 1) return fail(c, HX_ERR_BAD_ARG, "the grid has no 2:1 coarsening");
    const MvShape sh = vc_shape(c, precision == HX_FP32_IR, form == 4);
    const int kc = form == 4 ? smoothing_post_kchunk(sh.tile, sh.kchunk, precision == HX_FP64 ? 8 : 4) : sh.kchunk;
    if (2L * kVecBlocks + matvec_parity_blocks(c->g, sh.tile, kc) > kOpPartials)
        return fail(c, HX_ERR_BAD_ARG, "grid too large for the V-cycle's partial sums");
    if (int rc = ensure_fp32(c)) return rc;
    const bool was_timing = c->timing;
    c->timing = false;
    float ms = 0.f;
    const int rc = precision == HX_FP64 ? time_smoothing_t<double>(c, form, reps, &sh, 1, &ms)
                                        : time_smoothing_t<float>(c, form, reps, &sh, 1, &ms);
    c->timing = was_timing;
    if (rc) return rc;
    HXK();
    *avg_ms = (double)ms / reps;
    return HX_OK;
}

int hx_time_operator(hx_ctx* c, int precision, int nrhs, int reps, double* avg_ms) {
    return hx_time_operator_pc(c, precision, HX_PC_JACOBI, nrhs, reps, avg_ms);
}

int hx_time_operator_pc(hx_ctx* c, int precision, int precond, int nrhs, int reps, double* avg_ms) {
    if (!c || nrhs < 1 || nrhs > kMaxRhs || reps < 1 || !avg_ms) return HX_ERR_BAD_ARG;
    if (precision != HX_FP64 && precision != HX_FP32_IR) return HX_ERR_BAD_ARG;
    if (precond != HX_PC_JACOBI && precond != HX_PC_MULTILEVEL && precond != HX_PC_VCYCLE) return HX_ERR_BAD_ARG;
    const bool zform = precond != HX_PC_JACOBI;
    if (int rc = ensure_grid(c)) return rc;
    if (int rc = ensure_fp32(c)) return rc;
    const bool was_timing = c->timing;
    c->timing = false;
    launch_pcg_reset(c->st, 0.0, 0.0, 1 << 30, 1, c->stream, nrhs);
    float ms = 0.f;
    const int rc = precision == HX_FP64 ? time_operator_t<double>(c, nrhs, reps, &ms, zform)
                                        : time_operator_t<float>(c, nrhs, reps, &ms, zform);
    c->timing = was_timing;
    if (rc) return rc;
    HXK();
    *avg_ms = (double)ms / reps;
    return HX_OK;
}

int hx_tune_set(hx_ctx* c, const hx_tune_result* t) {
    if (!c || !t) return HX_ERR_BAD_ARG;
    auto shape_ok = [](int tile, int kchunk, int pd) {
        return tile >= 0 && tile <= 2 && kchunk >= 1 && (pd == 1 || pd == 2);
    };
    auto var_ok = [](int v) { return v == HX_MV_PARITY || v == HX_MV_DENSE || v == HX_MV_GATHER || v == HX_MV_MFMA; };
    if (!shape_ok(t->tile, t->kchunk, t->pdepth) || !shape_ok(t->tile32, t->kchunk32, t->pdepth32) ||
        !var_ok(t->variant64) || !var_ok(t->variant32) || t->variant32 == HX_MV_GATHER || t->variant64 == HX_MV_MFMA)
        return fail(c, HX_ERR_BAD_ARG, "invalid tuning result");
    // smoothing shapes: kchunk 0 leaves them unset (the operator's shape)
    if ((t->kchunk_pre != 0 && !shape_ok(t->tile_pre, t->kchunk_pre, t->pdepth_pre)) ||
        (t->kchunk_post != 0 && !shape_ok(t->tile_post, t->kchunk_post, t->pdepth_post)))
        return fail(c, HX_ERR_BAD_ARG, "invalid tuning result");
    c->shape64 = MvShape{t->tile, t->kchunk, t->pdepth};
    c->shape32 = MvShape{t->tile32, t->kchunk32, t->pdepth32};
    c->shape_pre = t->kchunk_pre != 0 ? MvShape{t->tile_pre, t->kchunk_pre, t->pdepth_pre} : MvShape{0, 0, 1};
    c->shape_post = t->kchunk_post != 0 ? MvShape{t->tile_post, t->kchunk_post, t->pdepth_post} : MvShape{0, 0, 1};
    c->mv_variant = t->variant64;
    c->mv32_variant = t->variant32;
    clear_graphs(c);
    return HX_OK;
}

int hx_test_device_alloc(hx_ctx* c, long bytes, void** ptr) {
    if (!c || bytes <= 0 || !ptr) return HX_ERR_BAD_ARG;
    HXC(hipSetDevice(c->device));
    HXC(hipMalloc(ptr, (size_t)bytes));
    return HX_OK;
}
int hx_test_device_free(hx_ctx* c, void* ptr) {
    if (!c) return HX_ERR_BAD_ARG;
    HXC(hipStreamSynchronize(c->stream));
    if (ptr) HXC(hipFree(ptr));
    return HX_OK;
}
int hx_test_memcpy(hx_ctx* c, void* dst, const void* src, long bytes, int to_device) {
    if (!c || !dst || !src || bytes < 0) return HX_ERR_BAD_ARG;
    HXC(hipMemcpyAsync(dst, src, (size_t)bytes, to_device ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost,
                       c->stream));
    HXC(hipStreamSynchronize(c->stream));
    return HX_OK;
}

int hx_test_copy(hx_ctx* c, long bytes, int reps, double* avg_ms) {
    if (!c || bytes <= 0 || bytes % 16 || reps < 1 || !avg_ms) return HX_ERR_BAD_ARG;
    DevBuf<uint8_t> a, b;
    HXA(a, bytes);
    HXA(b, bytes);
    HXC(hipMemsetAsync(a.p, 1, bytes, c->stream));
    hipEvent_t e0 = get_event(c), e1 = get_event(c);
    launch_copy16(bytes / 16, a.p, b.p, c->stream);  // warm-up
    (void)hipEventRecord(e0, c->stream);
    for (int r = 0; r < reps; ++r) launch_copy16(bytes / 16, a.p, b.p, c->stream);
    (void)hipEventRecord(e1, c->stream);
    HXC(hipEventSynchronize(e1));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    c->free_events.push_back(e0);
    c->free_events.push_back(e1);
    HXK();
    *avg_ms = (double)ms / reps;
    return HX_OK;
}

namespace {
int test_pc_apply(hx_ctx* c, const double* Ee, const double* r, double* z, int precision, bool vcycle, int* levels,
                  int* vlevels) {
    if (!c || !Ee || !r || !z) return HX_ERR_BAD_ARG;
    if (int rc = ensure_grid(c)) return rc;
    if (precision != HX_FP64 && precision != HX_FP32_IR) return fail(c, HX_ERR_BAD_ARG, "precision");
    const long n = ndof(c);
    int rc;
    if ((rc = upload(c, c->E.p, Ee, c->g.ne * sizeof(double))) || (rc = upload(c, c->r.p, r, n * sizeof(double))))
        return rc;
    launch_elastic_minv<double>(c->g, c->kd, c->E.p, c->has_fixed ? c->fixed.p : nullptr, 0.0, c->minv.p, c->stream);
    launch_pcg_reset(c->st, 1e-8, 0.0, 1, 1, c->stream, 1);
    const bool f32 = precision == HX_FP32_IR;
    if ((rc = f32 ? ensure_fp32(c) : HX_OK)) return rc;
    if (f32) {
        launch_convert<double, float>(c->g.ne, c->E.p, c->E32.p, c->stream);
        launch_convert<double, float>(n, c->minv.p, c->minv32.p, c->stream);
        launch_convert<double, float>(n, c->r.p, c->r32.p, c->stream);
    }
    if ((rc = ml_setup(c, f32, vcycle))) return rc;
    if (f32) {
        if ((rc = pc_apply32(c, vcycle, c->r32.p, 1))) return rc;
        launch_convert<float, double>(n, c->ml_z32.p, c->q.p, c->stream);
    } else {
        if ((rc = pc_apply64(c, vcycle, c->r.p, 1))) return rc;
        HXC(hipMemcpyAsync(c->q.p, c->ml_z.p, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    }
    HXK();
    if (levels) *levels = c->ml64.levels;
    if (vlevels) *vlevels = f32 ? vc_levels(c->ml32) : vc_levels(c->ml64);
    return download(c, z, c->q.p, n * sizeof(double));
}
}  // namespace

int hx_test_ml_apply(hx_ctx* c, const double* Ee, const double* r, double* z, int precision, int* levels) {
    return test_pc_apply(c, Ee, r, z, precision, false, levels, nullptr);
}

int hx_test_vc_apply(hx_ctx* c, const double* Ee, const double* r, double* z, int precision, int* levels,
                     int* vlevels) {
    return test_pc_apply(c, Ee, r, z, precision, true, levels, vlevels);
}

int hx_timing_enable(hx_ctx* c, int on) {
    if (!c) return HX_ERR_BAD_ARG;
    (void)hipStreamSynchronize(c->stream);
    flush_timing(c);
    c->timing = on != 0;
    for (int k = 0; k < HX_K_COUNT; ++k) {
        c->acc_ms[k] = 0.0;
        c->cnt[k] = 0;
    }
    return HX_OK;
}

int hx_timing_read(hx_ctx* c, int kernel, double* avg_ms, long* count) {
    if (!c || kernel < 0 || kernel >= HX_K_COUNT) return HX_ERR_BAD_ARG;
    (void)hipStreamSynchronize(c->stream);
    flush_timing(c);
    if (avg_ms) *avg_ms = c->cnt[kernel] ? c->acc_ms[kernel] / c->cnt[kernel] : 0.0;
    if (count) *count = c->cnt[kernel];
    return HX_OK;
}

int hx_device_info(hx_ctx* c, int* cu_count, int* clock_khz, double* mem_gb) {
    if (!c) return HX_ERR_BAD_ARG;
    hipDeviceProp_t pr;
    if (hipGetDeviceProperties(&pr, c->device) != hipSuccess) return HX_ERR_HIP;
    if (cu_count) *cu_count = pr.multiProcessorCount;
    if (clock_khz) *clock_khz = pr.clockRate;
    if (mem_gb) *mem_gb = pr.totalGlobalMem / 1073741824.0;
    return HX_OK;
}

int hx_device_memory(hx_ctx* c, double* used_gb, double* total_gb) {
    if (!c) return HX_ERR_BAD_ARG;
    (void)hipSetDevice(c->device);
    size_t fr = 0, tot = 0;
    HXC(hipMemGetInfo(&fr, &tot));
    if (used_gb) *used_gb = (double)(tot - fr) / 1073741824.0;
    if (total_gb) *total_gb = (double)tot / 1073741824.0;
    return HX_OK;
}

}  // extern "C"
