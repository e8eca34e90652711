// hx_launch.h -- host-side launch wrappers of libhexat's kernels.
#pragma once

#include "hx_internal.h"

namespace hx {

// ---- elasticity operator -------------------------------------------------
enum MvMode { MV_PLAIN = 0, MV_PCG = 1 };

// Launch shape of the z-marching operator: tile 0 = 16x16 element columns per
// workgroup (256 threads), 1 = 16x32 (512 threads, parity contraction only),
// 2 = 32x8 (256 threads, parity contraction only: longer row segments);
// kchunk z planes per workgroup; pdepth node planes in flight per workgroup
// (parity contraction: 1 or 2).
struct MvShape {
    int tile = 0;
    int kchunk = 16;
    int pdepth = 2;
};

template <typename T>
struct MvArgs {
    const T* x;        // input vector (3*nn)
    T* y;              // output vector (3*nn)
    const T* E;        // element moduli (ne)
    T shift;           // y += shift * x (only where the node is owned)
    int mode;          // MvMode
    PcgState* st;      // MV_PCG: pq -> alpha, breakdown test (one state per column)
    double* partials;  // MV_PCG: one double per block and column
    long cs;           // column stride of x / y (multi-RHS: blockIdx.y = column)
    // fused p-update (MV_PCG only): x holds p_old; the kernel uses
    // p = minv*r + beta*p_old, writes it to p_out and applies K to it
    const T* r;
    const T* minv;     // shared by all columns
    T* p_out;
    // host side only: MV_PCG without the alpha finaliser -- the PCG update
    // then reduces the p'q partials itself (launch_pcg_update's pq_partials)
    int defer_alpha = 0;
    // parity contraction, MV_PLAIN: 3 = pre-smoothing (x = S r, y = masked
    // r - K x), 4 = post-smoothing (x = S r + mask(P ec), y = x + S (r - K x),
    // r'y partials when partials is set); 6 = residual of the iterate in x
    // (y = masked r - K x), 7 = post-smoothing from the iterate in x
    // (x + mask(P ec), then as 4); S in minv; st (if set) skips a finished
    // solve
    int form = 0;
    const T* ec = nullptr;         // forms 4, 7: coarse correction on the next level's node grid
    int cNX = 0, cNY = 0, cNZ = 0;  // its node counts
    long ccs = 0;                   // forms 4, 7: column stride of ec (multi-RHS)
};

// tile: 0 -> 16x16 element columns (default), 1 -> 16x32, 2 -> 32x8.
// The launchers return the number of workgroups per column (= p'q partials).
template <typename T>
int launch_matvec_parity(const GridArgs& g, const ParityK<T>& K, const MvArgs<T>& a, const MvShape& sh,
                          hipStream_t s, int nrhs = 1);
template <typename T>
int launch_matvec_dense(const GridArgs& g, const T* Kbin, const MvArgs<T>& a, const MvShape& sh, hipStream_t s,
                         int nrhs = 1);
int launch_matvec_gather(const GridArgs& g, const double* Kbin, const MvArgs<double>& a, hipStream_t s,
                          int nrhs = 1);
// FP32 element contraction on MFMA: dense [16 elem x 24] x Ke^T tiles with
// v_mfma_f32_16x16x4f32, inside the same z-marching tile
int launch_matvec_mfma(const GridArgs& g, const float* Kbin, const MvArgs<float>& a, const MvShape& sh,
                        hipStream_t s, int nrhs = 1);
int matvec_parity_blocks(const GridArgs& g, int tile, int kchunk);
// z chunk the post-smoothing form (MvArgs::form 4) runs at for a requested
// one: its dynamic LDS, (kchunk + 2) coarse footprints, stays within 64 KiB
int smoothing_post_kchunk(int tile, int kchunk, int elem_bytes);
int matvec_dense_blocks(const GridArgs& g, int kchunk);
int matvec_gather_blocks(const GridArgs& g);


// Jacobi inverse diagonal: minv = free ? 1/(kd[c] * sum_{e ni n} E_e + shift) : 0
template <typename T>
void launch_elastic_minv(const GridArgs& g, const double kd[3], const T* E, const uint8_t* fixed, double shift,
                         T* minv, hipStream_t s);
void launch_elastic_diag(const GridArgs& g, const double kd[3], const double* E, double* diag, hipStream_t s);

// one-block finalisers (one block per column): alpha = rz / p'q with the
// breakdown test, from the per-block p'q partials of a PCG-mode operator
void launch_fin_alpha(PcgState* st, const double* partials, int nblocks, int nrhs, hipStream_t s);

// ---- PCG vector kernels ----------------------------------------------------
// All of them take nrhs columns (blockIdx.y) of length n at stride n; the
// Jacobi inverse diagonal is shared; each column has its own PcgState and
// partials block.
constexpr int kVecBlocks = 1024;  // blocks of the streaming kernels (multiple of 8)
// dst[i] = src[i] for n 16-byte words (counter calibration)
void launch_copy16(long n, const void* src, void* dst, hipStream_t s);

// ---- viewer/checkpoint payload (hx_export.hip) ----------------------------
// tiles of the ordered compaction (tile_count needs web_tiles(ne) entries,
// offs web_tiles(ne) + 1; offs[web_tiles(ne)] receives the kept count)
int web_tiles(long ne);
void launch_web_pack(const GridArgs& g, const float* density, const float* l1, const float* principal, float thr,
                     unsigned* tile_count, unsigned* offs, float* cent, float* dens_out, float* sl1_out,
                     hipStream_t s, unsigned* total_out = nullptr);
// legacy VTK words (big-endian FP32 / Int32) of ExportVTK.export_solution:
// point data of every node, cells and cell scalars of the elements with
// density >= thr in element order (*total receives their count)
void launch_vtk_pack(const GridArgs& g, const double* u, const double* f, const uint8_t* fixed, const float* density,
                     const float* l1, const float* vm, float thr, unsigned* tile_count, unsigned* offs,
                     unsigned* total, unsigned* pts, unsigned* disp, unsigned* frc, unsigned* bcs, unsigned* dmag,
                     unsigned* cells, unsigned* l1_out, unsigned* vm_out, unsigned* dens_out, hipStream_t s);
// ---- STL isosurface (hx_stl.hip) ----------------------------------------------
// the marching-cubes case table (part 0: int8 [256][16] edge ids, part 1:
// uint8 [256] triangle counts), generated once per process
const void* mc_table_host(int part);
// nodal mean density (nn) and the sampling grid of (sub nx + 3)(sub ny + 3)(sub nz + 3) points
void launch_stl_field(const GridArgs& g, int sub, const float* rho, float* nod, float* fine, hipStream_t s);
long mc_blocks(int FX, int FY, int FZ);
// per-block triangle counts and their exclusive scan (offs: mc_blocks + 1, the last = total)
void launch_mc_count(const float* fine, int FX, int FY, int FZ, float iso, const uint8_t* ntri, unsigned* blk,
                     unsigned* offs, hipStream_t s);
// triangles (12 floats each: unit normal, 3 vertices), at most cap
void launch_mc_write(const float* fine, int FX, int FY, int FZ, float iso, float x0, float y0, float z0, float hx,
                     float hy, float hz, const int8_t* tri, const uint8_t* ntri, const unsigned* offs, unsigned cap,
                     float* out, hipStream_t s);
// ---- boundary conditions, loads and regrid (hx_boundary.hip) ------------------
// one location: per component "every node" (all = 1) or a value and tolerance
struct LocSpec {
    int all[3];
    float val[3];
    float tol[3];
};
long boundary_blocks(const GridArgs& g);
// fixed[3n + d] = 1 for the matched nodes and the DoF bits (bit d = DoF d + 1)
void launch_bc_mark(const GridArgs& g, const LocSpec& s, int dof_bits, uint8_t* fixed, hipStream_t st);
// *count = matched nodes (part: boundary_blocks words)
void launch_loc_count(const GridArgs& g, const LocSpec& s, unsigned* part, unsigned* count, hipStream_t st);
// F += f / *count on the matched nodes
void launch_force_add(const GridArgs& g, const LocSpec& s, const unsigned* count, const double f[3], double* F,
                      hipStream_t st);
// the nearest node to t as a key (bit 63 set, then i, j, k in 21 bits each)
void launch_nearest(const GridArgs& g, const float t[3], float* dpart, float* dmin, unsigned long long* kpart,
                    unsigned long long* key, hipStream_t st);
void launch_force_point(const GridArgs& g, const unsigned long long* key, const double f[3], double* F,
                        hipStream_t st);
// nf stacked element fields of the old grid onto the new grid (nearest cell)
void launch_regrid(int onx, int ony, int onz, int nnx, int nny, int nnz, const float nh[3], const float oh[3], int nf,
                   const float* oldf, float* newf, hipStream_t st);
// the packing's scratch for hx_web_pack_device: kWebScratchTiles tile counts
// then kWebScratchTiles + 1 offsets (at most kWebScratchTiles * 2048 elements)
constexpr int kWebScratchTiles = 65536;
constexpr int kWebScratchDevices = 64;
// its device address on `device`, which must be current (looked up once per
// device); 0 on success
int web_scratch_address(int device, unsigned** out);
template <typename T>
void launch_pcg_init(long n, const T* b, T* x, const T* q, const T* minv, T* r, T* p, double shift,
                     PcgState* st, double* partials, hipStream_t s, int nrhs = 1);
// set_beta = false leaves the iteration's bookkeeping (||r||, beta, the
// convergence decision) to an explicit preconditioner's finaliser, which reads
// the update's partials (launch_ml_apply); no finaliser is launched here.
// pq_partials (npq per column): the operator's p'q partials of an MV_PCG
// launch with defer_alpha; every block reduces them to alpha itself (block 0
// records it and the breakdown), so no alpha finaliser runs in between.
// xbest (non-null, one column): the state's per-iteration snapshot (mode 1),
// x copied there before the update when the last finaliser flagged it
template <typename T>
void launch_pcg_update(long n, T* x, const T* p, T* r, const T* q, const T* minv, PcgState* st,
                       double* partials, hipStream_t s, int nrhs = 1, bool set_beta = true,
                       const double* pq_partials = nullptr, int npq = 0, T* xbest = nullptr);
// x += alpha p ; r -= alpha q without the mask or the r'z, r'r partials (the
// V-cycle's post-smoothing operator forms r'r); nrhs columns of n values, the
// operator's npq p'q partials per column
template <typename T>
void launch_pcg_update_free(long n, T* x, const T* p, T* r, const T* q, PcgState* st, const double* pq_partials,
                            int npq, hipStream_t s, T* xbest = nullptr, int nrhs = 1);
// r -= alpha q alone (the r half of launch_pcg_update_free, the same bits), for
// a V-cycle whose one-workgroup pass takes x += alpha p (launch_vc_apply's xup)
template <typename T>
void launch_pcg_update_r(long n, T* r, const T* q, PcgState* st, const double* pq_partials, int npq, hipStream_t s,
                         int nrhs = 1);
template <typename T>
void launch_pcg_pupdate(long n, const T* r, const T* minv, T* p, const PcgState* st, hipStream_t s, int nrhs = 1);
// snap_mode: PcgState::snap_mode of every column
void launch_pcg_reset(PcgState* st, double tol, double shift, int max_iter, int check_every, hipStream_t s,
                      int nrhs = 1, int snap_mode = 0);

// generic reductions / helpers
void launch_dot(long n, const double* a, const double* b, double* out, unsigned* ticket, double* partials,
                hipStream_t s);
void launch_fill(long n, double v, double* a, hipStream_t s);
// per column: st->scal[5] = r'z (FP64 sums), for k_ir_resume after a reliable update that keeps z
template <typename T>
void launch_rz_park(long n, const T* r, const T* z, PcgState* st, double* partials, hipStream_t s, int nrhs = 1);
void launch_fill32(long n, float v, float* a, hipStream_t s);
template <typename Tin, typename Tout>
void launch_convert(long n, const Tin* a, Tout* b, hipStream_t s);
// r = mask(b - q - shift*x) in FP64 with FP32 copy; rz, ||r||^2 (scal[4]) and
// rel; mode bit0 also sets normb, bit1 parks rz in scal[5]
void launch_true_residual(long n, const double* b, const double* q, const double* x, const double* minv64,
                          double shift, double* r64, float* r32, PcgState* st, double* partials, int mode,
                          hipStream_t s, int nrhs = 1);
// per column: resume unless the column has converged (rel < tol_final) or failed
void launch_ir_resume(PcgState* st, double tol_final, double delta, int first, hipStream_t s, int nrhs = 1);
// move each column's live search direction from pb into p (double-buffered p)
template <typename T>
void launch_fix_p(long n, T* p, const T* pb, const PcgState* st, hipStream_t s, int nrhs = 1);
template <typename T>
void launch_mask(long n, const T* minv, T* x, hipStream_t s, int nrhs = 1);
// x64 += x32; x32 = 0
void launch_accumulate(long n, double* x64, float* x32, hipStream_t s, int nrhs = 1);
// xbest = x when st->rel < min(1e-2, st->snap_rel); then snap_rel = rel
// (best-iterate snapshot of one column at a check point, mode 2)
void launch_snapshot(long n, const double* x, double* xbest, PcgState* st, hipStream_t s);
// xbest = x if the last finaliser flagged x (mode 1, end of the solve)
void launch_snapshot_pending(long n, const double* x, double* xbest, PcgState* st, hipStream_t s);
// out[0] = #(rho >= thr), out[1] = #(rho > thr_keep) (kVecBlocks x 2 partials)
void launch_count_active(long ne, const double* rho, double thr, double thr_keep, double* partials, double* out,
                         hipStream_t s);
// Eout = E rho where rho >= thr (and rho > thr_keep with prune), else 0
void launch_solve_moduli(long ne, const double* rho, double E, double thr, double thr_keep, int prune, double* Eout,
                         hipStream_t s);
// out[0] = max |diag| over the DOFs not fixed (kVecBlocks partials)
void launch_free_absmax(long n, const double* diag, const uint8_t* fixed, double* partials, double* out,
                        hipStream_t s);

// ---- flexible CG pieces (learned preconditioner) ----------------------------
// Truncated flexible CG: the new direction is A-orthogonalised against the
// last m (<= kFcgKeep) stored ones, P_j = P[slot[j]], AP_j = AP[slot[j]].
constexpr int kFcgKeep = 8;
struct FcgCoef {
    int m;
    int slot[kFcgKeep];
    double c[kFcgKeep];
};
// one pass for the iteration's dot products (u'Au, u'Ay, y'Ay, u'r, y'r, then
// Au'P_j, Ay'P_j for the cf.m stored directions) into out[0 .. 5 + 2 cf.m)
constexpr int kFcgDots = 5 + 2 * (kFcgKeep - 1);
void launch_fcg_dots(long n, const double* W, const double* AW, const double* r, const FcgCoef& cf, const double* P,
                     double* partials, double* out, hipStream_t s);
// W = [u | y], AW = [Au | Ay] (two columns of n): p = a u + b y - sum c_j P_j,
// Ap = a Au + b Ay - sum c_j AP_j; leaves scal[0] = p'Ap, scal[1] = p'r,
// scal[2] = alpha (0 and scal[3] = 1 on breakdown) on the device
void launch_fcg_direction2(long n, double a, double b, const double* W, const double* AW, const FcgCoef& cf,
                           const double* P, const double* AP, const double* r, double* p, double* Ap, double* partials,
                           double* scal, hipStream_t s);
// x += alpha p ; r = free ? r - alpha Ap : 0 with alpha = scal_alpha[2] read on the device; scal[4] = r'r after it
void launch_fcg_step2(long n, const double* scal_alpha, double* x, const double* p, double* r, const double* Ap,
                      const double* minv, double* partials, double* scal, hipStream_t s);
void launch_masked_diff(long n, const double* a, const double* b, const double* minv, double* e, hipStream_t s);
// out = a * b (elementwise)
void launch_mul(long n, const double* a, const double* b, double* out, hipStream_t s);

// ---- multilevel diagonal scaling (explicit preconditioner z = M^-1 r) ---------
constexpr int kMlMaxLevels = 12;
constexpr int kMlExactMax = 192;  // DOFs of a coarsest level solved exactly
template <typename T>
struct MlHierarchy {
    int levels = 0;                  // coarse levels L (0: Jacobi only)
    const T* ainv = nullptr;         // non-null: level L is solved exactly, y_L = A_L^-1 r_L (dense, 3 nn_L square)
    GridArgs g[kMlMaxLevels + 1];    // g[0] = fine grid, g[l] = level l
    T* r[kMlMaxLevels + 1] = {};     // restricted residual of level l >= 1
    T* y[kMlMaxLevels + 1] = {};     // level output D_l^-1 r_l + P y_{l+1}
    T* minv[kMlMaxLevels + 1] = {};  // level Jacobi inverse diagonal
    double w[kMlMaxLevels + 1] = {};  // level weights
    // V-cycle levels l < vc_levels(): smoother diagonal, residual after
    // pre-smoothing, operator moduli (level 0: the fine moduli)
    T* sinv[kMlMaxLevels + 1] = {};
    T* vres[kMlMaxLevels + 1] = {};
    const T* Eop[kMlMaxLevels + 1] = {};
    // workgroups a coarse level's launch aims for: the V-cycle's level
    // operators and every restriction take shallower z chunks until reached
    int min_blocks = 512;
    // the V-cycle's one-workgroup pass: moduli and smoother diagonal of each
    // level it smooths (E_l then S_l, level by level), one contiguous copy
    // made per hierarchy (launch_vc_coarse_stage), so the pass stages them
    // with coalesced loads from one address
    const T* cstage = nullptr;
    long cstage_n = 0;
};
// number of 2:1 coarsenings of g (all element counts even), at most max_levels
int ml_levels_for(const GridArgs& g, int max_levels);
GridArgs ml_coarse_grid(const GridArgs& g);
void launch_ml_coarsen_E(const GridArgs& gf, const double* Ef, const GridArgs& gc, double* Ec, float* Ec32,
                         hipStream_t s);
template <typename T>
void launch_ml_minv(const GridArgs& gc, const double kd[3], const double* Ec, const uint8_t* fixed_f, int stride,
                    const GridArgs& gfine, T* minv, hipStream_t s);
// z = M^-1 r (masked where minv0 = 0) and r'z into the state: fin_mode 0 = an
// iteration (records ||r|| from the update's partials at `partials`, r'r at
// index 2b+1 of kVecBlocks blocks, then beta and the convergence decision),
// 1 = start (rz, beta = 0), 2 = park in scal[5].  Its own partials go after
// the update's.
template <typename T>
void launch_ml_apply(const MlHierarchy<T>& h, const T* r, const T* minv0, T* z, PcgState* st, double* partials,
                     int fin_mode, hipStream_t s);
int ml_partial_blocks();
// A_L^-1 of the coarsest level gc (moduli Ec, element stiffness h Kbin, DOFs
// fixed where the fine DOF at stride * (I,J,K) is, or without stiffness, give
// zero rows and columns), built and inverted on the device in one workgroup
// (3 nn <= kMlExactMax)
template <typename T>
void launch_ml_exact_setup(const GridArgs& gc, const double* Ec, const double* Kbin, double h, const uint8_t* fixed_f,
                           int stride, const GridArgs& gfine, T* ainv, hipStream_t s);
void ml_prepare_kernels();  // once per process, outside graph capture
// V-cycle form: every level above the coarsest smoothed (vc_levels = levels),
// the coarsest solved exactly or by its diagonal
template <typename T>
int vc_levels(const MlHierarchy<T>& h);
// smoother diagonal of a level: omega / absolute row sums (0 where the DOF is
// not solved for); arguments as launch_ml_exact_setup's
template <typename T>
void launch_vc_smoother(const GridArgs& gc, const double* Ec, double h, const double* Kbin, const uint8_t* fixed_f,
                        int stride, const GridArgs& gfine, double omega, T* sinv, hipStream_t s);
template <typename T>
void launch_vc_scale_E(long ne, const double* Ec, double h, T* Eop, hipStream_t s);
// values of h.cstage for the hierarchy's one-workgroup pass (0: none)
template <typename T>
long vc_coarse_stage_values(const MlHierarchy<T>& h);
// fill buf (vc_coarse_stage_values(h) values) and point h.cstage at it;
// after the levels' Eop and sinv are set, outside graph capture
template <typename T>
int launch_vc_coarse_stage(MlHierarchy<T>& h, T* buf, hipStream_t s);
// z = V(r) with the parity operator K, its pre- and post-smoothing forms at
// shapes sh_pre / sh_post on the multi-block levels (coarse ones with
// shallower z chunks), the small levels in one workgroup per column; the r'z
// partials (one per level-0 post-smoothing block) go after the update's and
// the finaliser runs as in launch_ml_apply.  HX_ERR_BAD_ARG when those
// partials would pass max_partials.
// rr_post: the finaliser takes r'r from the fine post-smoothing operator
// (after a launch_pcg_update_free) instead of the update's partials; needs a
// coarse level.  nrhs columns (blockIdx.y) share the hierarchy: r, z and the
// level vectors hold nrhs columns each, one state per column; nrhs > 1 needs
// a coarse level.
// xup, pup (fin_mode 0, a one-workgroup pass present: vc_coarse_pass): the
// pass's launch carries spare workgroups that do xup += alpha pup per column
// (alpha from the state, as the update kernel before it left it), beside the
// pass.
template <typename T>
int launch_vc_apply(const MlHierarchy<T>& h, const ParityK<T>& K, const MvShape& sh_pre, const MvShape& sh_post,
                    const T* r, T* z, PcgState* st, double* partials, long max_partials, int fin_mode, hipStream_t s,
                    bool rr_post = false, int nrhs = 1, T* xup = nullptr, const T* pup = nullptr);
// the V-cycle of h runs a one-workgroup pass (levels small enough for it exist)
template <typename T>
bool vc_coarse_pass(const MlHierarchy<T>& h);

// ---- Helmholtz filter --------------------------------------------------------
struct HelmArgs {
    double Ae[64];  // reference node order (R^2 Ks + Ms)
    // the box it was formed on (element sides, the quirk's taller z
    // included) and R^2: with them the operator runs sum-factorised, after
    // a check that the factors reproduce every entry of Ae (h[0] = 0: not
    // given, the gather form runs)
    double h[3] = {0.0, 0.0, 0.0};
    double R2 = 0.0;
};
void launch_helm_rhs(const GridArgs& g, const double* in_e, double vol8, double* b, hipStream_t s);
void launch_helm_minv(const GridArgs& g, const HelmArgs& A, double* minv, hipStream_t s);
// y = A x; in MV_PCG mode also the per-block p'q partials (one per block,
// the returned count), from which the update forms alpha
int launch_helm_matvec(const GridArgs& g, const HelmArgs& A, const double* x, double* y, int mode, PcgState* st,
                       double* partials, hipStream_t s);
int helm_matvec_blocks(const GridArgs& g);
// true when A's sum-factorised form reproduces its element matrix (the
// operator then runs in that form, and the fused PCG operator below exists)
bool helm_sf_available(const HelmArgs& A);
// the PCG operator with the p-update fused: p_out = minv r + beta p_old and
// q = A p_out, with the per-block p'q partials; -1 (nothing launched) when A
// has no checked sum-factorised form
int launch_helm_matvec_fused(const GridArgs& g, const HelmArgs& A, const double* p_old, const double* r,
                             const double* minv, double* p_out, double* q, PcgState* st, double* partials,
                             hipStream_t s);
void launch_helm_extract(const GridArgs& g, const double* xn, double* out_e, hipStream_t s);

// ---- element kernels -----------------------------------------------------------
struct ElemArgs {
    const double* u;
    const double* Emod;      // stress modulus per element (nullable -> skip stress)
    const double* xphys;     // for dc (nullable -> skip dc)
    double penal, E0, Emin;
    double lam1, lam2, mu;   // unit-E material: c1 = (1-nu)f, c2 = nu f, c3 = (1-2nu)/2 f
    double* principal;       // nullable outputs
    double* vm;
    double* voigt;
    double* l1;
    double* dir;
    double* ce;
    double* dc;
    double* vm_max_partials; // nullable: per-block max vm
};
void launch_element_post(const GridArgs& g, const ParityK<double>& Kunit, const ElemArgs& a, hipStream_t s);
int element_post_blocks(const GridArgs& g);
void launch_simp_modulus(long ne, const double* xphys, double penal, double E0, double Emin, double* E, hipStream_t s);
// lguess > 0: the 8-point search starts from [lguess/16, 16 lguess] (widened on its first pass when the
// volume does not cross inside); otherwise [0, 1e9]
void launch_oc_init(PcgState* st, hipStream_t s, double lguess = 0.0);
// one pass of the 8-point search for the same multiplier (the device-resident
// loop's OC): bracket shrunk 9x per pass; partials: 8 per block
void launch_oc_eval_k(long ne, const double* x, const double* dc, const double* dv, double volfrac, double move,
                      double xmin, double rel_tol, PcgState* st, double* partials, hipStream_t s);
void launch_oc_eval(long ne, const double* x, const double* dc, const double* dv, double volfrac, double move,
                    double xmin, double rel_tol, PcgState* st, double* partials, hipStream_t s);
void launch_oc_apply(long ne, const double* x, const double* dc, const double* dv, double move, double xmin,
                     double* xnew, PcgState* st, double* partials, hipStream_t s);
void launch_max_reduce(long n, const double* partials_in, int nin, double* out, hipStream_t s);

}  // namespace hx
