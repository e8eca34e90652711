/*
 * hexat.h -- C ABI of libhexat.so, the MI355X (gfx950) hot path of the HEXAT
 * topology-optimisation loop.  Plain pointers and sizes only; every array is
 * caller-owned host memory unless a function says otherwise.  One context per
 * host thread; a context is not thread-safe.  All functions return an
 * hx_status (0 = ok); hx_last_error() gives the message of the last failure.
 *
 * Structured grid conventions (0-based; the reference is 1-based Julia):
 *   nx*ny*nz box elements of size dx*dy*dz, nodes NX=nx+1, NY, NZ,
 *   node(i,j,k) = i + NX*(j + NY*k), elem(i,j,k) = i + nx*(j + ny*k),
 *   DOF = 3*node + d (node-major, d = x,y,z).
 * This is exactly the layout of Mesh.generate_mesh (Mesh.jl:23-60) and of the
 * DOF numbering of GPUSolver.jl:50-57, so a Julia caller passes its arrays
 * unchanged (Float64 copies) and drops the explicit connectivity, which the
 * kernels compute from (i,j,k).
 */
#ifndef HEXAT_H
#define HEXAT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    HX_OK = 0,
    HX_ERR_NO_DEVICE = 1,     /* GPUSolver.jl:620 error("CUDA not found.") */
    HX_ERR_BAD_ARG = 2,       /* invalid size / pointer / option */
    HX_ERR_NOT_CONVERGED = 3, /* max_iter reached; best iterate returned (GPUSolver.jl:603-606) */
    HX_ERR_BREAKDOWN = 4,     /* |p'Ap| tiny or NaN (GPUSolver.jl:462-466) */
    HX_ERR_DIVERGED = 5,      /* NaN/Inf, > 1000*|b|, or > 100x best after 50 its (:480-490) */
    HX_ERR_NO_ACTIVE = 6,     /* no element with non-zero stiffness (GPUSolver.jl:132-134) */
    HX_ERR_HIP = 7,           /* HIP runtime failure */
    HX_ERR_STATE = 8,         /* call order (e.g. no grid set) */
    HX_ERR_OOM = 9,           /* device allocation failed */
    HX_ZERO_RHS = 10          /* informational: zero right-hand side, zero solution (GPUSolver.jl:373) */
} hx_status;

typedef struct hx_ctx hx_ctx;

/* ---- context ------------------------------------------------------------ */
int hx_version(void);
/* number of visible HIP devices (0 when none) */
int hx_device_count(int* count);
/* create a context on a device; HX_ERR_NO_DEVICE when there is none */
int hx_ctx_create(int device, hx_ctx** out);
void hx_ctx_destroy(hx_ctx* ctx);
const char* hx_last_error(const hx_ctx* ctx);

/* Set the structured grid and Poisson ratio.  Builds Ke (replaces
 * GPUSolver.setup_matrix_free_operator, GPUSolver.jl:124-150, whose dx,dy,dz
 * come from element 1) and (re)allocates the device workspace. */
int hx_grid_set(hx_ctx* ctx, int nx, int ny, int nz, double dx, double dy, double dz, double nu);
/* fixed-DOF mask, 3*NX*NY*NZ bytes, non-zero = fixed.  Replaces
 * GPUSolver.get_free_dofs (GPUSolver.jl:112-122) + free<->full maps (:81-97). */
int hx_set_fixed(hx_ctx* ctx, const uint8_t* fixed);

/* ---- element matrices (host, FP64) ---------------------------------------- */
/* 24x24 row-major H8 stiffness of a dx*dy*dz box (Element.jl:64-102) */
int hx_element_stiffness(double dx, double dy, double dz, double nu, double E, double* Ke);
/* 8x8 scalar Laplacian / mass (Element.jl:104-125); quirk=1 keeps the
 * reference's top face at z=+dz */
int hx_scalar_matrices(double dx, double dy, double dz, int quirk, double* Ks, double* Ms);

/* ---- elasticity ------------------------------------------------------------ */
enum { HX_FP64 = 0, HX_FP32_IR = 1 };           /* precision form */
/* preconditioner.  HX_PC_MULTILEVEL: additive multilevel diagonal scaling
 * z = D0^-1 r + sum_l P_l D_l^-1 P_l^T r over every 2:1 coarsening of the
 * grid (trilinear P_l, coarse moduli averaged over 2x2x2 children), the linear
 * symmetric form of a U-Net preconditioner; plain PCG, one right-hand side.
 * No reference counterpart (the reference's strong-preconditioner slot is the
 * learned one, and its CPU AMG branch, GPUSolver.jl:168-268). */
/* HX_PC_VCYCLE: the same hierarchy as a symmetric V-cycle: l1-Jacobi
 * smoothing (omega / absolute row sums of the rediscretised level stiffness,
 * one step before and one after the coarse correction) on every level above
 * the coarsest, which is solved exactly (or by its diagonal); SPD for every
 * moduli field; one right-hand side.  No reference counterpart. */
enum { HX_PC_JACOBI = 0, HX_PC_LEARNED = 1, HX_PC_MULTILEVEL = 2, HX_PC_VCYCLE = 3 };
enum { HX_MV_AUTO = 0, HX_MV_PARITY = 1, HX_MV_DENSE = 2, HX_MV_GATHER = 3, HX_MV_MFMA = 4 };

typedef struct {
    double tol;       /* relative residual ||r||/||b|| target */
    int max_iter;     /* PCG iteration cap */
    double shift;     /* absolute diagonal shift s (K_ff + s I), GPUSolver.jl:391 */
    int precision;    /* HX_FP64 | HX_FP32_IR */
    int precond;      /* HX_PC_JACOBI | HX_PC_LEARNED | HX_PC_MULTILEVEL | HX_PC_VCYCLE */
    int warm_start;   /* 1: u_inout holds the start vector (GPUSolver.jl:407-412) */
    int matvec;       /* HX_MV_* kernel variant (HX_MV_AUTO = tuned choice) */
    int batch;        /* iterations per host check (0 = default) */
} hx_pcg_opts;

typedef struct {
    int iterations;   /* PCG iterations (inner iterations for HX_FP32_IR) */
    int status;       /* hx_status of the solve */
    double rel_res;   /* final ||b - K u||/||b||, recomputed in FP64 */
    double seconds;   /* device time of the solve */
    int refinements;  /* FP64 residual replacements (HX_FP32_IR; HX_PC_LEARNED: its restarts) */
    int fallbacks;    /* learned -> Jacobi fallbacks */
} hx_pcg_stats;

/* y = sum_e Ee[e] Ke x_e over all DOFs (no mask); Ee[e] = 0 drops the element.
 * Replaces GPUSolver.apply_matrix_free_operator! (GPUSolver.jl:152-162). */
int hx_matvec(hx_ctx* ctx, const double* Ee, const double* x, double* y, int variant);
/* Jacobi diagonal sum_e Ee[e] diag(Ke) (GPUSolver.jl:301-317) */
int hx_diagonal(hx_ctx* ctx, const double* Ee, double* diag);
/* Solve (K_ff + s I) u = f on free DOFs; fixed DOFs of u are 0 on return.
 * Replaces GPUSolver.gpu_matrix_free_cg_solve (GPUSolver.jl:274-609). */
int hx_pcg_solve(hx_ctx* ctx, const double* Ee, const double* f, double* u_inout,
                 const hx_pcg_opts* opts, hx_pcg_stats* stats);
/* The solve policy of Solver.solve_system (Solver.jl:38-164) and
 * GPUSolver.gpu_matrix_free_cg_solve (GPUSolver.jl:274-609) on the device.
 * Elements with density >= min_stiffness_threshold enter the operator with
 * modulus E*density (GPUSolver.jl:128-147).  With prune_voids, when fewer than
 * 99 % of the elements have density > 1.01*threshold (the product taken in
 * FP32, Solver.jl:61-65), only those stay and the nodes no kept element
 * touches leave the system (MeshPruner.prune_system, MeshPruner.jl:18-138):
 * here their DOFs get a zero diagonal, are not solved for and return 0, and a
 * warm start keeps u_prev on the kept nodes (Solver.jl:71-83).
 * shift = shift_factor * max |diag| over free DOFs (GPUSolver.jl:315-316, 391).
 * Up to max_attempts PCG attempts (GPUSolver.jl:390-594): converged -> done;
 * not diverged with best residual < 5e-2 -> accepted (the attempt's best
 * iterate); diverged / broken down / best > 0.1 -> shift factor x10; else the
 * attempt's best iterate is kept if it is the best so far and the factor x5.
 * The reference's switch to CPU AMG ("jacobi_amg") is not built.  After all
 * attempts fail: the best kept iterate (zeros if none) and
 * HX_ERR_NOT_CONVERGED.  Zero right-hand side on the free DOFs: zeros,
 * stats->outcome HX_SOLVE_ZERO_RHS, return HX_OK (GPUSolver.jl:372-373).
 * No element in the operator: HX_ERR_NO_ACTIVE (GPUSolver.jl:132-134,
 * MeshPruner.jl:35-37).  density: Ne, f and u_inout: ndof (u_inout holds
 * u_prev when warm_start). */
typedef struct {
    double E;                        /* Young's modulus */
    double min_stiffness_threshold;  /* density >= this: in the operator */
    int prune_voids;                 /* Solver.jl:60-91 */
    double tol;                      /* relative residual target (after the caller's schedule) */
    int max_iter;
    double shift_factor;             /* diagonal_shift_factor */
    int precision;                   /* HX_FP64 | HX_FP32_IR */
    int precond;                     /* HX_PC_JACOBI | HX_PC_MULTILEVEL | HX_PC_VCYCLE */
    int max_attempts;                /* 3 in the reference (GPUSolver.jl:368) */
    int warm_start;                  /* first attempt starts from u_inout (GPUSolver.jl:407-412) */
} hx_solve_opts;

enum { HX_SOLVE_CONVERGED = 0, HX_SOLVE_ACCEPTED = 1, HX_SOLVE_FAILED = 2, HX_SOLVE_ZERO_RHS = 3 };

typedef struct {
    int status;         /* hx_status returned */
    int outcome;        /* HX_SOLVE_* */
    int attempts;       /* PCG attempts made */
    int iterations;     /* PCG iterations, all attempts */
    double rel_res;     /* recursive relative residual where the last attempt stopped
                           (the reference's final_rel_res_val) */
    double true_rel;    /* FP64 relative residual of the returned vector, this attempt's shift */
    double best_rel;    /* best recursive residual of the last attempt */
    double shift;       /* absolute shift of the last attempt */
    long n_active;      /* elements with density >= threshold */
    long n_kept;        /* elements in the operator */
    int pruned;         /* 1: void pruning applied */
    double seconds;     /* device time */
} hx_solve_stats;

int hx_solve_system(hx_ctx* ctx, const double* density, const double* f, double* u_inout,
                    const hx_solve_opts* opts, hx_solve_stats* stats);

/* nrhs independent systems with one shared operator, solved together
 * (fused multi-RHS kernels; blockIdx.y = column).  F, U are nrhs x ndof
 * row-major (RHS-major).  stats: nrhs entries.  Preconditioners: Jacobi, or
 * the V-cycle on grids with a coarse level (every V-cycle launch carries the
 * columns; each column takes its single solve's iterations and bits); the
 * additive multilevel and learned ones take one column.  No reference
 * counterpart (the reference solves one load case per call, Main.jl:414-418). */
int hx_block_pcg_solve(hx_ctx* ctx, int nrhs, const double* Ee, const double* F, double* U,
                       const hx_pcg_opts* opts, hx_pcg_stats* stats);

/* ---- Helmholtz PDE filter ------------------------------------------------- */
typedef struct {
    int quirk;        /* 1: reference element (top at +dz, constant -> x2/3) */
    double tol;       /* relative residual target (reference filter_tolerance 1e-5) */
    int max_iter;     /* reference: 200 (GPUHelmholtz.jl:191) */
    int check_every;  /* residual test period (reference: 10, GPUHelmholtz.jl:208) */
} hx_filter_opts;

typedef struct {
    int iterations;
    int status;
    double rel_res;
    double seconds;
} hx_filter_stats;

/* out_e = mean over the element's nodes of xn, (R^2 Ks + Ms) xn = sum_e in_e vol/8.
 * Replaces GPUHelmholtz.apply_gpu_filter! (GPUHelmholtz.jl:344-376) and
 * solve_helmholtz_on_gpu (:165-229). */
int hx_helmholtz_filter(hx_ctx* ctx, double R, const hx_filter_opts* opts, const double* in_e,
                        double* out_e, hx_filter_stats* stats);

/* ---- element post-processing ---------------------------------------------- */
/* Centroid stress of material with modulus Ee[e] (Ee = 0 -> zeros).  Any
 * output may be NULL.  principal: Ne x 3 descending; voigt: Ne x 6
 * (xx,yy,zz,xy,yz,xz); dir: Ne x 3 unit eigenvector of the largest principal
 * stress (sign arbitrary).  Replaces Stress.compute_stress_field (Stress.jl:57-100). */
int hx_element_stress(hx_ctx* ctx, const double* Ee, const double* u, double* principal,
                      double* vm, double* voigt, double* l1, double* dir);
/* ---- boundary conditions, loads, regrid on the device --------------------- */
/* A location of the YAML grammar resolved to this grid (Boundary.jl:79-130):
 * per component, every node (all = 1) or the FP32 coordinate value the
 * reference resolves ("50%", fractions in [0,1], absolute numbers); nodes
 * match within 1e-4 * max(1, extent) per component, in FP32. */
typedef struct {
    int all[3];
    float value[3];
} hx_location;
/* Boundary.get_bc_indicator (Boundary.jl:135-160): the fixed-DOF mask of the
 * specs (dof_bits[q]: bit d = DoF d+1), installed in the context as
 * hx_set_fixed would, and copied to fixed_out (3 nn bytes, nullable). */
int hx_bc_indicator(hx_ctx* ctx, int nspec, const hx_location* locs, const int* dof_bits, uint8_t* fixed_out);
/* Boundary.apply_external_forces! (Boundary.jl:232-291): F (ndof, in/out)
 * += each spec's force (3 doubles) split evenly over its matched nodes, the
 * specs in order; a spec with no match whose three components are values and
 * point_fallback[q] set puts the whole force on the nearest node (least FP32
 * squared distance; within 1e-9 of it the largest x, then y, then z,
 * Boundary.jl:195-224).  matched (nullable): nodes that took each spec. */
int hx_external_forces(hx_ctx* ctx, int nspec, const hx_location* locs, const double* force,
                       const int* point_fallback, double* F, uint32_t* matched);
/* MeshRefiner.refine_mesh_and_fields' field map (MeshRefiner.jl:131-150): each
 * element of the new nnx*nny*nnz grid of the same domain takes the values of
 * the old element its FP32 centroid falls in.  nfields stacked element fields
 * (old: nfields x onx*ony*onz, new: nfields x nnx*nny*nnz).  The grid of the
 * context is not changed. */
int hx_regrid(hx_ctx* ctx, int onx, int ony, int onz, int nnx, int nny, int nnz, float len_x, float len_y,
              float len_z, int nfields, const float* old_fields, float* new_fields);

/* Payload of the viewer/checkpoint file: the elements with
 * density[e] >= threshold, in element order; for each, its centroid (FP32 mean
 * of the 8 node coordinates, MeshUtilities.jl:18-32), its density and its L1
 * stress signed by the principal stress of largest magnitude (first on ties).
 * principal: Ne x 3 as hx_element_stress.  centroids (3 Ne), density_out and
 * signed_l1 (Ne each) are sized for the worst case; *count receives the number
 * kept (0: nothing is written, as the reference writes no file).  Replaces
 * the packing loop of Postprocessing.export_binary_for_web
 * (Postprocessing.jl:173-195); the file itself is written by the host. */
int hx_web_pack(hx_ctx* ctx, const float* density, const float* l1, const float* principal, float threshold,
                float* centroids, float* density_out, float* signed_l1, uint32_t* count);
/* hx_web_pack on caller-owned DEVICE buffers that allocates nothing: no
 * hipMalloc, no hipFree, no allocating library call.  No counterpart in the
 * reference; it is for callers that cannot allocate (code under graph
 * capture, tight serving loops, harnesses that forbid it).  Every pointer is
 * device memory: the inputs (density, l1: Ne; principal: Ne x 3), the outputs
 * sized for the worst case (centroids 3 Ne, density_out and signed_l1 Ne) and
 * the kept count (one uint32_t).  The tile counts live in a static device
 * array of 2 * 65536 + 1 words (512 KiB, held on each GPU the library runs on
 * until the process ends), so the grid may hold at most 65536 * 2048 =
 * 134,217,728 elements (HX_ERR_BAD_ARG beyond).  Calls that overlap on one GPU
 * race on that array.  Runs on the context's stream and synchronizes it. */
int hx_web_pack_device(hx_ctx* ctx, const float* density, const float* l1, const float* principal,
                       float threshold, float* centroids, float* density_out, float* signed_l1,
                       uint32_t* count);
/* The data blocks of the legacy binary VTK file of ExportVTK.export_solution
 * (ExportVTK.jl:20-172), every word an FP32 or Int32 in big-endian byte order
 * as the file stores it; the caller writes the ASCII section headers between
 * them.  Inputs: u, f (ndof, nullable: zeros), fixed (3 nn bytes, non-zero =
 * fixed; nullable: none), density, l1, vm (Ne FP32; l1, vm nullable: zeros),
 * threshold.  Outputs:
 *   points: 13 nn words -- coordinates (3 nn), displacement (3 nn), external
 *     force (3 nn), boundary-condition indicator (3 nn), displacement
 *     magnitude (nn); u, f and l1 pass through the reference's
 *     sanitize_data (non-finite -> 0, clamp to +-1e10);
 *   cells: 9 words per kept element (density >= threshold, element order):
 *     8 and its 0-based node ids in the reference's node order;
 *   cell_scalars: l1, von Mises, density of the kept elements, count words each;
 *   *count: the number kept.  cells and cell_scalars sized for the worst
 *   case (9 Ne and 3 Ne words). */
int hx_vtk_pack(hx_ctx* ctx, const double* u, const double* f, const uint8_t* fixed, const float* density,
                const float* l1, const float* vm, float threshold, uint32_t* points, uint32_t* cells,
                uint32_t* cell_scalars, uint32_t* count);
/* Isosurface of the density for the STL file of
 * Postprocessing.export_smooth_watertight_stl (Postprocessing.jl:53-148): the
 * nodal mean density, refined `subdivision` times per element by trilinear
 * interpolation with a layer of zeros around it, then marching cubes at
 * `threshold` (a generated, crack-free case table; the reference uses
 * MarchingCubes.jl, whose choice of triangles differs).  Writes up to
 * max_triangles triangles of 12 floats (unit normal or zeros, three vertices,
 * outward from the region >= threshold) and the total in *count; with
 * max_triangles 0 only the count is formed.  field (nullable): the sampling
 * grid, (s nx + 3)(s ny + 3)(s nz + 3) FP32 values, x fastest. */
int hx_stl_pack(hx_ctx* ctx, const float* density, int subdivision, float threshold, float* triangles,
                long max_triangles, long* count, float* field);
/* ce_e = u_e^T Ke(E=1) u_e and SIMP sensitivity
 * dc_e = -penal xphys^(penal-1) (E0-Emin) ce_e; compliance = f^T u.
 * No reference counterpart (the reference has no sensitivities). */
int hx_compliance_sens(hx_ctx* ctx, const double* xphys, double penal, double E0, double Emin,
                       const double* u, const double* f, double* ce, double* dc, double* compliance);
/* OC bisection (top88): xnew = clamp(x sqrt(-dc/(dv lmid)), max(xmin,x-move), min(1,x+move)),
 * volume tested on mean(xnew); stops at (l2-l1)/(l1+l2) <= rel_tol.
 * No reference counterpart. */
int hx_oc_update(hx_ctx* ctx, const double* x, const double* dc, const double* dv, double volfrac,
                 double move, double xmin, double rel_tol, double* xnew, double* lmid, int* steps,
                 double* change);

/* ---- device-resident SIMP/OC topology-optimisation loop -------------------- */
typedef struct {
    double E0, Emin, nu, penal, volfrac, move, xmin;
    double rmin;          /* filter radius r; the PDE filter uses R = r / (2 sqrt 3) */
    double pcg_tol;       /* elasticity relative residual target */
    int pcg_max_iter;
    int precision;        /* HX_FP64 | HX_FP32_IR */
    int precond;          /* HX_PC_JACOBI | HX_PC_LEARNED | HX_PC_MULTILEVEL | HX_PC_VCYCLE */
    double filter_tol;    /* Helmholtz filter relative residual target */
    int filter_max_iter;
    double oc_rel_tol;    /* OC bisection stop */
    int with_stress;      /* 1: von Mises at centroids every iteration */
} hx_to_params;

typedef struct {
    int iter;
    double compliance;    /* f^T u */
    double change;        /* max |x_new - x| */
    double volume;        /* mean(xphys) after the update */
    int pcg_iterations;
    double pcg_rel_res;
    int pcg_status;
    int filter_iterations;/* sum over the two filter solves */
    double max_vm;        /* max von Mises stress (with_stress) */
    double seconds;       /* device time of the iteration */
} hx_to_stats;

enum { HX_FIELD_X = 0, HX_FIELD_XPHYS = 1, HX_FIELD_U = 2, HX_FIELD_VM = 3, HX_FIELD_DC = 4 };

/* Upload f (ndof) and x0 (Ne); xphys = filter(x0). */
int hx_to_init(hx_ctx* ctx, const hx_to_params* params, const double* f, const double* x0);
/* One TO iteration on the device: E(xphys), PCG (warm start), compliance,
 * sensitivities (+ von Mises), sensitivity filter, OC, density filter. */
int hx_to_step(hx_ctx* ctx, hx_to_stats* stats);
/* copy a device field (HX_FIELD_*) to host (Ne or ndof doubles) */
int hx_to_read(hx_ctx* ctx, int field, double* out);

/* ---- learned preconditioner (BF16 3-D conv U-Net) --------------------------- */
/* Train the network on (density, residual, error) triples that the Jacobi-PCG
 * of the current grid generates; loss, epochs and timings in the stats.  The
 * loss is the fraction of the error's energy |e|_A^2 left outside the best
 * combination of D^-1 r and the network's direction (0 = exact, 1 = none);
 * jacobi_loss is the same fraction for D^-1 r alone on the same samples. */
typedef struct {
    int samples;      /* (rho, r, e) triples to harvest */
    int epochs;
    double lr;
    int seed;
    double final_loss;   /* mean loss of the last epoch */
    double seconds;
    double jacobi_loss;  /* mean loss of D^-1 r alone */
    /* 0: triples from reference solves of random-load problems (stored,
     * `epochs` passes; Ee/f give the first problem).  m = 1..4: triples
     * streamed from manufactured solutions: per random density field smooth
     * random displacements e* (multi-scale lattice noise), loads b = A e*,
     * Jacobi-PCG from zero; (r_k, e* - x_k) at the first m of the iterations
     * k = 0, 2, 8, 32 each take one training step at once (no storage, any
     * grid size); `samples` triples per pass, `epochs` passes over the same
     * seeded stream. */
    int manufactured;
} hx_learn_stats;
int hx_learned_train(hx_ctx* ctx, const double* Ee, const double* f, hx_learn_stats* stats);
/* copy of the current weights (BF16 bit patterns) and their count */
int hx_learned_weights(hx_ctx* ctx, uint16_t* out, int* count);

/* ---- autotuner ---------------------------------------------------------------- */
typedef struct {
    int tile;         /* 0: 16x16 element columns per workgroup, 1: 16x32, 2: 32x8 */
    int kchunk;       /* z planes per workgroup */
    int variant64;    /* chosen FP64 contraction (HX_MV_*) */
    int variant32;    /* chosen FP32 contraction (HX_MV_*) */
    double ms64, ms32;                       /* best times of one operator application */
    double ms32_parity, ms32_dense, ms32_mfma; /* FP32 contraction candidates */
    int tile32, kchunk32; /* launch shape of the FP32 operator (tile, kchunk: FP64) */
    int pdepth, pdepth32; /* node planes in flight per workgroup, FP64 / FP32 operator */
    /* launch shapes of the V-cycle's fine-level pre- and post-smoothing
     * operators (HX_PC_VCYCLE; kchunk 0 = unset: the operator's shape) and
     * their best FP32 times; 0 when the grid has no 2:1 coarsening */
    int tile_pre, kchunk_pre, pdepth_pre, tile_post, kchunk_post, pdepth_post;
    double ms_pre, ms_post;
} hx_tune_result;
/* Time the operator's launch shapes (tile, z-chunk) and element contractions on
 * the current grid for nrhs columns and keep the fastest (HX_MV_AUTO then uses
 * them); likewise the V-cycle's smoothing operators' shapes.  No reference
 * counterpart. */
int hx_autotune(hx_ctx* ctx, int nrhs, hx_tune_result* result);
/* Average duration (ms) of one application of the PCG operator kernel
 * (fused p-update, FP64 or FP32 contraction per `precision`) at the current
 * launch shape on the current grid: `reps` back-to-back launches for nrhs
 * columns between two HIP events on the context stream, rotating over vector
 * sets larger together than the last-level cache.  No reference
 * counterpart. */
int hx_time_operator(hx_ctx* ctx, int precision, int nrhs, int reps, double* avg_ms);
/* The same for the operator form a preconditioner uses: HX_PC_JACOBI (fused
 * p = D^-1 r + beta p_old, as hx_time_operator) or HX_PC_MULTILEVEL /
 * HX_PC_VCYCLE (fused p = z + beta p_old from an explicit z, parity
 * contraction).  No reference counterpart. */
int hx_time_operator_pc(hx_ctx* ctx, int precision, int precond, int nrhs, int reps, double* avg_ms);
/* The fine level's V-cycle smoothing operator (HX_PC_VCYCLE): form 3 =
 * pre-smoothing (x = S r, residual r - K x), 4 = post-smoothing with the
 * coarse correction's prolongation fused; `reps` back-to-back launches at the
 * current launch shape and moduli, rotating over vector sets larger together
 * than the last-level cache.  The grid needs one 2:1 coarsening.  No reference
 * counterpart. */
int hx_time_smoothing(hx_ctx* ctx, int precision, int form, int reps, double* avg_ms);
/* Apply a tuning result (e.g. one cached from an earlier hx_autotune) without
 * timing anything: its launch shapes and contractions.  No reference
 * counterpart. */
int hx_tune_set(hx_ctx* ctx, const hx_tune_result* result);

/* ---- kernel timing ---------------------------------------------------------- */
enum { HX_K_MATVEC = 0, HX_K_UPDATE = 1, HX_K_PUPDATE = 2, HX_K_HELM = 3, HX_K_ELEM = 4,
       HX_K_MATVEC32 = 5, HX_K_PRECOND = 6, HX_K_COUNT = 7 };
/* 1: bracket each hot-kernel launch with HIP events on the context stream */
int hx_timing_enable(hx_ctx* ctx, int on);
/* average launch duration (ms) and launch count of kernel class `kernel`
 * since the last hx_timing_enable(ctx, 1) */
int hx_timing_read(hx_ctx* ctx, int kernel, double* avg_ms, long* count);
/* hardware facts used for roofline reporting */
int hx_device_info(hx_ctx* ctx, int* cu_count, int* clock_khz, double* mem_gb);
/* device memory in use and in total (GiB): the VRAM column of the iteration
 * log (Diagnostics.jl:32-39) */
int hx_device_memory(hx_ctx* ctx, double* used_gb, double* total_gb);

#ifdef __cplusplus
}
#endif
#endif /* HEXAT_H */
