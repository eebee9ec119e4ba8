/*
 * hyqmom.h -- C ABI of libhyqmom_hip.so, the MI355X (gfx950) implementation of the
 * HyQMOM.jl 3D moment-solver time step.
 *
 * The reference (comp-physics/HyQMOM.jl) has no FFI of its own; its public boundary for
 * this path is the Julia API
 *     simulation_runner(params)                 src/simulation_runner.jl:63-759
 *     run_simulation_with_snapshots(params; ..) src/simulation_runner.jl:994-1011
 * whose step loop (src/simulation_runner.jl:346-560) these entry points replace.  A Julia
 * caller keeps that API and binds the functions below with `ccall`.
 *
 * Conventions
 *   - every function returns HYQ_OK (0) or a negative HYQ_E* code; nothing throws or
 *     aborts across the ABI; hyq_last_error() gives the text of the last failure on the
 *     calling thread;
 *   - host state buffers use the reference's own layout: Julia column-major
 *     (Nx, Ny, Nz, 35), x fastest, moment index slowest;
 *   - "batch" entry points take DEVICE pointers to structure-of-arrays planes
 *     ([35][n], plane m at ptr + m*n) and run on the default stream of the current
 *     device; they synchronize before returning;
 *   - one context per host thread; calls on one context must be serialized.
 *
 * The 35-moment order is the reference's (Flux_closure35_and_realizable_3D.jl:18-20):
 *   M000 M100 M200 M300 M400 M010 M110 M210 M310 M020 M120 M220 M030 M130 M040
 *   M001 M101 M201 M301 M002 M102 M202 M003 M103 M004 M011 M111 M211 M021 M121
 *   M031 M012 M112 M013 M022
 */
#ifndef HYQMOM_H
#define HYQMOM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* return codes */
#define HYQ_OK 0
#define HYQ_EINVAL (-1)      /* invalid argument (sizes, null pointers, Nmom != 35) */
#define HYQ_ENOMEM (-2)      /* device allocation failed */
#define HYQ_EHIP (-3)        /* a HIP runtime call failed */
#define HYQ_ENONFINITE (-4)  /* non-finite Jacobian in the z wave-speed solve: the
                                reference throws there (eigenvalues6z_hyperbolic_3D.jl:108,121) */
#define HYQ_ENOTCONV (-5)    /* an eigenvalue iteration did not converge */
#define HYQ_ENODEV (-6)      /* no HIP device */

/* per-cell branch flags (bitmask), accumulated over one pass-1 evaluation */
#define HYQ_FLAG_UNIVARIATE (1u << 0)   /* enforce_univariate changed S3/S4 */
#define HYQ_FLAG_EDGE_CORNER (1u << 1)  /* edge_corner_correction branch */
#define HYQ_FLAG_FACES (1u << 2)        /* realizable_3D faces branch (S2 <= 1e-12) */
#define HYQ_FLAG_FLAG220 (1u << 3)      /* the 6-pass flag220 loop ran */
#define HYQ_FLAG_ROOTS2D (1u << 4)      /* realizable_2D L3 < 0 root bound evaluated */
#define HYQ_FLAG_ROOTS4 (1u << 5)       /* a 4x4 minor of Delta2* was negative */
#define HYQ_FLAG_EIGCORR_X (1u << 6)    /* complex x eigenvalues -> moment correction */
#define HYQ_FLAG_EIGCORR_Y (1u << 7)
#define HYQ_FLAG_EIGCORR_Z (1u << 8)
#define HYQ_FLAG_ZNEGVAR (1u << 9)      /* negative-variance re-projection before z eigs */
#define HYQ_FLAG_NONFINITE (1u << 10)   /* non-finite z Jacobian */
#define HYQ_FLAG_RHOCLAMP (1u << 11)    /* density clamp in the z eigen path */
#define HYQ_FLAG_EIG_FALLBACK (1u << 12) /* learned warm start fell back to the full solve */
#define HYQ_FLAG_QR_NOCONV (1u << 13)   /* QR iteration limit reached */
/* bits 16..27: root-polishing (Laguerre) iterations summed over the cell's fast plane
   solves (precision HYQ_PREC_MIXED_NEWTON / HYQ_PREC_MIXED_LEARNED; saturates at 4095) */
#define HYQ_FLAG_NEWTON_SHIFT 16
#define HYQ_FLAG_NEWTON_MASK 0xFFFu

/* precision modes */
#define HYQ_PREC_FP64 0           /* everything FP64 */
#define HYQ_PREC_MIXED 1          /* FP32 fluxes, HLL update and BGK relaxation */
#define HYQ_PREC_MIXED_LEARNED 2  /* MIXED + extreme wave speeds by a Sturm test and Laguerre
                                     iteration seeded by the warm-start network */
#define HYQ_PREC_MIXED_NEWTON 3   /* the same seeded by the analytic root bounds (the learned
                                     mode without its network) */

/* Mirrors the params NamedTuple of simulation_runner (src/simulation_runner.jl:20-37,
 * 74-113).  Fields the device step does not need (snapshot, symmetry, debug, MPI) stay
 * on the host side. */
typedef struct hyq_params {
  int32_t Nx, Ny, Nz;
  int32_t Nmom;       /* must be 35 */
  int32_t flag2D;
  int32_t positivity; /* POSITIVITY_ENABLED (src/HyQMOM.jl:20) */
  int32_t precision;  /* HYQ_PREC_* */
  int32_t reserved;
  double tmax, Kn, Ma, CFL, dtmax;
  double xmin, xmax, ymin, ymax, zmin, zmax;
} hyq_params;

typedef struct hyq_ctx hyq_ctx;

/* Context: device SoA planes, stream and scratch for one run (at most 2^29 - 1 cells). */
int hyq_create(const hyq_params* params, hyq_ctx** out);
int hyq_destroy(hyq_ctx* ctx);

/* Upload / download the (Nx,Ny,Nz,35) state (host, Julia column-major). */
int hyq_set_state(hyq_ctx* ctx, const double* M);
int hyq_get_state(hyq_ctx* ctx, double* M);

/* One step of the reference loop body (simulation_runner.jl:346-560): pass 1 (closure,
 * wave speeds), dt = min(CFL*min(dx,dy,dz)/vmax, min(dtmax,Kn), tmax - t), x/y/z HLL
 * updates combined as Mx+My+Mz-2M, pass 2 realizability and BGK.  t and nn advance as
 * in the reference.  Returns HYQ_ENONFINITE (with t, nn and the state of the completed
 * device step, which then carries NaNs) where the reference's z eigen solve throws. */
int hyq_step(hyq_ctx* ctx, double* t_inout, double* dt_out, int64_t* nn_inout);

/* while t < tmax && nn < nnmax: hyq_step.  The loop condition of
 * simulation_runner.jl:346. */
int hyq_run(hyq_ctx* ctx, double tmax, int64_t nnmax, double* t_inout, int64_t* nn_inout);

/* Pass-1 outputs of the last step: wave speeds [6][N] (xmin,xmax,ymin,ymax,zmin,zmax; the
 * reference's vp{x,y,z}{min,max}, simulation_runner.jl:374-379) and branch flags [N]. */
int hyq_get_wave_speeds(hyq_ctx* ctx, double* vp);
int hyq_get_flags(hyq_ctx* ctx, uint32_t* flags);

/* Average device time per step of pass 1 (closure, wave speeds, fluxes; in the learned mode
 * also the network), of the x/y/z update and of pass 2 (realizability, BGK), over the steps
 * since the last reset (HIP events on the context stream).  Any output may be NULL. */
int hyq_kernel_times(hyq_ctx* ctx, double* ms_pass1, double* ms_update, double* ms_pass2,
                     int64_t* nsteps);
int hyq_reset_kernel_times(hyq_ctx* ctx);

/* Cells of the last step that the fast modes (mixed_newton, mixed_learned) did not finish in
 * their fast kernel, in pass 1 and pass 2: cells with a full plane solve plus cells on the
 * full path (eigenvalue corrections, z-path clamps, non-finite planes); 0 in the other modes.
 * No reference counterpart (the reference has a single path). */
int hyq_deferred_cells(hyq_ctx* ctx, int64_t* pass1, int64_t* pass2);

/* Why cells of the last step left the fast kernels (fast modes; zeros otherwise), counts[8]:
 * pass 1 [0] cells with planes the fast solve could not decide, [1] such plane solves (QR),
 * [2] cells a plane solve found complex (eigenvalue correction: full path), [3] cells on the
 * full path from the start (z density clamp or re-projection, non-finite plane); [4..7] the
 * same for pass 2.  No reference counterpart. */
int hyq_fast_path_counts(hyq_ctx* ctx, int64_t* counts);

/* The snapshot fields of the current state on the device: the central moments C
 * (compute_central_field) and the standardized moments S (compute_standardized_field,
 * src/utils/compute_standardized_field.jl:16-62; M2CS4_35 per cell), host outputs in the
 * state's (Nx,Ny,Nz,35) layout; either may be NULL.  Call between steps (the device
 * scratch it uses is rewritten by every step). */
int hyq_snapshot_fields(hyq_ctx* ctx, double* C, double* S);

/* The runner's symmetry diagnostic (simulation_runner.jl:565-628) on the current state:
 * the diagonal cells (i, i, 1), i <= Nd = min(Nx, Ny), against (Nd+1-i, Nd+1-i, 1) for the
 * first five moments (differences for moments 1, 3, 5, sums for 2 and 4); per moment
 * max(Diff) / (norm(Diff) + 1) into per_moment[5] (may be NULL), and the largest magnitude
 * of the five into *maxdiff.  A device reduction. */
int hyq_symmetry_maxdiff(hyq_ctx* ctx, double* maxdiff, double* per_moment);

/* The first-step diagnostics of simulation_runner.jl:404-494 on the last step's wave
 * speeds: vdir[3] = max |speed| in x, y and z (vx_max, vy_max, vz_max; NaN if any speed is
 * NaN), and worst[3] (may be NULL) = the 0-based (i, j, k) of the cell with the largest
 * |speed|, the first in the reference's search order (k, then i, then j), cells with a NaN
 * speed skipped as the reference's `>` test skips them (-1s if every cell has one).  Device
 * reductions. */
int hyq_vmax_breakdown(hyq_ctx* ctx, double* vdir, int64_t* worst);

/* Batched per-cell kernels on device SoA planes ([35][n] doubles). */

/* Flux_closure35_and_realizable_3D (src/numerics/Flux_closure35_and_realizable_3D.jl:33-234)
 * for n independent cells: realizable Mr and fluxes Fx, Fy, Fz (any output may be NULL). */
int hyq_closure_batch(const double* M, int64_t n, double Ma, int32_t flag2D, double* Mr,
                      double* Fx, double* Fy, double* Fz, uint32_t* flags);

/* Moments5_3D (src/moments/Moments5_3D.jl:16-138) for n independent cells: the HyQMOM
 * closure of the 5th-order moments without realizability projection.  M is [35][n]; M5, C5
 * and S5 are [21][n], rows in the reference's order 500 410 320 230 140 401 302 203 104 311
 * 221 131 212 113 122 050 041 032 023 014 005. */
int hyq_moments5_batch(const double* M, int64_t n, double* M5, double* C5, double* S5);

/* The pass-1 cell pipeline of simulation_runner.jl:352-380 for n independent cells:
 * closure -> x/y/z wave speeds (eigenvalues6_hyperbolic_3D.jl, eigenvalues6z_hyperbolic_3D.jl)
 * -> closure -> 1D closures (closure_and_eigenvalues.jl).  vp is [6][n]; Fx, Fy, Fz must be
 * consecutive [35][n] blocks.  precision HYQ_PREC_FP64: QR eigen solves;
 * HYQ_PREC_MIXED_NEWTON: the fast step's pass-1 kernels (Sturm test, 2x2 block test of a
 * nearly double pair, Laguerre extremes, QR where those do not decide, full path for the
 * eigenvalue corrections), with scratch the call allocates; needs flags.  Both give FP64
 * outputs. */
int hyq_wavespeed_batch(const double* M, int64_t n, double Ma, int32_t flag2D,
                        int32_t positivity, int32_t precision, double* Mr, double* Fx,
                        double* Fy, double* Fz, double* vp, uint32_t* flags);

/* hyq_wavespeed_batch without any allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  The fast mode's scratch lives in a fixed device array
 * that every GPU holds until the process ends; at most 65536 cells per call (HYQ_EINVAL
 * beyond), and calls that overlap on one GPU race.  Same arguments and results as
 * hyq_wavespeed_batch. */
int hyq_wavespeed_batch_noalloc(const double* M, int64_t n, double Ma, int32_t flag2D,
                                int32_t positivity, int32_t precision, double* Mr, double* Fx,
                                double* Fy, double* Fz, double* vp, uint32_t* flags);

/* eigvals of jacobian6 (src/autogen/jacobian6.jl, compute_jacobian_eigenvalues.jl:18-50)
 * for n 15-moment slices ([15][n]); outputs the 6 eigenvalues sorted by real part
 * ([6][n] real, [6][n] imaginary). */
int hyq_eig6_batch(const double* m15, int64_t n, double* wr, double* wi);

/* closure_and_eigenvalues (src/numerics/closure_and_eigenvalues.jl:17-78): min/max real
 * eigenvalue of the 1D HyQMOM Jacobi matrix of n 5-moment sets m5 [5][n]. */
int hyq_closure1d_batch(const double* m5, int64_t n, double* vmin, double* vmax);

/* pas_HLL (src/numerics/pas_HLL.jl:24-84 with flux_HLL.jl:20-32) on nlines independent lines
 * of np cells: M, F, Mout [35][nlines*np] (cell s of line l at l*np + s), wave-speed bounds
 * vmin, vmax [nlines*np], Mout = the line's updated state with the reference's BC copies;
 * np at most 2^29 - 1. */
int hyq_hll_lines(const double* M, const double* F, const double* vmin, const double* vmax,
                  int64_t nlines, int32_t np, double dt, double dx, double* Mout);

/* collision35 (src/numerics/collision35.jl:17-45) on n cells. */
int hyq_collision_batch(const double* M, int64_t n, double dt, double Kn, int32_t positivity,
                        double* Mout);

/* ---- learned warm start (a 15 -> H -> H -> H -> 2 GELU MLP on bf16 MFMA, H = 32 or 128) ---
 * No counterpart in the reference: the network predicts the extreme eigenvalues of the
 * standardized flux Jacobian of a plane from its 15 standardized moments; the predictions
 * only seed a Laguerre iteration whose result is exact to FP64 or falls back to the full solve.
 * Parameters travel as one FP32 blob of hyq_mlp_blob_size_hidden(H) floats (weights, biases,
 * input and label normalization); hyq_mlp_param_count_hidden(H) = 2 H^2 + 35 H + 16.  The
 * hidden width is chosen when a network is created; the calls without it mean H = 32. */
typedef struct hyq_mlp hyq_mlp;
int64_t hyq_mlp_blob_size(void);
int64_t hyq_mlp_param_count(void);
int64_t hyq_mlp_blob_size_hidden(int32_t hidden);   /* -1 for an unsupported width */
int64_t hyq_mlp_param_count_hidden(int32_t hidden); /* -1 for an unsupported width */
int hyq_mlp_create(const float* blob_host /* NULL: zeros */, hyq_mlp** out);
int hyq_mlp_create_hidden(int32_t hidden /* 32 or 128 */, const float* blob_host, hyq_mlp** out);
int32_t hyq_mlp_hidden(hyq_mlp* m);
int hyq_mlp_destroy(hyq_mlp* m);
int hyq_mlp_get_blob(hyq_mlp* m, float* blob_host);
int hyq_mlp_set_blob(hyq_mlp* m, const float* blob_host);
/* X [n][16] features (device), P [n][2] predictions (device): the fused inference kernel. */
int hyq_mlp_infer(hyq_mlp* m, const float* X, int64_t n, float* P);
/* the same through the layer-by-layer training kernels (bitwise equal; a cross-check) */
int hyq_mlp_infer_layers(hyq_mlp* m, const float* X, int64_t n, float* P);
/* launch shape of the fused inference: -1 (the default) times the shapes on the first
 * inference of 65536 rows or more and keeps the fastest; 0..2 fixes one (0: 16 waves x 1
 * row tile, 1: 8 waves x 2 row tiles, 2: 8 waves x 1 row tile).  All give the same values.
 * No reference counterpart. */
int hyq_mlp_set_shape(hyq_mlp* m, int32_t shape);
/* loss and gradients (grad_host: hyq_mlp_param_count() floats, may be NULL); T [n][2]. */
int hyq_mlp_grad(hyq_mlp* m, const float* X, const float* T, int64_t n, float* loss_out,
                 float* grad_host);
/* one Adam step (beta 0.9/0.999, eps 1e-8) on a batch */
int hyq_mlp_train_step(hyq_mlp* m, const float* X, const float* T, int64_t n, float lr,
                       float* loss_out);
const char* hyq_mlp_last_error(void);
/* attach a network to a HYQ_PREC_MIXED_LEARNED context (not owned; NULL detaches, which
 * makes the context use analytic starts) */
int hyq_set_mlp(hyq_ctx* ctx, hyq_mlp* mlp);
/* training data: features X [5n][16] of the five plane Jacobians (UV, UW, VU, VW, WU; sample
 * = plane*n + cell) of n states M [35][n], and optionally labels L [5n][2] (standardized
 * extreme eigenvalues from the full FP64 solve) with valid [5n] (1: real and finite). */
int hyq_plane_dataset(const double* M, int64_t n, float* X, float* L, uint8_t* valid);

/* ---- ensemble: independent runs advanced together --------------------------------------
 * Runs with their own grid, Ma, Kn, CFL and tmax (one precision for all: FP64, MIXED or
 * MIXED_NEWTON) share SoA planes and one HIP-graph-captured step; each run keeps its own
 * CFL dt and loop condition t < tmax.  No counterpart in the reference, which runs one
 * simulation per process (simulation_runner.jl:63). */
typedef struct hyq_ensemble hyq_ensemble;
int hyq_ensemble_create(const hyq_params* runs, int32_t nruns, hyq_ensemble** out);  // at most 2^29 - 1 cells in all
int hyq_ensemble_destroy(hyq_ensemble* e);
/* state of one run, host Julia layout (Nx,Ny,Nz,35); set_state resets its t and nn */
int hyq_ensemble_set_state(hyq_ensemble* e, int32_t run, const double* M);
int hyq_ensemble_get_state(hyq_ensemble* e, int32_t run, double* M);
/* per-run t, last dt, steps taken (arrays of nruns; any may be NULL) */
int hyq_ensemble_status(hyq_ensemble* e, double* t, double* dt, int64_t* nn);
/* nsteps steps (graph replays if use_graph, else direct launches); synchronizes */
int hyq_ensemble_run(hyq_ensemble* e, int64_t nsteps, int32_t use_graph);
/* block sizes of pass 1, update, pass 2 (each 64, 128 or 256) */
int hyq_ensemble_set_blocks(hyq_ensemble* e, const int32_t* blocks);
/* autotuner: times each block-size candidate over trial_steps steps from the current
 * state (restored afterwards) and keeps the fastest per kernel; blocks_out [3],
 * ms_out [3 candidates][3 kernels] (either may be NULL) */
int hyq_ensemble_tune(hyq_ensemble* e, int32_t trial_steps, int32_t* blocks_out, double* ms_out);
const char* hyq_ensemble_last_error(void);

/* Text of the last error on the calling thread ("" if none). */
const char* hyq_last_error(void);

/* Library version string. */
const char* hyq_version(void);

#ifdef __cplusplus
}
#endif
#endif /* HYQMOM_H */
