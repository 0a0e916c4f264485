/* C-ABI of the MI355X-native DiReBM step engine (libdirebm_hip.so, gfx950).
 *
 * Every entry point returns DRB_OK (0) or an error code; drb_last_error() gives the message of
 * the last failure on the calling thread.  Device pointers are plain HIP device addresses, sizes
 * are element counts, streams are hipStream_t passed as void*.  No entry point throws.
 *
 * Each declaration names the reference function (berfandil/DiReBM, file:line) it replaces.
 * The reference reaches these stages from Python through Warp launches; the drop-in binding is
 * the ctypes layer in direbm_amd/_lib.py.
 */
#ifndef DIREBM_HIP_H
#define DIREBM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  DRB_OK = 0,
  DRB_ERR_INVALID_ARG = 1, /* bad size, pointer or parameter */
  DRB_ERR_CAPACITY = 2,    /* a size limit of the engine was exceeded */
  DRB_ERR_KEY_RANGE = 3,   /* cell index range too wide for 32-bit keys */
  DRB_ERR_NONFINITE = 4,   /* NaN / Inf met in the state */
  DRB_ERR_HIP = 5,         /* a HIP runtime call failed */
  DRB_ERR_STATE = 6        /* call out of order (e.g. step before set_state) */
};

typedef void* drb_stream_t; /* hipStream_t */

/* Message of the last failure on this thread ("" if none). */
const char* drb_last_error(void);
/* ABI version (major * 100 + minor). */
int drb_version(void);

/* ---------------------------------------------------------------- physics stages
 * f is (n, 7) addressed as f[i * stride_i + k * stride_k] (row-major AoS: stride_i = 7,
 * stride_k = 1; SoA planes: stride_i = 1, stride_k = plane pitch). */

/* BGK collide in place, fp32.  Replaces _collide / collide(), direbm/warp/physics.py:45-62,78-85.
 * rho == 0 gives u = 0 (the guard of direbm/physics.py:24).  The rest equilibrium is the remainder
 * rho - sum_{k>0} f_eq_k (equal to the formula in exact arithmetic): the fp32 weights sum to
 * 1 + 1.5e-8, which would otherwise add 1.5e-8 rho / tau of mass per collide. */
int drb_collide(float* f, int64_t n, int64_t stride_i, int64_t stride_k, float tau, drb_stream_t s);

/* out[i, k] = f_eq(rho[i], u[i]).  u is (n, 2) row-major.
 * Replaces _equilibrium / equilibrium(), direbm/warp/physics.py:65-75,88-97. */
int drb_equilibrium(const float* rho, const float* u, int64_t n, float* out, int64_t stride_i,
                    int64_t stride_k, drb_stream_t s);

/* Adjoint of drb_collide: gin = (d collide / d f)^T gout and, when gtau_i is not NULL,
 * gtau_i[i] = gout_i . d collide_i / d tau (FP64, per point).  f is the pre-collision input; gout,
 * gin share f's strides.  No reference counterpart (the reference is not differentiated). */
int drb_collide_vjp(const float* f, int64_t n, int64_t stride_i, int64_t stride_k, float tau,
                    const float* gout, float* gin, double* gtau_i, drb_stream_t s);

/* Adjoint of drb_equilibrium: grho[i] = g_i . d f_eq / d rho, gu[i] = g_i . d f_eq / d u (either
 * may be NULL).  g shares out's strides.  No reference counterpart. */
int drb_equilibrium_vjp(const float* rho, const float* u, int64_t n, const float* g, int64_t stride_i,
                        int64_t stride_k, float* grho, float* gu, drb_stream_t s);

/* ------------------------------------------------------------ propagation stages */

/* Components of m moments: pos_c[7i+k] = pos_m[i] + C[k]*dx, val_c = f_m[i,k], dir_c = k.
 * pos_m, pos_c are (n, 2) row-major, f_m is (m, 7) row-major.
 * Replaces _disperse / disperse(), direbm/warp/propagation.py:21-36,115-127. */
int drb_disperse(const float* pos_m, const float* f_m, int64_t m, float dx, float* pos_c,
                 float* val_c, int32_t* dir_c, drb_stream_t s);

/* Workspace for the stage calls below (sort buffers, cell grids).  Not thread-safe. */
typedef struct drb_stage drb_stage;
int drb_stage_create(drb_stage** out);
int drb_stage_destroy(drb_stage* st);

/* Cell-thinning: one control point per occupied cell of size cs, at the fp32 centroid of its
 * components in ascending component order; control points in ascending cell-key order.
 * Writes *p_out (host) and cp_out[0 .. P) as (P, 2); fails with DRB_ERR_CAPACITY if P > cp_cap.
 * Replaces _cell_key/_centroid/create_control_points(), direbm/warp/propagation.py:39-68,130-153
 * (warp.utils.radix_sort_pairs / runlength_encode / array_scan). */
int drb_create_control_points(drb_stage* st, const float* pos_c, int64_t nc, float cs, float* cp_out,
                              int64_t cp_cap, int64_t* p_out, drb_stream_t s);

/* Stable LSD radix sort of (key, value) pairs by the low `bits` bits of the keys (1..32; keys must
 * lie below 2^bits), ties kept in input order.  keys and vals are device arrays of 2n elements,
 * as warp.utils.radix_sort_pairs takes them: the first n hold the input and receive the result,
 * the upper n are scratch.  The digit histograms live in the stage workspace.
 * Replaces warp.utils.radix_sort_pairs (warp-lang 1.14.0), called at
 * direbm/warp/propagation.py:138. */
int drb_radix_sort_pairs(drb_stage* st, uint32_t* keys, uint32_t* vals, int64_t n, int32_t bits, drb_stream_t s);

/* Size of the static device array behind drb_create_control_points_noalloc, per GPU. */
#define DRB_NOALLOC_SCRATCH_BYTES (64u << 20)

/* Allocation-free variant of drb_create_control_points: the same control points, bit for bit,
 * with no device memory allocated by the call (no hipMalloc / hipFree, no allocating library
 * routine).  No counterpart in the reference: it is for callers that cannot allocate device
 * memory -- tight serving loops and harnesses that forbid hipMalloc.  The caller passes the
 * inputs and outputs; every intermediate lives in one static device array of
 * DRB_NOALLOC_SCRATCH_BYTES (64 MiB), which each GPU the library runs on holds until the process
 * ends.  A call needs about 44 bytes per component plus 16 per cell of the components' bounding
 * grid (cell size cs): larger inputs fail with DRB_ERR_CAPACITY before any kernel runs on them.
 * Every call on one GPU shares that GPU's array, so calls that overlap on one GPU (from other
 * threads or streams) race; calls in sequence are independent (the call clears what must start
 * at zero).  Like drb_create_control_points it synchronises the stream to read the count, so it
 * cannot be captured in a HIP graph. */
int drb_create_control_points_noalloc(const float* pos_c, int64_t nc, float cs, float* cp_out, int64_t cp_cap,
                                      int64_t* p_out, drb_stream_t s);

/* Refine: kappa = distinct directions among components within dx; kappa <= kappa_hard keeps the
 * point, else it moves to the exp(val)-weighted centroid.  dir_c values must lie in 0..6
 * (DRB_ERR_INVALID_ARG otherwise).
 * Replaces _refine / refine_control_points(), direbm/warp/propagation.py:77-112,249-275. */
int drb_refine_control_points(drb_stage* st, const float* cp, int64_t p, const float* pos_c,
                              const float* val_c, const int32_t* dir_c, int64_t nc, float dx,
                              int kappa_hard, float* new_pos, drb_stream_t s);

/* Resample: deterministic scatter of component values into control points within dx (weights
 * (dx - d) / sum), then gap-fill of empty directions (< 1e-6) from f_rest and emission of one
 * moment per control point.  out_pos (p, 2), out_f (p, 7) row-major.  dir_c values must lie in
 * 0..6 (DRB_ERR_INVALID_ARG otherwise).
 * Replaces _scatter/_gapfill_emit/resample(), direbm/warp/propagation.py:156-246. */
int drb_resample(drb_stage* st, const float* pos_c, const float* val_c, const int32_t* dir_c,
                 int64_t nc, const float* cp, int64_t p, float dx, const float* f_rest, float* out_pos,
                 float* out_f, drb_stream_t s);

/* Mass and momentum of m moments binned onto an nb x nb grid with bin edges edges[0..nb] (FP64,
 * the same edges in x and y): mass[iy*nb + ix] = sum sum_k f_k, mom[2*(iy*nb + ix) + {0,1}] =
 * sum sum_k f_k c_k (FP64 c), summed in ascending moment order.  A coordinate's bin is the number
 * of edges <= it minus one, with the last edge in the last bin; moments outside are skipped.
 * pos (m, 2), f (m, 7) row-major fp32.  Replaces bin_fields, direbm/fields.py:17-41 (which then
 * divides by the cell area and forms u = mom / mass). */
int drb_bin_fields(drb_stage* st, const float* pos, const float* f, int64_t m, const double* edges, int32_t nb,
                   double* mass, double* mom, drb_stream_t s);

/* Adjoint of drb_bin_fields w.r.t. f: gf[i, k] = gmass[bin_i] + gmom[bin_i] . c_k (0 outside the
 * grid); gmass or gmom may be NULL.  No reference counterpart. */
int drb_bin_fields_vjp(const float* pos, int64_t m, const double* edges, int32_t nb, const double* gmass,
                       const double* gmom, float* gf, drb_stream_t s);

/* ------------------------------------------------------------------- simulator */

/* Step parameters.  The first five fields are the GpuSimulator constructor arguments
 * (direbm/warp/simulator.py:26-47); the rest extend it (0 = reference behaviour). */
typedef struct {
  float tau;        /* BGK relaxation time */
  float dx;         /* dispersion length */
  float alpha;      /* thinning cell = dx / alpha */
  int32_t kappa_hard;
  float rho_rest;   /* gap-fill rest density, f_rest = rho_rest * W */
  float r_res;      /* resample radius; 0 -> dx */
  float eps;        /* gap-fill emptiness threshold; 0 -> 1e-6 */
  int32_t periodic; /* 1: periodic square box [0, box)^2 */
  float box;
  int32_t two_level; /* 1: cells of 2*cs outside radius r_in of (cx, cy) */
  float cx, cy, r_in;
  /* > kappa_hard: the per-object solver's soft_outer spawn (direbm/reference/simulator.py:141-166):
   * a control point with kappa_hard < kappa <= kappa_soft adds one point at its pre-refine position
   * + n 2 (1 - sqrt(3)/2) dx, n the unit sum of the directions of its components within dx (none
   * when the sum is zero); spawned points follow the refined ones and are not refined.  0: off
   * (the reference GPU step treats soft_outer points as inner, direbm/warp/propagation.py:108).
   * 2D single simulations only (n_inst = 1), not with obstacles, graph mode or recorded steps. */
  int32_t kappa_soft;
} drb_params;

typedef struct drb_sim drb_sim;

/* A simulation of n_inst independent instances (n_inst = 1 is GpuSimulator).
 * Replaces GpuSimulator.__init__, direbm/warp/simulator.py:26-48. */
int drb_sim_create(drb_sim** out, const drb_params* prm, int32_t n_inst, drb_stream_t s);
int drb_sim_destroy(drb_sim* sim);

/* Load the state from host memory: pos (m, 2), f (m, 7) row-major fp32; counts[n_inst] gives the
 * moments of each instance, stored one instance after another (NULL when n_inst = 1). */
int drb_sim_set_state(drb_sim* sim, const float* pos, const float* f, int64_t m, const int64_t* counts);

/* As drb_sim_set_state / drb_sim_get_state with device pointers (row-major fp32).  No reference
 * counterpart (the reference moves state through host numpy arrays, simulator.py:31-36,75-77). */
int drb_sim_set_state_device(drb_sim* sim, const float* pos, const float* f, int64_t m, const int64_t* counts);
int drb_sim_get_state_device(drb_sim* sim, float* pos, float* f);

/* Advance nsteps steps (collide, disperse, thin, refine, resample); returns when they are complete.
 * Replaces GpuSimulator.step, direbm/warp/simulator.py:50-73.  Within one call each step's next
 * layout is read back while its resample still runs, so the host queues step t+1 behind step t
 * and the GPU does not wait on the host between them; a non-finite state stops the call at the
 * step whose read back shows it. */
int drb_sim_step(drb_sim* sim, int32_t nsteps);

/* Autotuner: advances one ordinary step, then times the refine and gather kernels on that step's
 * buffers at five tiles -- (lanes per point, candidates in flight per lane) = (2, 4), (4, 2),
 * (4, 4), (8, 2), (8, 4) -- and keeps the fastest of each for the following steps.  lanes[0],
 * lanes[1] receive the refine and gather lane choices (drb_sim_get_tiles gives the whole tiles).
 * The tile sets the order of each point's partial sums: a tuned simulation is deterministic, but
 * may differ from an untuned one in the last bits.  No reference counterpart. */
int drb_sim_autotune(drb_sim* sim, int32_t* lanes);
/* Set the lanes per point of the refine and gather kernels (2, 4 or 8; default 4), keeping their
 * depths, e.g. to reproduce a tuned run bit for bit.  No reference counterpart. */
int drb_sim_set_lanes(drb_sim* sim, int32_t refine, int32_t gather);
/* Set the whole tiles of the refine and gather kernels: lanes per point (2, 4 or 8) and candidates
 * in flight per lane (2 or 4; default 4).  No reference counterpart (tile parameters of this
 * engine; results depend on them only through the order of fp32 partial sums). */
int drb_sim_set_tiles(drb_sim* sim, int32_t refine_lanes, int32_t refine_depth, int32_t gather_lanes,
                      int32_t gather_depth);
/* The current tiles: out[4] = refine lanes, refine depth, gather lanes, gather depth.  No reference
 * counterpart. */
int drb_sim_get_tiles(drb_sim* sim, int32_t* out);

/* Graph mode: drb_sim_step replays captured HIP graphs of one step, with the live counts kept on
 * the device, so a call of n steps synchronises the host once.  Capacities: a periodic grid is
 * fixed (at most one control point per thinning cell, five with two-level thinning); an open
 * instance's grid follows its bounding box, so each step lays its grid out on the device, within a
 * cell capacity sized for the largest growth of drb_sim_set_graph_horizon steps (a moment moves at
 * most dx per step).  A step whose grid would exceed the capacity runs empty, and the call resumes
 * from it with a larger capacity (a new capture).  Results are bit-identical to plain steps.
 * Recording (drb_sim_tape) and stage timing fall back to plain steps; obstacles and the soft_outer
 * spawn are refused (DRB_ERR_INVALID_ARG).  Replaces the reference's per-step host read of the
 * point count (direbm/warp/propagation.py:144-146) for whole multi-step calls. */
int drb_sim_set_graph(drb_sim* sim, int32_t on);
/* Steps of the largest growth the open-domain graph capacity covers (default 16; at least 1):
 * more steps hold more memory and capture less often.  No reference counterpart. */
int drb_sim_set_graph_horizon(drb_sim* sim, int32_t steps);
/* Graph mode bookkeeping: out[4] = captures so far, replays resumed after an empty step, cells and
 * control points of the current capture's capacity.  No reference counterpart. */
int drb_sim_graph_info(drb_sim* sim, int64_t* out);

/* Circular obstacle of instance inst (fluid outside), centre (cx, cy), radius r; r <= 0 removes it.
 * A component k != 0 that would land inside bounces specularly: FP64 first crossing of the segment
 * from its moment, c_k reflected about the outward normal and split over the two adjacent
 * directions by angle, each part at its moment's position + c_j dx (the moment's position when
 * that is inside too), value f_k w_j.  Control points inside are removed after resampling.  A
 * periodic instance's circle must keep dx clear of the box edges.  Not supported in graph mode
 * (drb_sim_set_graph) or recorded steps (drb_sim_tape): DRB_ERR_INVALID_ARG.
 * Replaces the v1 obstacle handling, direbm/reference/simulator.py:79-80,90-117 with Circle,
 * reflect and split_direction, direbm/reference/boundary.py:20-70. */
int drb_sim_set_obstacle(drb_sim* sim, int32_t inst, float cx, float cy, float r);

/* Current moment count, total and per instance (counts may be NULL). */
int drb_sim_count(drb_sim* sim, int64_t* m, int64_t* counts);

/* Copy the state to host memory (pos (m, 2), f (m, 7) row-major).
 * Replaces GpuSimulator.moments, direbm/warp/simulator.py:75-77. */
int drb_sim_get_state(drb_sim* sim, float* pos, float* f);

/* FP64 totals of the current state per instance: out[5*b + {0..4}] = mass, momentum x, momentum y,
 * mass dropped by the last step (components with no control point in range), mass injected by
 * the last step's gap-fill.  No reference counterpart (the reference has no diagnostics). */
int drb_sim_diagnostics(drb_sim* sim, double* out);

/* HIP-event timing of the last step's stages in ms, recorded on the engine's stream when enabled
 * with drb_sim_set_timing(sim, 1): ms9 = {collide, keys, sort, thin (mark + scan + centroid),
 * refine, bin (the refined points binned by their cell), scatter normaliser, scatter gather,
 * gap-fill + emit}.  on = 0x200 | mask times only the stages in mask (bit i: ms9[i]; the others
 * read 0): an event between two kernels leaves the GPU idle for a few microseconds, so a run that
 * times one stage records two events per step instead of ten.  No reference counterpart. */
int drb_sim_set_timing(drb_sim* sim, int32_t on);
int drb_sim_stage_ms(drb_sim* sim, float* ms9);

/* Point-updates performed since creation: the moments entering each step, summed over steps (the
 * numerator of MPUPS over any run of steps, e.g. one drb_sim_step(sim, n) call).  No reference
 * counterpart. */
int drb_sim_updates(drb_sim* sim, int64_t* out);

/* ------------------------------------------------------------------- adjoint
 * No reference counterpart: the reference step is not differentiated.  The discrete choices of a
 * step (thinning runs, refine hard/soft decision, pairs within the radii, empty directions) are
 * held fixed as recorded; every continuous dependence is differentiated. */

/* Record the next nslots steps: step j after this call keeps its intermediates in slot j (device
 * memory of about 250 bytes per moment per slot; slots stay allocated for later recordings).
 * nslots = 0 stops recording and frees the slots. */
int drb_sim_tape(drb_sim* sim, int32_t nslots);

/* Moments entering (m_in) and leaving (p_out) the step recorded in a slot. */
int drb_sim_tape_info(drb_sim* sim, int32_t slot, int64_t* m_in, int64_t* p_out);

/* Vector-Jacobian product of the recorded step: given the cotangents of its output state
 * (gpos_out (p_out, 2), gf_out (p_out, 7), device, row-major; gpos_out may be NULL = 0), writes
 * the cotangents of its input state (gpos_in (m_in, 2) may be NULL, gf_in (m_in, 7)) and adds the
 * d/d tau of each instance to gtau[n_inst] (host, FP64; may be NULL). */
int drb_sim_step_vjp(drb_sim* sim, int32_t slot, const float* gpos_out, const float* gf_out, float* gpos_in,
                     float* gf_in, double* gtau);

/* Vector-Jacobian product of all steps recorded since drb_sim_tape, last to first, with one host
 * synchronisation: cotangents of the last recorded step's output in, of the first recorded step's
 * input out (shapes and NULL rules as drb_sim_step_vjp). */
int drb_sim_tape_vjp(drb_sim* sim, const float* gpos_out, const float* gf_out, float* gpos_in, float* gf_in,
                     double* gtau);

/* Evaluate the step recorded in `slot` at another input state (pos_in (m_in, 2), f_in (m_in, 7),
 * device, row-major; tau > 0 replaces the recorded tau) with the recorded discrete choices held
 * fixed -- the function whose derivative drb_sim_step_vjp computes.  Writes pos_out (p_out, 2)
 * (unwrapped, near the recorded positions) and f_out (p_out, 7).  Feeding a slot's output into the
 * next slot replays a recorded rollout, e.g. for central differences of its gradient. */
int drb_sim_replay(drb_sim* sim, int32_t slot, float tau, const float* pos_in, const float* f_in, float* pos_out,
                   float* f_out);

/* drb_sim_replay evaluated in FP64: the same function of the input state (the recorded discrete
 * choices held fixed, the engine's fp32 velocity set and weights), every value computed in double.
 * Arrays are FP64 with the shapes of drb_sim_replay.  Central differences of this function are
 * free of fp32 rounding, so they check the adjoint's gradient at small steps (eps 1e-3).
 * No reference counterpart. */
int drb_sim_replay64(drb_sim* sim, int32_t slot, double tau, const double* pos_in, const double* f_in,
                     double* pos_out, double* f_out);

/* ------------------------------------------------------------------- 3D (D3Q13) */
/* The step in 3D on the icosahedral D3Q13 lattice (rest + 12 unit vertex directions, W = [2/5,
 * 1/20 x 12], cs^2 = 1/5; direbm/lattices.py:44-61).  The reference runs 3D only in its per-object
 * CPU solver (direbm/reference/simulator.py:30-198 with lattice=D3Q13); this is the GPU step it plans
 * for 3D (HANDOFF.md:135-172): 13 components per moment, cell keys (ix, iy, iz) x-major, the same
 * refine / scatter / gap-fill with 3D distances.  One simulation per handle; open domain or
 * periodic cube [0, box)^3 (params.periodic); two-level thinning is not supported. */
typedef struct drb_sim3 drb_sim3;

/* The fp32 D3Q13 velocity set (13 x 3, row-major) and weights the 3D kernels use. */
int drb_d3q13(float* C39, float* W13);

int drb_sim3_create(drb_sim3** out, const drb_params* prm, drb_stream_t s);
int drb_sim3_destroy(drb_sim3* sim);
/* Spherical obstacle, fluid outside (r <= 0 removes it): replaces the reference's Sphere obstacle
 * (direbm/reference/boundary.py:50-51) as the per-object solver applies it
 * (direbm/reference/simulator.py:79-80,90-117).  A component k != 0 whose fp32 position is inside
 * (((dx^2 + dy^2) + dz^2) < fp32(r * r)) is replaced by its bounce: FP64 ray hit from its moment,
 * specular reflection of c_k, split over the 12 directions with weights max(d.c_j, 0)^2 / sum
 * (> 1e-9 kept), each part at its moment's j-th component position, or at the moment itself when
 * that is inside too.  Parts follow the regular components within a cell.  Control points inside
 * the obstacle after resampling are dropped.  In a periodic box the sphere must lie 2 dx inside. */
int drb_sim3_set_sphere(drb_sim3* sim, float cx, float cy, float cz, float r);
/* pos (m, 3), f (m, 13): row-major fp32 host arrays. */
int drb_sim3_set_state(drb_sim3* sim, const float* pos, const float* f, int64_t m);
int drb_sim3_step(drb_sim3* sim, int32_t nsteps);
/* Lanes per control point of the 3D refine and gather kernels: 4, 8 or 16 (default 4).  No
 * counterpart in the reference (a tile parameter of this engine); results do not depend on it
 * beyond the order of the fp32 sums. */
int drb_sim3_set_lanes(drb_sim3* sim, int32_t refine_lanes, int32_t gather_lanes);
int drb_sim3_count(drb_sim3* sim, int64_t* m);
int drb_sim3_get_state(drb_sim3* sim, float* pos, float* f);
/* FP64 mass and momentum (x, y, z) of the current state. */
int drb_sim3_diagnostics(drb_sim3* sim, double* out4);

/* ------------------------------------------------------------------- hex LBM */

/* One or more steps of the D2Q7 hex LBM on SoA planes f[k][j][i] (plane pitch nj*ni):
 * collide then periodic pull stream (fused, push form), ping-pong between f and out; after an odd
 * number of steps the result is in out, after an even number in f.
 * Replaces _lbm_collide/_lbm_stream/GpuHexLBM.step, direbm/warp/lbm.py:24-57,95-105. */
int drb_lbm_step(float* f, float* out, int32_t ni, int32_t nj, float tau, int32_t nsteps, drb_stream_t s);

/* f[k][cj][ci] = rho * W[k].  Replaces _lbm_set_center / set_pulse, direbm/warp/lbm.py:60-69,86-93. */
int drb_lbm_set_node(float* f, int32_t ni, int32_t nj, int32_t ci, int32_t cj, float rho, drb_stream_t s);

/* f = f_eq(rho[j][i], 0) on SoA planes.  Replaces the numpy equilibrium init, direbm/warp/lbm.py:78. */
int drb_lbm_init(float* f, const float* rho, int32_t ni, int32_t nj, drb_stream_t s);

#ifdef __cplusplus
}
#endif

#endif /* DIREBM_HIP_H */
