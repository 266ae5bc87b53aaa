/*
 * nmhd.h -- C ABI of the MI355X-native ideal-MHD step (libnmhd.so).
 *
 * The reference (nushio3/nmhd) has no FFI: its "operator API" is the C++
 * class MHDSolver (library/mhd_solver.h:514-640, template
 * library/mhd_solver.hrb:150-276) whose methods launch CUDA kernels.  This
 * header is the thin C layer that the C++ facade (include/nmhd/mhd_solver.h)
 * and any other host language call; each entry point names the reference
 * interface it replaces.  Paths are relative to the reference's src/ tree;
 * INL = library/mhd_solver.inl.
 *
 * Conventions
 *  - Every function returns an int status: NMHD_OK (0) or an NMHD_E* code;
 *    nmhd_last_error() gives the message of the last failure on a context.
 *  - Field arrays are structure-of-arrays FP32, one array per field in the
 *    order rho, vx, vy, vz, p, Bx, By, Bz (the order of the reference's .bin
 *    files, INL:1226-1244).  Each array holds the whole box including ghost
 *    layers: addr = (z*ny + y)*nx + x, x fastest (crystal.h:31-50).  Face
 *    B_d[i] is stored at the lower d-face of cell i.
 *  - A context owns its device buffers and one HIP stream.  Calls on one
 *    context must be serialised by the caller; different contexts may be
 *    used concurrently.  Host arrays passed in are read or written before
 *    the call returns.
 *  - Shearing box orientation is the reference's: radial = Y (ghost layers
 *    along Y), azimuth = X, vertical = Z (application/kepler_disk_demo).
 */
#ifndef NMHD_H
#define NMHD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum nmhd_status {
  NMHD_OK = 0,
  NMHD_EINVAL = 1,   /* bad argument or descriptor                 */
  NMHD_EHIP = 2,     /* a HIP runtime call failed                  */
  NMHD_ENOMEM = 3,   /* device or host allocation failed           */
  NMHD_EIO = 4,      /* file could not be opened / sizes mismatch  */
  NMHD_ESTATE = 5    /* call not valid in the context's state      */
};

enum nmhd_source {
  NMHD_SOURCE_NONE = 0,
  /* Coriolis + tidal acceleration of the shearing box, the hooks
   * external_acceleration / relative_velocity / initial_displacement of
   * application/kepler_disk_demo/mhd_solver.curb:46-81 */
  NMHD_SOURCE_SHEARING_BOX = 1
};

enum nmhd_recon {
  NMHD_RECON_MINBEE2 = 0,        /* minmod MUSCL, INL:68-84            */
  NMHD_RECON_PIECEWISECONST1 = 1 /* first order, mhd_solver.inlrb:128-135 */
};

enum nmhd_integrator {
  NMHD_MIDPOINT2 = 0, /* two-stage midpoint, mhd_solver.inlrb:904-978 */
  NMHD_EULER1 = 1     /* forward Euler, mhd_solver.inlrb:966-978      */
};

enum nmhd_path {
  NMHD_PATH_AUTO = 0,     /* fastest path valid for the descriptor (TILED)  */
  NMHD_PATH_UNFUSED = 1,  /* per-direction flux kernels + one update pass   */
  /* per RK stage: one z-marching LDS-tiled kernel for all face fluxes and
   * one cell kernel for CT + update + primitive recovery (+ the CFL
   * reduction of the new state in the last stage) */
  NMHD_PATH_TILED = 2,
  /* verification path: the unfused kernels with the reference's rounding
   * (no FMA contraction, IEEE division and square root, the FP64-promoted
   * HLLD literals); bit-identical to the CPU restatement of the reference */
  NMHD_PATH_STRICT = 3
};

/* Corrections of reference behaviour (nmhd_desc.fixes, a bit set).  The
 * default, 0, reproduces the reference exactly, quirks included; each bit
 * replaces one quirk by what the reference's code evidently intends. */
enum nmhd_fix {
  /* CFL signal speed |v_d| + c instead of |v_d + c| (INL:715-720), which
   * underestimates the speed where v_d < 0 */
  NMHD_FIX_CFL_SPEED = 1,
  /* shearing-periodic fill interpolates linearly between the two source
   * cells by the fractional displacement, with the weights the gather
   * kernel computes (INL:100-110: val_inc1 / val_dec1, read but unused),
   * instead of copying the cell at the floor of the displacement */
  NMHD_FIX_SHEAR_INTERP = 2,
  /* face predictor's tidal term 2 q Omega^2 Y at the radial face's own
   * position instead of the centre of its cell, half a cell further out
   * (`Bit & 2 == 0` parses as `Bit & (2 == 0)`, crystal.h:86-104, used by
   * INL:793-797) */
  NMHD_FIX_FACE_POSITION = 4,
  NMHD_FIX_ALL = 7
};

/* Tile shapes (TX x TY columns marched along z) of the tiled kernels. */
enum nmhd_tile {
  NMHD_TILE_64x4 = 0,
  NMHD_TILE_32x8 = 1,
  NMHD_TILE_16x16 = 2
};

/* Box description; the compile-time constants of the reference's
 * simulation_config.h (simulation_config_default.h:25-122) as run-time
 * values.  Derived constants (kExtent*, kD*, kLocalExtent*) are computed in
 * FP32 exactly as the reference's constant expressions are. */
typedef struct nmhd_desc {
  int global_size[3];      /* interior cells per axis (kGlobalSizeX..)          */
  int margin[3];           /* ghost layers per side (kMarginSizeX..); [0] == 0  */
  float global_extent[3];  /* physical extent of the interior (kGlobalExtentX..) */
  float gamma;             /* kGamma                                            */
  float cfl_factor;        /* kCflFactor                                        */
  int source;              /* enum nmhd_source                                  */
  float omega;             /* kOmega (shearing box)                             */
  float qshear;            /* kDifferentialRotationIndex (shearing box)         */
  int shear_y;             /* Y ghost fill: 0 periodic copy, 1 shearing-periodic */
  int recon;               /* enum nmhd_recon                                   */
  int integrator;          /* enum nmhd_integrator                              */
  int path;                /* enum nmhd_path                                    */
  int fixes;               /* enum nmhd_fix bits; 0 = the reference's behaviour */
} nmhd_desc;

/* Derived geometry (LocalSetup members, mhd_solver.hrb:93-101). */
typedef struct nmhd_geometry {
  int size[3];            /* array size incl. ghosts (kSizeX..)   */
  size_t ncell;           /* size[0]*size[1]*size[2] (kSize)      */
  float extent[3];        /* array extent incl. ghosts (kExtentX..) */
  float local_extent[3];  /* interior extent (kLocalExtentX..)    */
  float d[3];             /* cell spacing (kDX..)                 */
} nmhd_geometry;

typedef struct nmhd_ctx nmhd_ctx;

/* MHDSolver::MHDSolver (INL:14-41): allocate the device state for one box on
 * the current HIP device.  Limits: fewer than 2^31 cells including ghosts,
 * and fewer than 2^30 in one XY plane (the kernels index with 32-bit
 * offsets). */
int nmhd_create(const nmhd_desc *desc, nmhd_ctx **out);
/* ~MHDSolver: free the context and its device memory. */
int nmhd_destroy(nmhd_ctx *ctx);
/* Message of the last failure on ctx (or, when ctx is NULL, of the calling
 * thread's last failed nmhd_create). */
const char *nmhd_last_error(const nmhd_ctx *ctx);

int nmhd_get_geometry(const nmhd_ctx *ctx, nmhd_geometry *geo);

/* MHDSolver::read / write (INL:1207-1244) without the file: copy the eight
 * state fields in from / out to host memory. */
int nmhd_set_state(nmhd_ctx *ctx, const float *const fields[8]);
int nmhd_get_state(nmhd_ctx *ctx, float *const fields[8]);
/* Same with device pointers (ncell floats each), e.g. torch tensors. */
int nmhd_set_state_device(nmhd_ctx *ctx, const float *const dev_fields[8]);
int nmhd_get_state_device(nmhd_ctx *ctx, float *const dev_fields[8]);

/* GlobalSetup::current_time / generation (mhd_solver.hrb:47-59). */
int nmhd_set_time(nmhd_ctx *ctx, double t, int generation);
int nmhd_get_time(nmhd_ctx *ctx, double *t, int *generation);

/* MHDSolver::communicate (INL:307-711) for one rank: shearing-periodic (or
 * periodic) Y ghost fill with relative_displacement (INL:288-295) at the
 * context's current time, then the periodic Z ghost copy. */
int nmhd_fill_ghosts(nmhd_ctx *ctx);

/* relative_displacement (INL:288-295) at time t with the shearing-box hooks
 * (application/kepler_disk_demo/mhd_solver.curb:46-52). */
int nmhd_relative_displacement(const nmhd_ctx *ctx, double t, float *normalized,
                               int *int_displacement, float *minor_displacement);

/* cfl_condition_kernel + thrust::min_element (INL:714-739, 1106-1116): the
 * minimum over every cell (ghosts included) of the CFL time, at most
 * goal_time.  Synchronises. */
int nmhd_cfl_dt(nmhd_ctx *ctx, float goal_time, float *dt);

/* clear_pace_kernel + add_flux_kernel_caller (INL:741-755, 918-988): the
 * right-hand side P(W) of the current state with stage step h (used by the
 * face predictor 0.5*h*a), written to host arrays in the order rho, mx, my,
 * mz, E, Bx, By, Bz.  Synchronises. */
int nmhd_rhs(nmhd_ctx *ctx, float h, float *const pace[8]);

/* One iteration of MHDSolver::proceed's loop (INL:1094-1201): ghost fill,
 * CFL dt clipped to goal_time, the two RK stages, t += dt.  Synchronises and
 * reports the dt used and whether the step reached goal_time. */
int nmhd_step(nmhd_ctx *ctx, float goal_time, float *dt, int *goal_reached);

/* MHDSolver::proceed (INL:1086-1204) without the monitor/wallclock hooks
 * (the C++ facade adds them): run at most max_steps steps toward goal_time,
 * enqueued in chunks with dt kept on the device.  Synchronises. */
int nmhd_run(nmhd_ctx *ctx, float goal_time, int max_steps, int *steps_taken,
             int *goal_reached);

/* Asynchronous form: nmhd_reset_run starts a new run (clears the
 * goal-reached latch and the step counter); nmhd_enqueue_steps enqueues
 * nsteps iterations of the proceed loop without host round trips (dt is
 * computed, clipped and consumed on the device; once a step reaches
 * goal_time the remaining enqueued steps do nothing) and returns at once;
 * nmhd_synchronize waits and reports the steps executed since the reset and
 * whether goal_time was reached. */
int nmhd_reset_run(nmhd_ctx *ctx);
int nmhd_enqueue_steps(nmhd_ctx *ctx, float goal_time, int nsteps);
int nmhd_synchronize(nmhd_ctx *ctx, int *steps_taken, int *goal_reached);

/* The device HLLD solver (hlld.inl:7-284) on n host face states, with the
 * fast arithmetic or, for path == NMHD_PATH_STRICT, the reference's rounding:
 * in[15] = Bx, rhoL, rhoR, vxL, vxR, vyL, vyR, vzL, vzR, pL, pR, ByL, ByR,
 * BzL, BzR; out[7] = F_rho, F_mx, F_my, F_mz, F_E, F_By, F_Bz. */
int nmhd_hlld_batch(size_t n, float gamma, int path, const float *const in[15],
                    float *const out[7]);

/* Run diagnostics reduced on the device (no array leaves the GPU): the
 * quantities of the Kepler-disk monitor (application/kepler_disk_demo/
 * mhd_solver.curb:107-130, which copies B_x and B_y to the host each step)
 * and conservation checks.  Synchronises. */
typedef struct nmhd_diagnostics {
  double b2_mean[3]; /* <B_d^2> over every face of the array (ghosts included),
                        squares in FP32 summed in FP64 as the monitor does */
  double total[8];   /* interior sums of rho, m_x, m_y, m_z, E (with the
                        face-averaged B^2), B_x, B_y, B_z */
  double divb_max;   /* max over the interior of |div B| * min(d) */
} nmhd_diagnostics;
int nmhd_get_diagnostics(nmhd_ctx *ctx, nmhd_diagnostics *out);

/* Per-kernel timing with HIP events on the context's stream. */
int nmhd_timing_enable(nmhd_ctx *ctx, int enable);
int nmhd_timing_reset(nmhd_ctx *ctx);
/* Kernel slot i: name, total milliseconds and launches since reset;
 * NMHD_EINVAL past the last slot. */
int nmhd_timing_get(nmhd_ctx *ctx, int i, const char **name, double *total_ms,
                    long *launches);

/* Select the kernel path (enum nmhd_path) for later steps. */
int nmhd_set_path(nmhd_ctx *ctx, int path);
int nmhd_get_path(const nmhd_ctx *ctx, int *path);
/* Tile shapes (enum nmhd_tile) of the face and update kernels of
 * NMHD_PATH_TILED (default NMHD_TILE_32x8 and NMHD_TILE_64x4); results do
 * not depend on them. */
int nmhd_set_tiles(nmhd_ctx *ctx, int faces_tile, int update_tile);
int nmhd_get_tiles(const nmhd_ctx *ctx, int *faces_tile, int *update_tile);

/* ------------------------------------------------------------------------
 * Batches: many independent boxes of uneven sizes advanced together.  No reference counterpart: the reference runs one
 * compile-time-sized box per process (mhd_solver.h:514-640), so its
 * equivalent is one MHDSolver::proceed (INL:1086-1204) per box, run one
 * after another.  Every box keeps its own time, CFL dt and goal clipping;
 * its results are bit-identical to the same box advanced alone by an
 * nmhd_ctx on NMHD_PATH_TILED.  Each RK stage is two persistent launches
 * over a work table of tiles from all boxes; runs replay the step as a HIP
 * graph.  Any descriptor a context accepts may be batched (periodic and
 * shearing boxes, ghost layers, sources and quirk corrections per box), at
 * most 32767 cells per axis including ghosts; all boxes share one
 * integrator.
 * ---------------------------------------------------------------------- */
typedef struct nmhd_batch nmhd_batch;

int nmhd_batch_create(const nmhd_desc *descs, int nbox, nmhd_batch **out);
int nmhd_batch_destroy(nmhd_batch *batch);
/* Message of the last failure on batch (or, when batch is NULL, of the
 * calling thread's last failed create). */
const char *nmhd_batch_last_error(const nmhd_batch *batch);
int nmhd_batch_get_geometry(const nmhd_batch *batch, int box, nmhd_geometry *geo);

/* State and time of one box, as nmhd_set_state / nmhd_set_time. */
int nmhd_batch_set_state(nmhd_batch *batch, int box, const float *const fields[8]);
int nmhd_batch_get_state(nmhd_batch *batch, int box, float *const fields[8]);
int nmhd_batch_set_time(nmhd_batch *batch, int box, double t, int generation);
int nmhd_batch_get_time(nmhd_batch *batch, int box, double *t, int *generation);
/* Steps taken by a box in the last nmhd_batch_run, whether it reached the
 * goal, and its last dt. */
int nmhd_batch_get_status(nmhd_batch *batch, int box, int *steps_taken, int *goal_reached,
                          float *last_dt);

/* Tiling: tile shape (enum nmhd_tile) of box's face and update kernels
 * (box = -1: every box); z chunk (planes per work item) and workgroups per
 * CU of the two persistent launches, where 0 (the default) launches one
 * workgroup per work item instead, one launch per tile shape.  Results do
 * not depend on the tiling or the launch mode. */
int nmhd_batch_set_tiles(nmhd_batch *batch, int box, int faces_tile, int update_tile);
int nmhd_batch_set_launch(nmhd_batch *batch, int faces_zchunk, int update_zchunk,
                          int faces_wg_per_cu, int update_wg_per_cu);
int nmhd_batch_get_launch(const nmhd_batch *batch, int *faces_zchunk, int *update_zchunk,
                          int *faces_wg_per_cu, int *update_wg_per_cu);
/* Autotuner: per distinct box shape, time every tile shape of both kernels
 * (one workgroup walking that shape's tiles, i.e. the workgroup-time the
 * shape costs in a persistent launch) and keep the fastest; then time the
 * z chunk and workgroups per CU of the whole batch.  Results are cached in
 * the text file cache_path (NULL: $NMHD_TUNE_CACHE, else
 * $HOME/.cache/nmhd_amd/batch_tuning.txt), keyed by the GPU architecture
 * and the box shape (and, for the launch parameters, the batch's shapes);
 * cached entries are reused without timing.  *timed_shapes (may be NULL)
 * returns how many shapes were timed.  Overwrites the batch's scratch
 * buffers, not the boxes' states. */
int nmhd_batch_autotune(nmhd_batch *batch, const char *cache_path, int *timed_shapes);

/* MHDSolver::proceed for every box: at most nsteps steps toward goal_time
 * each (a box that reaches goal_time stops there).  use_graph != 0 replays
 * the step as a HIP graph (the first step of a run after set_state or
 * set_time is always enqueued directly, with a full CFL pass).
 * Synchronises; *min_steps = fewest steps taken by any box, *all_reached =
 * every box reached goal_time (either may be NULL). */
int nmhd_batch_run(nmhd_batch *batch, float goal_time, int nsteps, int use_graph,
                   int *min_steps, int *all_reached);

/* ------------------------------------------------------------------------
 * Adjoint: the step S(W; dt) of a periodic box with dt frozen (no CFL), and
 * its hand-written reverse lam_in = (dS/dW)^T lam_out.  No reference
 * counterpart (the reference has no derivatives); S is one iteration of
 * proceed()'s loop (INL:1138-1199) for margin {0,0,0}, source NONE.  State
 * and adjoint arrays are caller-owned DEVICE arrays (8 fields, ncell floats
 * each, same layout as nmhd_set_state); work is enqueued on the stream set
 * by nmhd_adj_set_stream and is asynchronous unless stated, so the caller's
 * own work on that stream is ordered with it.  The reverse of the minmod limiter, of the upwind CT choice
 * and of the HLLD fan follow the branches the forward takes at w_in.
 * ---------------------------------------------------------------------- */
typedef struct nmhd_adj nmhd_adj;

int nmhd_adj_create(const nmhd_desc *desc, nmhd_adj **out);
int nmhd_adj_destroy(nmhd_adj *adj);
/* Message of the last failure on adj (or, for adj == NULL, of the calling
 * thread's last failed nmhd_adj_create or nmhd_adj_forward_noalloc). */
const char *nmhd_adj_last_error(const nmhd_adj *adj);
int nmhd_adj_get_geometry(const nmhd_adj *adj, nmhd_geometry *geo);
/* Launch later work on `stream`, a hipStream_t such as
 * torch.cuda.current_stream().cuda_stream; NULL selects HIP's null stream
 * (torch's default stream).  Until the first call the object uses a stream
 * of its own, which the first call releases. */
int nmhd_adj_set_stream(nmhd_adj *adj, void *stream);
int nmhd_adj_synchronize(nmhd_adj *adj);
/* The CFL dt of a device state (as nmhd_cfl_dt).  Synchronises. */
int nmhd_adj_cfl_dt(nmhd_adj *adj, const float *const w[8], float goal_time, float *dt);
/* w_out = S(w_in; dt) with the production stage kernels. */
int nmhd_adj_forward(nmhd_adj *adj, const float *const w_in[8], float dt,
                     float *const w_out[8]);
/* lam_in = (dS/dW at w_in)^T lam_out; recomputes the stage states from w_in. */
int nmhd_adj_reverse(nmhd_adj *adj, const float *const w_in[8], float dt,
                     const float *const lam_out[8], float *const lam_in[8]);
/* No reference counterpart (the reference has no derivatives).  As
 * nmhd_adj_forward, also storing the midpoint stage state W* into w_stage
 * (unused, may be NULL, for NMHD_EULER1). */
int nmhd_adj_forward_stages(nmhd_adj *adj, const float *const w_in[8], float dt,
                            float *const w_stage[8], float *const w_out[8]);
/* As nmhd_adj_reverse, with this step's stage states supplied instead of
 * recomputed: w_stage and w_out as nmhd_adj_forward_stages wrote them for the
 * same w_in and dt (w_stage unused, may be NULL, for NMHD_EULER1).  The
 * result is bitwise that of nmhd_adj_reverse; the forward kernels are not
 * run again. */
int nmhd_adj_reverse_stages(nmhd_adj *adj, const float *const w_in[8],
                            const float *const w_stage[8], const float *const w_out[8],
                            float dt, const float *const lam_out[8], float *const lam_in[8]);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (work under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): w_out = S(w_in; dt) as nmhd_adj_forward computes it,
 * with the same kernels and bit-identical results, but with no nmhd_adj
 * object and no allocation.  The step's scratch (step record, midpoint
 * stage state, the face-to-update hand-off) is a fixed device array of the
 * library, one per GPU, held until the process ends (76 MiB): boxes of at
 * most NMHD_NOALLOC_MAX_CELLS cells; larger ones are rejected (NMHD_EINVAL).
 * Periodic boxes only, as nmhd_adj_create.  Work is enqueued on `stream` (a
 * hipStream_t; NULL = HIP's null stream) and the call returns at once.
 * Every call on one GPU shares that GPU's scratch: calls that overlap on
 * one GPU (on different streams, or from threads without ordering) race;
 * calls on one stream are ordered by it.  The first call on a device looks
 * the scratch up (hipGetSymbolAddress): make it outside stream capture.  A
 * failure's message is nmhd_adj_last_error(NULL). */
#define NMHD_NOALLOC_MAX_CELLS (1 << 20)
int nmhd_adj_forward_noalloc(const nmhd_desc *desc, const float *const w_in[8], float dt,
                             float *const w_out[8], void *stream);

#ifdef __cplusplus
}
#endif
#endif /* NMHD_H */
