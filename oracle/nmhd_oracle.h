/*
 * nmhd_oracle.h -- CPU restatement of the reference MHD step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * product path (nmhd_amd/csrc).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker or as the
 * timed CPU baseline -- never as a fallback for the product.
 *
 * Every function follows the reference text cited beside it (paths relative
 * to the reference repository's src/ tree; INL = library/mhd_solver.inl, the
 * generated shearing-box snapshot; INLRB = library/mhd_solver.inlrb, its
 * template; HLLD = library/hlld.inl).  Arithmetic is FP32 with the same
 * FP64-promoted literals as the reference; build with -ffp-contract=off so
 * that no fused multiply-add changes the rounding.
 *
 * Each function is pinned by the checks of tests/test_oracle_*.py (golden
 * HLLD vectors from the reference's own hlld.inl, closed forms, conservation,
 * symmetries, convergence against exact solutions); none is left unpinned.
 *
 * Data layout (identical to the reference, crystal.h:31-50): one array per
 * field, addr = (z*ny + y)*nx + x, x fastest; face-centred B_d[i] sits on the
 * lower d-face of cell i; every axis wraps periodically over the whole array
 * (ghosts included), as CRYSTAL_PERIODIC_{X,Y,Z} do.
 */
#ifndef NMHD_ORACLE_H
#define NMHD_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Field order of the state, the .bin file and every 8-array argument:
 * rho, vx, vy, vz, p, Bx, By, Bz (INL:1226-1244). */
enum { OR_RHO = 0, OR_VX, OR_VY, OR_VZ, OR_P, OR_BX, OR_BY, OR_BZ, OR_NFIELD };

typedef struct or_grid {
  int n[3];            /* array size per axis including ghosts (kSizeX..)      */
  int m[3];            /* ghost layers per side (kMarginSizeX..); m[0] must be 0 */
  float d[3];          /* cell spacing (kDX.., simulation_config_default.h:101) */
  float ext[3];        /* array extent incl. ghosts (kExtentX.., :79-81)       */
  float local_ext[3];  /* interior extent (kLocalExtentX.., :87-89)           */
  float gamma;         /* kGamma                                              */
  float cfl_factor;    /* kCflFactor                                          */
  int source;          /* 0: none; 1: shearing box, radial = Y, azimuth = X    */
  float omega;         /* kOmega                                              */
  float qshear;        /* kDifferentialRotationIndex                          */
  int shear_y;         /* Y ghost fill: 0 periodic copy, 1 shearing-periodic   */
  int recon;           /* 0: minbee2 (minmod MUSCL), 1: piecewiseconst1        */
  int integrator;      /* 0: midpoint2, 1: euler1                             */
  int fixes;           /* corrections of reference quirks, a bit set (0: none):
                        * 1 CFL speed |v| + c, 2 interpolating shear fill,
                        * 4 face predictor at the face position             */
} or_grid;

typedef struct or_ctx or_ctx;

or_ctx *or_create(const or_grid *g);
void or_destroy(or_ctx *c);
void or_set_threads(int n);

void or_set_state(or_ctx *c, float *const fields[8]);
void or_get_state(const or_ctx *c, float *fields[8]);
void or_get_predictor(const or_ctx *c, float *fields[8]);
void or_set_time(or_ctx *c, double t, int generation);
double or_get_time(const or_ctx *c);
int or_get_generation(const or_ctx *c);

/* relative_displacement, INL:288-295 (KEP hooks: D0 = 0, relvel = -q*Omega*Ly). */
void or_relative_displacement(const or_grid *g, double t, float *normalized,
                              int *int_disp, float *minor);
/* communicate(): Y ghost fill (shearing or periodic) then Z periodic copy,
 * INL:307-711, INLRB:168-226; uses the context's current time. */
void or_fill_ghosts(or_ctx *c);
/* cfl_condition_kernel + thrust::min_element, INL:714-739, 1106-1116. */
float or_cfl_dt(or_ctx *c, float goal_time);
/* Per-cell CFL time into out[ncell] (INL:722-739). */
void or_cfl_field(or_ctx *c, float goal_time, float *out);
/* clear_pace + add_flux_kernel_caller (INL:741-755, 918-988): the RHS P(W)
 * of the current state with predictor step h, copied to pace[8]
 * (rho, mx, my, mz, E, Bx, By, Bz). */
void or_rhs(or_ctx *c, float h, float *pace[8]);
/* calc_flux_kernel<EX> (INL:760-816) of sweep d (0 = X, 1 = Y, 2 = Z) on the
 * current state with predictor step h: out[7] = mass, momentum X, Y, Z,
 * energy, B_{d+1}, B_{d+2} fluxes through the lower d-face of every cell. */
void or_face_flux(or_ctx *c, int d, float h, float *out[7]);
/* One iteration of MHDSolver::proceed's loop (INL:1094-1201) without the
 * monitor/wallclock hooks.  Returns the dt used; *goal_reached set when the
 * step was clipped to goal_time. */
float or_step(or_ctx *c, float goal_time, int *goal_reached);
/* proceed() loop for at most max_steps steps; returns the steps taken. */
int or_run(or_ctx *c, float goal_time, int max_steps, int *goal_reached);

/* Function-level entry points for golden tables. */
void or_hlld_batch(size_t n, float gamma, float *const in[15], float *out[7]);
void or_interpolate_batch(size_t n, int recon, const float *x0, const float *x1,
                          const float *x2, const float *x3, float *left,
                          float *right);

#ifdef __cplusplus
}
#endif
#endif
