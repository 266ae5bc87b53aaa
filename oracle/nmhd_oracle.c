/*
 * nmhd_oracle.c -- CPU restatement of the reference MHD time step.
 *
 * TEST INFRASTRUCTURE ONLY (see nmhd_oracle.h).  Paths cited are relative to
 * the reference's src/ tree: INL = library/mhd_solver.inl (generated
 * shearing-box snapshot), INLRB = library/mhd_solver.inlrb (template),
 * HLLD = library/hlld.inl, KEP = application/kepler_disk_demo/.
 *
 * Each reference kernel is one function here that loops over every cell of
 * the array (ghosts included) and writes only its own address, as the
 * reference's CRYSTAL_MAP kernels do; OpenMP splits the loop, so results do
 * not depend on the thread count.
 *
 * Build: gcc -O2 -fopenmp -ffp-contract=off (no FMA: the reference's host
 * rounding has none; literals such as 0.5 and 1.0 stay double so that the
 * same sub-expressions are promoted to FP64 as in the reference).
 */
#include "nmhd_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

struct or_ctx {
  or_grid g;
  long ncell;
  float *w[8];     /* state W: rho, v(3), p, B(3)         (INL:17-18)  */
  float *pred[8];  /* predictor W*                       (INL:29-30)  */
  float *mom[3];   /* momentum                           (INL:19)     */
  float *energy;   /* energy                             (INL:19)     */
  float *pace[8];  /* RHS: rho, m(3), E, B(3)            (INL:23-24)  */
  float *flux[7];  /* rho, m(3), E, B_ey, B_ez per face  (INL:26-27)  */
  float *cfl;      /* per-cell CFL time                  (INL:21)     */
  double t;        /* GlobalSetup::current_time (PreciseReal)          */
  int generation;
};

/* ---------------------------------------------------------------- helpers */

static inline float sqf(float x) { return x * x; }          /* math.h:24-25 */
static inline float avgf(float a, float b) { return 0.5f * (a + b); } /* math.h:27-28 */
static inline float minf_(float a, float b) { return b < a ? b : a; }
static inline float maxf_(float a, float b) { return a < b ? b : a; }

/* RCo::calc_addr with CRYSTAL_PERIODIC_{X,Y,Z}: (i + N) % N, crystal.h:31-50 */
static inline int wrapi(int i, int n) { return (i + n) % n; }

static inline long addr3(const or_grid *g, int x, int y, int z) {
  return ((long)wrapi(z, g->n[2]) * g->n[1] + wrapi(y, g->n[1])) * g->n[0] +
         wrapi(x, g->n[0]);
}

/* addr of (idx + off) where off is a 3-vector of small offsets */
static inline long addr_off(const or_grid *g, const int *idx, int ox, int oy,
                            int oz) {
  return addr3(g, idx[0] + ox, idx[1] + oy, idx[2] + oz);
}

static inline void decode(const or_grid *g, long a, int *idx) {
  idx[0] = (int)(a % g->n[0]);
  idx[1] = (int)((a / g->n[0]) % g->n[1]);
  idx[2] = (int)((a / g->n[0]) / g->n[1]);
}

/* offset vector of +1 along axis d, scaled by s */
static inline void axis_off(int d, int s, int *o) {
  o[0] = o[1] = o[2] = 0;
  o[d] = s;
}

void or_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* -------------------------------------------------------- HLLD (HLLD:7-284) */

static void hlld_flux(float gamma, float Bx, float dens_L, float dens_R,
                      float velx_L, float velx_R, float vely_L, float vely_R,
                      float velz_L, float velz_R, float pres_L, float pres_R,
                      float By_L, float By_R, float Bz_L, float Bz_R,
                      float *Fdens, float *Fmomx, float *Fmomy, float *Fmomz,
                      float *Fetot, float *F_By, float *F_Bz) {
  const float TINY = 1.0e-10; /* HLLD:23, a Real constant */
  float signBx = 0;           /* HLLD:25-26 */
  if (fabsf(Bx) > 0) signBx = Bx / fabsf(Bx);

  float momx_L = dens_L * velx_L, momy_L = dens_L * vely_L,
        momz_L = dens_L * velz_L;
  float momx_R = dens_R * velx_R, momy_R = dens_R * vely_R,
        momz_R = dens_R * velz_R;

  /* HLLD:36-42: the 0.5 literals are double, so these sums round once to float */
  float B2_L = sqf(Bx) + sqf(By_L) + sqf(Bz_L);
  float v2_L = sqf(velx_L) + sqf(vely_L) + sqf(velz_L);
  float etot_L = pres_L / (gamma - 1) + 0.5 * dens_L * v2_L + 0.5 * B2_L;
  float B2_R = sqf(Bx) + sqf(By_R) + sqf(Bz_R);
  float v2_R = sqf(velx_R) + sqf(vely_R) + sqf(velz_R);
  float etot_R = pres_R / (gamma - 1) + 0.5 * dens_R * v2_R + 0.5 * B2_R;

  float gpl = gamma * pres_L, gpr = gamma * pres_R; /* HLLD:44-47 */
  float gpbl = gpl + B2_L, gpbr = gpr + B2_R;

  const float fsig = sqrtf(2.0); /* HLLD:49 */
  /* HLLD:55-62, the active "#if 1" branch: fast magnetosonic speeds */
  float cfl = sqrtf((gpbl + sqrtf(sqf(gpbl) - 4 * gpl * sqf(Bx))) / (2.0 * dens_L));
  float cfr = sqrtf((gpbr + sqrtf(sqf(gpbr) - 4 * gpr * sqf(Bx))) / (2.0 * dens_R));
  float cfmax = fsig * fmaxf(cfl, cfr);
  float S_L = fminf(velx_L, velx_R) - cfmax;
  float S_R = fmaxf(velx_L, velx_R) + cfmax;

  float pT_L = pres_L + 0.5 * B2_L; /* HLLD:128-129 */
  float pT_R = pres_R + 0.5 * B2_R;

  /* HLLD:131-136 */
  float S_M = (S_R - velx_R) * momx_R - (S_L - velx_L) * momx_L - pT_R + pT_L;
  S_M *= 1.0 / ((S_R - velx_R) * dens_R - (S_L - velx_L) * dens_L);
  float pT_s = (S_R - velx_R) * dens_R * pT_L - (S_L - velx_L) * dens_L * pT_R;
  pT_s += dens_L * dens_R * (S_R - velx_R) * (S_L - velx_L) * (velx_R - velx_L);
  pT_s *= 1.0 / ((S_R - velx_R) * dens_R - (S_L - velx_L) * dens_L);

  float velx_L_s = S_M, velx_L_ss = S_M, velx_R_s = S_M, velx_R_ss = S_M;
  float B2x = Bx * Bx;

  /* HLLD:146-156: single-star states */
  float dens_L_s = dens_L * (S_L - velx_L) / (S_L - S_M);
  float vely_L_s = vely_L - Bx * By_L * (S_M - velx_L) /
                                (dens_L * (S_L - velx_L) * (S_L - S_M) - B2x + TINY);
  float velz_L_s = velz_L - Bx * Bz_L * (S_M - velx_L) /
                                (dens_L * (S_L - velx_L) * (S_L - S_M) - B2x + TINY);
  float By_L_s = By_L * (dens_L * sqf(S_L - velx_L) - B2x) /
                 (dens_L * (S_L - velx_L) * (S_L - S_M) - B2x + TINY);
  float Bz_L_s = Bz_L * (dens_L * sqf(S_L - velx_L) - B2x) /
                 (dens_L * (S_L - velx_L) * (S_L - S_M) - B2x + TINY);

  float dens_R_s = dens_R * (S_R - velx_R) / (S_R - S_M);
  float vely_R_s = vely_R - Bx * By_R * (S_M - velx_R) /
                                (dens_R * (S_R - velx_R) * (S_R - S_M) - B2x + TINY);
  float velz_R_s = velz_R - Bx * Bz_R * (S_M - velx_R) /
                                (dens_R * (S_R - velx_R) * (S_R - S_M) - B2x + TINY);
  float By_R_s = By_R * (dens_R * sqf(S_R - velx_R) - B2x) /
                 (dens_R * (S_R - velx_R) * (S_R - S_M) - B2x + TINY);
  float Bz_R_s = Bz_R * (dens_R * sqf(S_R - velx_R) - B2x) /
                 (dens_R * (S_R - velx_R) * (S_R - S_M) - B2x + TINY);

  /* HLLD:158-164 */
  float vB_L = velx_L * Bx + vely_L * By_L + velz_L * Bz_L;
  float vB_L_s = velx_L_s * Bx + vely_L_s * By_L_s + velz_L_s * Bz_L_s;
  float etot_L_s = ((S_L - velx_L) * etot_L - pT_L * velx_L + pT_s * S_M +
                    Bx * (vB_L - vB_L_s)) / (S_L - S_M);
  float vB_R = velx_R * Bx + vely_R * By_R + velz_R * Bz_R;
  float vB_R_s = velx_R_s * Bx + vely_R_s * By_R_s + velz_R_s * Bz_R_s;
  float etot_R_s = ((S_R - velx_R) * etot_R - pT_R * velx_R + pT_s * S_M +
                    Bx * (vB_R - vB_R_s)) / (S_R - S_M);

  /* HLLD:166-193: double-star (Alfven) states */
  float dens_L_ss = dens_L_s;
  float S_L_s = S_M - fabsf(Bx) / sqrtf(dens_L_s);
  float dens_R_ss = dens_R_s;
  float S_R_s = S_M + fabsf(Bx) / sqrtf(dens_R_s);

  float f = 1.0 / (sqrtf(dens_L_s) + sqrtf(dens_R_s));
  float vely_ss = f * (sqrtf(dens_L_s) * vely_L_s + sqrtf(dens_R_s) * vely_R_s +
                       (By_R_s - By_L_s) * signBx);
  float velz_ss = f * (sqrtf(dens_L_s) * velz_L_s + sqrtf(dens_R_s) * velz_R_s +
                       (Bz_R_s - Bz_L_s) * signBx);
  float By_ss = f * (sqrtf(dens_L_s) * By_R_s + sqrtf(dens_R_s) * By_L_s +
                     sqrtf(dens_L_s * dens_R_s) * (vely_R_s - vely_L_s) * signBx);
  float Bz_ss = f * (sqrtf(dens_L_s) * Bz_R_s + sqrtf(dens_R_s) * Bz_L_s +
                     sqrtf(dens_L_s * dens_R_s) * (velz_R_s - velz_L_s) * signBx);

  float vely_L_ss = vely_ss, velz_L_ss = velz_ss, By_L_ss = By_ss, Bz_L_ss = Bz_ss;
  float vely_R_ss = vely_ss, velz_R_ss = velz_ss, By_R_ss = By_ss, Bz_R_ss = Bz_ss;

  float vB_L_ss = velx_L_ss * Bx + vely_L_ss * By_L_ss + velz_L_ss * Bz_L_ss;
  float etot_L_ss = etot_L_s - sqrtf(dens_L_s) * (vB_L_s - vB_L_ss) * signBx;
  float vB_R_ss = velx_R_ss * Bx + vely_R_ss * By_R_ss + velz_R_ss * Bz_R_ss;
  float etot_R_ss = etot_R_s + sqrtf(dens_R_s) * (vB_R_s - vB_R_ss) * signBx;

  /* HLLD:195-209: physical fluxes of the outer states */
  float Fdens_L = dens_L * velx_L;
  float Fmomx_L = momx_L * velx_L + pT_L - B2x;
  float Fmomy_L = momy_L * velx_L - Bx * By_L;
  float Fmomz_L = momz_L * velx_L - Bx * Bz_L;
  float Fetot_L = etot_L * velx_L + pT_L * velx_L - Bx * vB_L;
  float Fby_L = By_L * velx_L - Bx * vely_L;
  float Fbz_L = Bz_L * velx_L - Bx * velz_L;

  float Fdens_R = dens_R * velx_R;
  float Fmomx_R = momx_R * velx_R + pT_R - B2x;
  float Fmomy_R = momy_R * velx_R - Bx * By_R;
  float Fmomz_R = momz_R * velx_R - Bx * Bz_R;
  float Fetot_R = etot_R * velx_R + pT_R * velx_R - Bx * vB_R;
  float Fby_R = By_R * velx_R - Bx * vely_R;
  float Fbz_R = Bz_R * velx_R - Bx * velz_R;

  /* HLLD:211-225 */
  float momx_L_s = dens_L_s * velx_L_s, momy_L_s = dens_L_s * vely_L_s,
        momz_L_s = dens_L_s * velz_L_s;
  float momx_L_ss = dens_L_ss * velx_L_ss, momy_L_ss = dens_L_ss * vely_L_ss,
        momz_L_ss = dens_L_ss * velz_L_ss;
  float momx_R_s = dens_R_s * velx_R_s, momy_R_s = dens_R_s * vely_R_s,
        momz_R_s = dens_R_s * velz_R_s;
  float momx_R_ss = dens_R_ss * velx_R_ss, momy_R_ss = dens_R_ss * vely_R_ss,
        momz_R_ss = dens_R_ss * velz_R_ss;

  /* HLLD:228-282: six-way fan selection */
  if (S_L > 0) {
    *Fdens = Fdens_L; *Fmomx = Fmomx_L; *Fmomy = Fmomy_L; *Fmomz = Fmomz_L;
    *Fetot = Fetot_L; *F_By = Fby_L; *F_Bz = Fbz_L;
  } else if (S_L <= 0 && 0 <= S_L_s) {
    *Fdens = Fdens_L + S_L * dens_L_s - S_L * dens_L;
    *Fmomx = Fmomx_L + S_L * momx_L_s - S_L * momx_L;
    *Fmomy = Fmomy_L + S_L * momy_L_s - S_L * momy_L;
    *Fmomz = Fmomz_L + S_L * momz_L_s - S_L * momz_L;
    *Fetot = Fetot_L + S_L * etot_L_s - S_L * etot_L;
    *F_By = Fby_L + S_L * By_L_s - S_L * By_L;
    *F_Bz = Fbz_L + S_L * Bz_L_s - S_L * Bz_L;
  } else if (S_L_s <= 0 && 0 <= S_M) {
    *Fdens = Fdens_L + S_L_s * dens_L_ss - (S_L_s - S_L) * dens_L_s - S_L * dens_L;
    *Fmomx = Fmomx_L + S_L_s * momx_L_ss - (S_L_s - S_L) * momx_L_s - S_L * momx_L;
    *Fmomy = Fmomy_L + S_L_s * momy_L_ss - (S_L_s - S_L) * momy_L_s - S_L * momy_L;
    *Fmomz = Fmomz_L + S_L_s * momz_L_ss - (S_L_s - S_L) * momz_L_s - S_L * momz_L;
    *Fetot = Fetot_L + S_L_s * etot_L_ss - (S_L_s - S_L) * etot_L_s - S_L * etot_L;
    *F_By = Fby_L + S_L_s * By_L_ss - (S_L_s - S_L) * By_L_s - S_L * By_L;
    *F_Bz = Fbz_L + S_L_s * Bz_L_ss - (S_L_s - S_L) * Bz_L_s - S_L * Bz_L;
  } else if (S_M <= 0 && 0 <= S_R_s) {
    *Fdens = Fdens_R + S_R_s * dens_R_ss - (S_R_s - S_R) * dens_R_s - S_R * dens_R;
    *Fmomx = Fmomx_R + S_R_s * momx_R_ss - (S_R_s - S_R) * momx_R_s - S_R * momx_R;
    *Fmomy = Fmomy_R + S_R_s * momy_R_ss - (S_R_s - S_R) * momy_R_s - S_R * momy_R;
    *Fmomz = Fmomz_R + S_R_s * momz_R_ss - (S_R_s - S_R) * momz_R_s - S_R * momz_R;
    *Fetot = Fetot_R + S_R_s * etot_R_ss - (S_R_s - S_R) * etot_R_s - S_R * etot_R;
    *F_By = Fby_R + S_R_s * By_R_ss - (S_R_s - S_R) * By_R_s - S_R * By_R;
    *F_Bz = Fbz_R + S_R_s * Bz_R_ss - (S_R_s - S_R) * Bz_R_s - S_R * Bz_R;
  } else if (S_R_s <= 0 && 0 <= S_R) {
    *Fdens = Fdens_R + S_R * dens_R_s - S_R * dens_R;
    *Fmomx = Fmomx_R + S_R * momx_R_s - S_R * momx_R;
    *Fmomy = Fmomy_R + S_R * momy_R_s - S_R * momy_R;
    *Fmomz = Fmomz_R + S_R * momz_R_s - S_R * momz_R;
    *Fetot = Fetot_R + S_R * etot_R_s - S_R * etot_R;
    *F_By = Fby_R + S_R * By_R_s - S_R * By_R;
    *F_Bz = Fbz_R + S_R * Bz_R_s - S_R * Bz_R;
  } else {
    *Fdens = Fdens_R; *Fmomx = Fmomx_R; *Fmomy = Fmomy_R; *Fmomz = Fmomz_R;
    *Fetot = Fetot_R; *F_By = Fby_R; *F_Bz = Fbz_R;
  }
}

void or_hlld_batch(size_t n, float gamma, float *const in[15], float *out[7]) {
  long i;
#pragma omp parallel for schedule(static)
  for (i = 0; i < (long)n; ++i) {
    hlld_flux(gamma, in[0][i], in[1][i], in[2][i], in[3][i], in[4][i], in[5][i],
              in[6][i], in[7][i], in[8][i], in[9][i], in[10][i], in[11][i],
              in[12][i], in[13][i], in[14][i], &out[0][i], &out[1][i],
              &out[2][i], &out[3][i], &out[4][i], &out[5][i], &out[6][i]);
  }
}

/* ------------------------------------------- reconstruction (INL:68-84) */

static inline float absolute_smaller(float a, float b) {
  return a * b <= 0 ? 0 : fabsf(a) < fabsf(b) ? a : b;
}

/* minbee2: INL:75-83 */
static inline void interpolate4(float x0, float x1, float x2, float x3,
                                float *left, float *right) {
  const float d01 = x1 - x0, d12 = x2 - x1, d23 = x3 - x2;
  const float d1 = absolute_smaller(d01, d12);
  const float d2 = absolute_smaller(d12, d23);
  *left = x1 + d1 / 2;
  *right = x2 - d2 / 2;
}

void or_interpolate_batch(size_t n, int recon, const float *x0, const float *x1,
                          const float *x2, const float *x3, float *left,
                          float *right) {
  for (size_t i = 0; i < n; ++i) {
    if (recon == 1) { /* piecewiseconst1, INLRB:128-135: stencil (c1, c2) */
      left[i] = x1[i];
      right[i] = x2[i];
    } else {
      interpolate4(x0[i], x1[i], x2[i], x3[i], &left[i], &right[i]);
    }
  }
}

/* ------------------------------------------ external acceleration (KEP) */

/* RCo::position with the corner (-ext/2): every staggered position gets the
 * +0.5 offset because `Bit & 2 == 0` parses as `Bit & (2 == 0)`
 * (crystal.h:86-104).  Only the radial (Y) component is used.  off is the
 * offset in cells: 0.5 everywhere in the reference; 0 for a radial face's
 * own position (fix bit 4). */
static inline float local_position_y_at(const or_grid *g, int y_wrapped, float off) {
  const float corner = -g->ext[1] / 2;
  return corner + g->d[1] * ((float)y_wrapped + off);
}

/* KEP/mhd_solver.curb:64-81: the acceleration at radial position offset off
 * of row y (0.5: the cell centre, and the reference's face position) */
static inline void accel_at(const or_grid *g, float vX, float vY, int y, float off,
                            float *a) {
  const float Om = g->omega, q = g->qshear;
  a[0] = -2.0f * Om * vY;
  a[1] = 2.0f * Om * vX + 2.0f * q * sqf(Om) * local_position_y_at(g, y, off);
  a[2] = 0;
}

/* KEP/mhd_solver.curb:64-81, cell-centred overload (Bit 7) */
static inline void accel_cell(const or_grid *g, float vX, float vY, int y,
                              float *a) {
  accel_at(g, vX, vY, y, 0.5f, a);
}

/* ------------------------------------------------------------ kernels */

/* cfl_condition_kernel, INL:714-739 */
void or_cfl_field(or_ctx *c, float goal_time, float *out) {
  const or_grid *g = &c->g;
  float **w = c->w;
  long a;
#pragma omp parallel for schedule(static)
  for (a = 0; a < c->ncell; ++a) {
    int idx[3];
    decode(g, a, idx);
    float min_time = goal_time;
    const float b2 = maxf_(sqf(w[OR_BX][a]), sqf(w[OR_BX][addr_off(g, idx, 1, 0, 0)])) +
                     maxf_(sqf(w[OR_BY][a]), sqf(w[OR_BY][addr_off(g, idx, 0, 1, 0)])) +
                     maxf_(sqf(w[OR_BZ][a]), sqf(w[OR_BZ][addr_off(g, idx, 0, 0, 1)]));
    for (int d = 0; d < 3; ++d) {
      const float len = g->d[d];
      /* |v + c| (INL:715-720); |v| + c with fix bit 1 */
      const float v = (g->fixes & 1) ? fabsf(w[OR_VX + d][a]) : w[OR_VX + d][a];
      const float speed = fabsf(v + sqrtf((g->gamma * w[OR_P][a] + b2) / w[OR_RHO][a]));
      min_time = minf_(min_time, g->cfl_factor * len / speed);
    }
    out[a] = min_time;
  }
}

float or_cfl_dt(or_ctx *c, float goal_time) {
  or_cfl_field(c, goal_time, c->cfl);
  /* thrust::min_element (INL:1111-1114): exact, order-independent */
  float m = c->cfl[0];
  for (long a = 1; a < c->ncell; ++a)
    if (c->cfl[a] < m) m = c->cfl[a];
  return m;
}

/* clear_pace_kernel, INL:741-755 */
static void clear_pace(or_ctx *c) {
  for (int f = 0; f < 8; ++f) memset(c->pace[f], 0, sizeof(float) * c->ncell);
}

/* calc_flux_kernel<EX>, INL:760-816 (INLRB:562-636).  W = stage input,
 * h = stage step (the predictor term is 0.5*h*a).  Face d of cell i is the
 * lower d-face; stencil cells are i-2, i-1 | i, i+1 along d. */
static void calc_flux(or_ctx *c, int d, float h, float *const *W) {
  const or_grid *g = &c->g;
  const int ey = (d + 1) % 3, ez = (d + 2) % 3; /* EX::Next, EX::Prev */
  long a;
#pragma omp parallel for schedule(static)
  for (a = 0; a < c->ncell; ++a) {
    int idx[3], o[3], oy[3], oz[3];
    decode(g, a, idx);
    axis_off(d, 1, o);
    axis_off(ey, 1, oy);
    axis_off(ez, 1, oz);
    long cc[4];
    for (int s = 0; s < 4; ++s)
      cc[s] = addr_off(g, idx, (s - 2) * o[0], (s - 2) * o[1], (s - 2) * o[2]);
    float L[7], R[7];
    const int q[5] = {OR_RHO, OR_VX + d, OR_VX + ey, OR_VX + ez, OR_P};
    for (int k = 0; k < 5; ++k) {
      const float *f = W[q[k]];
      if (g->recon == 1)
        L[k] = f[cc[1]], R[k] = f[cc[2]];
      else
        interpolate4(f[cc[0]], f[cc[1]], f[cc[2]], f[cc[3]], &L[k], &R[k]);
    }
    /* transverse B averaged from the cell's two faces (INL:789-791) */
    float bey[4], bez[4];
    for (int s = 0; s < 4; ++s) {
      int ci[3];
      decode(g, cc[s], ci);
      bey[s] = avgf(W[OR_BX + ey][cc[s]], W[OR_BX + ey][addr_off(g, ci, oy[0], oy[1], oy[2])]);
      bez[s] = avgf(W[OR_BX + ez][cc[s]], W[OR_BX + ez][addr_off(g, ci, oz[0], oz[1], oz[2])]);
    }
    if (g->recon == 1) {
      L[5] = bey[1], R[5] = bey[2];
      L[6] = bez[1], R[6] = bez[2];
    } else {
      interpolate4(bey[0], bey[1], bey[2], bey[3], &L[5], &R[5]);
      interpolate4(bez[0], bez[1], bez[2], bez[3], &L[6], &R[6]);
    }
    if (g->source == 1 && d != 2) {
      /* INL:793-797 + the face overload of KEP external_acceleration:
       * v averaged over the two cells of the face, Y position at the centre
       * of the face's own cell index (the +0.5 quirk), or at the face
       * itself with fix bit 4. */
      const float vX = avgf(W[OR_VX][cc[1]], W[OR_VX][cc[2]]);
      const float vY = avgf(W[OR_VY][cc[1]], W[OR_VY][cc[2]]);
      float acc[3];
      accel_at(g, vX, vY, idx[1], (g->fixes & 4) ? 0.0f : 0.5f, acc);
      const float mbv = 0.5f * h * acc[d];
      L[1] += mbv;
      R[1] += mbv;
    } else if (g->source == 1) {
      const float mbv = 0.5f * h * 0.0f; /* a_Z = 0 (KEP:78-80) */
      L[1] += mbv;
      R[1] += mbv;
    }
    float F[7];
    hlld_flux(g->gamma, W[OR_BX + d][a], L[0], R[0], L[1], R[1], L[2], R[2], L[3],
              R[3], L[4], R[4], L[5], R[5], L[6], R[6], &F[0], &F[1], &F[2],
              &F[3], &F[4], &F[5], &F[6]);
    c->flux[0][a] = F[0];
    c->flux[1 + d][a] = F[1];  /* momentum_flux[ex] */
    c->flux[1 + ey][a] = F[2]; /* momentum_flux[ey] */
    c->flux[1 + ez][a] = F[3]; /* momentum_flux[ez] */
    c->flux[4][a] = F[4];
    c->flux[5][a] = F[5]; /* magnet_flux_ey */
    c->flux[6][a] = F[6]; /* magnet_flux_ez */
  }
}

/* corner_magnet_flux, INL:819-840.  Edge at (d-face a, e-face b) given by
 * the cell index idx of the edge's upper cells; wind = v_e; flux = the
 * d-flux of B_e.  Upwinds along e by the sign of the 4 surrounding v_e. */
static inline float corner_flux(const or_grid *g, const int *edge, int d, int e,
                                const float *wind, const float *flux) {
  int od[3], oe[3];
  axis_off(d, 1, od);
  axis_off(e, 1, oe);
  /* wind[ey][corner+hex+hey] + [+hex-hey] + [-hex+hey] + [-hex-hey] */
  const float s = wind[addr_off(g, edge, 0, 0, 0)] +
                  wind[addr_off(g, edge, -oe[0], -oe[1], -oe[2])] +
                  wind[addr_off(g, edge, -od[0], -od[1], -od[2])] +
                  wind[addr_off(g, edge, -od[0] - oe[0], -od[1] - oe[1], -od[2] - oe[2])];
  const float f0 = flux[addr_off(g, edge, -oe[0], -oe[1], -oe[2])];
  const float f1 = flux[addr_off(g, edge, 0, 0, 0)];
  return s > 0 ? f0 : s < 0 ? f1 : avgf(f0, f1);
}

/* add_flux_kernel<EX>, INL:851-897 */
static void add_flux(or_ctx *c, int d, float *const *wind3) {
  const or_grid *g = &c->g;
  const int ey = (d + 1) % 3, ez = (d + 2) % 3;
  const float magic_factor = 0.5f, black_magic_factor = 1.0f;
  const float dd = g->d[d];
  long a;
#pragma omp parallel for schedule(static)
  for (a = 0; a < c->ncell; ++a) {
    int idx[3], o[3];
    decode(g, a, idx);
    axis_off(d, 1, o);
    const long r = addr_off(g, idx, o[0], o[1], o[2]);
    c->pace[0][a] += (c->flux[0][a] - c->flux[0][r]) / dd;
    for (int k = 1; k <= 3; ++k) c->pace[k][a] += (c->flux[k][a] - c->flux[k][r]) / dd;
    c->pace[4][a] += (c->flux[4][a] - c->flux[4][r]) / dd;

    int up_d[3] = {idx[0] + o[0], idx[1] + o[1], idx[2] + o[2]};
    int up_ey[3] = {idx[0], idx[1], idx[2]};
    int up_ez[3] = {idx[0], idx[1], idx[2]};
    up_ey[ey] += 1;
    up_ez[ez] += 1;
    /* oy: lower ey-face of the cell; edges at d-faces i+1 and i */
    {
      const float diff = corner_flux(g, up_d, d, ey, wind3[ey], c->flux[5]) -
                         corner_flux(g, idx, d, ey, wind3[ey], c->flux[5]);
      c->pace[5 + ey][a] -= magic_factor * diff / dd;
    }
    /* oz: lower ez-face; edges at d-faces i+1 and i */
    {
      const float diff = corner_flux(g, up_d, d, ez, wind3[ez], c->flux[6]) -
                         corner_flux(g, idx, d, ez, wind3[ez], c->flux[6]);
      c->pace[5 + ez][a] -= magic_factor * diff / dd;
    }
    /* ox: lower d-face; edges at ey-faces j+1, j and ez-faces k+1, k */
    {
      const float dy_ = corner_flux(g, up_ey, d, ey, wind3[ey], c->flux[5]) -
                        corner_flux(g, idx, d, ey, wind3[ey], c->flux[5]);
      const float dz_ = corner_flux(g, up_ez, d, ez, wind3[ez], c->flux[6]) -
                        corner_flux(g, idx, d, ez, wind3[ez], c->flux[6]);
      c->pace[5 + d][a] += magic_factor * black_magic_factor *
                           (dy_ / g->d[ey] + dz_ / g->d[ez]);
    }
  }
}

/* add_flux_kernel_caller, INL:918-988: sweeps X, Y, Z in that order */
static void rhs(or_ctx *c, float h, float *const *W) {
  clear_pace(c);
  for (int d = 0; d < 3; ++d) {
    calc_flux(c, d, h, W);
    add_flux(c, d, W + OR_VX);
  }
}

/* to_conserved_kernel, INL:990-1007 */
static void to_conserved(or_ctx *c, float *const *W) {
  const or_grid *g = &c->g;
  const float gm1 = g->gamma - 1;
  long a;
#pragma omp parallel for schedule(static)
  for (a = 0; a < c->ncell; ++a) {
    int idx[3];
    decode(g, a, idx);
    const float rho = W[OR_RHO][a], vx = W[OR_VX][a], vy = W[OR_VY][a],
                vz = W[OR_VZ][a];
    c->mom[0][a] = rho * vx;
    c->mom[1][a] = rho * vy;
    c->mom[2][a] = rho * vz;
    c->energy[a] = W[OR_P][a] / gm1 + 0.5f * rho * (sqf(vx) + sqf(vy) + sqf(vz)) +
                   0.25f * (sqf(W[OR_BX][a]) + sqf(W[OR_BX][addr_off(g, idx, 1, 0, 0)]) +
                            sqf(W[OR_BY][a]) + sqf(W[OR_BY][addr_off(g, idx, 0, 1, 0)]) +
                            sqf(W[OR_BZ][a]) + sqf(W[OR_BZ][addr_off(g, idx, 0, 0, 1)]));
  }
}

/* to_primitive_kernel, INL:1009-1027: (rho, mom, energy, B) -> v, p of W */
static void to_primitive(or_ctx *c, float *const *W) {
  const or_grid *g = &c->g;
  const float gm1 = g->gamma - 1;
  long a;
#pragma omp parallel for schedule(static)
  for (a = 0; a < c->ncell; ++a) {
    int idx[3];
    decode(g, a, idx);
    const float rho = W[OR_RHO][a];
    const float mx = c->mom[0][a], my = c->mom[1][a], mz = c->mom[2][a];
    W[OR_VX][a] = mx / rho;
    W[OR_VY][a] = my / rho;
    W[OR_VZ][a] = mz / rho;
    const float internal_energy =
        c->energy[a] - 0.5f * (sqf(mx) + sqf(my) + sqf(mz)) / rho -
        0.25f * (sqf(W[OR_BX][a]) + sqf(W[OR_BX][addr_off(g, idx, 1, 0, 0)]) +
                 sqf(W[OR_BY][a]) + sqf(W[OR_BY][addr_off(g, idx, 0, 1, 0)]) +
                 sqf(W[OR_BZ][a]) + sqf(W[OR_BZ][addr_off(g, idx, 0, 0, 1)]));
    W[OR_P][a] = internal_energy * gm1;
  }
}

/* add_pace_kernel, INL:1046-1071: rho_new = rho_old + h P; U += h P;
 * B_new = B_old + h P */
static void add_pace(or_ctx *c, float h, float *const *Wold, float *const *Wnew) {
  long a;
#pragma omp parallel for schedule(static)
  for (a = 0; a < c->ncell; ++a) {
    Wnew[OR_RHO][a] = Wold[OR_RHO][a] + h * c->pace[0][a];
    c->mom[0][a] += h * c->pace[1][a];
    c->mom[1][a] += h * c->pace[2][a];
    c->mom[2][a] += h * c->pace[3][a];
    c->energy[a] += h * c->pace[4][a];
    Wnew[OR_BX][a] = Wold[OR_BX][a] + h * c->pace[5][a];
    Wnew[OR_BY][a] = Wold[OR_BY][a] + h * c->pace[6][a];
    Wnew[OR_BZ][a] = Wold[OR_BZ][a] + h * c->pace[7][a];
  }
}

/* update_kernel, INL:1029-1044 (in place) */
static void update(or_ctx *c, float h, float *const *W) {
  long a;
#pragma omp parallel for schedule(static)
  for (a = 0; a < c->ncell; ++a) {
    W[OR_RHO][a] += h * c->pace[0][a];
    c->mom[0][a] += h * c->pace[1][a];
    c->mom[1][a] += h * c->pace[2][a];
    c->mom[2][a] += h * c->pace[3][a];
    c->energy[a] += h * c->pace[4][a];
    W[OR_BX][a] += h * c->pace[5][a];
    W[OR_BY][a] += h * c->pace[6][a];
    W[OR_BZ][a] += h * c->pace[7][a];
  }
}

/* external_acceleration_kernel, INL:1073-1084 (KEP cell overload) */
static void external_acceleration(or_ctx *c, float h, float *const *W) {
  const or_grid *g = &c->g;
  long a;
#pragma omp parallel for schedule(static)
  for (a = 0; a < c->ncell; ++a) {
    int idx[3];
    decode(g, a, idx);
    float acc[3];
    accel_cell(g, W[OR_VX][a], W[OR_VY][a], idx[1], acc);
    W[OR_VX][a] += h * acc[0];
    W[OR_VY][a] += h * acc[1];
    W[OR_VZ][a] += h * acc[2];
  }
}

/* ------------------------------------------------ shearing ghost fill */

/* relative_displacement, INL:288-295 with KEP's hooks (KEP:46-52):
 * initial_displacement = 0, relative_velocity = -q*Omega*kLocalExtentY. */
void or_relative_displacement(const or_grid *g, double t, float *normalized,
                              int *int_disp, float *minor) {
  const float relvel = -g->qshear * g->omega * g->local_ext[1];
  const double d0 = 0.0 + t * (double)relvel;
  const int fold = (int)floor(d0 / (double)g->ext[0]);
  const double d1 = d0 - (double)((float)fold * g->ext[0]); /* int*Real is float */
  *normalized = (float)d1;
  *int_disp = (int)floor(d1 / (double)g->d[0]);
  *minor = (float)(d1 / (double)g->d[0] - *int_disp);
}

/* communicate_gather_kernel_y + communicate_scatter_kernel_y, INL:90-286,
 * INLRB:168-226, for a single rank (both neighbours are this rank).  The
 * gather reads only interior rows and the scatter writes only ghost rows, so
 * doing both per ghost cell is the same as the two-kernel sequence. */
static void fill_y(or_ctx *c) {
  const or_grid *g = &c->g;
  const int nx = g->n[0], ny = g->n[1], nz = g->n[2], M = g->m[1];
  if (M <= 0) return;
  float nd = 0, dr = 0;
  int n = 0;
  float relvel = 0;
  if (g->shear_y) {
    or_relative_displacement(g, c->t, &nd, &n, &dr);
    relvel = -g->qshear * g->omega * g->local_ext[1];
  }
  const long unit = (long)nx * M * nz;
  long a;
#pragma omp parallel for schedule(static)
  for (a = 0; a < unit; ++a) {
    const int sx = (int)(a % nx), sy = (int)((a / nx) % M), sz = (int)(a / nx / M);
    int inc_x0 = sx, dec_x0 = sx, inc_x1 = sx, dec_x1 = sx;
    if (g->shear_y) {
      /* the reference's `% gSizeX` assumes 0 <= n < nx; wrap fully here */
      inc_x0 = ((sx + n) % nx + nx) % nx;
      dec_x0 = ((sx - n - 1) % nx + nx) % nx;
      inc_x1 = ((sx + n + 1) % nx + nx) % nx; /* INL:101, 103 */
      dec_x1 = ((sx - n) % nx + nx) % nx;
    }
    const long row_inc = ((long)sz * ny + (ny - 2 * M + sy)) * nx;
    const long row_dec = ((long)sz * ny + (M + sy)) * nx;
    const long dst_inc = ((long)sz * ny + sy) * nx + sx;
    const long dst_dec = ((long)sz * ny + (ny - M + sy)) * nx + sx;
    const int interp = g->shear_y && (g->fixes & 2);
    for (int f = 0; f < 8; ++f) {
      float vi = c->w[f][row_inc + inc_x0], vd = c->w[f][row_dec + dec_x0];
      if (g->shear_y) {
        /* the reference reads val_inc1 / val_dec1 (INL:104-107) but weights
         * val_inc0 / val_dec0 twice (INL:108-110); fix bit 2 uses them */
        const float vi1 = interp ? c->w[f][row_inc + inc_x1] : vi;
        const float vd1 = interp ? c->w[f][row_dec + dec_x1] : vd;
        float bi = (1.0f - dr) * vi + dr * vi1;
        float bd = dr * vd + (1.0f - dr) * vd1;
        if (f == OR_VX) {
          bi = bi - relvel;
          bd = bd + relvel;
        }
        vi = bi;
        vd = bd;
      }
      c->w[f][dst_inc] = vi;
      if (!(f == OR_BY && sy == 0)) c->w[f][dst_dec] = vd; /* INL:277-278 */
    }
  }
}

/* Z halo exchange with itself, INL:446-708: plain periodic plane copies */
static void fill_z(or_ctx *c) {
  const or_grid *g = &c->g;
  const int M = g->m[2];
  if (M <= 0) return;
  const long unit = (long)g->n[0] * g->n[1] * M;
  for (int f = 0; f < 8; ++f) {
    float *p = c->w[f];
    memcpy(p, p + (c->ncell - 2 * unit), sizeof(float) * unit);
    memcpy(p + (c->ncell - unit), p + unit, sizeof(float) * unit);
  }
}

void or_fill_ghosts(or_ctx *c) {
  fill_y(c);
  fill_z(c);
}

/* ------------------------------------------------------------- stepping */

void or_rhs(or_ctx *c, float h, float *pace[8]) {
  rhs(c, h, c->w);
  for (int f = 0; f < 8; ++f)
    memcpy(pace[f], c->pace[f], sizeof(float) * c->ncell);
}

/* calc_flux_kernel<EX> alone (INL:760-816) for sweep d of the current state:
 * the 7 face-flux arrays of the lower d-faces (mass, momentum along X, Y, Z,
 * energy, B along d+1, B along d+2). */
void or_face_flux(or_ctx *c, int d, float h, float *out[7]) {
  calc_flux(c, d, h, c->w);
  for (int f = 0; f < 7; ++f) memcpy(out[f], c->flux[f], sizeof(float) * c->ncell);
}

/* MHDSolver::proceed loop body, INL:1094-1201 (INLRB:858-990) */
float or_step(or_ctx *c, float goal_time, int *goal_reached) {
  *goal_reached = 0;
  if (c->g.m[1] > 0 || c->g.m[2] > 0) or_fill_ghosts(c); /* communicate() */
  /* boundary_condition(): a no-op in every application (KEP:35-43) */
  float dt = or_cfl_dt(c, goal_time);
  const double next_time = c->t + dt;
  if (next_time > goal_time) { /* INL:1133-1137 */
    dt = (float)(goal_time - c->t);
    *goal_reached = 1;
  }
  const int src = (c->g.source == 1);
  if (c->g.integrator == 1) {
    /* euler1: integrator['dt','','',''] (INLRB:966-978) */
    rhs(c, dt, c->w);
    to_conserved(c, c->w);
    update(c, dt, c->w);
    to_primitive(c, c->w);
    if (src) external_acceleration(c, dt, c->w);
  } else {
    /* midpoint2 stage 1: integrator['dt/2','','','predictor_'] */
    const float h = dt / 2;
    rhs(c, h, c->w);
    to_conserved(c, c->w);
    add_pace(c, h, c->w, c->pred);
    to_primitive(c, c->pred);
    if (src) external_acceleration(c, h, c->pred);
    /* stage 2: integrator['dt','predictor_','',''] */
    rhs(c, dt, c->pred);
    to_conserved(c, c->w); /* conserved base recomputed from W^n (INL:1181) */
    update(c, dt, c->w);
    to_primitive(c, c->w);
    if (src) external_acceleration(c, dt, c->w);
  }
  c->t += dt; /* GlobalSetup::proceed, mhd_solver.hrb:56-59 */
  c->generation += 1;
  return dt;
}

int or_run(or_ctx *c, float goal_time, int max_steps, int *goal_reached) {
  int k = 0;
  *goal_reached = 0;
  while (k < max_steps && !*goal_reached) {
    or_step(c, goal_time, goal_reached);
    ++k;
  }
  return k;
}

/* ------------------------------------------------------------ lifecycle */

or_ctx *or_create(const or_grid *g) {
  if (g->n[0] <= 0 || g->n[1] <= 0 || g->n[2] <= 0 || g->m[0] != 0) return NULL;
  or_ctx *c = (or_ctx *)calloc(1, sizeof(or_ctx));
  if (!c) return NULL;
  c->g = *g;
  c->ncell = (long)g->n[0] * g->n[1] * g->n[2];
  float **all[] = {c->w, c->pred, c->pace};
  for (int s = 0; s < 3; ++s)
    for (int f = 0; f < 8; ++f) all[s][f] = (float *)calloc(c->ncell, sizeof(float));
  for (int f = 0; f < 3; ++f) c->mom[f] = (float *)calloc(c->ncell, sizeof(float));
  for (int f = 0; f < 7; ++f) c->flux[f] = (float *)calloc(c->ncell, sizeof(float));
  c->energy = (float *)calloc(c->ncell, sizeof(float));
  c->cfl = (float *)calloc(c->ncell, sizeof(float));
  return c;
}

void or_destroy(or_ctx *c) {
  if (!c) return;
  for (int f = 0; f < 8; ++f) {
    free(c->w[f]);
    free(c->pred[f]);
    free(c->pace[f]);
  }
  for (int f = 0; f < 3; ++f) free(c->mom[f]);
  for (int f = 0; f < 7; ++f) free(c->flux[f]);
  free(c->energy);
  free(c->cfl);
  free(c);
}

void or_set_state(or_ctx *c, float *const fields[8]) {
  for (int f = 0; f < 8; ++f) memcpy(c->w[f], fields[f], sizeof(float) * c->ncell);
}
void or_get_state(const or_ctx *c, float *fields[8]) {
  for (int f = 0; f < 8; ++f) memcpy(fields[f], c->w[f], sizeof(float) * c->ncell);
}
void or_get_predictor(const or_ctx *c, float *fields[8]) {
  for (int f = 0; f < 8; ++f) memcpy(fields[f], c->pred[f], sizeof(float) * c->ncell);
}
void or_set_time(or_ctx *c, double t, int generation) {
  c->t = t;
  c->generation = generation;
}
double or_get_time(const or_ctx *c) { return c->t; }
int or_get_generation(const or_ctx *c) { return c->generation; }
