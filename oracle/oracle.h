/*
 * oracle.h -- CPU restatement of the reference tissue time-stepper.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is linked or
 * loaded by tests/, by __graft_entry__.smoke() and by bench.py's CPU-baseline
 * leg, and by nothing else.  The product path (libyolohtli_amd.so) never
 * calls it.
 *
 * Every function restates a reference function (file:line cited at each
 * definition) in plain C with the reference's operation order.  The FP64
 * variants (suffix _d) follow the reference's REAL=double
 * (typeDefinition.cuh:4); the FP32 twins (suffix _f) use the same order in
 * float.  The Fenton-Karma model and the adjoint have no reference
 * counterpart; they restate the published equations (Fenton & Karma, Chaos
 * 8:20, 1998) and the discrete adjoint of this file's own forward step.
 */
#ifndef YH_ORACLE_H
#define YH_ORACLE_H

#include "../include/yolohtli.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_coef {
  double hx, hy, dt, Dxx, Dyy, rx, ry, invdx, invdy;
} orc_coef;

/* saveFiles.cu:140-170 and main.cu:148-150. */
void orc_derive(const yh_params* p, orc_coef* c);
/* coord_i / coord_j, helper_functions.cu:69-79. */
int orc_coord(int k, int n);

/* One explicit-Euler step (reactionDiffusion.cu:107-199,498-561).
 * in/out: u,v (2V) or u,v,w (FK).  vt may be NULL (velTan not written). */
void orc_rd_step_d(const yh_params* p, const uint8_t* solid,
                   const double* const in[3], double* const out[3], double* const vt[3]);
void orc_rd_step_f(const yh_params* p, const uint8_t* solid,
                   const float* const in[3], float* const out[3], float* const vt[3]);
/* nsteps of the plain loop (main.cu:876-885) in place on st[], OpenMP rows. */
void orc_run_d(const yh_params* p, const uint8_t* solid, double* const st[3],
               int nsteps, int nthreads);
void orc_run_f(const yh_params* p, const uint8_t* solid, float* const st[3],
               int nsteps, int nthreads);

/* orc_rd_step_d with imposed Heaviside decisions (FK): force_p[c] and
 * force_q[c] = 0 or 1 replace STEP(u - u_c) and STEP(u - u_v) of cell c, -1
 * keeps the cell's own.  A run that takes another implementation's
 * decisions where they differ from its own ("shadowing") separates the
 * effect of threshold ties from that of the update arithmetic. */
void orc_rd_step_forced_d(const yh_params* p, const uint8_t* solid,
                          const double* const in[3], double* const out[3],
                          const int8_t* force_p, const int8_t* force_q, int nthreads);

/* spiralTip_kernel + tipRecordPlane (tipTracker.cu:20-241), row-major cell
 * order, clockwise root before counter-clockwise root.  Returns the number of
 * tips found (entries beyond cap are not written). keys = 2*cell+root. */
int orc_tips_d(const yh_params* p, const double* g_past, const double* g_present,
               double time, yh_tip* out, int* keys, int cap);

/* ---------------------------------------------------- symmetry reduction */
/* Slice scheme 2 for one field (symmetryReduction.cu:72-116,167-210,224-262):
 * fx,fy upwind by the sign of the advection field built from (ac, aphi)
 * (symmetryReduction.cu:45-56); ft = hx*i*fy - hy*j*fx; *0 = template with the
 * 4th-order central differences.  Zero outside the disk centred at
 * (nx/2+cx, ny/2+cy) of radius^2 tipOffsetX*tipOffsetY. Any output may be NULL. */
void orc_slice_field_d(const yh_params* p, const double* f, int cx, int cy,
                       const double ac[3], double aphi,
                       double* fx, double* fy, double* ft,
                       double* fx0, double* fy0, double* ft0);
/* The 12 disk inner products (integralTrapz.cu:17-184), summed over all model
 * variables: I_k = hx*hy*sum_disk sum_var f*g. */
void orc_sr_integrals_d(const yh_params* p, const double* const X[3],
                        const double* const vt[3], int cx, int cy,
                        const double ac[3], double aphi, double I[12]);
/* solve_matrix (symmetryReduction.cu:386-418), unpivoted elimination. */
void orc_solve_matrix(const double phi[3], const double I[12], double cout[3]);
/* Race-free three-phase BFECC (advFDBFECC.cu:18-165) of every model
 * variable by the advection field of (c, phi_t) (Cxy_field_kernel,
 * symmetryReduction.cu:20-62). */
void orc_bfecc_d(const yh_params* p, const uint8_t* solid, const double* const in[3],
                 double* const out[3], const double c[3], double phit);

typedef struct orc_sr_state {
  double c[3], phi[3];       /* current c and phi (main.cu:158-161) */
  double ac[3], aphi;        /* (c, phi_t) of the last Cxy_field call */
  int cx, cy;                /* disk centre offset from (nx/2, ny/2) */
  float tip0x, tip0y;        /* tipx0_d, tipy0_d (main.cu:373-376) */
  long count;                /* param.count */
  int ntips;                 /* tips found in the last step */
  double I[12];              /* integrals of the last step */
} orc_sr_state;

void orc_sr_init(orc_sr_state* s, float tip0x, float tip0y);
/* One symmetry-reduction step (main.cu:894-953) in place on X.  Its pieces
 * (RD step, tips, slices, integrals, solve, Cxy, BFECC) are pinned one by one
 * in tests/, and so is the solve-then-advance order (c is the solve of the
 * step's own integrals with the phi from before the step; phi then advances
 * by dt*c).  Which state each remaining piece sees (RD output vs old state
 * for tips, slices and BFECC) is restated from main.cu:894-953 with no check
 * of its own: parity unpinned for that part. */
int orc_sr_step_d(const yh_params* p, const uint8_t* solid, double* const X[3],
                  orc_sr_state* s);

/* The whole reactionDiffusion_kernel (reactionDiffusion.cu:36-561): Neumann
 * or Dirichlet, square or masked, lap4, anisotropy, Euler/RK2/RK4 and the
 * mouse stimulus (stim_on at the point (stim_x, stim_y)), race-free (each
 * stage from the previous stage's field; lap4's neighbour currents from the
 * neighbour's stage state).  FP64.  vt may be NULL.  Returns 0, or -1 for an
 * unknown timeIntOrder. */
int orc_rd_step_general_d(const yh_params* p, const uint8_t* solid, int stim_on, int stim_x,
                          int stim_y, const double* const in[3], double* const out[3],
                          double* const vt[3]);

/* countour_kernel (spaceAPD.cu:18-153): mode 1 zero-crossings of f2 with
 * linear interpolation, mode 2 single and mode 3 double threshold contours
 * of f2 where f1 < th1; area (may be NULL = everywhere) is stimArea.  out
 * receives (x, y, t) per point, keys 2*cell+sub, in cell order.  Returns the
 * number of points (beyond cap not written), -1 for an unknown mode. */
int orc_contour_d(int nx, int ny, int mode, double th1, double th2, double th3,
                  const double* f1, const double* f2, const uint8_t* area, float t,
                  float* out, int* keys, int cap);

/* ------------------------------------------------------------------ APD */
/* sAPD_kernel without stimulus area (spaceAPD.cu:278-374, the
 * stimulate==false branch), threshold 0.15. */
void orc_apd_update_d(int ncell, double dt, long count, const double* uold,
                      const double* unew, double* apd1, double* apd2, double* front,
                      double* back, uint8_t* first);

/* orc_run_d with the sAPD update after every step, crossings in time order
 * (old state, new state) and threshold th; apd arrays updated in place. */
void orc_run_apd_d(const yh_params* p, const uint8_t* solid, double* const st[3], int nsteps,
                   long count0, double th, double* apd1, double* apd2, double* front,
                   double* back, uint8_t* first);

/* ---------------------------------------------------------------- adjoint */
/* FK forward with a spatial diffusion field D (cm^2/ms per cell), used by the
 * gradient path: u' = u + tc*(L_D u - dt*I), with
 * L_D u = dt/hx^2 sum_x c_dir*0.5*(D_c+D_n)*(u_n-u_c) + (same in y). */
void orc_fk_dfield_step_d(const yh_params* p, const uint8_t* solid, const double* D,
                          const double* const in[3], double* const out[3]);
/* Loss = sum_{n=1..nsteps, n % rec == 0} sum_s (u^n(sensor_s) - target)^2.
 * targets[(n/rec - 1)*ns + s].  Returns the loss; traces (may be NULL)
 * receives u at the sensors with the same layout. */
double orc_fk_loss_d(const yh_params* p, const uint8_t* solid, const double* D,
                     const double* const X0[3], int nsteps, int rec, int ns,
                     const int* sensors, const double* targets, double* traces);
/* Discrete adjoint of orc_fk_loss_d: gD[ncell] and gtau[6] =
 * dL/d(tau_d, tau_r, tau_si, tau_0, tau_vp, tau_wp).  Stores every forward
 * state (test sizes only). */
double orc_fk_adjoint_d(const yh_params* p, const uint8_t* solid, const double* D,
                        const double* const X0[3], int nsteps, int rec, int ns,
                        const int* sensors, const double* targets,
                        double* gD, double gtau[6]);

#ifdef __cplusplus
}
#endif
#endif
