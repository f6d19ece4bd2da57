/*
 * yh_shim.c -- the reference-side binding of yh_shim.h (see there).
 * Replaces reactionDiffusion_wrapper + swapSoA (main.cu:876-888),
 * tip_wrapper (main.cu:961-964, tipTracker.cu:569-611) and the SR step's
 * slice/trapz/solve_matrix/Cxy/advFDBFECC calls (main.cu:894-953) as
 * declared at hostPrototypes.h:22-54.
 */
#include "yh_shim.h"

#include <stdlib.h>

static yh_plan* g_plan;
static int g_sr_started;

int yh_shim_init(const paramVar* param, const bool* solid_h, stateVar g_h) {
  yh_params p;
  yh_params_default(&p);
  p.nx = param->nx;
  p.ny = param->ny;
  p.Lx = param->Lx;
  p.Ly = param->Ly;
  p.dt = param->dt; /* the library halves it when reduceSym (main.cu:148) */
  p.diff_par = param->diff_par;
  p.diff_per = param->diff_per;
  p.rscale = param->rscale;
  p.tc = param->tc;
  p.alpha = param->alpha;
  p.beta = param->beta;
  p.gamma = param->gamma;
  p.delta = param->delta;
  p.eps = param->eps;
  p.mu = param->mu;
  p.theta = param->theta;
  p.neumannBC = param->neumannBC;
  p.solidSwitch = param->solidSwitch;
  p.gateDiff = param->gateDiff;
  p.reduceSym = param->reduceSym;
  p.timeIntOrder = param->timeIntOrder; /* 1: the fast kernels; 2, 4: the stage path */
  p.lap4 = param->lap4 ? 1 : 0;
  p.anisotropy = param->anisotropy;
  p.degrad = param->degrad;
  p.boundaryVal = param->boundaryVal;
  p.Uth = param->Uth;
  p.tipOffsetX = param->tipOffsetX;
  p.tipOffsetY = param->tipOffsetY;
  yh_shim_close();
  int rc = yh_plan_create(&p, &g_plan);
  if (rc) return rc;
  if (p.solidSwitch) {
    const size_t n = (size_t)p.nx * (size_t)p.ny;
    uint8_t* m = (uint8_t*)malloc(n);
    if (!m) return YH_ENOMEM;
    for (size_t k = 0; k < n; ++k) m[k] = solid_h[k] ? 1 : 0;
    rc = yh_set_mask(g_plan, m);
    free(m);
    if (rc) return rc;
  }
  return yh_set_state(g_plan, g_h.u, g_h.v, NULL);
}

int yh_shim_plain(int sampleIt, bool plotTip) {
  int rc = yh_step(g_plan, sampleIt, NULL);
  if (!rc && plotTip) rc = yh_track_tips(g_plan, NULL);
  return rc;
}

int yh_shim_sr(int nsteps, float tipx0, float tipy0, double* c_phi) {
  if (!g_sr_started) {
    const int rc = yh_sr_begin(g_plan, tipx0, tipy0, NULL);
    if (rc) return rc;
    g_sr_started = 1;
  }
  return yh_sr_step(g_plan, nsteps, NULL, c_phi);
}

int yh_shim_read(stateVar g_h, vec5dyn* tips, int cap, int* ntips) {
  const int rc = yh_get_state(g_plan, g_h.u, g_h.v, NULL);
  if (rc) return rc;
  /* yh_tip is {x, y, vx, vy, t}: the layout of vec5dyn */
  return yh_read_tips(g_plan, (yh_tip*)tips, cap, ntips);
}

void yh_shim_close(void) {
  if (g_plan) yh_plan_destroy(g_plan);
  g_plan = NULL;
  g_sr_started = 0;
}

int yh_example_batch(const yh_params* per_tissue, int nt, uint8_t** masks, float** u,
                     float** v, float** w, int nsteps) {
  yh_batch* b;
  int rc = yh_batch_create(per_tissue, nt, &b);
  if (rc) return rc;
  for (int t = 0; !rc && t < nt; ++t) {
    if (per_tissue[t].solidSwitch) rc = yh_batch_set_mask(b, t, masks[t]);
    if (!rc) rc = yh_batch_set_state(b, t, u[t], v[t], w[t]);
  }
  if (!rc) rc = yh_batch_step(b, nsteps, NULL); /* graph-captured on the first call */
  for (int t = 0; !rc && t < nt; ++t) rc = yh_batch_get_state(b, t, u[t], v[t], w[t]);
  yh_batch_destroy(b);
  return rc;
}

int yh_example_gradient(const yh_params* p, int nsteps, const int* sensors, int ns,
                        const float* u0, const float* v0, const float* w0,
                        const float* targets, const float* d_D, const double tau6[6],
                        double* d_gD, double gtau[6], double* loss) {
  yh_adjoint* a;
  int rc = yh_adjoint_create(p, nsteps, 1, 100, sensors, ns, &a);
  if (rc) return rc;
  rc = yh_adjoint_set_initial(a, u0, v0, w0);
  if (!rc) rc = yh_adjoint_set_targets(a, targets); /* nsteps x ns */
  if (!rc) rc = yh_adjoint_forward(a, d_D, tau6, NULL, loss, NULL);
  if (!rc) rc = yh_adjoint_backward(a, NULL, d_gD, gtau);
  yh_adjoint_destroy(a);
  return rc;
}
