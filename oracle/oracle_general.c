/*
 * oracle_general.c -- the reference's whole reactionDiffusion_kernel
 * (reactionDiffusion.cu:36-561) in FP64: Neumann and Dirichlet boundaries,
 * square and masked domains, the 4th-order compact Laplacian, the anisotropic
 * cross term, Euler / RK2 / RK4 and the circular mouse stimulus.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The reference evaluates its Runge-Kutta stages in place (g_in is
 * overwritten with the stage state while neighbours read it, :114-119) and
 * reads the lap4 neighbour currents J from the same launch (:221-236): both
 * race.  Restated race-free: every stage is evaluated for all cells from the
 * stage field y_s = y0 + ki[s]*d_{s-1} (d = the previous stage's increments
 * of all cells), and a neighbour's current is evaluated from that
 * neighbour's stage state.  Boundary-condition reads the reference makes
 * outside the array (the masked Dirichlet branch at the domain edge,
 * :364-373) count as non-tissue here.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define ORC_PI (3.14159265359) /* globalVariables.cuh:27 */

typedef struct gk {
  int nx, ny, model, masked, neumann, lap4, aniso, gateDiff, order;
  double rx, ry, rxy, rbx, rby, qx4, qy4, fx4, fy4, dt, tc, rscale, bval;
  double mu, alpha, beta, gamma_, delta, eps, theta;
  const yh_fk_params* fk;
  int stim_on, stim_x, stim_y;
} gk;

static void gk_load(const yh_params* p, int stim_on, int stim_x, int stim_y, gk* k) {
  /* parameterSetup (saveFiles.cu:144-170) and main.cu:148-155 */
  const double hx = p->Lx / (p->nx - 1.0), hy = p->Ly / (p->ny - 1.0);
  const double th = p->degrad * ORC_PI / 180.0;
  const double Dxx = p->diff_par * cos(th) * cos(th) + p->diff_per * sin(th) * sin(th);
  const double Dyy = p->diff_par * sin(th) * sin(th) + p->diff_per * cos(th) * cos(th);
  const double Dxy = (p->diff_par - p->diff_per) * sin(th) * cos(th);
  const double dt = p->reduceSym ? 0.5 * p->dt : p->dt;
  k->nx = p->nx;
  k->ny = p->ny;
  k->model = p->model;
  k->masked = p->solidSwitch;
  k->neumann = p->neumannBC;
  k->lap4 = p->lap4;
  k->aniso = p->anisotropy;
  k->gateDiff = p->gateDiff;
  k->order = p->timeIntOrder;
  k->rx = dt * Dxx / (hx * hx);
  k->ry = dt * Dyy / (hy * hy);
  k->rxy = 2.0 * Dxy * dt / (4.0 * hx * hy);
  k->rbx = hx * Dxy / (Dxx * hy);
  k->rby = hy * Dxy / (Dyy * hx);
  k->qx4 = dt * Dyy / (hy * hy * 12.0);
  k->qy4 = dt * Dxx / (hx * hx * 12.0);
  k->fx4 = dt / 12.0;
  k->fy4 = dt / 12.0;
  k->dt = dt;
  k->tc = p->tc;
  k->rscale = p->rscale;
  k->bval = p->boundaryVal;
  k->mu = p->mu;
  k->alpha = p->alpha;
  k->beta = p->beta;
  k->gamma_ = p->gamma;
  k->delta = p->delta;
  k->eps = p->eps;
  k->theta = p->theta;
  k->fk = &p->fk;
  k->stim_on = stim_on;
  k->stim_x = stim_x;
  k->stim_y = stim_y;
}

#define I2D(i, j) ((i) + k->nx * (j))

/* Circular mouse stimulus (reactionDiffusion.cu:54-61): cells within
 * radius^2 < 400 of the point while the stimulus is on. */
static int g_scs(const gk* k, int i, int j) {
  const int ic = i - k->nx / 2, jc = j - k->ny / 2;
  const int cx = k->stim_x - k->nx / 2, cy = k->stim_y - k->ny / 2;
  return k->stim_on && ((ic - cx) * (ic - cx) + (jc - cy) * (jc - cy) < 400);
}

/* Ionic currents at cell (i,j) of the stage field y: Iu (I_sum, with the
 * mouse stimulus, :131-134), Iv (I_v, :141) and, for Fenton-Karma, Iw.  The
 * update is d -= dt*I (:498-499). */
static void g_currents(const gk* k, const double* const y[3], int i, int j, double* Iu,
                       double* Iv, double* Iw) {
  const int c = I2D(i, j);
  const double u = y[0][c], v = y[1][c];
  const double st = g_scs(k, i, j) ? 24.7 : 0.0;
  if (k->model == YH_MODEL_2V_REF) {
    *Iu = -(k->mu * u * (1.0 - u) * (u - k->alpha) - u * v) - st;
    *Iv = -(k->eps * (k->delta * (u - k->gamma_) * (k->beta - u) - v - k->theta));
    *Iw = 0.0;
  } else {
    /* Fenton & Karma 1998 (sharp Heaviside); the same stimulus current. */
    const yh_fk_params* f = k->fk;
    const double w = y[2][c];
    const double p = u >= f->u_c ? 1.0 : 0.0, q = u >= f->u_v ? 1.0 : 0.0;
    const double J_fi = -v * p * (1.0 - u) * (u - f->u_c) / f->tau_d;
    const double J_so = u * (1.0 - p) / f->tau_0 + p / f->tau_r;
    const double J_si = -w * (1.0 + tanh(f->k * (u - f->u_csi))) / (2.0 * f->tau_si);
    const double tvm = (1.0 - q) * f->tau_v1m + q * f->tau_v2m;
    *Iu = J_fi + J_so + J_si - st;
    *Iv = -((1.0 - p) * (1.0 - v) / tvm - p * v / f->tau_vp);
    *Iw = -((1.0 - p) * (1.0 - w) / f->tau_wm - p * w / f->tau_wp);
  }
}

/* Diagonal neighbours of the Neumann square branch (:207-221, :250-264). */
static void g_diag(const gk* k, int i, int j, int* SW, int* SE, int* NW, int* NE) {
  const int nx = k->nx, ny = k->ny;
  *SW = (i > 0 && j > 0) ? I2D(i - 1, j - 1)
        : ((i == 0 && j > 0) ? I2D(i + 1, j - 1)
                             : ((i > 0 && j == 0) ? I2D(i - 1, j + 1) : I2D(i + 1, j + 1)));
  *SE = (i < nx - 1 && j > 0) ? I2D(i + 1, j - 1)
        : ((i == nx - 1 && j > 0) ? I2D(i - 1, j - 1)
                                  : ((i < nx - 1 && j == 0) ? I2D(i + 1, j + 1) : I2D(i - 1, j + 1)));
  *NW = (i > 0 && j < ny - 1) ? I2D(i - 1, j + 1)
        : ((i == 0 && j < ny - 1) ? I2D(i + 1, j + 1)
                                  : ((i > 0 && j == ny - 1) ? I2D(i - 1, j - 1) : I2D(i + 1, j - 1)));
  *NE = (i < nx - 1 && j < ny - 1) ? I2D(i + 1, j + 1)
        : ((i == nx - 1 && j < ny - 1) ? I2D(i - 1, j + 1)
                                       : ((i < nx - 1 && j == ny - 1) ? I2D(i + 1, j - 1) : I2D(i - 1, j - 1)));
}

/* Anisotropic term of one field f on the Neumann square branch (:266-313),
 * scaled by s (1 for u, rscale for v). */
static double g_aniso_neumann(const gk* k, const double* f, int i, int j, double s) {
  const int nx = k->nx, ny = k->ny, c = I2D(i, j);
  int SW, SE, NW, NE;
  g_diag(k, i, j, &SW, &SE, &NW, &NE);
  const double b_S = (j > 0) ? 0.0 : ((j == 0 && (i == 0 || i == nx - 1)) ? 0.0
                     : k->rby * (f[I2D(i + 1, j)] - f[I2D(i - 1, j)]));
  const double b_N = (j < ny - 1) ? 0.0 : ((j == ny - 1 && (i == 0 || i == nx - 1)) ? 0.0
                     : -k->rby * (f[I2D(i + 1, j)] - f[I2D(i - 1, j)]));
  const double b_W = (i > 0) ? 0.0 : ((i == 0 && (j == 0 || j == ny - 1)) ? 0.0
                     : k->rbx * (f[I2D(i, j + 1)] - f[I2D(i, j - 1)]));
  const double b_E = (i < nx - 1) ? 0.0 : ((i == nx - 1 && (j == 0 || j == ny - 1)) ? 0.0
                     : -k->rbx * (f[I2D(i, j + 1)] - f[I2D(i, j - 1)]));
  double d = ((b_S + b_N) * k->ry + (b_W + b_E) * k->rx) * s;
  const double b_SW = (i > 0 && j > 0) ? 0.0
                      : ((i == 0 && j > 1) ? k->rbx * (f[c] - f[I2D(i, j - 2)])
                         : ((i > 1 && j == 0) ? k->rby * (f[c] - f[I2D(i - 2, j)]) : 0.0));
  const double b_SE = (i < nx - 1 && j > 0) ? 0.0
                      : ((i == nx - 1 && j > 1) ? -k->rbx * (f[c] - f[I2D(i, j - 2)])
                         : ((i < nx - 2 && j == 0) ? k->rby * (f[I2D(i + 2, j)] - f[c]) : 0.0));
  const double b_NW = (i > 0 && j < ny - 1) ? 0.0
                      : ((i == 0 && j < ny - 2) ? k->rbx * (f[I2D(i, j + 2)] - f[c])
                         : ((i > 1 && j == ny - 1) ? -k->rby * (f[c] - f[I2D(i - 2, j)]) : 0.0));
  const double b_NE = (i < nx - 1 && j < ny - 1) ? 0.0
                      : ((i == nx - 1 && j < ny - 2) ? -k->rbx * (f[I2D(i, j + 2)] - f[c])
                         : ((i < nx - 2 && j == ny - 1) ? -k->rby * (f[I2D(i + 2, j)] - f[c]) : 0.0));
  d += k->rxy * ((f[SW] + b_SW) + (f[NE] + b_NE) - (f[SE] + b_SE) - (f[NW] + b_NW)) * s;
  return d;
}

/* Neighbour value for the Dirichlet branch: the field where the neighbour
 * (ni, nj) is inside the domain and, with a mask, it and the cell are tissue
 * (:362-373, :399-408, :434-437); the boundary value otherwise. */
static double g_dir(const gk* k, const uint8_t* solid, const double* f, int sc, int ni, int nj) {
  if (ni < 0 || ni >= k->nx || nj < 0 || nj >= k->ny) return k->bval;
  const int n = I2D(ni, nj);
  if (k->masked && !(sc && solid[n])) return k->bval;
  return f[n];
}

/* Increments d = F(y) of one stage at cell (i,j) (:114-503, one pass of the
 * rk4idx loop, before the sumrk4 accumulation). */
static void g_stage_cell(const gk* k, const uint8_t* solid, const double* const y[3], int i,
                         int j, double d[3]) {
  const int nx = k->nx, ny = k->ny, c = I2D(i, j);
  const double* gu = y[0];
  const double* gv = y[1];
  const double u = gu[c], v = gv[c];
  const int fk = k->model == YH_MODEL_3V_FK;
  const int vdiff = !fk && k->gateDiff;
  double Iu, Iv, Iw;
  g_currents(k, y, i, j, &Iu, &Iv, &Iw);
  double du = 0.0, dv = 0.0;
  if (k->neumann) {
    const int S = I2D(i, orc_coord(j - 1, ny)), N = I2D(i, orc_coord(j + 1, ny));
    const int W = I2D(orc_coord(i - 1, nx), j), E = I2D(orc_coord(i + 1, nx), j);
    if (k->masked) {
      /* :154-185 (lap4 and the anisotropic term are not applied here, as in the reference) */
      const int sc = solid[c] != 0, sw = solid[W] != 0, se = solid[E] != 0;
      const int sn = solid[N] != 0, ss = solid[S] != 0;
      const double cxx = (sw && se) && (sw && sc) ? 1.0 : ((sw && sc) ? 2.0 : 0.0);
      const double cxy = sc ? ((sw || se) ? 2.0 : 0.0) : 0.0;
      const double cxz = (sw && se) && (sc && se) ? 1.0 : ((sc && se) ? 2.0 : 0.0);
      const double cyx = (sn && ss) && (sn && sc) ? 1.0 : ((sn && sc) ? 2.0 : 0.0);
      const double cyy = sc ? ((sn || ss) ? 2.0 : 0.0) : 0.0;
      const double cyz = (sn && ss) && (sc && ss) ? 1.0 : ((sc && ss) ? 2.0 : 0.0);
      du = (cxx * gu[W] - cxy * u + cxz * gu[E]) * k->rx + (cyx * gu[N] - cyy * u + cyz * gu[S]) * k->ry;
      if (vdiff)
        dv = (cxx * gv[W] - cxy * v + cxz * gv[E]) * k->rx * k->rscale
           + (cyx * gv[N] - cyy * v + cyz * gv[S]) * k->ry * k->rscale;
    } else {
      /* :188-199 */
      du = (gu[W] - 2.0 * u + gu[E]) * k->rx + (gu[N] - 2.0 * u + gu[S]) * k->ry;
      if (vdiff)
        dv = (gv[W] - 2.0 * v + gv[E]) * k->rx * k->rscale + (gv[N] - 2.0 * v + gv[S]) * k->ry * k->rscale;
      if (k->lap4) {
        /* :201-247; J of a neighbour = its current in this stage */
        int SW, SE, NW, NE;
        g_diag(k, i, j, &SW, &SE, &NW, &NE);
        double JW[3], JE[3], JN[3], JS[3];
        g_currents(k, y, orc_coord(i - 1, nx), j, &JW[0], &JW[1], &JW[2]);
        g_currents(k, y, orc_coord(i + 1, nx), j, &JE[0], &JE[1], &JE[2]);
        g_currents(k, y, i, orc_coord(j + 1, ny), &JN[0], &JN[1], &JN[2]);
        g_currents(k, y, i, orc_coord(j - 1, ny), &JS[0], &JS[1], &JS[2]);
        du += -2.0 * (k->qx4 + k->qy4) * ((gu[W] - u + gu[E]) + (gu[N] - u + gu[S]));
        du += (k->qx4 + k->qy4) * (gu[SW] + gu[SE] + gu[NW] + gu[NE]);
        du -= ((JW[0] - 2.0 * Iu + JE[0]) * k->fx4 + (JN[0] - 2.0 * Iu + JS[0]) * k->fy4);
        if (vdiff) {
          dv += -k->rscale * 2.0 * (k->qx4 + k->qy4) * ((gv[W] - v + gv[E]) + (gv[N] - v + gv[S]));
          dv += k->rscale * (k->qx4 + k->qy4) * (gv[SW] + gv[SE] + gv[NW] + gv[NE]);
          dv -= ((JW[1] - 2.0 * Iv + JE[1]) * k->fx4 + (JN[1] - 2.0 * Iv + JS[1]) * k->fy4);
        }
      }
      if (k->aniso) {
        /* :249-356 */
        du += g_aniso_neumann(k, gu, i, j, 1.0);
        if (vdiff) dv += g_aniso_neumann(k, gv, i, j, k->rscale);
      }
    }
  } else {
    /* Dirichlet (:360-490) */
    const int sc = k->masked ? solid[c] != 0 : 1;
    const double uS = g_dir(k, solid, gu, sc, i, j - 1), uN = g_dir(k, solid, gu, sc, i, j + 1);
    const double uW = g_dir(k, solid, gu, sc, i - 1, j), uE = g_dir(k, solid, gu, sc, i + 1, j);
    du = (uW - 2.0 * u + uE) * k->rx + (uN - 2.0 * u + uS) * k->ry;
    if (vdiff) {
      const double vS = g_dir(k, solid, gv, sc, i, j - 1), vN = g_dir(k, solid, gv, sc, i, j + 1);
      const double vW = g_dir(k, solid, gv, sc, i - 1, j), vE = g_dir(k, solid, gv, sc, i + 1, j);
      dv = (vW - 2.0 * v + vE) * k->rx * k->rscale + (vN - 2.0 * v + vS) * k->ry * k->rscale;
    }
    if (k->aniso) {
      const double aSW = g_dir(k, solid, gu, sc, i - 1, j - 1), aSE = g_dir(k, solid, gu, sc, i + 1, j - 1);
      const double aNW = g_dir(k, solid, gu, sc, i - 1, j + 1), aNE = g_dir(k, solid, gu, sc, i + 1, j + 1);
      du += k->rxy * (aSW + aNE - aSE - aNW);
      if (vdiff) {
        const double bSW = g_dir(k, solid, gv, sc, i - 1, j - 1), bSE = g_dir(k, solid, gv, sc, i + 1, j - 1);
        const double bNW = g_dir(k, solid, gv, sc, i - 1, j + 1), bNE = g_dir(k, solid, gv, sc, i + 1, j + 1);
        dv += k->rxy * (bSW + bNE - bSE - bNW) * k->rscale;
      }
    }
  }
  /* :498-499 */
  d[0] = du - k->dt * Iu;
  d[1] = dv - k->dt * Iv;
  d[2] = fk ? -k->dt * Iw : 0.0;
}

int orc_rd_step_general_d(const yh_params* p, const uint8_t* solid, int stim_on, int stim_x,
                          int stim_y, const double* const in[3], double* const out[3],
                          double* const vt[3]) {
  gk kk, *k = &kk;
  gk_load(p, stim_on, stim_x, stim_y, k);
  /* :70-95 (RK4 weights as the reference writes them) */
  double ki[4] = {0, 0, 0, 0}, sw[4] = {0, 0, 0, 0};
  switch (k->order) {
    case 1: sw[0] = 1.0; break;
    case 2: ki[1] = 0.5; sw[1] = 1.0; break;
    case 4:
      ki[1] = 0.5; ki[2] = 0.5; ki[3] = 1.0;
      sw[0] = 0.166666666666667; sw[1] = 0.333333333333333;
      sw[2] = 0.333333333333333; sw[3] = 0.166666666666667;
      break;
    default: return -1;
  }
  const int nv = p->model == YH_MODEL_3V_FK ? 3 : 2;
  const size_t n = (size_t)k->nx * k->ny;
  double* buf = (double*)calloc(n * 3 * 3, sizeof(double));
  if (!buf) return -2;
  double* ys[3] = {buf, buf + n, buf + 2 * n};          /* stage field */
  double* dp[3] = {buf + 3 * n, buf + 4 * n, buf + 5 * n}; /* previous stage d */
  double* rhs[3] = {buf + 6 * n, buf + 7 * n, buf + 8 * n};
  const double* ysc[3] = {ys[0], ys[1], ys[2]};
  for (int s = 0; s < k->order; ++s) {
    for (int a = 0; a < nv; ++a)
      for (size_t q = 0; q < n; ++q) ys[a][q] = in[a][q] + ki[s] * dp[a][q]; /* :116-117 */
    for (int j = 0; j < k->ny; ++j)
      for (int i = 0; i < k->nx; ++i) {
        double d[3];
        g_stage_cell(k, solid, ysc, i, j, d);
        const int c = I2D(i, j);
        for (int a = 0; a < nv; ++a) {
          rhs[a][c] += sw[s] * d[a]; /* :501-502 */
          dp[a][c] = d[a];
        }
      }
  }
  /* :506-561 */
  for (int j = 0; j < k->ny; ++j)
    for (int i = 0; i < k->nx; ++i) {
      const int c = I2D(i, j);
      const int sc = k->masked ? solid[c] != 0 : 1;
      for (int a = 0; a < nv; ++a) {
        const double x = in[a][c] + k->tc * rhs[a][c];
        out[a][c] = sc ? x : 0.0;
        if (vt) vt[a][c] = sc ? rhs[a][c] / k->dt : 0.0;
      }
    }
  free(buf);
  return 0;
}

/* ------------------------------------------------------------- contours */
/* countour_kernel (spaceAPD.cu:18-153) for one cell, points appended in
 * (cell, sub) order.  The reference reads field2 at (i+1, j) and (i, j+1)
 * even on the last column / row (a wrapped or out-of-bounds element, used by
 * mode 1's interpolation and mode 3's band test); here the cell's own value
 * stands in, which makes that direction crossing-free as the zpm tests
 * already assume.  mode 1's condition keeps the reference's precedence:
 * zpmx < 0 || (zpmy < 0 && sc). */
int orc_contour_d(int nx, int ny, int mode, double th1, double th2, double th3,
                  const double* f1, const double* f2, const uint8_t* area, float t,
                  float* out, int* keys, int cap) {
  const double eps = 2.220446049250313e-16; /* MACHINE_EPS, globalVariables.cuh:26 */
  int n = 0;
#define ORC_PUSH(px, py, sub)                                        \
  do {                                                               \
    if (n < cap) {                                                   \
      out[3 * n] = (float)(px);                                      \
      out[3 * n + 1] = (float)(py);                                  \
      out[3 * n + 2] = t;                                            \
      keys[n] = 2 * c + (sub);                                       \
    }                                                                \
    ++n;                                                             \
  } while (0)
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const int c = i + nx * j;
      const int sc = area ? area[c] != 0 : 1;
      const double f2c = f2[c];
      const double f2x = i < nx - 1 ? f2[c + 1] : f2c;
      const double f2y = j < ny - 1 ? f2[c + nx] : f2c;
      if (mode == 1) {
        const double v0 = f2c, v1x = f2x, v1y = f2y;
        const double zpmx = i < nx - 1 ? v0 * v1x : v0 * v0;
        const double zpmy = j < ny - 1 ? v0 * v1y : v0 * v0;
        if ((zpmx < 0.0) || ((zpmy < 0.0) && sc)) {
          const double ppx = fabs(v0 - v1x) > eps ? i + v0 / (v0 - v1x) : i;
          const double ppy = fabs(v0 - v1y) > eps ? j + v0 / (v0 - v1y) : j;
          ORC_PUSH(ppx, ppy, 0);
        }
      } else if (mode == 2) {
        if ((f1[c] < th1) && sc) {
          const double v0 = f2c - th2, v1x = f2x - th2, v1y = f2y - th2;
          const double zpmx = i < nx - 1 ? v0 * v1x : v0 * v0;
          const double zpmy = j < ny - 1 ? v0 * v1y : v0 * v0;
          if ((zpmx < 0.0) || (zpmy < 0.0)) ORC_PUSH(i, j, 0);
        }
      } else if (mode == 3) {
        const double V0 = f2c - th2, V1x = f2x - th2, V1y = f2y - th2;
        if ((f1[c] < th1) && (fabs(V0) < (th3 + 0.1)) && (fabs(V1x) < (th3 + 0.1)) &&
            (fabs(V1y) < (th3 + 0.1)) && sc) {
          double v0 = V0 - th2, v1x = V1x - th2, v1y = V1y - th2;
          double zpmx = i < nx - 1 ? v0 * v1x : v0 * v0;
          double zpmy = j < ny - 1 ? v0 * v1y : v0 * v0;
          if ((zpmx < 0.0) || (zpmy < 0.0)) ORC_PUSH(i, j, 0);
          v0 = V0 + th3;
          v1x = V1x + th3;
          v1y = V1y + th3;
          zpmx = i < nx - 1 ? v0 * v1x : v0 * v0;
          zpmy = j < ny - 1 ? v0 * v1y : v0 * v0;
          if ((zpmx < 0.0) || (zpmy < 0.0)) ORC_PUSH(i, j, 1);
        }
      } else {
        return -1;
      }
    }
#undef ORC_PUSH
  return n;
}
