/*
 * oracle.c -- CPU restatement of the reference tissue time-stepper.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * CPU-baseline leg of bench.py as the checker.  The product library never
 * links it.  See oracle.h for the list of restated reference functions; each
 * definition below cites the reference lines it follows.
 *
 * Build: oracle/Makefile (gcc, -ffp-contract=off so the FP64 results follow
 * the written operation order; OpenMP for the row loop of orc_run_*).
 */
#include "oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

/* pi as the reference defines it (globalVariables.cuh:27). */
#define ORC_PI (3.14159265359)

void orc_derive(const yh_params* p, orc_coef* c) {
  /* saveFiles.cu:144-150: hx = Lx/(nx-1), Dxx/Dyy from the fibre angle. */
  c->hx = p->Lx / (p->nx - 1.0);
  c->hy = p->Ly / (p->ny - 1.0);
  const double th = p->degrad * ORC_PI / 180.0;
  c->Dxx = p->diff_par * cos(th) * cos(th) + p->diff_per * sin(th) * sin(th);
  c->Dyy = p->diff_par * sin(th) * sin(th) + p->diff_per * cos(th) * cos(th);
  /* main.cu:148-150: dt halved in symmetry-reduction mode. */
  c->dt = p->reduceSym ? 0.5 * p->dt : p->dt;
  c->rx = c->dt * c->Dxx / (c->hx * c->hx);
  c->ry = c->dt * c->Dyy / (c->hy * c->hy);
  /* saveFiles.cu:164-165 */
  c->invdx = 0.5 / c->hx;
  c->invdy = 0.5 / c->hy;
}

/* helper_functions.cu:69-79 */
int orc_coord(int k, int n) {
  return (int)((k >= 0) && (k < n)) * k + (int)(k < 0) * (-k) + (int)(k >= n) * (2 * (n - 1) - k);
}

#define REAL double
#define SFX d
#include "oracle_rd.inc"
#undef REAL
#undef SFX

#define REAL float
#define SFX f
#include "oracle_rd.inc"
#undef REAL
#undef SFX

/* ------------------------------------------------------------------ tips */

/* tipRecordPlane, tipTracker.cu:210-241 (tipGrad off: vx = vy = 0). */
static void orc_tip_record(int i, int j, float tx, float ty, double t, int key,
                           yh_tip* out, int* keys, int cap, int* n) {
  if (((tx > 0.0) && (tx < 1.0)) && ((ty > 0.0) && (ty < 1.0))) {
    if (*n < cap) {
      yh_tip d;
      d.x = (float)((float)i + tx);
      d.y = (float)((float)j + ty);
      d.vx = 0.0f;
      d.vy = 0.0f;
      d.t = (float)t;
      out[*n] = d;
      if (keys) keys[*n] = key;
    }
    (*n)++;
  }
}

/* spiralTip_kernel, tipTracker.cu:20-208. */
int orc_tips_d(const yh_params* p, const double* g_past, const double* g_present,
               double time, yh_tip* out, int* keys, int cap) {
  const int nx = p->nx, ny = p->ny;
  const double Uth = p->Uth;
  int n = 0;
  for (int j = 0; j < ny; ++j) {
    for (int i = 0; i < nx; ++i) {
      int inside;
      if (p->solidSwitch) {
        const int ic = i - nx / 2, jc = j - ny / 2;
        inside = (ic * ic + jc * jc) < p->tipOffsetX * p->tipOffsetY;
      } else {
        inside = (i >= 1) && (i < (nx - 2)) && (j >= 1) && (j < (ny - 2));
      }
      if (!inside) continue;
      const int s0 = i + nx * j;
      const int sx = (i < (nx - 1)) ? (i + 1) + nx * j : s0;
      const int sy = (j < (ny - 1)) ? i + nx * (j + 1) : s0;
      const int sxy = ((j < (ny - 1)) && (i < (nx - 1))) ? (i + 1) + nx * (j + 1) : s0;
      const double x1 = g_present[s0], x2 = g_present[sx], x4 = g_present[sy], x3 = g_present[sxy];
      const double y1 = g_past[s0], y2 = g_past[sx], y4 = g_past[sy], y3 = g_past[sxy];
      const double x3y1 = x3 * y1, x4y1 = x4 * y1, x3y2 = x3 * y2, x4y2 = x4 * y2;
      const double x1y3 = x1 * y3, x2y3 = x2 * y3, x1y4 = x1 * y4, x2y4 = x2 * y4;
      const double x2y1 = x2 * y1, x1y2 = x1 * y2, x4y3 = x4 * y3, x3y4 = x3 * y4;
      const double den1 = 2.0 * (x3y1 - x4y1 - x3y2 + x4y2 - x1y3 + x2y3 + x1y4 - x2y4);
      const double den2 = 2.0 * (x2y1 - x3y1 - x1y2 + x4y2 + x1y3 - x4y3 - x2y4 + x3y4);
      const double ctn1 = x1 - x2 + x3 - x4 - y1 + y2 - y3 + y4;
      const double ctn2 = x3y1 - 2.0 * x4y1 + x4y2 - x1y3 + 2.0 * x1y4 - x2y4;
      const double disc = sqrt(4.0 * (x3y1 - x3y2 - x4y1 + x4y2 - x1y3 + x1y4 + x2y3 - x2y4)
                               * (x4y1 - x1y4 + Uth * (x1 - x4 - y1 + y4))
                               + (-ctn2 + Uth * ctn1) * (-ctn2 + Uth * ctn1));
      const double px = ctn2 - Uth * ctn1;
      const double py = Uth * ctn1 - x3y1 + x4y2 + x1y3 - x2y4 + 2.0 * (x2y1 - x1y2);
      const int check_disc = !p->solidSwitch; /* tipTracker.cu:185,197 */
      if (!check_disc || disc >= 0.0) {
        const float tx = (float)((px + disc) / den1);
        const float ty = (float)((py + disc) / den2);
        orc_tip_record(i, j, tx, ty, time, 2 * s0, out, keys, cap, &n);
      }
      if (!check_disc || disc >= 0.0) {
        const float tx = (float)((px - disc) / den1);
        const float ty = (float)((py - disc) / den2);
        orc_tip_record(i, j, tx, ty, time, 2 * s0 + 1, out, keys, cap, &n);
      }
    }
  }
  return n;
}

/* ---------------------------------------------------- symmetry reduction */

/* Advection field of Cxy_field_kernel, symmetryReduction.cu:45-56 (x = i, y = j). */
static inline void orc_adv(const orc_coef* cf, int i, int j, const double c[3], double phit,
                           double* ax, double* ay) {
  const double x = (double)i, y = (double)j;
  *ax = cf->hy * y * c[2] - c[0] * cos(phit) + c[1] * sin(phit);
  *ay = -cf->hx * x * c[2] - c[0] * sin(phit) - c[1] * cos(phit);
}

static inline int orc_in_disk(const yh_params* p, int i, int j, int cx, int cy) {
  /* symmetryReduction.cu:95-107, integralTrapz.cu:39-51 */
  const int ic = i - p->nx / 2, jc = j - p->ny / 2;
  return ((ic - cx) * (ic - cx) + (jc - cy) * (jc - cy)) < p->tipOffsetX * p->tipOffsetY;
}

typedef struct orc_sl {
  double fx, fy, ft, fx0, fy0, ft0;
} orc_sl;

/* slice_kernel scheme 2 at one cell (symmetryReduction.cu:167-202) with
 * convFB2ndOX/Y (:242-262) and convCentral2X/Y (:224-240). */
static orc_sl orc_slice_cell(const yh_params* p, const orc_coef* cf, const double* f,
                             int i, int j, int cx, int cy, const double ac[3], double aphi) {
  orc_sl r = {0, 0, 0, 0, 0, 0};
  if (!orc_in_disk(p, i, j, cx, cy)) return r;
  const int nx = p->nx, ny = p->ny;
  const int C = i + nx * j;
  const int E = orc_coord(i + 1, nx) + nx * j, W = orc_coord(i - 1, nx) + nx * j;
  const int EE = orc_coord(i + 2, nx) + nx * j, WW = orc_coord(i - 2, nx) + nx * j;
  const int N = i + nx * orc_coord(j + 1, ny), S = i + nx * orc_coord(j - 1, ny);
  const int NN = i + nx * orc_coord(j + 2, ny), SS = i + nx * orc_coord(j - 2, ny);
  double ax, ay;
  orc_adv(cf, i, j, ac, aphi, &ax, &ay);
  r.fx = (ax > 0.0) ? (-3.0 * f[C] + 4.0 * f[E] - f[EE]) * cf->invdx
                    : (3.0 * f[C] - 4.0 * f[W] + f[WW]) * cf->invdx;
  r.fy = (ay > 0.0) ? (-3.0 * f[C] + 4.0 * f[N] - f[NN]) * cf->invdy
                    : (3.0 * f[C] - 4.0 * f[S] + f[SS]) * cf->invdy;
  const double x = (double)i, y = (double)j;
  r.ft = cf->hx * x * r.fy - cf->hy * y * r.fx;
  r.fx0 = (f[EE] - 8.0 * f[E] + 8.0 * f[W] - f[WW]) * cf->invdx * (1.0 / 6.0);
  r.fy0 = (f[NN] - 8.0 * f[N] + 8.0 * f[S] - f[SS]) * cf->invdy * (1.0 / 6.0);
  r.ft0 = cf->hx * x * r.fy0 - cf->hy * y * r.fx0;
  return r;
}

void orc_slice_field_d(const yh_params* p, const double* f, int cx, int cy,
                       const double ac[3], double aphi,
                       double* fx, double* fy, double* ft,
                       double* fx0, double* fy0, double* ft0) {
  orc_coef cf;
  orc_derive(p, &cf);
  for (int j = 0; j < p->ny; ++j)
    for (int i = 0; i < p->nx; ++i) {
      const orc_sl s = orc_slice_cell(p, &cf, f, i, j, cx, cy, ac, aphi);
      const int c = i + p->nx * j;
      if (fx) fx[c] = s.fx;
      if (fy) fy[c] = s.fy;
      if (ft) ft[c] = s.ft;
      if (fx0) fx0[c] = s.fx0;
      if (fy0) fy0[c] = s.fy0;
      if (ft0) ft0[c] = s.ft0;
    }
}

/* trapz_wrapper's 12 quadruples (integralTrapz.cu:97-177): template derivative
 * index a (x,y,t of slice0) times slice index b (x,y,t of slice) for k<9,
 * and template a times velTan for k = 9..11.  Disk weight hx*hy (:55-57,79). */
void orc_sr_integrals_d(const yh_params* p, const double* const X[3],
                        const double* const vt[3], int cx, int cy,
                        const double ac[3], double aphi, double I[12]) {
  orc_coef cf;
  orc_derive(p, &cf);
  const int nv = p->model == YH_MODEL_3V_FK ? 3 : 2;
  for (int k = 0; k < 12; ++k) I[k] = 0.0;
  for (int j = 0; j < p->ny; ++j)
    for (int i = 0; i < p->nx; ++i) {
      if (!orc_in_disk(p, i, j, cx, cy)) continue;
      const int c = i + p->nx * j;
      double acc[12] = {0};
      for (int a = 0; a < nv; ++a) {
        const orc_sl s = orc_slice_cell(p, &cf, X[a], i, j, cx, cy, ac, aphi);
        const double t0[3] = {s.fx0, s.fy0, s.ft0};
        const double t1[3] = {s.fx, s.fy, s.ft};
        for (int r = 0; r < 3; ++r) {
          for (int q = 0; q < 3; ++q) acc[3 * r + q] += t0[r] * t1[q];
          acc[9 + r] += t0[r] * vt[a][c];
        }
      }
      for (int k = 0; k < 12; ++k) I[k] += acc[k];
    }
  for (int k = 0; k < 12; ++k) I[k] *= cf.hx * cf.hy;
}

/* solve_matrix, symmetryReduction.cu:386-418. */
void orc_solve_matrix(const double phi[3], const double Int[12], double cout[3]) {
  const double cp = cos(phi[2]), sp = sin(phi[2]);
  const double a1 = Int[0] * cp + Int[1] * sp, a2 = Int[1] * cp - Int[0] * sp, a3 = Int[2];
  const double b1 = Int[3] * cp + Int[4] * sp, b2 = Int[4] * cp - Int[3] * sp, b3 = Int[5];
  const double C1 = Int[6] * cp + Int[7] * sp, C2 = Int[7] * cp - Int[6] * sp, C3 = Int[8];
  const double d1 = Int[9], d2 = Int[10], d3 = Int[11];
  const double b2p = a1 / b1 * b2 - a2;
  const double b3p = a1 / b1 * b3 - a3;
  const double d2p = a1 / b1 * d2 - d1;
  const double c2p = a1 / C1 * C2 - a2;
  const double c3p = a1 / C1 * C3 - a3;
  const double d3p = a1 / C1 * d3 - d1;
  const double c3pp = b2p / c2p * c3p - b3p;
  const double d3pp = b2p / c2p * d3p - d2p;
  const double x3 = d3pp / c3pp;
  const double x2 = (d2p - b3p * x3) / b2p;
  const double x1 = (d1 - a2 * x2 - a3 * x3) / a1;
  cout[0] = x1;
  cout[1] = x2;
  cout[2] = x3;
}

/* advFDBFECC_kernel, Neumann branch (advFDBFECC.cu:30-165), split into its
 * three dependent phases so every neighbour read sees a completed field.  The
 * masked branch (:38-103) is restated with the mask of the cell actually
 * examined (the reference tests solid[j-1] where it means j+1 and vice versa,
 * :47-48,54-55); a neighbour outside the domain counts as a hole, which makes
 * the masked rule equal the square-boundary rule (:112-126) on an all-tissue
 * mask, and a substitute neighbour outside the domain reads as a hole (0). */
typedef struct orc_bf_idx {
  int W, E, S, N, iW, W2, iE, E2, iS, S2, iN, N2, sc;
} orc_bf_idx;

static orc_bf_idx orc_bf_index(const yh_params* p, const uint8_t* solid, int i, int j) {
  const int nx = p->nx, ny = p->ny;
  const int c = i + nx * j;
  int sc = 1, sw, se, ss, sn;
  if (p->solidSwitch) {
    sc = solid[c] != 0;
    sw = (i > 0) && solid[c - 1];
    se = (i < nx - 1) && solid[c + 1];
    ss = (j > 0) && solid[c - nx];
    sn = (j < ny - 1) && solid[c + nx];
  } else {
    sw = i > 0;
    se = i < nx - 1;
    ss = j > 0;
    sn = j < ny - 1;
  }
  /* the substitute on the other side, or -1 where that lies outside the
   * domain (a hole on one side at the domain edge): it reads as a hole, 0.
   * The reference takes the linear index i2d +- 1 (+- nx) there, the first
   * cell of the next row or past the array. */
  const int oE = i < nx - 1 ? c + 1 : -1, oW = i > 0 ? c - 1 : -1;
  const int oN = j < ny - 1 ? c + nx : -1, oS = j > 0 ? c - nx : -1;
  orc_bf_idx r;
  r.sc = sc;
  r.W = sc ? (sw ? c - 1 : oE) : c;
  r.E = sc ? (se ? c + 1 : oW) : c;
  r.S = sc ? (ss ? c - nx : oN) : c;
  r.N = sc ? (sn ? c + nx : oS) : c;
  r.iW = sc ? (sw ? c : oE) : c;
  r.W2 = sc ? (sw ? c - 1 : c) : c;
  r.iE = sc ? (se ? c : oW) : c;
  r.E2 = sc ? (se ? c + 1 : c) : c;
  r.iS = sc ? (ss ? c : oN) : c;
  r.S2 = sc ? (ss ? c - nx : c) : c;
  r.iN = sc ? (sn ? c : oS) : c;
  r.N2 = sc ? (sn ? c + nx : c) : c;
  return r;
}

/* A field value, 0 for an index outside the domain (-1). */
static inline double orc_bfv(const double* f, int k) { return k < 0 ? 0.0 : f[k]; }

void orc_bfecc_d(const yh_params* p, const uint8_t* solid, const double* const in[3],
                 double* const out[3], const double c[3], double phit) {
  orc_coef cf;
  orc_derive(p, &cf);
  const int nx = p->nx, ny = p->ny;
  const size_t n = (size_t)nx * ny;
  const int nv = p->model == YH_MODEL_3V_FK ? 3 : 2;
  const double tc = p->tc;
  double* uf = (double*)malloc(n * sizeof(double));
  double* ue = (double*)malloc(n * sizeof(double));
  double* Rx = (double*)malloc(n * sizeof(double));
  double* Ry = (double*)malloc(n * sizeof(double));
  unsigned char* px = (unsigned char*)malloc(n);
  unsigned char* py = (unsigned char*)malloc(n);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      double ax, ay;
      orc_adv(&cf, i, j, c, phit, &ax, &ay);
      const int k = i + nx * j;
      if (p->solidSwitch && !solid[k]) ax = ay = 0.0; /* symmetryReduction.cu:50-51 */
      const double cx = -ax, cy = -ay; /* advFDBFECC.cu:30-34 */
      const double sgx = cx < 0.0 ? -1.0 : (cx > 0.0 ? 1.0 : 0.0);
      const double sgy = cy < 0.0 ? -1.0 : (cy > 0.0 ? 1.0 : 0.0);
      Rx[k] = sgx * cx * cf.dt / cf.hx;
      Ry[k] = sgy * cy * cf.dt / cf.hy;
      px[k] = cx > 0.0;
      py[k] = cy > 0.0;
    }
  for (int a = 0; a < nv; ++a) {
    const double* g = in[a];
    for (int j = 0; j < ny; ++j) /* phase 1: forward step (:68-71, :128-131) */
      for (int i = 0; i < nx; ++i) {
        const int k = i + nx * j;
        const orc_bf_idx x = orc_bf_index(p, solid, i, j);
        const double FDx = px[k] ? g[k] - orc_bfv(g, x.W) : g[k] - orc_bfv(g, x.E);
        const double FDy = py[k] ? g[k] - orc_bfv(g, x.S) : g[k] - orc_bfv(g, x.N);
        uf[k] = g[k] - tc * (Rx[k] * FDx + Ry[k] * FDy);
      }
    for (int j = 0; j < ny; ++j) /* phase 2: backward step + correction (:73-78) */
      for (int i = 0; i < nx; ++i) {
        const int k = i + nx * j;
        const orc_bf_idx x = orc_bf_index(p, solid, i, j);
        const double FDx = px[k] ? orc_bfv(uf, x.iE) - uf[x.E2] : orc_bfv(uf, x.iW) - uf[x.W2];
        const double FDy = py[k] ? orc_bfv(uf, x.iN) - uf[x.N2] : orc_bfv(uf, x.iS) - uf[x.S2];
        const double ub = uf[k] - tc * (Rx[k] * FDx + Ry[k] * FDy);
        ue[k] = g[k] - 0.5 * (ub - g[k]);
      }
    for (int j = 0; j < ny; ++j) /* phase 3: forward step from ue (:80-83) */
      for (int i = 0; i < nx; ++i) {
        const int k = i + nx * j;
        const orc_bf_idx x = orc_bf_index(p, solid, i, j);
        const double FDx = px[k] ? ue[k] - orc_bfv(ue, x.W) : ue[k] - orc_bfv(ue, x.E);
        const double FDy = py[k] ? ue[k] - orc_bfv(ue, x.S) : ue[k] - orc_bfv(ue, x.N);
        const double o = ue[k] - tc * (Rx[k] * FDx + Ry[k] * FDy);
        out[a][k] = x.sc ? o : 0.0;
      }
  }
  free(uf); free(ue); free(Rx); free(Ry); free(px); free(py);
}

void orc_sr_init(orc_sr_state* s, float tip0x, float tip0y) {
  memset(s, 0, sizeof(*s));
  s->tip0x = tip0x;
  s->tip0y = tip0y;
}

/* One symmetry-reduction step, main.cu:894-953:
 *   RD (+velTan) -> tips(g_past = RD output, g_present = old state)
 *   -> slice with the disk at the last tip (tipx0 at count 0)
 *   -> [count 0: integrals, solve, Cxy, slice again] -> integrals -> solve
 *   -> Cxy -> BFECC of the RD output -> phi += dt*c (Euler: reduceSymStart is
 *   never cleared, main.cu:935-941).
 * The last tip is the last one in (cell, root) order; with no tip the previous
 * centre is kept (the reference would read tip_vector[-1]). */
int orc_sr_step_d(const yh_params* p, const uint8_t* solid, double* const X[3],
                  orc_sr_state* s) {
  orc_coef cf;
  orc_derive(p, &cf);
  const size_t n = (size_t)p->nx * p->ny;
  const int nv = p->model == YH_MODEL_3V_FK ? 3 : 2;
  double* Xo[3] = {0, 0, 0};
  double* vt[3] = {0, 0, 0};
  for (int a = 0; a < nv; ++a) {
    Xo[a] = (double*)malloc(n * sizeof(double));
    vt[a] = (double*)malloc(n * sizeof(double));
  }
  const double* cin[3] = {X[0], X[1], nv == 3 ? X[2] : 0};
  orc_rd_step_d(p, solid, cin, Xo, vt);

  const double time = cf.dt * (double)s->count;
  const int cap = 4096;
  yh_tip* tips = (yh_tip*)malloc(cap * sizeof(yh_tip));
  const int nt = orc_tips_d(p, Xo[0], X[0], time, tips, NULL, cap);
  s->ntips = nt;
  if (s->count == 0) {
    s->cx = (int)lrintf(s->tip0x - (float)(p->nx / 2));
    s->cy = (int)lrintf(s->tip0y - (float)(p->ny / 2));
  } else if (nt > 0) {
    const yh_tip* last = &tips[(nt < cap ? nt : cap) - 1];
    s->cx = (int)lrintf(last->x - (float)(p->nx / 2));
    s->cy = (int)lrintf(last->y - (float)(p->ny / 2));
  }
  free(tips);

  const double* Xc[3] = {X[0], X[1], nv == 3 ? X[2] : 0};
  const double* vtc[3] = {vt[0], vt[1], vt[2]};
  int rc = YH_OK;
  if (s->count == 0) {
    orc_sr_integrals_d(p, Xc, vtc, s->cx, s->cy, s->ac, s->aphi, s->I);
    orc_solve_matrix(s->phi, s->I, s->c);
    s->ac[0] = s->c[0]; s->ac[1] = s->c[1]; s->ac[2] = s->c[2];
    s->aphi = s->phi[2];
  }
  orc_sr_integrals_d(p, Xc, vtc, s->cx, s->cy, s->ac, s->aphi, s->I);
  orc_solve_matrix(s->phi, s->I, s->c);
  if (!isfinite(s->c[0]) || !isfinite(s->c[1]) || !isfinite(s->c[2])) rc = YH_ESINGULAR;
  s->ac[0] = s->c[0]; s->ac[1] = s->c[1]; s->ac[2] = s->c[2];
  s->aphi = s->phi[2];

  const double* Xoc[3] = {Xo[0], Xo[1], Xo[2]};
  orc_bfecc_d(p, solid, Xoc, X, s->c, s->phi[2]);
  for (int q = 0; q < 3; ++q) s->phi[q] = s->phi[q] + s->c[q] * cf.dt;
  s->count++;
  for (int a = 0; a < nv; ++a) {
    free(Xo[a]);
    free(vt[a]);
  }
  return rc;
}

/* ------------------------------------------------------------------ APD */
/* spaceAPD.cu:328-365 (stimulate == false). */
void orc_apd_update_d(int ncell, double dt, long count, const double* uold,
                      const double* unew, double* apd1, double* apd2, double* front,
                      double* back, uint8_t* first) {
  const double apdTh = 0.15; /* spaceAPD.cu:288 */
  for (int k = 0; k < ncell; ++k) {
    const double uo = uold[k], un = unew[k];
    if ((un > apdTh) && (uo < apdTh)) front[k] = dt * (count - (un - apdTh) / (un - uo));
    if ((un < apdTh) && (uo > apdTh)) back[k] = dt * (count - (un - apdTh) / (un - uo));
    if ((back[k] > 0.0) && (front[k] > 0.0) && (first[k] == 0)) {
      apd1[k] = back[k] - front[k];
      front[k] = 0.0;
      back[k] = 0.0;
      first[k] = 1;
    }
    if ((back[k] > 0.0) && (front[k] > 0.0) && first[k]) {
      apd2[k] = back[k] - front[k];
      front[k] = 0.0;
      back[k] = 0.0;
      first[k] = 0;
    }
  }
}

/* ---------------------------------------------------------------- adjoint */

typedef struct orc_fkd {
  double p, dp, q, dq, f, I, Gv, Gw;
  double dIdu, dIdv, dIdw, dGvdu, dGvdv, dGwdu, dGwdw;
  double dI_dtau[4];  /* tau_d, tau_r, tau_si, tau_0 */
  double dGv_dtvp, dGw_dtwp;
} orc_fkd;

static void orc_heav_d(double x, double w, double* h, double* dh) {
  if (w > 0.0) {
    const double t = tanh(x / w);
    *h = 0.5 * (1.0 + t);
    *dh = 0.5 / w * (1.0 - t * t);
  } else {
    *h = x >= 0.0 ? 1.0 : 0.0;
    *dh = 0.0;
  }
}

/* FK currents and their derivatives (Fenton & Karma 1998 eqs. 2-8). */
static orc_fkd orc_fk_eval(const yh_fk_params* fk, double hw, double u, double v, double w) {
  orc_fkd r;
  orc_heav_d(u - fk->u_c, hw, &r.p, &r.dp);
  orc_heav_d(u - fk->u_v, hw, &r.q, &r.dq);
  const double f = (1.0 - u) * (u - fk->u_c);
  const double df = 1.0 + fk->u_c - 2.0 * u;
  const double Jfi = -v * r.p * f / fk->tau_d;
  const double Jso = u * (1.0 - r.p) / fk->tau_0 + r.p / fk->tau_r;
  const double th = tanh(fk->k * (u - fk->u_csi));
  const double Jsi = -w * (1.0 + th) / (2.0 * fk->tau_si);
  r.I = Jfi + Jso + Jsi;
  r.dIdu = -v * (r.dp * f + r.p * df) / fk->tau_d
           + (1.0 - r.p) / fk->tau_0 - u * r.dp / fk->tau_0 + r.dp / fk->tau_r
           - w * fk->k * (1.0 - th * th) / (2.0 * fk->tau_si);
  r.dIdv = -r.p * f / fk->tau_d;
  r.dIdw = -(1.0 + th) / (2.0 * fk->tau_si);
  const double tvm = (1.0 - r.q) * fk->tau_v1m + r.q * fk->tau_v2m;
  const double dtvm = r.dq * (fk->tau_v2m - fk->tau_v1m);
  const double A = (1.0 - r.p) * (1.0 - v) / tvm - r.p * v / fk->tau_vp;
  const double dAdu = -r.dp * (1.0 - v) / tvm - (1.0 - r.p) * (1.0 - v) * dtvm / (tvm * tvm)
                      - r.dp * v / fk->tau_vp;
  const double dAdv = -(1.0 - r.p) / tvm - r.p / fk->tau_vp;
  const double B = (1.0 - r.p) * (1.0 - w) / fk->tau_wm - r.p * w / fk->tau_wp;
  const double dBdu = -r.dp * (1.0 - w) / fk->tau_wm - r.dp * w / fk->tau_wp;
  const double dBdw = -(1.0 - r.p) / fk->tau_wm - r.p / fk->tau_wp;
  r.Gv = -A;
  r.Gw = -B;
  r.dGvdu = -dAdu;
  r.dGvdv = -dAdv;
  r.dGwdu = -dBdu;
  r.dGwdw = -dBdw;
  r.dI_dtau[0] = -Jfi / fk->tau_d;
  r.dI_dtau[1] = -r.p / (fk->tau_r * fk->tau_r);
  r.dI_dtau[2] = -Jsi / fk->tau_si;
  r.dI_dtau[3] = -u * (1.0 - r.p) / (fk->tau_0 * fk->tau_0);
  r.dGv_dtvp = -r.p * v / (fk->tau_vp * fk->tau_vp);
  r.dGw_dtwp = -r.p * w / (fk->tau_wp * fk->tau_wp);
  r.f = f;
  return r;
}

/* Masked coefficients of reactionDiffusion.cu:162-169 in the form
 * c_W (u_W-u) + c_E (u_E-u) (the centre coefficient is always c_W + c_E). */
static void orc_mask_coef(const yh_params* p, const uint8_t* solid, int i, int j,
                          int* cW, int* cE, int* cN, int* cS, int* sc) {
  const int nx = p->nx, ny = p->ny;
  if (!p->solidSwitch) {
    *cW = *cE = *cN = *cS = 1;
    *sc = 1;
    return;
  }
  const int c = i + nx * j;
  const int s = solid[c] != 0;
  const int sw = solid[orc_coord(i - 1, nx) + nx * j] != 0;
  const int se = solid[orc_coord(i + 1, nx) + nx * j] != 0;
  const int sn = solid[i + nx * orc_coord(j + 1, ny)] != 0;
  const int ss = solid[i + nx * orc_coord(j - 1, ny)] != 0;
  *cW = (sw && se && s) ? 1 : ((sw && s) ? 2 : 0);
  *cE = (sw && se && s) ? 1 : ((s && se) ? 2 : 0);
  *cN = (sn && ss && s) ? 1 : ((sn && s) ? 2 : 0);
  *cS = (sn && ss && s) ? 1 : ((s && ss) ? 2 : 0);
  *sc = s;
}

void orc_fk_dfield_step_d(const yh_params* p, const uint8_t* solid, const double* D,
                          const double* const in[3], double* const out[3]) {
  orc_coef cf;
  orc_derive(p, &cf);
  const int nx = p->nx, ny = p->ny;
  const double rxh = cf.dt / (cf.hx * cf.hx), ryh = cf.dt / (cf.hy * cf.hy);
#pragma omp parallel for schedule(static) if ((long)nx * ny >= 16384)
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const int c = i + nx * j;
      int cW, cE, cN, cS, sc;
      orc_mask_coef(p, solid, i, j, &cW, &cE, &cN, &cS, &sc);
      const int W = orc_coord(i - 1, nx) + nx * j, E = orc_coord(i + 1, nx) + nx * j;
      const int N = i + nx * orc_coord(j + 1, ny), S = i + nx * orc_coord(j - 1, ny);
      const double u = in[0][c], v = in[1][c], w = in[2][c];
      const double lap = rxh * (cW * 0.5 * (D[c] + D[W]) * (in[0][W] - u)
                                + cE * 0.5 * (D[c] + D[E]) * (in[0][E] - u))
                       + ryh * (cN * 0.5 * (D[c] + D[N]) * (in[0][N] - u)
                                + cS * 0.5 * (D[c] + D[S]) * (in[0][S] - u));
      const orc_fkd r = orc_fk_eval(&p->fk, p->heaviside_width, u, v, w);
      out[0][c] = sc ? u + p->tc * (lap - cf.dt * r.I) : 0.0;
      out[1][c] = sc ? v + p->tc * (-cf.dt * r.Gv) : 0.0;
      out[2][c] = sc ? w + p->tc * (-cf.dt * r.Gw) : 0.0;
    }
}

double orc_fk_loss_d(const yh_params* p, const uint8_t* solid, const double* D,
                     const double* const X0[3], int nsteps, int rec, int ns,
                     const int* sensors, const double* targets, double* traces) {
  const size_t n = (size_t)p->nx * p->ny;
  double* A[3];
  double* B[3];
  for (int a = 0; a < 3; ++a) {
    A[a] = (double*)malloc(n * sizeof(double));
    B[a] = (double*)malloc(n * sizeof(double));
    memcpy(A[a], X0[a], n * sizeof(double));
  }
  double L = 0.0;
  for (int s = 1; s <= nsteps; ++s) {
    const double* cin[3] = {A[0], A[1], A[2]};
    orc_fk_dfield_step_d(p, solid, D, cin, B);
    for (int a = 0; a < 3; ++a) {
      double* t = A[a];
      A[a] = B[a];
      B[a] = t;
    }
    if (s % rec == 0) {
      const int r = s / rec - 1;
      for (int q = 0; q < ns; ++q) {
        const double val = A[0][sensors[q]];
        if (traces) traces[(size_t)r * ns + q] = val;
        const double d = val - (targets ? targets[(size_t)r * ns + q] : 0.0);
        L += d * d;
      }
    }
  }
  for (int a = 0; a < 3; ++a) {
    free(A[a]);
    free(B[a]);
  }
  return L;
}

double orc_fk_adjoint_d(const yh_params* p, const uint8_t* solid, const double* D,
                        const double* const X0[3], int nsteps, int rec, int ns,
                        const int* sensors, const double* targets,
                        double* gD, double gtau[6]) {
  orc_coef cf;
  orc_derive(p, &cf);
  const int nx = p->nx, ny = p->ny;
  const size_t n = (size_t)nx * ny;
  const double rxh = cf.dt / (cf.hx * cf.hx), ryh = cf.dt / (cf.hy * cf.hy);
  const double tc = p->tc, dt = cf.dt;
  /* Forward, keeping every state. */
  double* traj = (double*)malloc((size_t)(nsteps + 1) * 3 * n * sizeof(double));
  for (int a = 0; a < 3; ++a) memcpy(traj + a * n, X0[a], n * sizeof(double));
  double L = 0.0;
  for (int s = 1; s <= nsteps; ++s) {
    double* prev = traj + (size_t)(s - 1) * 3 * n;
    double* cur = traj + (size_t)s * 3 * n;
    const double* cin[3] = {prev, prev + n, prev + 2 * n};
    double* co[3] = {cur, cur + n, cur + 2 * n};
    orc_fk_dfield_step_d(p, solid, D, cin, co);
    if (s % rec == 0) {
      const int r = s / rec - 1;
      for (int q = 0; q < ns; ++q) {
        const double d = cur[sensors[q]] - targets[(size_t)r * ns + q];
        L += d * d;
      }
    }
  }
  /* Reverse sweep. */
  double* lam = (double*)calloc(3 * n, sizeof(double)); /* adjoint of X^{s} */
  double* nl = (double*)calloc(3 * n, sizeof(double));
  for (size_t c = 0; c < n; ++c) gD[c] = 0.0;
  for (int k = 0; k < 6; ++k) gtau[k] = 0.0;
  for (int s = nsteps; s >= 1; --s) {
    double* cur = traj + (size_t)s * 3 * n;
    if (s % rec == 0) { /* source dL/du^s */
      const int r = s / rec - 1;
      for (int q = 0; q < ns; ++q)
        lam[sensors[q]] += 2.0 * (cur[sensors[q]] - targets[(size_t)r * ns + q]);
    }
    /* lam holds the adjoint of X^s; propagate through X^s = F(X^{s-1}). */
    const double* X = traj + (size_t)(s - 1) * 3 * n;
    memset(nl, 0, 3 * n * sizeof(double));
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const int c = i + nx * j;
        int cW, cE, cN, cS, sc;
        orc_mask_coef(p, solid, i, j, &cW, &cE, &cN, &cS, &sc);
        if (!sc) continue;
        const double mu = lam[c], mv = lam[n + c], mw = lam[2 * n + c];
        const double u = X[c], v = X[n + c], w = X[2 * n + c];
        const int nb[4] = {orc_coord(i - 1, nx) + nx * j, orc_coord(i + 1, nx) + nx * j,
                           i + nx * orc_coord(j + 1, ny), i + nx * orc_coord(j - 1, ny)};
        const int co[4] = {cW, cE, cN, cS};
        const double rr[4] = {rxh, rxh, ryh, ryh};
        for (int d = 0; d < 4; ++d) {
          const double a = rr[d] * co[d] * 0.5 * (D[c] + D[nb[d]]);
          nl[nb[d]] += tc * mu * a;
          nl[c] -= tc * mu * a;
          const double g = tc * mu * rr[d] * co[d] * 0.5 * (X[nb[d]] - u);
          gD[c] += g;
          gD[nb[d]] += g;
        }
        const orc_fkd r = orc_fk_eval(&p->fk, p->heaviside_width, u, v, w);
        nl[c] += mu * (1.0 - tc * dt * r.dIdu) + mv * (-tc * dt * r.dGvdu)
                 + mw * (-tc * dt * r.dGwdu);
        nl[n + c] += mu * (-tc * dt * r.dIdv) + mv * (1.0 - tc * dt * r.dGvdv);
        nl[2 * n + c] += mu * (-tc * dt * r.dIdw) + mw * (1.0 - tc * dt * r.dGwdw);
        for (int k = 0; k < 4; ++k) gtau[k] += mu * (-tc * dt * r.dI_dtau[k]);
        gtau[4] += mv * (-tc * dt * r.dGv_dtvp);
        gtau[5] += mw * (-tc * dt * r.dGw_dtwp);
      }
    double* t = lam;
    lam = nl;
    nl = t;
  }
  free(lam);
  free(nl);
  free(traj);
  return L;
}

/* nsteps of the plain loop with the sAPD update after every step
 * (spaceAPD.cu:328-358, stimulus off), crossings taken in time order
 * (uold = state before the step, unew = after), threshold th. */
void orc_run_apd_d(const yh_params* p, const uint8_t* solid, double* const st[3], int nsteps,
                   long count0, double th, double* apd1, double* apd2, double* front,
                   double* back, uint8_t* first) {
  orc_coef cf;
  orc_derive(p, &cf);
  const size_t n = (size_t)p->nx * p->ny;
  const int nv = p->model == YH_MODEL_3V_FK ? 3 : 2;
  double* nw[3] = {0, 0, 0};
  for (int a = 0; a < nv; ++a) nw[a] = (double*)malloc(n * sizeof(double));
  const int nthreads = omp_get_max_threads();
  for (int s = 0; s < nsteps; ++s) {
    const double* cin[3] = {st[0], st[1], nv == 3 ? st[2] : 0};
    /* the plain step (no imposed decisions), OpenMP over rows */
    orc_rd_step_forced_d(p, solid, cin, nw, 0, 0, nthreads);
    const long count = count0 + s + 1;
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (size_t k = 0; k < n; ++k) {
      if (p->solidSwitch && !solid[k]) continue;
      const double uo = st[0][k], un = nw[0][k];
      if ((un > th) && (uo < th)) front[k] = cf.dt * (count - (un - th) / (un - uo));
      if ((un < th) && (uo > th)) back[k] = cf.dt * (count - (un - th) / (un - uo));
      if ((back[k] > 0.0) && (front[k] > 0.0) && (first[k] == 0)) {
        apd1[k] = back[k] - front[k];
        front[k] = 0.0;
        back[k] = 0.0;
        first[k] = 1;
      }
      if ((back[k] > 0.0) && (front[k] > 0.0) && first[k]) {
        apd2[k] = back[k] - front[k];
        front[k] = 0.0;
        back[k] = 0.0;
        first[k] = 0;
      }
    }
    for (int a = 0; a < nv; ++a) memcpy(st[a], nw[a], n * sizeof(double));
  }
  for (int a = 0; a < nv; ++a) free(nw[a]);
}
