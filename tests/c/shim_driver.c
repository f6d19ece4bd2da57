/*
 * shim_driver.c -- a compiled C caller of include/yolohtli.h through the
 * reference-side shim (yolohtli_amd/shim/yh_shim.c).
 *   shim_driver host DIR   host-only calls: defaults, derive, solve_matrix,
 *                          the parameter CSV round trip (text on stdout)
 *   shim_driver plain DIR  128^2 2V from the reference's cross-field start:
 *                          4 x (50 steps + tips) through yh_shim_plain
 *   shim_driver sr DIR     the same start, 1000 plain steps, then 200 SR steps
 *                          through yh_shim_sr from the last tip
 * The GPU modes write the state (float32 u then v), the tip log and the
 * c/phi history as raw binary files into DIR.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../yolohtli_amd/shim/yh_shim.h"

#define N 128

static int fail(const char* what, int rc) {
  fprintf(stderr, "%s: %s (%d)\n", what, yh_strerror(rc), rc);
  return 1;
}

static int dump(const char* dir, const char* name, const void* data, size_t bytes) {
  char path[1024];
  snprintf(path, sizeof path, "%s/%s", dir, name);
  FILE* f = fopen(path, "wb");
  if (!f) return 1;
  const size_t w = fwrite(data, 1, bytes, f);
  fclose(f);
  return w != bytes;
}

/* parameterSetup defaults (through yh_params_default) on a 128^2 grid of
 * spacing 0.025 cm, as the reference's paramVar */
static paramVar reference_params(int reduce_sym) {
  yh_params d;
  yh_params_default(&d);
  paramVar q;
  memset(&q, 0, sizeof q);
  q.nx = N;
  q.ny = N;
  q.Lx = (REAL)(0.025 * (N - 1));
  q.Ly = (REAL)(0.025 * (N - 1));
  q.dt = (REAL)d.dt;
  q.diff_par = (REAL)d.diff_par;
  q.diff_per = (REAL)d.diff_per;
  q.rscale = (REAL)d.rscale;
  q.degrad = (REAL)d.degrad;
  q.boundaryVal = (REAL)d.boundaryVal;
  q.Uth = (REAL)d.Uth;
  q.tc = (REAL)d.tc;
  q.alpha = (REAL)d.alpha;
  q.beta = (REAL)d.beta;
  q.gamma = (REAL)d.gamma;
  q.delta = (REAL)d.delta;
  q.eps = (REAL)d.eps;
  q.mu = (REAL)d.mu;
  q.theta = (REAL)d.theta;
  q.neumannBC = d.neumannBC != 0;
  q.solidSwitch = false;
  q.gateDiff = d.gateDiff != 0;
  q.reduceSym = reduce_sym != 0;
  q.anisotropy = false;
  q.lap4 = 0;
  q.timeIntOrder = 1;
  q.tipOffsetX = d.tipOffsetX;
  q.tipOffsetY = d.tipOffsetY;
  return q;
}

static int host_mode(const char* dir) {
  yh_params p;
  yh_params_default(&p);
  printf("default %d %d %.17g %.17g %.17g %.17g %.17g %d %d\n", p.nx, p.ny, p.Lx, p.dt, p.Uth,
         p.fk.tau_v1m, p.fk.tau_v2m, p.tipOffsetX, p.timeIntOrder);
  p.nx = p.ny = 512;
  p.Lx = p.Ly = 0.025 * 511;
  yh_derived d;
  int rc = yh_params_derive(&p, &d);
  if (rc) return fail("yh_params_derive", rc);
  printf("derive %.17g %.17g %.17g\n", d.hx, d.rx, d.dt);
  const double phi[3] = {0.1, -0.2, 0.3};
  const double I[12] = {3, 0.1, 0.2, 0.1, 3, 0.3, 0.2, 0.3, 3, 0.5, -0.4, 0.7};
  double c[3];
  rc = yh_solve_matrix(phi, I, c);
  if (rc) return fail("yh_solve_matrix", rc);
  printf("solve %.17g %.17g %.17g\n", c[0], c[1], c[2]);
  char path[1024];
  snprintf(path, sizeof path, "%s/dataparamcsv.csv", dir);
  p.dt = 0.015;
  p.tipOffsetX = 77;
  rc = yh_params_write_csv(path, &p, NULL, NULL, 1.5f, 2.5f);
  if (rc) return fail("yh_params_write_csv", rc);
  yh_params back;
  yh_params_default(&back);
  rc = yh_params_read_csv(path, &back);
  if (rc) return fail("yh_params_read_csv", rc);
  printf("csv %d %d %.9g %d\n", back.nx, back.ny, back.dt, back.tipOffsetX);
  return 0;
}

static int gpu_mode(const char* dir, int sr) {
  static float u[N * N], v[N * N];
  for (int j = 0; j < N; ++j)
    for (int i = 0; i < N; ++i) { /* initGates, main.cu:606-618 */
      u[i + N * j] = i < N / 8 ? 1.0f : 0.0f;
      v[i + N * j] = j >= N / 2 ? 1.0f : 0.0f;
    }
  const paramVar q = reference_params(0);
  stateVar g = {u, v};
  int rc = yh_shim_init(&q, NULL, g);
  if (rc) return fail("yh_shim_init", rc);
  static vec5dyn tips[4096];
  int nt = 0;
  if (!sr) {
    for (int k = 0; k < 4 && !rc; ++k) rc = yh_shim_plain(50, true);
    if (rc) return fail("yh_shim_plain", rc);
  } else {
    /* a formed wave first: 1000 plain steps, then the SR disk starts at the
     * last tip found (the centre if none) */
    rc = yh_shim_plain(1000, true);
    if (!rc) rc = yh_shim_read(g, tips, 4096, &nt);
    if (rc && rc != YH_EOVERFLOW) return fail("warm-up", rc);
    const int last = (nt < 4096 ? nt : 4096) - 1;
    const float tx = last >= 0 ? tips[last].x : N / 2, ty = last >= 0 ? tips[last].y : N / 2;
    const paramVar qs = reference_params(1);
    rc = yh_shim_init(&qs, NULL, g);
    if (rc) return fail("yh_shim_init (SR)", rc);
    static double cphi[6 * 200];
    rc = yh_shim_sr(200, tx, ty, cphi);
    if (rc) return fail("yh_shim_sr", rc);
    if (dump(dir, "sr_cphi.bin", cphi, sizeof cphi)) return 1;
  }
  rc = yh_shim_read(g, tips, 4096, &nt);
  if (rc && rc != YH_EOVERFLOW) return fail("yh_shim_read", rc);
  yh_shim_close();
  const char* pre = sr ? "sr" : "plain";
  char name[64];
  snprintf(name, sizeof name, "%s_state.bin", pre);
  float* both = (float*)malloc(sizeof(float) * 2 * N * N);
  if (!both) return 1;
  memcpy(both, u, sizeof u);
  memcpy(both + N * N, v, sizeof v);
  rc = dump(dir, name, both, sizeof(float) * 2 * N * N);
  free(both);
  snprintf(name, sizeof name, "%s_tips.bin", pre);
  rc |= dump(dir, name, tips, sizeof(vec5dyn) * (size_t)(nt < 4096 ? nt : 4096));
  printf("%s ok tips %d\n", pre, nt);
  return rc;
}

int main(int argc, char** argv) {
  if (argc != 3) {
    fprintf(stderr, "usage: shim_driver host|plain|sr DIR\n");
    return 2;
  }
  if (!strcmp(argv[1], "host")) return host_mode(argv[2]);
  if (!strcmp(argv[1], "plain")) return gpu_mode(argv[2], 0);
  if (!strcmp(argv[1], "sr")) return gpu_mode(argv[2], 1);
  return 2;
}
