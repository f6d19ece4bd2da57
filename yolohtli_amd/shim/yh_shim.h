/*
 * yh_shim.h -- the reference-side binding: what a maintainer adds to the
 * reference's host code (main.cu's display(), main.cu:862-953) to call the
 * drop-in instead of its own kernels.  Plain C over include/yolohtli.h.
 *
 * Inside the reference, define YH_SHIM_REFERENCE_TYPES and include its
 * typeDefinition.cuh (built with `typedef float REAL`) before this header.
 * Standalone, the block below declares the reference's type names with only
 * the paramVar fields the shim reads (typeDefinition.cuh:3-20,35-125).
 */
#ifndef YH_SHIM_H
#define YH_SHIM_H

#include <stdbool.h>

#include "../../include/yolohtli.h"

#ifndef YH_SHIM_REFERENCE_TYPES
typedef float REAL;
typedef struct stateVar { REAL *u, *v; } stateVar;
typedef struct vec5dyn { float x, y, vx, vy, t; } vec5dyn;
typedef struct paramVar {
  bool reduceSym, solidSwitch, neumannBC, gateDiff, anisotropy;
  int lap4, timeIntOrder, nx, ny, tipOffsetX, tipOffsetY;
  REAL Lx, Ly, dt, diff_par, diff_per, rscale, degrad, boundaryVal;
  REAL Uth, tc, alpha, beta, gamma, delta, eps, mu, theta;
} paramVar;
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* At startup, after parameterSetup(), domainObjects() and initGates():
 * creates the plan (replacing the allocation and cudaMemcpyToSymbol of
 * main.cu:213-402), the mask and the state.  The plan keeps
 * YH_KERNEL_AUTO; a caller may add yh_autotune(plan, sampleIt, ...) here
 * (tuned for calls of sampleIt steps, the plain branch's call size). */
int yh_shim_init(const paramVar* param, const bool* solid_h, stateVar g_h);
/* The plain branch of display() (main.cu:876-888): sampleIt steps, then the
 * tip test (main.cu:961-964) when plotTip. */
int yh_shim_plain(int sampleIt, bool plotTip);
/* The reduceSym branch (main.cu:894-953); c_phi (6*nsteps doubles, may be
 * NULL) receives clist/philist.  tipx0/tipy0 seed the disk at count 0. */
int yh_shim_sr(int nsteps, float tipx0, float tipy0, double* c_phi);
/* State and tip log for the writers (raw_data.dat, dataTip.dat). */
int yh_shim_read(stateVar g_h, vec5dyn* tips, int cap, int* ntips);
/* At exit. */
void yh_shim_close(void);

/* The two paths with no reference counterpart, same conventions. */
int yh_example_batch(const yh_params* per_tissue, int nt, uint8_t** masks, float** u,
                     float** v, float** w, int nsteps);
int yh_example_gradient(const yh_params* p, int nsteps, const int* sensors, int ns,
                        const float* u0, const float* v0, const float* w0,
                        const float* targets, const float* d_D, const double tau6[6],
                        double* d_gD, double gtau[6], double* loss);

#ifdef __cplusplus
}
#endif
#endif
