// direct.h -- one explicit-Euler step of one cell (device code), shared by
// the one-thread-per-cell kernels (stepper.hip) and the fused RD + tips
// kernel of the symmetry-reduction step (sr.hip).
#pragma once

#include "yh_internal.h"

namespace yh {

// One step, one thread per cell.  Diffusion in the coefficient form
// sum_dir c_dir*(f_dir - f) with c in {0,1,2} (reactionDiffusion.cu:162-173);
// c_W = c_E = c_N = c_S = 1 without a mask (:188-197).
template <int MODEL>
__device__ __forceinline__ void cell_update(const KParams& k, const float* const* f, int c, int W,
                                            int E, int N, int S, float cW, float cE, float cN,
                                            float cS, bool sc, float* xo, float* inc) {
  using M = Model<MODEL>;
  float x[3] = {0.f, 0.f, 0.f};
#pragma unroll
  for (int a = 0; a < M::NV; ++a) x[a] = f[a][c];
  float lap[2] = {0.f, 0.f};
#pragma unroll
  for (int a = 0; a < M::ND; ++a) {
    const float fc = x[a];
    const float lx = cW * (f[a][W] - fc) + cE * (f[a][E] - fc);
    const float ly = cN * (f[a][N] - fc) + cS * (f[a][S] - fc);
    lap[a] = k.trx * lx + k.try_ * ly;
  }
  float r[3] = {0.f, 0.f, 0.f};
  M::react(k, x[0], x[1], x[2], r[0], r[1], r[2]);
  inc[0] = lap[0] + r[0];
  inc[1] = (MODEL == M2V ? k.rs * lap[M::ND - 1] : 0.0f) + r[1];
  inc[2] = r[2];
#pragma unroll
  for (int a = 0; a < M::NV; ++a) xo[a] = sc ? x[a] + inc[a] : 0.0f;
}

// The same from the stencil codes of cell (i, j): xo = new state (0 in holes).
template <int MODEL>
__device__ __forceinline__ bool cell_update_codes(const KParams& k, const float* const* f,
                                                  const uint8_t* codes, int i, int j, float* xo,
                                                  float* inc) {
  const int nx = k.nx, ny = k.ny;
  const int c = i + nx * j;
  const int W = mirror(i - 1, nx) + nx * j, E = mirror(i + 1, nx) + nx * j;
  const int N = i + nx * mirror(j + 1, ny), S = i + nx * mirror(j - 1, ny);
  float cW = 1.f, cE = 1.f, cN = 1.f, cS = 1.f;
  bool sc = true;
  if (codes) {
    const uint32_t b = codes[c];
    sc = b != kHole;
    const uint32_t q = sc ? b : 0u;
    cW = (float)(q & 3u);
    cE = (float)((q >> 2) & 3u);
    cN = (float)((q >> 4) & 3u);
    cS = (float)((q >> 6) & 3u);
  }
  cell_update<MODEL>(k, f, c, W, E, N, S, cW, cE, cN, cS, sc, xo, inc);
  return sc;
}

}  // namespace yh
