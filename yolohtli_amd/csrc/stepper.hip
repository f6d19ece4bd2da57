// stepper.hip -- plain tissue time-stepper kernels for gfx950 (CDNA4).
//
// The update is the explicit-Euler, 5-point, no-flux step of
// reactionDiffusion_kernel (reactionDiffusion.cu:107-199,498-561) with the
// ionic model as a compile-time functor (Model<M2V>: the reference's 2V
// reaction; Model<MFK>: Fenton-Karma 3V).
//
// Kernels
//   k_codes        mask -> per-cell 2-bit stencil coefficients (once per mask)
//   k_direct       one step, one thread per cell, reads the raw mask; writes
//                  velTan when asked (the reactionDiffusion_wrapper surface)
//   k_direct_codes one step from stencil codes (the k=1 plan path)
//   k_tiled        K fused steps per launch (temporal blocking): a workgroup
//                  owns a 64 x (NW*R) region, each lane one column, each wave
//                  R consecutive rows kept in registers.  West/east neighbours
//                  come from ds_bpermute (no LDS memory), north/south from the
//                  lane's own registers; only the first/last row of each wave
//                  crosses waves through a small double-buffered LDS slot.
//                  Halo cells outside the domain are loaded by reflection, so
//                  the no-flux boundary costs nothing inside the substep loop.
#include <algorithm>

#include "direct.h"
#include "tile.h"
#include "yh_internal.h"

namespace yh {

// ------------------------------------------------------------------ codes
// reactionDiffusion.cu:156-169 with the centre coefficient folded in
// (coeffx.y == coeffx.x + coeffx.z for every mask pattern).
__global__ void k_codes(const uint8_t* __restrict__ solid, uint8_t* __restrict__ codes,
                        int nx, int ny) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= nx || j >= ny) return;
  const int c = i + nx * j;
  const bool sc = solid[c] != 0;
  if (!sc) {
    codes[c] = kHole;
    return;
  }
  const bool sw = solid[mirror(i - 1, nx) + nx * j] != 0;
  const bool se = solid[mirror(i + 1, nx) + nx * j] != 0;
  const bool sn = solid[i + nx * mirror(j + 1, ny)] != 0;
  const bool ss = solid[i + nx * mirror(j - 1, ny)] != 0;
  const int cW = (sw && se) ? 1 : (sw ? 2 : 0);
  const int cE = (sw && se) ? 1 : (se ? 2 : 0);
  const int cN = (sn && ss) ? 1 : (sn ? 2 : 0);
  const int cS = (sn && ss) ? 1 : (ss ? 2 : 0);
  codes[c] = (uint8_t)(cW | (cE << 2) | (cN << 4) | (cS << 6));
}

hipError_t launch_codes(const uint8_t* solid, uint8_t* codes, int nx, int ny, hipStream_t s) {
  dim3 b(64, 4), g((nx + 63) / 64, (ny + 3) / 4);
  hipLaunchKernelGGL(k_codes, g, b, 0, s, solid, codes, nx, ny);
  return hipGetLastError();
}

// ------------------------------------------------------------------ direct
template <int MODEL, bool MASKED>
__global__ __launch_bounds__(256) void k_direct(const KParams k, CFields in, Fields out,
                                                Fields vt, const uint8_t* __restrict__ solid) {
  using M = Model<MODEL>;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= k.nx || j >= k.ny) return;
  const int nx = k.nx, ny = k.ny;
  const int c = i + nx * j;
  const int W = mirror(i - 1, nx) + nx * j, E = mirror(i + 1, nx) + nx * j;
  const int N = i + nx * mirror(j + 1, ny), S = i + nx * mirror(j - 1, ny);
  float cW = 1.f, cE = 1.f, cN = 1.f, cS = 1.f;
  bool sc = true;
  if (MASKED) {
    sc = solid[c] != 0;
    const bool sw = solid[W] != 0, se = solid[E] != 0, sn = solid[N] != 0, ss = solid[S] != 0;
    cW = (sw && se && sc) ? 1.f : ((sw && sc) ? 2.f : 0.f);
    cE = (sw && se && sc) ? 1.f : ((se && sc) ? 2.f : 0.f);
    cN = (sn && ss && sc) ? 1.f : ((sn && sc) ? 2.f : 0.f);
    cS = (sn && ss && sc) ? 1.f : ((ss && sc) ? 2.f : 0.f);
  }
  float xo[3], inc[3];
  cell_update<MODEL>(k, in.f, c, W, E, N, S, cW, cE, cN, cS, sc, xo, inc);
#pragma unroll
  for (int a = 0; a < M::NV; ++a) {
    out.f[a][c] = xo[a];
    if (vt.f[0]) vt.f[a][c] = sc ? inc[a] / k.tcdt : 0.0f;  // rhs/dt (:529-530,551-552)
  }
}

template <int MODEL, bool APD>
__global__ __launch_bounds__(256) void k_direct_codes(const KParams k, CFields in, Fields out,
                                                      const uint8_t* __restrict__ codes,
                                                      const ApdState apd) {
  using M = Model<MODEL>;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= k.nx || j >= k.ny) return;
  const int c = i + k.nx * j;
  float xo[3], inc[3];
  const bool sc = cell_update_codes<MODEL>(k, in.f, codes, i, j, xo, inc);
#pragma unroll
  for (int a = 0; a < M::NV; ++a) out.f[a][c] = xo[a];
  if (APD && sc) {
    float fr = apd.front[c], bk = apd.back[c], d1 = apd.apd1[c], d2 = apd.apd2[c];
    uint32_t fs = apd.first[c];
    apd_cell(in.f[0][c], xo[0], apd.count0 + 1, apd.dt, apd.th, fr, bk, d1, d2, fs);
    apd.front[c] = fr;
    apd.back[c] = bk;
    apd.apd1[c] = d1;
    apd.apd2[c] = d2;
    apd.first[c] = (uint8_t)fs;
  }
}

hipError_t launch_direct(int model, const KParams& k, CFields in, Fields out, Fields vt,
                         const uint8_t* solid, hipStream_t s) {
  dim3 b(64, 4), g((k.nx + 63) / 64, (k.ny + 3) / 4);
  if (model == M2V) {
    if (solid) hipLaunchKernelGGL((k_direct<M2V, true>), g, b, 0, s, k, in, out, vt, solid);
    else hipLaunchKernelGGL((k_direct<M2V, false>), g, b, 0, s, k, in, out, vt, solid);
  } else {
    if (solid) hipLaunchKernelGGL((k_direct<MFK, true>), g, b, 0, s, k, in, out, vt, solid);
    else hipLaunchKernelGGL((k_direct<MFK, false>), g, b, 0, s, k, in, out, vt, solid);
  }
  return hipGetLastError();
}

hipError_t launch_direct_codes(int model, const KParams& k, CFields in, Fields out,
                               const uint8_t* codes, hipStream_t s, const ApdState* apd) {
  dim3 b(64, 4), g((k.nx + 63) / 64, (k.ny + 3) / 4);
  const ApdState a = apd ? *apd : ApdState();
  if (model == M2V) {
    if (apd) hipLaunchKernelGGL((k_direct_codes<M2V, true>), g, b, 0, s, k, in, out, codes, a);
    else hipLaunchKernelGGL((k_direct_codes<M2V, false>), g, b, 0, s, k, in, out, codes, a);
  } else {
    if (apd) hipLaunchKernelGGL((k_direct_codes<MFK, true>), g, b, 0, s, k, in, out, codes, a);
    else hipLaunchKernelGGL((k_direct_codes<MFK, false>), g, b, 0, s, k, in, out, codes, a);
  }
  return hipGetLastError();
}

// ------------------------------------------------------------------- tiled
template <int MODEL, bool MASKED, int R, int NW, int CW, bool APD, bool PREV, int ISO = -1,
          bool WT = false>
__global__ __launch_bounds__(64 * NW) void k_tiled(const KParams k, CFields in, Fields out,
                                                   const uint8_t* __restrict__ codes, int K,
                                                   int tiles_x, int ntiles, const ApdState apd,
                                                   float* __restrict__ prev) {
  __shared__ TileLds<MODEL, NW, CW> lds;
  const int tile = xcd_tile(blockIdx.x, ntiles);
  const int ty = tile / tiles_x;
  tile_advance<MODEL, MASKED, R, NW, CW, APD, PREV, ISO, WT>(k, in, out, codes, K,
                                                             tile - ty * tiles_x, ty, lds, apd,
                                                             prev);
}

bool tiled_supported(int rows, int waves, int cols) {
  const bool rw = (rows == 4 && (waves == 4 || waves == 8 || waves == 16)) ||
                  (rows == 8 && (waves == 4 || waves == 8 || waves == 16)) ||
                  (rows == 16 && (waves == 4 || waves == 8)) ||
                  (rows == 7 && waves == 16 && cols == 1) ||  // 64 x 112: 253 tiles at 1024^2, K=9
                  (rows == 5 && waves == 16 && cols == 1);    // 4 x 16 merged with a remainder
  // two columns per lane double the registers: 16 rows, or 8 rows in 16 waves, spill
  return rw && (cols == 1 || (cols == 2 && rows * waves <= 64 && rows <= 8));
}

// Shapes built with the fused APD update (one column per lane, at most 8 rows).
bool tiled_apd_supported(int rows, int waves, int cols) {
  return cols == 1 && (rows == 4 || rows == 8) && (waves == 8 || waves == 16);
}

// Shapes built with the previous-u output (the APD set, 7 x 16 and 5 x 16).
bool tiled_prev_supported(int rows, int waves, int cols) {
  return tiled_apd_supported(rows, waves, cols) ||
         (cols == 1 && (rows == 7 || rows == 5) && waves == 16);
}

template <int MODEL, bool MASKED, int R, int NW, int CW, bool APD, bool PREV = false,
          int ISO = -1, bool WT = false>
static hipError_t launch_tiled_t(int K, const KParams& k, CFields in, Fields out,
                                 const uint8_t* codes, hipStream_t s, const ApdState& apd,
                                 float* prev = nullptr, int* occ = nullptr) {
  if (occ)  // the workgroups of this instantiation one CU holds at once; no launch
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(
        occ, k_tiled<MODEL, MASKED, R, NW, CW, APD, PREV, ISO, WT>, 64 * NW, 0);
  const int wout = 64 * CW - 2 * K, hout = NW * R - 2 * K;
  if (K < 1 || wout <= 0 || hout <= 0) return hipErrorInvalidValue;
  const int tiles_x = (k.nx + wout - 1) / wout;
  const int tiles_y = (k.ny + hout - 1) / hout;
  const int ntiles = tiles_x * tiles_y;
  hipLaunchKernelGGL((k_tiled<MODEL, MASKED, R, NW, CW, APD, PREV, ISO, WT>), dim3(ntiles),
                     dim3(64 * NW), 0, s, k, in, out, codes, K, tiles_x, ntiles, apd, prev);
  return hipGetLastError();
}

template <int MODEL, bool MASKED>
static hipError_t launch_tiled_m(int rows, int waves, int cols, int K, const KParams& k,
                                 CFields in, Fields out, const uint8_t* codes, hipStream_t s,
                                 const ApdState* apd, float* prev, int* occ) {
  const ApdState a = apd ? *apd : ApdState();
  if (apd) {
    if (prev) return hipErrorInvalidValue;
#define YH_TILE_APD(RR, WW)                    \
  if (rows == RR && waves == WW && cols == 1) \
    return launch_tiled_t<MODEL, MASKED, RR, WW, 1, true>(K, k, in, out, codes, s, a, nullptr, occ);
    YH_TILE_APD(4, 8)
    YH_TILE_APD(8, 8)
    YH_TILE_APD(4, 16)
    YH_TILE_APD(8, 16)
#undef YH_TILE_APD
    return hipErrorInvalidValue;
  }
  // the shapes the automatic choice takes (7 x 16 and the 8 x 16 its
  // remainders merge into on 1024^2, 8 x 8 on grids with more tiles than CUs,
  // 4 x 16 on grids with fewer, and the 5 x 16 its remainders merge into)
  // also come with isotropic diffusion fixed at compile time, and with
  // write-through stores (k.wt) for launches of two workgroups per CU
  if (k.iso && cols == 1) {
#define YH_TILE_ISO_W(RR, WW, W)                                              \
  return prev ? launch_tiled_t<MODEL, MASKED, RR, WW, 1, false, true, 1, W>(   \
                    K, k, in, out, codes, s, a, prev, occ)                    \
              : launch_tiled_t<MODEL, MASKED, RR, WW, 1, false, false, 1, W>(  \
                    K, k, in, out, codes, s, a, nullptr, occ);
#define YH_TILE_ISO(RR, WW)               \
  if (rows == RR && waves == WW) {        \
    if (k.wt) YH_TILE_ISO_W(RR, WW, true) \
    YH_TILE_ISO_W(RR, WW, false)          \
  }
    YH_TILE_ISO(7, 16)
    YH_TILE_ISO(8, 16)
    YH_TILE_ISO(8, 8)
    YH_TILE_ISO(4, 16)
    YH_TILE_ISO(5, 16)
#undef YH_TILE_ISO
#undef YH_TILE_ISO_W
  }
  if (prev) {
#define YH_TILE_PREV(RR, WW)                   \
  if (rows == RR && waves == WW && cols == 1) \
    return launch_tiled_t<MODEL, MASKED, RR, WW, 1, false, true>(K, k, in, out, codes, s, a, prev, \
                                                                 occ);
    YH_TILE_PREV(4, 8)
    YH_TILE_PREV(8, 8)
    YH_TILE_PREV(4, 16)
    YH_TILE_PREV(8, 16)
    YH_TILE_PREV(7, 16)
    YH_TILE_PREV(5, 16)
#undef YH_TILE_PREV
    return hipErrorInvalidValue;
  }
#define YH_TILE_CASE(RR, WW, CC)                 \
  if (rows == RR && waves == WW && cols == CC) \
    return launch_tiled_t<MODEL, MASKED, RR, WW, CC, false>(K, k, in, out, codes, s, a, nullptr, \
                                                            occ);
  YH_TILE_CASE(4, 4, 1)
  YH_TILE_CASE(8, 4, 1)
  YH_TILE_CASE(16, 4, 1)
  YH_TILE_CASE(4, 8, 1)
  YH_TILE_CASE(8, 8, 1)
  YH_TILE_CASE(16, 8, 1)
  YH_TILE_CASE(4, 16, 1)
  YH_TILE_CASE(8, 16, 1)
  YH_TILE_CASE(7, 16, 1)
  YH_TILE_CASE(5, 16, 1)
  YH_TILE_CASE(4, 4, 2)
  YH_TILE_CASE(8, 4, 2)
  YH_TILE_CASE(4, 8, 2)
  YH_TILE_CASE(8, 8, 2)
  YH_TILE_CASE(4, 16, 2)
#undef YH_TILE_CASE
  return hipErrorInvalidValue;
}

hipError_t launch_tiled(int model, bool masked, int rows, int waves, int cols, int K,
                        const KParams& k, CFields in, Fields out, const uint8_t* codes,
                        hipStream_t s, const ApdState* apd, float* prev, int* occ) {
  if (model == M2V) {
    return masked
               ? launch_tiled_m<M2V, true>(rows, waves, cols, K, k, in, out, codes, s, apd, prev, occ)
               : launch_tiled_m<M2V, false>(rows, waves, cols, K, k, in, out, codes, s, apd, prev,
                                            occ);
  }
  return masked
             ? launch_tiled_m<MFK, true>(rows, waves, cols, K, k, in, out, codes, s, apd, prev, occ)
             : launch_tiled_m<MFK, false>(rows, waves, cols, K, k, in, out, codes, s, apd, prev, occ);
}

// -------------------------------------------------------------------- tips
// Launched over the cells the tip test can accept: from (i0, j0), the
// centred disk's bounding box with a mask, the whole grid without.
__global__ void k_tips(int nx, int ny, int i0, int j0, int masked_disk, int r2, double Uth,
                       const float* __restrict__ past, const float* __restrict__ present,
                       float t_host, const long long* d_count, double dt, yh_tip* tips,
                       int* keys, int* count, int cap) {
  const int i = i0 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = j0 + blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= nx || j >= ny) return;
  // physicalTime = dt*count (main.cu:885), from the device step counter when given
  const float t = d_count ? (float)(dt * (double)(*d_count)) : t_host;
  tip_cell(i, j, nx, ny, masked_disk != 0, r2, Uth, past, present, t, tips, keys, count, cap);
}

hipError_t launch_tips(const KParams& k, int masked_disk, int tipOffsetX, int tipOffsetY,
                       double Uth, const float* past, const float* present, double time,
                       yh_tip* tips, int* keys, int* count, int cap, hipStream_t s) {
  const int r2 = tipOffsetX * tipOffsetY;
  int i0 = 0, j0 = 0, i1 = k.nx, j1 = k.ny;  // [i0, i1) x [j0, j1)
  if (masked_disk) {
    if (r2 <= 0) return hipSuccess;  // the disk is empty
    // (i - nx/2)^2 + (j - ny/2)^2 < r2 needs |i - nx/2| <= R with R^2 < r2
    int R = 0;
    while ((long long)(R + 1) * (R + 1) < r2) ++R;
    i0 = std::max(0, k.nx / 2 - R);
    j0 = std::max(0, k.ny / 2 - R);
    i1 = std::min(k.nx, k.nx / 2 + R + 1);
    j1 = std::min(k.ny, k.ny / 2 + R + 1);
  }
  dim3 b(64, 4), g((i1 - i0 + 63) / 64, (j1 - j0 + 3) / 4);
  hipLaunchKernelGGL(k_tips, g, b, 0, s, k.nx, k.ny, i0, j0, masked_disk, r2, Uth, past, present,
                     (float)time, (const long long*)nullptr, 0.0, tips, keys, count, cap);
  return hipGetLastError();
}

// --------------------------------------------------------------- finite
__global__ void k_finite(const float* __restrict__ a, const float* __restrict__ b,
                         const float* __restrict__ c, size_t n, int* flag) {
  bool bad = false;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n;
       t += (size_t)gridDim.x * blockDim.x) {
    bad |= !isfinite(a[t]);
    if (b) bad |= !isfinite(b[t]);
    if (c) bad |= !isfinite(c[t]);
  }
  if (bad) atomicOr(flag, 1);
}

hipError_t launch_finite(const float* const* f, int nf, size_t n, int* flag, hipStream_t s) {
  hipLaunchKernelGGL(k_finite, dim3(1024), dim3(256), 0, s, f[0], nf > 1 ? f[1] : nullptr,
                     nf > 2 ? f[2] : nullptr, n, flag);
  return hipGetLastError();
}

}  // namespace yh
