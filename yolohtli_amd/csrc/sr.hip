// sr.hip -- symmetry-reduction (co-moving frame) step on the device.
//
// One SR step follows display() (main.cu:894-953).  Three kernels make it:
//   1 k_sr_rdtips   RD X -> Xo (reactionDiffusion_kernel) and the tips
//                   between Xo.u and X.u (tip_wrapper, main.cu:900); blocks
//                   overlap by one column and row, so every corner of an
//                   owned cell is updated in the block, by the same code
//                   as in the block that owns it
//   2 k_sr_slices   every workgroup picks the disk centre (last tip of this
//                   step in (cell, root) order, tipx0 at count 0,
//                   symmetryReduction.cu:95-107), then over the disk's
//                   bounding box: slices, template and velTan in FP64, the
//                   12 disk inner products as per-workgroup partial sums
//                   (slice_kernel + trapz_kernel fused: no slice arrays)
//   3 k_bf_fused    every workgroup reduces the partials in one fixed order
//                   and solves solve_matrix (symmetryReduction.cu:386-418)
//                   itself, so all agree bit for bit; then BFECC of Xo into X
//                   (advFDBFECC.cu:18-165), the three dependent phases through
//                   an LDS tile with a 3-cell halo.  Workgroup 0 commits the
//                   step: c, phi += dt*c, centre, count, the tips to the log.
// By default a steady step is two launches: k_sr_slices, then k_bf_rd, which
// is k_bf_fused on a 5-cell halo followed by the next step's RD and tip test
// (what k_sr_rdtips would compute, bit for bit) on the same tile; the first
// step of a run, or of one after the state changed from outside, starts with
// k_sr_rdtips.  YH_OPT_SR_LAUNCHES 3 keeps the three separate launches.
// The SR state has two slots: a step reads slot count&1 and workgroup 0 of
// the BFECC launch writes slot (count+1)&1, so no workgroup reads what
// another writes in the same launch.  At count 0 a preliminary slices + solve
// pass runs first (main.cu:910-921).  Nothing leaves the device; steady
// steps are captured into hipGraphs (8 steps to a graph where at least 8
// remain) and replayed.  Results cross workgroups only at kernel
// boundaries: on this 8-XCD part a device-scope fence writes back the whole
// L2 of its XCD, so a last-workgroup election in every workgroup costs far
// more than a launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "direct.h"
#include "plan.h"

namespace yh {

struct SrDev {
  double c[3], phi[3];
  double ac2, axc, ayc;  // advection field = (hy*j*ac2 + axc, -hx*i*ac2 + ayc)
  double aphi;
  double I[12];
  long long count;
  int cx, cy;
  float tip0x, tip0y;
  int status;  // bit 0: singular solve
  int ntips;   // tips found in the last step
};

struct SrGeom {
  int nx, ny, r2, nv;
  double hx, hy, invdx, invdy, dt, tc;
  double dtx, dty;  // dt/hx, dt/hy (advection Courant factors)
};

// Buffers of one step: the state slots, the step's tips, the partials, the
// plan's tip log and the (c, phi) history.
struct SrArgs {
  const SrDev* st_in;
  SrDev* st_out;
  const double* partials;
  int nparts;
  int record;  // push (c, phi) before the solve into hist
  double* hist;
  long long histcap;
  const yh_tip* step_tips;
  const int* step_keys;
  int* step_count;
  yh_tip* log_tips;
  int* log_keys;
  int* log_count;
  int logcap;
  yh_tip* next_tips;  // the next step's tips (k_bf_rd)
  int* next_keys;
  int* next_count;
};

constexpr int kSrTipCap = 4096;
constexpr int kSrChunk = 8;  // steady SR steps per captured graph
constexpr int kSliceBX = 64, kSliceBY = 8;

__device__ __forceinline__ bool in_disk(const SrGeom& g, int i, int j, int cx, int cy) {
  const int a = i - g.nx / 2 - cx, b = j - g.ny / 2 - cy;
  return a * a + b * b < g.r2;
}

// ------------------------------------------------------------- RD + tips
// Blocks of 64 x 4 threads overlap by one column and one row: a block owns
// its first 63 x 3 cells and computes the RD step of all 64 x 4 (every thread
// exactly one), so the east, north and north-east corners of each owned cell
// are in the block's LDS.  A cell in the overlap is computed by two blocks
// with the same code, bit for bit, and stored by its owner.  Then the tip
// test between the new and the old u (g_past = RD output, g_present = old
// state: main.cu:900, tipTracker.cu:586).
constexpr int kRtBX = 64, kRtBY = 4;          // threads per block
constexpr int kRtOX = kRtBX - 1, kRtOY = kRtBY - 1;  // cells owned per block

template <int MODEL, bool MASKED>
__global__ __launch_bounds__(kRtBX* kRtBY) void k_sr_rdtips(const KParams k, const SrGeom g,
                                                           CFields X, Fields Xo,
                                                           const uint8_t* __restrict__ codes,
                                                           double Uth,
                                                           const SrDev* __restrict__ st,
                                                           yh_tip* tips, int* keys, int* count) {
  using M = Model<MODEL>;
  __shared__ float s_u[kRtBY * kRtBX];
  const int nx = g.nx, ny = g.ny;
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int i = blockIdx.x * kRtOX + tx, j = blockIdx.y * kRtOY + ty;
  const bool valid = i < nx && j < ny;
  const bool owned = valid && tx < kRtOX && ty < kRtOY;
  const int l0 = tx + kRtBX * ty;
  if (valid) {
    float xo[3], inc[3];
    cell_update_codes<MODEL>(k, X.f, MASKED ? codes : nullptr, i, j, xo, inc);
    if (owned) {
#pragma unroll
      for (int a = 0; a < M::NV; ++a) Xo.f[a][i + nx * j] = xo[a];
    }
    s_u[l0] = xo[0];
  }
  __syncthreads();
  if (!owned || !tip_inside(i, j, nx, ny, MASKED, g.r2)) return;
  // corners past the last column/row are the cell itself (tipTracker.cu:58-63)
  const bool ex = i < nx - 1, ey = j < ny - 1;
  const int lx = ex ? l0 + 1 : l0, ly = ey ? l0 + kRtBX : l0;
  const int lxy = (ex && ey) ? l0 + kRtBX + 1 : l0;
  const int c = i + nx * j;
  const int sx = ex ? c + 1 : c, sy = ey ? c + nx : c, sxy = (ex && ey) ? c + nx + 1 : c;
  const float* pr = X.f[0];
  // physicalTime = dt*count (main.cu:885)
  const float t = (float)(g.dt * (double)st->count);
  tip_solve(i, j, nx, MASKED, Uth, pr[c], pr[sx], pr[sxy], pr[sy], s_u[l0], s_u[lx], s_u[lxy],
            s_u[ly], t, tips, keys, count, kSrTipCap);
}

// ---------------------------------------------------------------- centre
// Disk centre of this step, the same in every wave that asks: tipx0 at
// count 0 (symmetryReduction.cu:98-100), else the last tip of this step in
// (cell, root) order (:102-103), else the previous centre.  Keys are unique,
// so the wave maximum names one slot.
__device__ __forceinline__ void sr_centre(const SrGeom& g, const SrDev* st, const yh_tip* tips,
                                          const int* keys, const int* count, int& cx, int& cy) {
  const int lane = __lane_id();
  // the step count, the tip count, the first 64 keys and their tips are read
  // together (slots past the tip count are ignored), so the centre costs one
  // memory round trip when a step has at most 64 tips
  const long long step = st->count;
  const int k0 = keys[lane];
  const yh_tip t0 = tips[lane];
  const int cnt = *count;
  if (step == 0) {
    cx = __float2int_rn(st->tip0x - (float)(g.nx / 2));
    cy = __float2int_rn(st->tip0y - (float)(g.ny / 2));
    return;
  }
  const int n = min(cnt, kSrTipCap);
  int bk = -1, bs = -1;
  float bx = 0.0f, by = 0.0f;
  if (lane < n) {
    bk = k0;
    bs = lane;
    bx = t0.x;
    by = t0.y;
  }
  for (int q = lane + 64; q < n; q += 64) {
    const int kq = keys[q];
    if (kq > bk) {
      bk = kq;
      bs = q;
      bx = tips[q].x;
      by = tips[q].y;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    const int ok = __shfl_xor(bk, off, 64), os = __shfl_xor(bs, off, 64);
    const float ox = __shfl_xor(bx, off, 64), oy = __shfl_xor(by, off, 64);
    if (ok > bk) {
      bk = ok;
      bs = os;
      bx = ox;
      by = oy;
    }
  }
  if (bs >= 0) {
    cx = __float2int_rn(bx - (float)(g.nx / 2));
    cy = __float2int_rn(by - (float)(g.ny / 2));
  } else {
    cx = st->cx;
    cy = st->cy;
  }
}

// ---------------------------------------------------------------- slices
// Slices, template and velTan in FP64 for the cells of the disk around this
// step's centre (launched over the disk's bounding box, reach R), the 12
// inner products as per-workgroup partial sums.
template <int MODEL, bool MASKED>
__global__ __launch_bounds__(kSliceBX* kSliceBY) void k_sr_slices(
    const KParams k, const SrGeom g, int R, CFields X, const uint8_t* __restrict__ codes,
    const SrDev* __restrict__ st, const yh_tip* __restrict__ step_tips,
    const int* __restrict__ step_keys, const int* __restrict__ step_count,
    double* __restrict__ partials) {
  using M = Model<MODEL>;
  constexpr int NV = M::NV;
  int cx, cy;
  sr_centre(g, st, step_tips, step_keys, step_count, cx, cy);
  const int i = g.nx / 2 + cx - R + blockIdx.x * kSliceBX + threadIdx.x;
  const int j = g.ny / 2 + cy - R + blockIdx.y * kSliceBY + threadIdx.y;
  double acc[12];
#pragma unroll
  for (int q = 0; q < 12; ++q) acc[q] = 0.0;
  if (i >= 0 && i < g.nx && j >= 0 && j < g.ny && in_disk(g, i, j, cx, cy)) {
    const int nx = g.nx, ny = g.ny;
    const int C = i + nx * j;
    const int E = mirror(i + 1, nx) + nx * j, W = mirror(i - 1, nx) + nx * j;
    const int EE = mirror(i + 2, nx) + nx * j, WW = mirror(i - 2, nx) + nx * j;
    const int N = i + nx * mirror(j + 1, ny), S = i + nx * mirror(j - 1, ny);
    const int NN = i + nx * mirror(j + 2, ny), SS = i + nx * mirror(j - 2, ny);
    const double ax = g.hy * (double)j * st->ac2 + st->axc;
    const double ay = -g.hx * (double)i * st->ac2 + st->ayc;
    // velTan = rhs/dt of the RD step of X at this cell (reactionDiffusion.cu:551-552)
    float cW = 1.f, cE = 1.f, cN = 1.f, cS = 1.f;
    bool sc = true;
    if (MASKED) {
      const uint32_t b = codes[C];
      sc = b != kHole;
      const uint32_t q = sc ? b : 0u;
      cW = (float)(q & 3u);
      cE = (float)((q >> 2) & 3u);
      cN = (float)((q >> 4) & 3u);
      cS = (float)((q >> 6) & 3u);
    }
    float x[3] = {0.f, 0.f, 0.f}, d[3] = {0.f, 0.f, 0.f}, r[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int a = 0; a < NV; ++a) x[a] = X.f[a][C];
#pragma unroll
    for (int a = 0; a < M::ND; ++a) {
      const float* f = X.f[a];
      const float fc = x[a];
      d[a] = k.rx * (cW * (f[W] - fc) + cE * (f[E] - fc)) + k.ry * (cN * (f[N] - fc) + cS * (f[S] - fc));
    }
    M::react(k, x[0], x[1], x[2], r[0], r[1], r[2]);
    const float scale[3] = {1.0f, MODEL == M2V ? k.rs : 0.0f, 0.0f};
#pragma unroll
    for (int a = 0; a < NV; ++a) {
      const float* f = X.f[a];
      const double fC = f[C], fE = f[E], fEE = f[EE], fW = f[W], fWW = f[WW];
      const double fN = f[N], fNN = f[NN], fS = f[S], fSS = f[SS];
      // convFB2ndOX/Y (symmetryReduction.cu:242-262)
      const double fx = (ax > 0.0) ? (-3.0 * fC + 4.0 * fE - fEE) * g.invdx
                                   : (3.0 * fC - 4.0 * fW + fWW) * g.invdx;
      const double fy = (ay > 0.0) ? (-3.0 * fC + 4.0 * fN - fNN) * g.invdy
                                   : (3.0 * fC - 4.0 * fS + fSS) * g.invdy;
      const double ft = g.hx * (double)i * fy - g.hy * (double)j * fx;
      // convCentral2X/Y (symmetryReduction.cu:224-240)
      const double fx0 = (fEE - 8.0 * fE + 8.0 * fW - fWW) * g.invdx * (1.0 / 6.0);
      const double fy0 = (fNN - 8.0 * fN + 8.0 * fS - fSS) * g.invdy * (1.0 / 6.0);
      const double ft0 = g.hx * (double)i * fy0 - g.hy * (double)j * fx0;
      const float inc = k.tc * scale[a] * d[a] + r[a];
      const double vt = sc ? (double)(inc / k.tcdt) : 0.0;
      const double t0[3] = {fx0, fy0, ft0};
      const double t1[3] = {fx, fy, ft};
#pragma unroll
      for (int rr = 0; rr < 3; ++rr) {
#pragma unroll
        for (int qq = 0; qq < 3; ++qq) acc[3 * rr + qq] += t0[rr] * t1[qq];
        acc[9 + rr] += t0[rr] * vt;
      }
    }
  }
  // block reduction: wave shuffles, then across the 4 waves through LDS
  __shared__ double red[kSliceBX * kSliceBY / 64][12];
  const int tid = threadIdx.y * kSliceBX + threadIdx.x;
  const int lane = tid & 63, wv = tid >> 6;
  // Sums over the wave's 64 lanes with the pairs of a shfl_down tree (at
  // each level lane i's partial meets lane i + off's, so every sum equals
  // that tree's bit for bit: IEEE addition commutes) but fewer exchanges:
  // at the first level the lower half of the wave keeps values 0-5 and the
  // upper half 6-11, at the second each quarter keeps three, so 6 + 3 +
  // 4 x 3 exchanges instead of 6 x 12.  Values 3r..3r+2 end in lane 16r.
  const bool up32 = (lane & 32) != 0, up16 = (lane & 16) != 0;
  double h[6], t[3];
#pragma unroll
  for (int p = 0; p < 6; ++p) {
    const double give = up32 ? acc[p] : acc[6 + p];
    h[p] = (up32 ? acc[6 + p] : acc[p]) + __shfl_xor(give, 32, 64);
  }
#pragma unroll
  for (int p = 0; p < 3; ++p) {
    const double give = up16 ? h[p] : h[3 + p];
    t[p] = (up16 ? h[3 + p] : h[p]) + __shfl_xor(give, 16, 64);
  }
#pragma unroll
  for (int off = 8; off > 0; off >>= 1) {
#pragma unroll
    for (int p = 0; p < 3; ++p) t[p] += __shfl_xor(t[p], off, 64);
  }
  if ((lane & 15) == 0) {
#pragma unroll
    for (int p = 0; p < 3; ++p) red[wv][3 * (lane >> 4) + p] = t[p];
  }
  __syncthreads();
  if (tid < 12) {
    double s = 0.0;
    for (int w = 0; w < kSliceBX * kSliceBY / 64; ++w) s += red[w][tid];
    partials[(size_t)tid * gridDim.x * gridDim.y + blockIdx.y * gridDim.x + blockIdx.x] = s;
  }
}

// ----------------------------------------------------------------- solve
// Column q of the partials (stored column-major, [12][nparts]) summed by one
// wave in a fixed order (lane-strided, then a shuffle tree), so the integrals do not depend on scheduling and
// every wave that sums a column gets the same bits.  The total is in lane 0.
__device__ __forceinline__ double sr_column_sum(const double* partials, int nparts, int q) {
  const int lane = __lane_id();
  double s = 0.0;
  const double* col = partials + (size_t)q * nparts;
  // four loads in flight per lane before the adds, which keep the order
  // lane, lane + 64, lane + 128, ... of a one-at-a-time loop
  for (int b0 = 0; b0 < nparts; b0 += 256) {
    double v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int b = b0 + lane + 64 * u;
      v[u] = b < nparts ? col[b] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (b0 + lane + 64 * u < nparts) s += v[u];
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  return s;
}

// One thread: the integrals sums*hx*hy, solve_matrix
// (symmetryReduction.cu:386-418), the advection parameters of Cxy_field
// (:45-56) with the new c and the current phi, and, when advance, the Euler
// update phi += dt*c (reduceSymStart stays set, main.cu:935-941) and the
// step count.
__device__ void sr_solve(const SrGeom& g, const double* sums, SrDev& o, bool advance, double sp,
                         double cp) {
  double I[12];
  for (int p = 0; p < 12; ++p) {
    I[p] = sums[p] * (g.hx * g.hy);
    o.I[p] = I[p];
  }
  const double a1 = I[0] * cp + I[1] * sp, a2 = I[1] * cp - I[0] * sp, a3 = I[2];
  const double b1 = I[3] * cp + I[4] * sp, b2 = I[4] * cp - I[3] * sp, b3 = I[5];
  const double C1 = I[6] * cp + I[7] * sp, C2 = I[7] * cp - I[6] * sp, C3 = I[8];
  const double d1 = I[9], d2 = I[10], d3 = I[11];
  const double b2p = a1 / b1 * b2 - a2, b3p = a1 / b1 * b3 - a3, d2p = a1 / b1 * d2 - d1;
  const double c2p = a1 / C1 * C2 - a2, c3p = a1 / C1 * C3 - a3, d3p = a1 / C1 * d3 - d1;
  const double c3pp = b2p / c2p * c3p - b3p, d3pp = b2p / c2p * d3p - d2p;
  const double x3 = d3pp / c3pp;
  const double x2 = (d2p - b3p * x3) / b2p;
  const double x1 = (d1 - a2 * x2 - a3 * x3) / a1;
  if (!(isfinite(x1) && isfinite(x2) && isfinite(x3))) o.status |= 1;
  o.c[0] = x1;
  o.c[1] = x2;
  o.c[2] = x3;
  o.aphi = o.phi[2];
  o.ac2 = x3;
  o.axc = -x1 * cp + x2 * sp;
  o.ayc = -x1 * sp - x2 * cp;
  if (advance) {
    for (int p = 0; p < 3; ++p) o.phi[p] = o.phi[p] + o.c[p] * g.dt;
    o.count += 1;
  }
}

// (c, phi) of the state before this step's first solve: the clist/philist
// push (main.cu:902-903); at count 0 that precedes the preliminary solve
// (main.cu:910-914).
__device__ __forceinline__ void sr_record(const SrDev& o, double* hist, long long histcap) {
  if (!hist || o.count >= histcap) return;
  double* h = hist + (size_t)o.count * 6;
  for (int p = 0; p < 3; ++p) {
    h[p] = o.c[p];
    h[3 + p] = o.phi[p];
  }
}

// The preliminary solve of count 0 (main.cu:910-921), in place: one
// workgroup of 12 waves, wave q sums column q.
__global__ __launch_bounds__(768) void k_sr_solve_pre(SrGeom g, SrDev* st, const double* partials,
                                                      int nparts, double* hist,
                                                      long long histcap) {
  __shared__ double Is[12];
  const int q = threadIdx.x >> 6;
  const double s = sr_column_sum(partials, nparts, q);
  if (__lane_id() == 0) Is[q] = s;
  __syncthreads();
  if (threadIdx.x != 0) return;
  SrDev o = *st;
  sr_record(o, hist, histcap);
  double sp, cp;
  sincos(o.phi[2], &sp, &cp);
  sr_solve(g, Is, o, false, sp, cp);
  *st = o;
}

// ----------------------------------------------------------------- BFECC
// advFDBFECC.cu:44-66 (masked) and :112-126 (square): which neighbours a cell
// uses.  A neighbour outside the domain counts as a hole, and each cell tests
// the mask of the neighbour it actually uses; a substitute neighbour outside
// the domain (a hole on one side at the domain edge) reads as a hole, 0 (the
// reference takes the linear index i2d +- 1 there, the next row's first cell).  Bits: 0 sc (cell is tissue),
// 1 sw, 2 se, 3 ss, 4 sn (that neighbour is usable).
template <bool MASKED>
__device__ __forceinline__ uint32_t bf_flags(int i, int j, int nx, int ny,
                                             const uint8_t* __restrict__ solid) {
  const int gc = i + nx * j;
  bool sc = true, sw, se, ss, sn;
  if (MASKED) {
    sc = solid[gc] != 0;
    sw = (i > 0) && solid[gc - 1];
    se = (i < nx - 1) && solid[gc + 1];
    ss = (j > 0) && solid[gc - nx];
    sn = (j < ny - 1) && solid[gc + nx];
  } else {
    sw = i > 0;
    se = i < nx - 1;
    ss = j > 0;
    sn = j < ny - 1;
  }
  return (sc ? 1u : 0u) | (sw ? 2u : 0u) | (se ? 4u : 0u) | (ss ? 8u : 0u) | (sn ? 16u : 0u);
}

struct BfIdx {
  int W, E, S, N, iW, W2, iE, E2, iS, S2, iN, N2;
};

// Neighbour indices around c (row stride `stride`) for flags f.
__device__ __forceinline__ BfIdx bf_index(uint32_t f, int c, int stride) {
  const bool sc = f & 1u, sw = f & 2u, se = f & 4u, ss = f & 8u, sn = f & 16u;
  BfIdx r;
  r.W = sc ? (sw ? c - 1 : c + 1) : c;
  r.E = sc ? (se ? c + 1 : c - 1) : c;
  r.S = sc ? (ss ? c - stride : c + stride) : c;
  r.N = sc ? (sn ? c + stride : c - stride) : c;
  r.iW = sc ? (sw ? c : c + 1) : c;
  r.W2 = sc ? (sw ? c - 1 : c) : c;
  r.iE = sc ? (se ? c : c - 1) : c;
  r.E2 = sc ? (se ? c + 1 : c) : c;
  r.iS = sc ? (ss ? c : c + stride) : c;
  r.S2 = sc ? (ss ? c - stride : c) : c;
  r.iN = sc ? (sn ? c : c - stride) : c;
  r.N2 = sc ? (sn ? c + stride : c) : c;
  return r;
}

// The three dependent BFECC phases (advFDBFECC.cu:128-144) in one launch:
// each phase reads only the previous phase at distance 1, so a workgroup
// loads its kBfTX x kBfTY output tile with a halo of 3 into LDS and computes
// phase p on the tile grown by 3 - p cells.  Each cell's advection
// coefficients (Cxy_field_kernel, symmetryReduction.cu:45-56, negated as at
// advFDBFECC.cu:30-31) and neighbour flags are computed once.  Cells outside
// the domain are never read: the boundary rules substitute an inside
// neighbour.
constexpr int kBfTX = 64, kBfTY = 16, kBfHalo = 3, kBfThreads = 1024;
// The fused step (k_bf_rd) also takes the RD step of the new state on the
// tile grown by one cell and the tip test on the tile, which needs the new
// state on the tile grown by two: a halo of 3 + 2.
constexpr int kBfHaloRd = 5;
template <int HALO>
struct BfTile {
  static constexpr int LW = kBfTX + 2 * HALO, LH = kBfTY + 2 * HALO, LN = LW * LH;
};
constexpr int kBfLN = BfTile<kBfHalo>::LN;

// The new state and the RD output are written through to memory
// (store_through): the next kernel of the step reads these fields from
// workgroups on every XCD, so dirty lines would only be written back at the
// end of the launch, between the two kernels of a step (a step measured
// 14.9 against 16.0 us at 512^2).

// Cells of a thread in the three phases: the phase-1 region (the output
// tile grown by HALO - 1 cells), cell l = tid + it * kBfThreads of it, so a
// thread keeps its cells' flags, Courant factors and own values in
// registers from phase to phase and reads only neighbours from LDS.
template <int HALO>
struct BfCells {
  static constexpr int W1 = BfTile<HALO>::LW - 2, H1 = BfTile<HALO>::LH - 2, N1 = W1 * H1;
  static constexpr int IT = (N1 + kBfThreads - 1) / kBfThreads;
};

// Per-thread state of its cells: flags (bit 7: inside the domain, bits 5-6:
// the advection's signs), |c| dt/h per axis, and per variable the input q,
// the forward step uf and the corrected field ue of the cell itself.
template <int NV, int HALO>
struct BfOwn {
  static constexpr int IT = BfCells<HALO>::IT;
  uint32_t fl[IT];
  float rx[IT], ry[IT];
  float q[IT][NV], f[IT][NV], e[IT][NV];
};

// One phase over its region (the output tile grown by HALO - PH cells).
// Phase 3 writes the tile's cells to `out`; with KEEP it also keeps every
// cell of its region (0 in holes) in s_f for what follows in the launch.
template <int PH, int NV, int HALO = kBfHalo, bool KEEP = false>
__device__ __forceinline__ void bf_phase(int tid, float tc, float (*s_q)[BfTile<HALO>::LN],
                                         float (*s_f)[BfTile<HALO>::LN], BfOwn<NV, HALO>& o,
                                         Fields out, int i0, int j0, int nx) {
  constexpr int LW = BfTile<HALO>::LW, LH = BfTile<HALO>::LH;
  constexpr int W1 = BfCells<HALO>::W1, N1 = BfCells<HALO>::N1, IT = BfCells<HALO>::IT;
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int l = tid + it * kBfThreads;
    const int li = 1 + l % W1, lj = 1 + l / W1;
    const bool in_region = l < N1 && li >= PH && li < LW - PH && lj >= PH && lj < LH - PH;
    const uint32_t fl = o.fl[it];
    if (in_region && (fl & 128u)) {  // inside the domain
      const int L = li + LW * lj;
      const BfIdx x = bf_index(fl, L, LW);
      const float Rx = o.rx[it], Ry = o.ry[it];
      const bool px = fl & 32u, py = fl & 64u;
#pragma unroll
      for (int a = 0; a < NV; ++a) {
        if (PH == 1) {  // forward step (advFDBFECC.cu:128-131)
          const float* q = s_q[a];
          const float qL = o.q[it][a];
          const float FDx = px ? qL - q[x.W] : qL - q[x.E];
          const float FDy = py ? qL - q[x.S] : qL - q[x.N];
          const float uf = qL - tc * (Rx * FDx + Ry * FDy);
          s_f[a][L] = uf;
          o.f[it][a] = uf;
        } else if (PH == 2) {  // backward step and error correction (:134-139)
          const float* f = s_f[a];
          const float FDx = px ? f[x.iE] - f[x.E2] : f[x.iW] - f[x.W2];
          const float FDy = py ? f[x.iN] - f[x.N2] : f[x.iS] - f[x.S2];
          const float ub = o.f[it][a] - tc * (Rx * FDx + Ry * FDy);
          const float g0 = o.q[it][a];
          const float ue = g0 - 0.5f * (ub - g0);
          s_q[a][L] = ue;  // over the input: only cell L read it there
          o.e[it][a] = ue;
        } else {  // forward step from the corrected field (:141-144)
          const float* e = s_q[a];
          const float eL = o.e[it][a];
          const float FDx = px ? eL - e[x.W] : eL - e[x.E];
          const float FDy = py ? eL - e[x.S] : eL - e[x.N];
          const float r = eL - tc * (Rx * FDx + Ry * FDy);
          const float m = (fl & 1u) ? r : 0.0f;
          if (KEEP) s_f[a][L] = m;  // uf is dead after phase 2
          if (li >= HALO && li < HALO + kBfTX && lj >= HALO && lj < HALO + kBfTY)
            store_through(out.f[a], (uint32_t)((i0 + li) + nx * (j0 + lj)), m);
        }
      }
    }
  }
}

// The solve and the three BFECC phases of one step on a tile with a halo of
// HALO cells: s_q holds the state advected, then the corrected field ue;
// s_f the forward step uf (with KEEP, then phase 3's output).  A thread's
// cells carry their flags bf_flags | px << 5 | py << 6 | inside << 7 in
// registers (BfOwn).
// Returns the step count of the state it advanced (from LDS).
template <bool MASKED, int NV, int HALO, bool KEEP>
__device__ __forceinline__ long long bf_advect(const SrGeom& g, const SrArgs& sa, CFields in,
                                          Fields out, const uint8_t* __restrict__ solid,
                                          float (*s_q)[BfTile<HALO>::LN],
                                          float (*s_f)[BfTile<HALO>::LN]) {
  constexpr int LW = BfTile<HALO>::LW, LN = BfTile<HALO>::LN;
  constexpr int W1 = BfCells<HALO>::W1, N1 = BfCells<HALO>::N1, IC = BfCells<HALO>::IT;
  __shared__ double s_I[12];
  __shared__ double s_adv[3];  // ac2, axc, ayc
  __shared__ int s_cx, s_cy, s_found, s_base;
  __shared__ SrDev s_st;      // this step's SR state
  __shared__ double s_sc[2];  // sine and cosine of its angle phi[2]
  static_assert(sizeof(SrDev) % 8 == 0 && sizeof(SrDev) / 8 <= 64, "SrDev copy by one wave");
  static_assert(kBfThreads / 64 >= 14, "waves 12 and 13 fetch the state");
  const int tid = threadIdx.x;
  const int lane = tid & 63, wv = tid >> 6;
  const int nx = g.nx, ny = g.ny;
  const int i0 = blockIdx.x * kBfTX - HALO, j0 = blockIdx.y * kBfTY - HALO;
  const bool commit = blockIdx.x == 0 && blockIdx.y == 0;
  // 1 the state tile and the neighbour flags, which do not depend on the
  //   solve; the integrals (wave w sums column w, w < 12 of the 16 waves);
  //   in workgroup 0 the centre and the tip log slots.  The tile's loads
  //   are issued first, so they and the partials' loads are in flight
  //   together.
  constexpr int IT = (LN + kBfThreads - 1) / kBfThreads;
  float tv[IT][NV];
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int l = tid + it * kBfThreads;
    if (l < LN) {
      const int gi = i0 + l % LW, gj = j0 + l / LW;
      const bool ok = gi >= 0 && gi < nx && gj >= 0 && gj < ny;
      // cells outside the domain hold 0 in every phase: the masked rule's
      // substitute neighbour can lie there (a hole on one side at the domain
      // edge), and reads as a hole
#pragma unroll
      for (int a = 0; a < NV; ++a) tv[it][a] = ok ? in.f[a][gi + (size_t)nx * gj] : 0.0f;
    }
  }
  // the flags of this thread's cells (the phase-1 region)
  BfOwn<NV, HALO> o;
#pragma unroll
  for (int it = 0; it < IC; ++it) {
    const int l = tid + it * kBfThreads;
    const int gi = i0 + 1 + l % W1, gj = j0 + 1 + l / W1;
    const bool ok = l < N1 && gi >= 0 && gi < nx && gj >= 0 && gj < ny;
    o.fl[it] = ok ? (bf_flags<MASKED>(gi, gj, nx, ny, solid) | 128u) : 0u;
  }
  for (int q = wv; q < 12; q += kBfThreads / 64) {
    const double v = sr_column_sum(sa.partials, sa.nparts, q);
    if (lane == 0) s_I[q] = v;
  }
  // the state of this step and the sine and cosine of its angle, which do
  // not depend on the integrals, by two of the waves that sum no column
  if (wv == 12) {
    if (lane < (int)(sizeof(SrDev) / 8))
      reinterpret_cast<unsigned long long*>(&s_st)[lane] =
          reinterpret_cast<const unsigned long long*>(sa.st_in)[lane];
  } else if (wv == 13 && lane == 0) {
    double sp, cp;
    sincos(sa.st_in->phi[2], &sp, &cp);
    s_sc[0] = sp;
    s_sc[1] = cp;
  }
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int l = tid + it * kBfThreads;
    if (l < LN) {
#pragma unroll
      for (int a = 0; a < NV; ++a) {
        s_q[a][l] = tv[it][a];
        s_f[a][l] = 0.0f;
      }
    }
  }
  if (commit && wv == 0) {
    int cx, cy;
    sr_centre(g, sa.st_in, sa.step_tips, sa.step_keys, sa.step_count, cx, cy);
    if (lane == 0) {
      const int found = *sa.step_count;
      s_cx = cx;
      s_cy = cy;
      s_found = found;
      s_base = atomicAdd(sa.log_count, found);
    }
  }
  __syncthreads();
  // 2 the solve, in every workgroup from the same sums; workgroup 0 writes
  //   the next state slot and moves this step's tips to the log
  if (tid == 0) {
    SrDev o = s_st;
    if (commit && sa.record) sr_record(o, sa.hist, sa.histcap);
    sr_solve(g, s_I, o, true, s_sc[0], s_sc[1]);
    s_adv[0] = o.ac2;
    s_adv[1] = o.axc;
    s_adv[2] = o.ayc;
    if (commit) {
      o.cx = s_cx;
      o.cy = s_cy;
      o.ntips = s_found;
      *sa.st_out = o;
    }
  }
  if (commit) {
    const int n = min(s_found, kSrTipCap);
    for (int q = tid; q < n; q += kBfThreads) {
      const int slot = s_base + q;
      if (slot < sa.logcap) {
        sa.log_tips[slot] = sa.step_tips[q];
        sa.log_keys[slot] = sa.step_keys[q];
      }
    }
  }
  __syncthreads();
  if (commit && tid == 0) *sa.step_count = 0;
  // 3 advection coefficients of this thread's cells (Cxy_field_kernel,
  //   symmetryReduction.cu:45-56, negated as at advFDBFECC.cu:30-31) and
  //   their input values; phase 1 reads nothing else another thread wrote
  //   after the barrier above
  const double ac2 = s_adv[0], axc = s_adv[1], ayc = s_adv[2];
#pragma unroll
  for (int it = 0; it < IC; ++it) {
    const int l = tid + it * kBfThreads;
    const int li = 1 + l % W1, lj = 1 + l / W1;
    uint32_t fl = o.fl[it];
    float rx = 0.0f, ry = 0.0f;
    if (fl & 128u) {
      const int gi = i0 + li, gj = j0 + lj;
      double ax = g.hy * (double)gj * ac2 + axc;
      double ay = -g.hx * (double)gi * ac2 + ayc;
      if (MASKED && !(fl & 1u)) ax = ay = 0.0;  // Cxy_field_kernel :50-51
      const double cx = -ax, cy = -ay;          // advFDBFECC.cu:30-31
      rx = (float)(fabs(cx) * g.dtx);
      ry = (float)(fabs(cy) * g.dty);
      fl |= (cx > 0.0 ? 32u : 0u) | (cy > 0.0 ? 64u : 0u);
#pragma unroll
      for (int a = 0; a < NV; ++a) o.q[it][a] = s_q[a][li + LW * lj];
    }
    o.rx[it] = rx;
    o.ry[it] = ry;
    o.fl[it] = fl;
  }
  const float tc = (float)g.tc;
  bf_phase<1, NV, HALO>(tid, tc, s_q, s_f, o, out, i0, j0, nx);
  __syncthreads();
  bf_phase<2, NV, HALO>(tid, tc, s_q, s_f, o, out, i0, j0, nx);
  __syncthreads();
  bf_phase<3, NV, HALO, KEEP>(tid, tc, s_q, s_f, o, out, i0, j0, nx);
  return s_st.count;
}

template <bool MASKED, int NV>
__global__ __launch_bounds__(kBfThreads) void k_bf_fused(SrGeom g, SrArgs sa, CFields in, Fields out,
                                                  const uint8_t* __restrict__ solid) {
  __shared__ float s_q[NV][kBfLN];
  __shared__ float s_f[NV][kBfLN];
  bf_advect<MASKED, NV, kBfHalo, false>(g, sa, in, out, solid, s_q, s_f);
}

// The fused step: the solve and BFECC of this step (as k_bf_fused, on a
// halo of kBfHaloRd, keeping the new state of the tile grown by two cells in
// LDS), then the RD step and the tip test of the next step, which
// k_sr_rdtips would otherwise take in a launch of its own:
//   RD of the new state on the tile grown by one, with cell_update (the
//   arithmetic of k_sr_rdtips, bit for bit), its output of the tile to
//   xo_out and its u to LDS;
//   the tips of the tile between the new state (g_present) and its RD output
//   (g_past) into the next step's tip buffer, stamped with the next count.
template <int MODEL, bool MASKED>
__global__ __launch_bounds__(kBfThreads) void k_bf_rd(const KParams k, SrGeom g, SrArgs sa,
                                                       CFields in, Fields out, Fields xo_out,
                                                       const uint8_t* __restrict__ solid,
                                                       const uint8_t* __restrict__ codes,
                                                       double Uth) {
  constexpr int NV = Model<MODEL>::NV;
  constexpr int H = kBfHaloRd, LW = BfTile<H>::LW, LN = BfTile<H>::LN;
  __shared__ float s_q[NV][LN];
  __shared__ float s_f[NV][LN];
  const long long count =
      bf_advect<MASKED, NV, H, true>(g, sa, in, out, solid, s_q, s_f);
  __syncthreads();
  float* s_ro = s_q[0];  // ue is dead after phase 3: u of the RD output
  const int tid = threadIdx.x;
  const int nx = g.nx, ny = g.ny;
  const int i0 = blockIdx.x * kBfTX - H, j0 = blockIdx.y * kBfTY - H;
  const float* f[3] = {s_f[0], s_f[1], s_f[NV - 1]};
  constexpr int RW = kBfTX + 2, RN = RW * (kBfTY + 2);
  for (int l = tid; l < RN; l += kBfThreads) {
    const int li = H - 1 + l % RW, lj = H - 1 + l / RW;
    const int gi = i0 + li, gj = j0 + lj;
    if (gi < 0 || gi >= nx || gj < 0 || gj >= ny) continue;
    const int L = li + LW * lj;
    const int Lw = (mirror(gi - 1, nx) - i0) + LW * lj, Le = (mirror(gi + 1, nx) - i0) + LW * lj;
    const int Ln = li + LW * (mirror(gj + 1, ny) - j0), Ls = li + LW * (mirror(gj - 1, ny) - j0);
    float cW = 1.f, cE = 1.f, cN = 1.f, cS = 1.f;
    bool sc = true;
    if (MASKED) {  // cell_update_codes
      const uint32_t b = codes[gi + (size_t)nx * gj];
      sc = b != kHole;
      const uint32_t q = sc ? b : 0u;
      cW = (float)(q & 3u);
      cE = (float)((q >> 2) & 3u);
      cN = (float)((q >> 4) & 3u);
      cS = (float)((q >> 6) & 3u);
    }
    float xo[3], inc[3];
    cell_update<MODEL>(k, f, L, Lw, Le, Ln, Ls, cW, cE, cN, cS, sc, xo, inc);
    s_ro[L] = xo[0];
    if (li >= H && li < H + kBfTX && lj >= H && lj < H + kBfTY) {
#pragma unroll
      for (int a = 0; a < NV; ++a) store_through(xo_out.f[a], (uint32_t)(gi + nx * gj), xo[a]);
    }
  }
  __syncthreads();
  // physicalTime = dt*count of the next step (main.cu:885)
  const float t = (float)(g.dt * (double)(count + 1));
  for (int l = tid; l < kBfTX * kBfTY; l += kBfThreads) {
    const int li = H + l % kBfTX, lj = H + l / kBfTX;
    const int gi = i0 + li, gj = j0 + lj;
    if (gi >= nx || gj >= ny || !tip_inside(gi, gj, nx, ny, MASKED, g.r2)) continue;
    // corners past the last column/row are the cell itself (tipTracker.cu:58-63)
    const int L = li + LW * lj;
    const bool ex = gi < nx - 1, ey = gj < ny - 1;
    const int lx = ex ? L + 1 : L, ly = ey ? L + LW : L, lxy = (ex && ey) ? L + LW + 1 : L;
    const float* pr = s_f[0];
    tip_solve(gi, gj, nx, MASKED, Uth, pr[L], pr[lx], pr[lxy], pr[ly], s_ro[L], s_ro[lx],
              s_ro[lxy], s_ro[ly], t, sa.next_tips, sa.next_keys, sa.next_count, kSrTipCap);
  }
}

// ----------------------------------------------------------------- host
}  // namespace yh

struct yh_sr {
  yh::SrDev* st = nullptr;  // two slots: a step reads st[count & 1], writes the other
  long long count = 0;      // steps taken since yh_sr_begin (the host's copy)
  double* partials = nullptr;
  int nparts = 0;
  // tips of a step, two buffers: with two launches per step the launch that
  // commits step n's tips (buffer n & 1) also finds step n+1's (the other)
  yh_tip* step_tips = nullptr;
  int* step_keys = nullptr;
  int* step_count = nullptr;
  float* xo2[3] = {nullptr, nullptr, nullptr};  // RD output of odd steps (two launches)
  bool pending = false;      // the RD output and tips of step `count` are computed ...
  long long epoch = -1;      // ... from the state of this plan state epoch
  int reach = 0;        // largest R with R^2 < r2: the disks span (2R+1)^2 cells
  double* hist = nullptr;
  long long histcap = 0;
  bool begun = false;
  bool use_graph = true;
  // steady steps captured per starting slot: [slot][0] one step,
  // [slot][1] kSrChunk steps (one graph launch for several steps)
  hipGraph_t graph[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
  hipGraphExec_t exec[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
  int graph_cur[2][2] = {{-1, -1}, {-1, -1}};
  hipStream_t cap_stream = nullptr;
};

namespace yh {

static void sr_drop_graphs(yh_sr* sr) {
  for (int q = 0; q < 2; ++q)
    for (int c = 0; c < 2; ++c) {
      if (sr->exec[q][c]) (void)hipGraphExecDestroy(sr->exec[q][c]);
      if (sr->graph[q][c]) (void)hipGraphDestroy(sr->graph[q][c]);
      sr->exec[q][c] = nullptr;
      sr->graph[q][c] = nullptr;
      sr->graph_cur[q][c] = -1;
    }
}

void sr_free(yh_sr* sr) {
  if (!sr) return;
  sr_drop_graphs(sr);
  if (sr->cap_stream) (void)hipStreamDestroy(sr->cap_stream);
  if (sr->st) (void)hipFree(sr->st);
  if (sr->partials) (void)hipFree(sr->partials);
  if (sr->step_tips) (void)hipFree(sr->step_tips);
  if (sr->step_keys) (void)hipFree(sr->step_keys);
  if (sr->step_count) (void)hipFree(sr->step_count);
  for (float* q : sr->xo2)
    if (q) (void)hipFree(q);
  if (sr->hist) (void)hipFree(sr->hist);
  delete sr;
}

static SrGeom sr_geom(const yh_plan* pl) {
  SrGeom g;
  g.nx = pl->p.nx;
  g.ny = pl->p.ny;
  g.r2 = pl->p.tipOffsetX * pl->p.tipOffsetY;
  g.nv = pl->nv;
  g.hx = pl->d.hx;
  g.hy = pl->d.hy;
  g.invdx = pl->d.invdx;
  g.invdy = pl->d.invdy;
  g.dt = pl->d.dt;
  g.tc = pl->p.tc;
  g.dtx = g.dt / g.hx;
  g.dty = g.dt / g.hy;
  return g;
}

// Largest R with R^2 < r2 (0 when r2 <= 1): a disk a^2 + b^2 < r2 lies in
// the (2R+1)^2 box around its centre.
static int disk_reach(int r2) {
  if (r2 <= 1) return 0;
  int R = (int)std::sqrt((double)r2);
  while (R > 0 && (long long)R * R >= r2) --R;
  while ((long long)(R + 1) * (R + 1) < r2) ++R;
  return R;
}

template <int MODEL, bool MASKED>
static void sr_launch_slices(const yh_plan* pl, const SrGeom& g, int R, CFields X, const SrDev* st,
                             const yh_tip* tips, const int* keys, const int* cnt, dim3 sg,
                             dim3 sb, hipStream_t s) {
  hipLaunchKernelGGL((k_sr_slices<MODEL, MASKED>), sg, sb, 0, s, pl->k, g, R, X, pl->codes_d, st,
                     tips, keys, cnt, pl->sr->partials);
}

// Clears a step's tip count with a kernel (capturable, unlike a memset).
__global__ void k_zero_count_sr(int* count) { *count = 0; }

// Enqueue one SR step on stream s, reading state slot `slot` (first: the
// count-0 step with its preliminary slices + solve).
// Three launches (YH_OPT_SR_LAUNCHES 3): RD + tips; slices; BFECC with the
// solve.  Two (the default): slices; BFECC with the solve, then RD + tips of
// the next step (k_bf_rd); the RD output and tips of this step come from the
// previous step's second launch, or, with need_rd (the first step, or a
// state changed since), from k_sr_rdtips first.  The RD output alternates
// between buf[cur ^ 1] (even steps) and xo2 (odd steps), because a
// workgroup's halo reads neighbours' RD output while they write the next.
static int sr_enqueue(yh_plan* pl, bool first, bool need_rd, int slot, hipStream_t s) {
  yh_sr* sr = pl->sr;
  const SrGeom g = sr_geom(pl);
  const bool masked = pl->p.solidSwitch != 0;
  const bool two = pl->sr_launches == 2;
  const int b = pl->cur;
  CFields X{{pl->buf[b][0], pl->buf[b][1], pl->buf[b][2]}};
  Fields other{{pl->buf[b ^ 1][0], pl->buf[b ^ 1][1], pl->buf[b ^ 1][2]}};
  Fields odd{{sr->xo2[0], sr->xo2[1], sr->xo2[2]}};
  const Fields Xo = (two && slot == 1) ? odd : other;
  const Fields Xo_next = slot == 1 ? other : odd;
  CFields Xoc{{Xo.f[0], Xo.f[1], Xo.f[2]}};
  Fields Xw{{pl->buf[b][0], pl->buf[b][1], pl->buf[b][2]}};
  SrDev* st_in = sr->st + slot;
  SrDev* st_out = sr->st + (slot ^ 1);
  const int tb = two ? slot : 0;  // this step's tip buffer
  yh_tip* tips = sr->step_tips + (size_t)tb * kSrTipCap;
  int* keys = sr->step_keys + (size_t)tb * kSrTipCap;
  int* cnt = sr->step_count + tb;
  const int R = sr->reach, box = 2 * R + 1;
  const int model = pl->p.model;
  // 1 RD + tips
  if (!two || need_rd) {
    if (two && !first) hipLaunchKernelGGL(k_zero_count_sr, dim3(1), dim3(1), 0, s, cnt);
    dim3 rb(kRtBX, kRtBY), rg((g.nx + kRtOX - 1) / kRtOX, (g.ny + kRtOY - 1) / kRtOY);
    const uint8_t* codes = pl->codes_d;
    const double Uth = pl->p.Uth;
    if (model == M2V) {
      if (masked) hipLaunchKernelGGL((k_sr_rdtips<M2V, true>), rg, rb, 0, s, pl->k, g, X, Xo, codes, Uth, st_in, tips, keys, cnt);
      else hipLaunchKernelGGL((k_sr_rdtips<M2V, false>), rg, rb, 0, s, pl->k, g, X, Xo, codes, Uth, st_in, tips, keys, cnt);
    } else {
      if (masked) hipLaunchKernelGGL((k_sr_rdtips<MFK, true>), rg, rb, 0, s, pl->k, g, X, Xo, codes, Uth, st_in, tips, keys, cnt);
      else hipLaunchKernelGGL((k_sr_rdtips<MFK, false>), rg, rb, 0, s, pl->k, g, X, Xo, codes, Uth, st_in, tips, keys, cnt);
    }
    YH_HIP(hipGetLastError());
  }
  // 2 slices + inner products over the moving disk's box (at count 0 first
  //   the preliminary pass and its solve, in place)
  dim3 sb(kSliceBX, kSliceBY), sg((box + kSliceBX - 1) / kSliceBX, (box + kSliceBY - 1) / kSliceBY);
  for (int pass = first ? 0 : 1; pass < 2; ++pass) {
    if (model == M2V) {
      if (masked) sr_launch_slices<M2V, true>(pl, g, R, X, st_in, tips, keys, cnt, sg, sb, s);
      else sr_launch_slices<M2V, false>(pl, g, R, X, st_in, tips, keys, cnt, sg, sb, s);
    } else {
      if (masked) sr_launch_slices<MFK, true>(pl, g, R, X, st_in, tips, keys, cnt, sg, sb, s);
      else sr_launch_slices<MFK, false>(pl, g, R, X, st_in, tips, keys, cnt, sg, sb, s);
    }
    YH_HIP(hipGetLastError());
    if (pass == 0) {
      hipLaunchKernelGGL(k_sr_solve_pre, dim3(1), dim3(768), 0, s, g, st_in,
                         (const double*)sr->partials, sr->nparts, sr->hist, sr->histcap);
      YH_HIP(hipGetLastError());
    }
  }
  // 3 solve + BFECC of Xo back into X, the three phases in one launch (with
  //   two launches, then RD + tips of the next step)
  SrArgs sa;
  sa.st_in = st_in;
  sa.st_out = st_out;
  sa.partials = sr->partials;
  sa.nparts = sr->nparts;
  sa.record = first ? 0 : 1;  // at count 0 the preliminary solve recorded
  sa.hist = sr->hist;
  sa.histcap = sr->histcap;
  sa.step_tips = tips;
  sa.step_keys = keys;
  sa.step_count = cnt;
  sa.log_tips = pl->tips_d;
  sa.log_keys = pl->keys_d;
  sa.log_count = pl->tipcount_d;
  sa.logcap = pl->tipcap;
  sa.next_tips = sr->step_tips + (size_t)(tb ^ 1) * kSrTipCap;
  sa.next_keys = sr->step_keys + (size_t)(tb ^ 1) * kSrTipCap;
  sa.next_count = sr->step_count + (tb ^ 1);
  dim3 bg((g.nx + kBfTX - 1) / kBfTX, (g.ny + kBfTY - 1) / kBfTY);
  if (two) {
    const uint8_t* codes = pl->codes_d;
    const double Uth = pl->p.Uth;
    if (model == M2V) {
      if (masked) hipLaunchKernelGGL((k_bf_rd<M2V, true>), bg, dim3(kBfThreads), 0, s, pl->k, g, sa, Xoc, Xw, Xo_next, pl->solid_d, codes, Uth);
      else hipLaunchKernelGGL((k_bf_rd<M2V, false>), bg, dim3(kBfThreads), 0, s, pl->k, g, sa, Xoc, Xw, Xo_next, pl->solid_d, codes, Uth);
    } else {
      if (masked) hipLaunchKernelGGL((k_bf_rd<MFK, true>), bg, dim3(kBfThreads), 0, s, pl->k, g, sa, Xoc, Xw, Xo_next, pl->solid_d, codes, Uth);
      else hipLaunchKernelGGL((k_bf_rd<MFK, false>), bg, dim3(kBfThreads), 0, s, pl->k, g, sa, Xoc, Xw, Xo_next, pl->solid_d, codes, Uth);
    }
  } else if (g.nv == 3) {
    if (masked) hipLaunchKernelGGL((k_bf_fused<true, 3>), bg, dim3(kBfThreads), 0, s, g, sa, Xoc, Xw, pl->solid_d);
    else hipLaunchKernelGGL((k_bf_fused<false, 3>), bg, dim3(kBfThreads), 0, s, g, sa, Xoc, Xw, pl->solid_d);
  } else {
    if (masked) hipLaunchKernelGGL((k_bf_fused<true, 2>), bg, dim3(kBfThreads), 0, s, g, sa, Xoc, Xw, pl->solid_d);
    else hipLaunchKernelGGL((k_bf_fused<false, 2>), bg, dim3(kBfThreads), 0, s, g, sa, Xoc, Xw, pl->solid_d);
  }
  YH_HIP(hipGetLastError());
  return YH_OK;
}

}  // namespace yh

using namespace yh;

extern "C" {

int yh_sr_begin(yh_plan* pl, float tip0x, float tip0y, void* stream) {
  if (!pl) return YH_EINVAL;
  if (!pl->p.reduceSym) return YH_ESTATE;  // dt must be the halved SR step (main.cu:148)
  if (!standard_branch(pl->p)) return YH_EUNSUPPORTED;
  if (pl->p.solidSwitch && !pl->mask_set) return YH_ESTATE;
  hipStream_t s = (hipStream_t)stream;
  if (!pl->sr) {
    yh_sr* sr = new (std::nothrow) yh_sr();
    if (!sr) return YH_ENOMEM;
    pl->sr = sr;
    sr->reach = disk_reach(pl->p.tipOffsetX * pl->p.tipOffsetY);
    const int box = 2 * sr->reach + 1;
    const int gx = (box + kSliceBX - 1) / kSliceBX, gy = (box + kSliceBY - 1) / kSliceBY;
    sr->nparts = gx * gy;
  }
  yh_sr* sr = pl->sr;
  // each buffer on its own, so a call after a failed allocation completes it
  if (!sr->st) YH_HIP(hipMalloc(&sr->st, 2 * sizeof(SrDev)));
  if (!sr->partials) YH_HIP(hipMalloc(&sr->partials, sizeof(double) * 12 * sr->nparts));
  if (!sr->step_tips) YH_HIP(hipMalloc(&sr->step_tips, 2 * sizeof(yh_tip) * kSrTipCap));
  if (!sr->step_keys) YH_HIP(hipMalloc(&sr->step_keys, 2 * sizeof(int) * kSrTipCap));
  if (!sr->step_count) YH_HIP(hipMalloc(&sr->step_count, 2 * sizeof(int)));
  for (int a = 0; a < pl->nv; ++a)
    if (!sr->xo2[a]) YH_HIP(hipMalloc(&sr->xo2[a], pl->n * sizeof(float)));
  SrDev h[2];
  std::memset(h, 0, sizeof h);
  h[0].tip0x = h[1].tip0x = tip0x;
  h[0].tip0y = h[1].tip0y = tip0y;
  YH_HIP(hipMemcpyAsync(sr->st, h, sizeof h, hipMemcpyHostToDevice, s));
  YH_HIP(hipMemsetAsync(sr->step_count, 0, 2 * sizeof(int), s));
  YH_HIP(hipStreamSynchronize(s));
  sr->count = 0;
  sr->begun = true;
  sr->pending = false;
  sr_drop_graphs(sr);
  pl->have_prev = false;
  return YH_OK;
}

int yh_sr_set_graph(yh_plan* pl, int on) {
  if (!pl || !pl->sr) return YH_ESTATE;
  pl->sr->use_graph = on != 0;
  return YH_OK;
}

int yh_sr_step(yh_plan* pl, int nsteps, void* stream, double* c_phi_out) {
  if (!pl || nsteps < 0) return YH_EINVAL;
  if (!pl->sr || !pl->sr->begun) return YH_ESTATE;
  if (pl->stim_on) return YH_EUNSUPPORTED;  // the SR step runs the standard branch
  yh_sr* sr = pl->sr;
  hipStream_t s = (hipStream_t)stream;
  const long long count0 = sr->count;
  // history buffer for (c, phi) of each step
  if (c_phi_out && count0 + nsteps > sr->histcap) {
    long long cap = std::max<long long>(count0 + nsteps, 2 * sr->histcap);
    double* nh = nullptr;
    YH_HIP(hipMalloc(&nh, sizeof(double) * 6 * cap));
    if (sr->hist) {
      YH_HIP(hipMemcpyAsync(nh, sr->hist, sizeof(double) * 6 * sr->histcap, hipMemcpyDeviceToDevice, s));
      YH_HIP(hipStreamSynchronize(s));
      (void)hipFree(sr->hist);
    }
    sr->hist = nh;
    sr->histcap = cap;
    sr_drop_graphs(sr);  // the captured steps hold the old history pointer
  }
  int rc = YH_OK;
  const bool two = pl->sr_launches == 2;
  // with two launches a step takes its RD output and tips from the step
  // before; after yh_sr_begin, three-launch steps or a change of the state
  // (plain steps, yh_set_state, yh_set_mask) they are computed first
  bool need_rd = two && !(sr->pending && sr->epoch == pl->state_epoch);
  for (int q = 0; q < nsteps; ++q) {
    const long long cnt = count0 + q;
    const int slot = (int)(cnt & 1);
    if (cnt == 0 || need_rd || !sr->use_graph) {
      rc = sr_enqueue(pl, cnt == 0, need_rd || cnt == 0, slot, s);
      if (rc) return rc;
      sr->count = cnt + 1;
      need_rd = false;
      continue;
    }
    // kSrChunk steps per graph launch while that many remain, else one
    const int chunk = (nsteps - q >= kSrChunk) ? 1 : 0;
    const int len = chunk ? kSrChunk : 1;
    const int gkey = pl->cur | (two ? 2 : 0);
    if (!sr->exec[slot][chunk] || sr->graph_cur[slot][chunk] != gkey) {
      if (sr->exec[slot][chunk]) {
        (void)hipGraphExecDestroy(sr->exec[slot][chunk]);
        (void)hipGraphDestroy(sr->graph[slot][chunk]);
        sr->exec[slot][chunk] = nullptr;
        sr->graph[slot][chunk] = nullptr;
      }
      if (!sr->cap_stream) YH_HIP(hipStreamCreateWithFlags(&sr->cap_stream, hipStreamNonBlocking));
      YH_HIP(hipStreamSynchronize(s));
      YH_HIP(hipStreamBeginCapture(sr->cap_stream, hipStreamCaptureModeThreadLocal));
      for (int k = 0; k < len && !rc; ++k)
        rc = sr_enqueue(pl, false, false, (slot + k) & 1, sr->cap_stream);
      hipGraph_t gph = nullptr;
      hipError_t e = hipStreamEndCapture(sr->cap_stream, &gph);
      if (rc) {
        if (gph) (void)hipGraphDestroy(gph);
        return rc;
      }
      YH_HIP(e);
      sr->graph[slot][chunk] = gph;
      YH_HIP(hipGraphInstantiate(&sr->exec[slot][chunk], sr->graph[slot][chunk], nullptr, nullptr,
                                 0));
      sr->graph_cur[slot][chunk] = gkey;
    }
    YH_HIP(hipGraphLaunch(sr->exec[slot][chunk], s));
    sr->count = cnt + len;
    q += len - 1;
  }
  pl->count += nsteps;
  pl->have_prev = false;
  if (nsteps > 0) {
    sr->pending = two;
    sr->epoch = pl->state_epoch;
  }
  if (c_phi_out && nsteps > 0) {
    YH_HIP(hipMemcpyAsync(c_phi_out, sr->hist + (size_t)count0 * 6, sizeof(double) * 6 * nsteps,
                          hipMemcpyDeviceToHost, s));
    YH_HIP(hipStreamSynchronize(s));
  }
  return YH_OK;
}

int yh_sr_info_get(yh_plan* pl, yh_sr_info* out) {
  if (!pl || !out) return YH_EINVAL;
  if (!pl->sr || !pl->sr->begun) return YH_ESTATE;
  SrDev h;
  YH_HIP(hipDeviceSynchronize());
  YH_HIP(hipMemcpy(&h, pl->sr->st + (pl->sr->count & 1), sizeof h, hipMemcpyDeviceToHost));
  for (int q = 0; q < 3; ++q) {
    out->c[q] = h.c[q];
    out->phi[q] = h.phi[q];
  }
  for (int q = 0; q < 12; ++q) out->I[q] = h.I[q];
  out->cx = h.cx;
  out->cy = h.cy;
  out->count = h.count;
  out->status = h.status;
  out->ntips = h.ntips;
  return (h.status & 1) ? YH_ESINGULAR : YH_OK;
}

int yh_solve_matrix(const double phi[3], const double I[12], double c[3]) {
  if (!phi || !I || !c) return YH_EINVAL;
  // symmetryReduction.cu:386-418 (host), unpivoted elimination
  const double cp = std::cos(phi[2]), sp = std::sin(phi[2]);
  const double a1 = I[0] * cp + I[1] * sp, a2 = I[1] * cp - I[0] * sp, a3 = I[2];
  const double b1 = I[3] * cp + I[4] * sp, b2 = I[4] * cp - I[3] * sp, b3 = I[5];
  const double C1 = I[6] * cp + I[7] * sp, C2 = I[7] * cp - I[6] * sp, C3 = I[8];
  const double d1 = I[9], d2 = I[10], d3 = I[11];
  const double b2p = a1 / b1 * b2 - a2, b3p = a1 / b1 * b3 - a3, d2p = a1 / b1 * d2 - d1;
  const double c2p = a1 / C1 * C2 - a2, c3p = a1 / C1 * C3 - a3, d3p = a1 / C1 * d3 - d1;
  const double c3pp = b2p / c2p * c3p - b3p, d3pp = b2p / c2p * d3p - d2p;
  c[2] = d3pp / c3pp;
  c[1] = (d2p - b3p * c[2]) / b2p;
  c[0] = (d1 - a2 * c[1] - a3 * c[2]) / a1;
  return (std::isfinite(c[0]) && std::isfinite(c[1]) && std::isfinite(c[2])) ? YH_OK : YH_ESINGULAR;
}

}  // extern "C"
