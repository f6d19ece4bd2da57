// sr.hip -- symmetry-reduction (co-moving frame) step on the device.
//
// One SR step follows display() (main.cu:894-953):
//   1 RD            X -> Xo                         (reactionDiffusion_kernel)
//   2 tips          g_past = Xo.u, g_present = X.u   (tip_wrapper, main.cu:900)
//   3 k_sr_center   disk centre = last tip of this step, tipx0 at count 0
//   4 k_sr_slices   slices, template and velTan in FP64, the 12 disk inner
//                   products as per-block partial sums (slice_kernel +
//                   trapz_kernel fused: no slice arrays, no atomics)
//   5 k_sr_solve    one workgroup: fixed-order reduction of the partials,
//                   solve_matrix (symmetryReduction.cu:386-418), advection
//                   parameters for Cxy_field (:45-56), phi += dt*c, tip log
//   6 k_bf_phase1/2/3  BFECC of Xo into X (advFDBFECC.cu:18-165), one kernel
//                   per dependent phase so no neighbour read races
// At count 0 steps 4-5 run twice (main.cu:910-921).  Nothing leaves the
// device; a steady step is captured once into a hipGraph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

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
};

constexpr int kSrTipCap = 4096;
constexpr int kSliceBX = 64, kSliceBY = 4;

__device__ __forceinline__ bool in_disk(const SrGeom& g, int i, int j, int cx, int cy) {
  const int a = i - g.nx / 2 - cx, b = j - g.ny / 2 - cy;
  return a * a + b * b < g.r2;
}

// ---------------------------------------------------------------- centre
__global__ void k_sr_center(SrGeom g, SrDev* st, const yh_tip* tips, const int* keys,
                            const int* count) {
  __shared__ int best_key[256];
  __shared__ int best_slot[256];
  const int t = threadIdx.x;
  const int n = min(*count, kSrTipCap);
  int bk = -1, bs = -1;
  for (int q = t; q < n; q += blockDim.x)
    if (keys[q] > bk) {
      bk = keys[q];
      bs = q;
    }
  best_key[t] = bk;
  best_slot[t] = bs;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (t < w && best_key[t + w] > best_key[t]) {
      best_key[t] = best_key[t + w];
      best_slot[t] = best_slot[t + w];
    }
    __syncthreads();
  }
  if (t == 0) {
    st->ntips = *count;
    if (st->count == 0) {
      // tipx0_d at count 0 (symmetryReduction.cu:98-100)
      st->cx = __float2int_rn(st->tip0x - (float)(g.nx / 2));
      st->cy = __float2int_rn(st->tip0y - (float)(g.ny / 2));
    } else if (best_slot[0] >= 0) {
      // last tip of this step (symmetryReduction.cu:102-103); none: keep the centre
      const yh_tip tp = tips[best_slot[0]];
      st->cx = __float2int_rn(tp.x - (float)(g.nx / 2));
      st->cy = __float2int_rn(tp.y - (float)(g.ny / 2));
    }
  }
}

// ---------------------------------------------------------------- slices
template <int MODEL, bool MASKED>
__global__ __launch_bounds__(kSliceBX* kSliceBY) void k_sr_slices(
    const KParams k, const SrGeom g, CFields X, const uint8_t* __restrict__ codes,
    const SrDev* __restrict__ st, double* __restrict__ partials) {
  using M = Model<MODEL>;
  constexpr int NV = M::NV;
  const int i = blockIdx.x * kSliceBX + threadIdx.x;
  const int j = blockIdx.y * kSliceBY + threadIdx.y;
  double acc[12];
#pragma unroll
  for (int q = 0; q < 12; ++q) acc[q] = 0.0;
  const int cx = st->cx, cy = st->cy;
  if (i < g.nx && j < g.ny && in_disk(g, i, j, cx, cy)) {
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
#pragma unroll
  for (int q = 0; q < 12; ++q) {
    double v = acc[q];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if (lane == 0) red[wv][q] = v;
  }
  __syncthreads();
  if (tid < 12) {
    double s = 0.0;
    for (int w = 0; w < kSliceBX * kSliceBY / 64; ++w) s += red[w][tid];
    partials[(size_t)(blockIdx.y * gridDim.x + blockIdx.x) * 12 + tid] = s;
  }
}

// ----------------------------------------------------------------- solve
__global__ __launch_bounds__(256) void k_sr_solve(SrGeom g, SrDev* st, const double* partials,
                                                  int nparts, int pre, int record,
                                                  yh_tip* step_tips,
                                                  int* step_keys, int* step_count,
                                                  yh_tip* log_tips, int* log_keys,
                                                  int* log_count, int logcap, double* hist,
                                                  long long histcap) {
  __shared__ double red[256][13];
  const int t = threadIdx.x;
  for (int q = 0; q < 12; ++q) {
    double s = 0.0;
    for (int b = t; b < nparts; b += 256) s += partials[(size_t)b * 12 + q];
    red[t][q] = s;
  }
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w)
      for (int q = 0; q < 12; ++q) red[t][q] += red[t + w][q];
    __syncthreads();
  }
  if (t == 0) {
    double I[12];
    for (int q = 0; q < 12; ++q) {
      I[q] = red[0][q] * (g.hx * g.hy);
      st->I[q] = I[q];
    }
    if (record && hist && st->count < histcap) {
      // clist/philist push (main.cu:902-903): c and phi before this step's
      // first solve (at count 0 that is the preliminary solve, main.cu:910-914)
      double* h = hist + (size_t)st->count * 6;
      for (int q = 0; q < 3; ++q) {
        h[q] = st->c[q];
        h[3 + q] = st->phi[q];
      }
    }
    // solve_matrix, symmetryReduction.cu:386-418
    const double cp = cos(st->phi[2]), sp = sin(st->phi[2]);
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
    if (!(isfinite(x1) && isfinite(x2) && isfinite(x3))) st->status |= 1;
    st->c[0] = x1;
    st->c[1] = x2;
    st->c[2] = x3;
    // Cxy_field_kernel (symmetryReduction.cu:45-56) with the new c, current phi
    const double ph = st->phi[2];
    st->aphi = ph;
    st->ac2 = x3;
    st->axc = -x1 * cos(ph) + x2 * sin(ph);
    st->ayc = -x1 * sin(ph) - x2 * cos(ph);
    if (!pre) {
      // phi += dt*c (Euler: reduceSymStart stays set, main.cu:935-941)
      for (int q = 0; q < 3; ++q) st->phi[q] = st->phi[q] + st->c[q] * g.dt;
    }
  }
  if (!pre) {
    // move this step's tips to the plan's log, then reset the step buffer
    __syncthreads();
    const int n = min(*step_count, kSrTipCap);
    __shared__ int base;
    if (t == 0) base = atomicAdd(log_count, *step_count);
    __syncthreads();
    for (int q = t; q < n; q += 256) {
      const int slot = base + q;
      if (slot < logcap) {
        log_tips[slot] = step_tips[q];
        log_keys[slot] = step_keys[q];
      }
    }
    __syncthreads();
    if (t == 0) {
      *step_count = 0;
      st->count += 1;
    }
  }
}

// ----------------------------------------------------------------- BFECC
struct BfIdx {
  int W, E, S, N, iW, W2, iE, E2, iS, S2, iN, N2;
  bool sc;
};

// advFDBFECC.cu:44-66 (masked) and :112-126 (square); a neighbour outside
// the domain counts as a hole, and each cell tests the mask of the neighbour
// it actually uses.
template <bool MASKED>
__device__ __forceinline__ BfIdx bf_index(int i, int j, int nx, int ny, const uint8_t* solid) {
  const int c = i + nx * j;
  bool sc = true, sw, se, ss, sn;
  if (MASKED) {
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
  BfIdx r;
  r.sc = sc;
  r.W = sc ? (sw ? c - 1 : c + 1) : c;
  r.E = sc ? (se ? c + 1 : c - 1) : c;
  r.S = sc ? (ss ? c - nx : c + nx) : c;
  r.N = sc ? (sn ? c + nx : c - nx) : c;
  r.iW = sc ? (sw ? c : c + 1) : c;
  r.W2 = sc ? (sw ? c - 1 : c) : c;
  r.iE = sc ? (se ? c : c - 1) : c;
  r.E2 = sc ? (se ? c + 1 : c) : c;
  r.iS = sc ? (ss ? c : c + nx) : c;
  r.S2 = sc ? (ss ? c - nx : c) : c;
  r.iN = sc ? (sn ? c : c - nx) : c;
  r.N2 = sc ? (sn ? c + nx : c) : c;
  return r;
}

struct BfCoef {
  float Rx, Ry;
  bool px, py;
};

template <bool MASKED>
__device__ __forceinline__ BfCoef bf_coef(const SrGeom& g, const SrDev* st, int i, int j,
                                          const uint8_t* solid) {
  double ax = g.hy * (double)j * st->ac2 + st->axc;
  double ay = -g.hx * (double)i * st->ac2 + st->ayc;
  if (MASKED && !solid[i + g.nx * j]) ax = ay = 0.0;  // Cxy_field_kernel :50-51
  const double cx = -ax, cy = -ay;                      // advFDBFECC.cu:30-31
  BfCoef r;
  r.Rx = (float)(fabs(cx) * g.dt / g.hx);
  r.Ry = (float)(fabs(cy) * g.dt / g.hy);
  r.px = cx > 0.0;
  r.py = cy > 0.0;
  return r;
}

template <bool MASKED>
__global__ __launch_bounds__(256) void k_bf_phase(int phase, SrGeom g, const SrDev* st,
                                                  CFields in, Fields uf, Fields ue, Fields out,
                                                  const uint8_t* __restrict__ solid) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= g.nx || j >= g.ny) return;
  const int k = i + g.nx * j;
  const BfIdx x = bf_index<MASKED>(i, j, g.nx, g.ny, solid);
  const BfCoef cf = bf_coef<MASKED>(g, st, i, j, solid);
  const float tc = (float)g.tc;
  for (int a = 0; a < g.nv; ++a) {
    if (phase == 1) {  // forward step (advFDBFECC.cu:128-131)
      const float* q = in.f[a];
      const float FDx = cf.px ? q[k] - q[x.W] : q[k] - q[x.E];
      const float FDy = cf.py ? q[k] - q[x.S] : q[k] - q[x.N];
      uf.f[a][k] = q[k] - tc * (cf.Rx * FDx + cf.Ry * FDy);
    } else if (phase == 2) {  // backward step and error correction (:134-139)
      const float* f = uf.f[a];
      const float FDx = cf.px ? f[x.iE] - f[x.E2] : f[x.iW] - f[x.W2];
      const float FDy = cf.py ? f[x.iN] - f[x.N2] : f[x.iS] - f[x.S2];
      const float ub = f[k] - tc * (cf.Rx * FDx + cf.Ry * FDy);
      const float g0 = in.f[a][k];
      ue.f[a][k] = g0 - 0.5f * (ub - g0);
    } else {  // forward step from the corrected field (:141-144)
      const float* e = ue.f[a];
      const float FDx = cf.px ? e[k] - e[x.W] : e[k] - e[x.E];
      const float FDy = cf.py ? e[k] - e[x.S] : e[k] - e[x.N];
      const float o = e[k] - tc * (cf.Rx * FDx + cf.Ry * FDy);
      out.f[a][k] = x.sc ? o : 0.0f;
    }
  }
}

// ----------------------------------------------------------------- host
}  // namespace yh

struct yh_sr {
  yh::SrDev* st = nullptr;
  double* partials = nullptr;
  int nparts = 0;
  yh_tip* step_tips = nullptr;
  int* step_keys = nullptr;
  int* step_count = nullptr;
  float* uf[3] = {nullptr, nullptr, nullptr};
  float* ue[3] = {nullptr, nullptr, nullptr};
  double* hist = nullptr;
  long long histcap = 0;
  bool begun = false;
  bool use_graph = true;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t cap_stream = nullptr;
  int graph_cur = -1;
  double* hist_graph = nullptr;
};

namespace yh {

void sr_free(yh_sr* sr) {
  if (!sr) return;
  if (sr->exec) (void)hipGraphExecDestroy(sr->exec);
  if (sr->graph) (void)hipGraphDestroy(sr->graph);
  if (sr->cap_stream) (void)hipStreamDestroy(sr->cap_stream);
  if (sr->st) (void)hipFree(sr->st);
  if (sr->partials) (void)hipFree(sr->partials);
  if (sr->step_tips) (void)hipFree(sr->step_tips);
  if (sr->step_keys) (void)hipFree(sr->step_keys);
  if (sr->step_count) (void)hipFree(sr->step_count);
  for (int a = 0; a < 3; ++a) {
    if (sr->uf[a]) (void)hipFree(sr->uf[a]);
    if (sr->ue[a]) (void)hipFree(sr->ue[a]);
  }
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
  return g;
}

hipError_t launch_tips_dev(const KParams& k, int masked_disk, int r2, double Uth,
                           const float* past, const float* present, double dt,
                           const long long* d_count, yh_tip* tips, int* keys, int* count, int cap,
                           hipStream_t s);

// Enqueue one SR step (pre: the count-0 double solve) on stream s.
static int sr_enqueue(yh_plan* pl, bool first, hipStream_t s) {
  yh_sr* sr = pl->sr;
  const SrGeom g = sr_geom(pl);
  const bool masked = pl->p.solidSwitch != 0;
  const int b = pl->cur;
  CFields X{{pl->buf[b][0], pl->buf[b][1], pl->buf[b][2]}};
  Fields Xo{{pl->buf[b ^ 1][0], pl->buf[b ^ 1][1], pl->buf[b ^ 1][2]}};
  CFields Xoc{{Xo.f[0], Xo.f[1], Xo.f[2]}};
  Fields Xw{{pl->buf[b][0], pl->buf[b][1], pl->buf[b][2]}};
  // 1 RD
  YH_HIP(launch_direct_codes(pl->p.model, pl->k, X, Xo, masked ? pl->codes_d : nullptr, s));
  // 2 tips: g_past = RD output, g_present = old state (main.cu:900, tipTracker.cu:586)
  YH_HIP(launch_tips_dev(pl->k, pl->p.solidSwitch, g.r2, pl->p.Uth, Xo.f[0], X.f[0], pl->d.dt,
                         &sr->st->count, sr->step_tips, sr->step_keys, sr->step_count, kSrTipCap,
                         s));
  // 3 centre
  hipLaunchKernelGGL(k_sr_center, dim3(1), dim3(256), 0, s, g, sr->st, sr->step_tips,
                     sr->step_keys, sr->step_count);
  YH_HIP(hipGetLastError());
  dim3 sb(kSliceBX, kSliceBY), sg((g.nx + kSliceBX - 1) / kSliceBX, (g.ny + kSliceBY - 1) / kSliceBY);
  for (int pass = first ? 0 : 1; pass < 2; ++pass) {
    const int pre = pass == 0;
    const int record = pass == (first ? 0 : 1);
    // 4 slices + inner products
    if (pl->p.model == M2V) {
      if (masked) hipLaunchKernelGGL((k_sr_slices<M2V, true>), sg, sb, 0, s, pl->k, g, X, pl->codes_d, sr->st, sr->partials);
      else hipLaunchKernelGGL((k_sr_slices<M2V, false>), sg, sb, 0, s, pl->k, g, X, pl->codes_d, sr->st, sr->partials);
    } else {
      if (masked) hipLaunchKernelGGL((k_sr_slices<MFK, true>), sg, sb, 0, s, pl->k, g, X, pl->codes_d, sr->st, sr->partials);
      else hipLaunchKernelGGL((k_sr_slices<MFK, false>), sg, sb, 0, s, pl->k, g, X, pl->codes_d, sr->st, sr->partials);
    }
    YH_HIP(hipGetLastError());
    // 5 solve
    hipLaunchKernelGGL(k_sr_solve, dim3(1), dim3(256), 0, s, g, sr->st, sr->partials, sr->nparts,
                       pre, record, sr->step_tips, sr->step_keys, sr->step_count, pl->tips_d,
                       pl->keys_d,
                       pl->tipcount_d, pl->tipcap, sr->hist, sr->histcap);
    YH_HIP(hipGetLastError());
  }
  // 6 BFECC of Xo back into X
  dim3 bb(64, 4), bg((g.nx + 63) / 64, (g.ny + 3) / 4);
  Fields uf{{sr->uf[0], sr->uf[1], sr->uf[2]}}, ue{{sr->ue[0], sr->ue[1], sr->ue[2]}};
  for (int ph = 1; ph <= 3; ++ph) {
    if (masked) hipLaunchKernelGGL((k_bf_phase<true>), bg, bb, 0, s, ph, g, sr->st, Xoc, uf, ue, Xw, pl->solid_d);
    else hipLaunchKernelGGL((k_bf_phase<false>), bg, bb, 0, s, ph, g, sr->st, Xoc, uf, ue, Xw, pl->solid_d);
    YH_HIP(hipGetLastError());
  }
  return YH_OK;
}

}  // namespace yh

using namespace yh;

extern "C" {

int yh_sr_begin(yh_plan* pl, float tip0x, float tip0y, void* stream) {
  if (!pl) return YH_EINVAL;
  if (!pl->p.reduceSym) return YH_ESTATE;  // dt must be the halved SR step (main.cu:148)
  if (pl->p.solidSwitch && !pl->mask_set) return YH_ESTATE;
  hipStream_t s = (hipStream_t)stream;
  if (!pl->sr) {
    yh_sr* sr = new (std::nothrow) yh_sr();
    if (!sr) return YH_ENOMEM;
    pl->sr = sr;
    const int gx = (pl->p.nx + kSliceBX - 1) / kSliceBX, gy = (pl->p.ny + kSliceBY - 1) / kSliceBY;
    sr->nparts = gx * gy;
    YH_HIP(hipMalloc(&sr->st, sizeof(SrDev)));
    YH_HIP(hipMalloc(&sr->partials, sizeof(double) * 12 * sr->nparts));
    YH_HIP(hipMalloc(&sr->step_tips, sizeof(yh_tip) * kSrTipCap));
    YH_HIP(hipMalloc(&sr->step_keys, sizeof(int) * kSrTipCap));
    YH_HIP(hipMalloc(&sr->step_count, sizeof(int)));
    for (int a = 0; a < pl->nv; ++a) {
      YH_HIP(hipMalloc(&sr->uf[a], sizeof(float) * pl->n));
      YH_HIP(hipMalloc(&sr->ue[a], sizeof(float) * pl->n));
    }
  }
  yh_sr* sr = pl->sr;
  SrDev h;
  std::memset(&h, 0, sizeof h);
  h.tip0x = tip0x;
  h.tip0y = tip0y;
  YH_HIP(hipMemcpyAsync(sr->st, &h, sizeof h, hipMemcpyHostToDevice, s));
  YH_HIP(hipMemsetAsync(sr->step_count, 0, sizeof(int), s));
  YH_HIP(hipStreamSynchronize(s));
  sr->begun = true;
  if (sr->exec) {
    (void)hipGraphExecDestroy(sr->exec);
    sr->exec = nullptr;
  }
  if (sr->graph) {
    (void)hipGraphDestroy(sr->graph);
    sr->graph = nullptr;
  }
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
  yh_sr* sr = pl->sr;
  hipStream_t s = (hipStream_t)stream;
  // history buffer for (c, phi) of each step
  long long count0 = 0;
  {
    SrDev h;
    YH_HIP(hipMemcpyAsync(&h, sr->st, sizeof h, hipMemcpyDeviceToHost, s));
    YH_HIP(hipStreamSynchronize(s));
    count0 = h.count;
  }
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
    if (sr->exec) {  // the captured step holds the old history pointer
      (void)hipGraphExecDestroy(sr->exec);
      sr->exec = nullptr;
      (void)hipGraphDestroy(sr->graph);
      sr->graph = nullptr;
    }
  }
  int rc = YH_OK;
  for (int q = 0; q < nsteps; ++q) {
    const bool first = (count0 + q) == 0;
    if (first || !sr->use_graph) {
      rc = sr_enqueue(pl, first, s);
      if (rc) return rc;
      continue;
    }
    if (!sr->exec || sr->graph_cur != pl->cur) {
      if (sr->exec) {
        (void)hipGraphExecDestroy(sr->exec);
        (void)hipGraphDestroy(sr->graph);
        sr->exec = nullptr;
        sr->graph = nullptr;
      }
      if (!sr->cap_stream) YH_HIP(hipStreamCreateWithFlags(&sr->cap_stream, hipStreamNonBlocking));
      YH_HIP(hipStreamSynchronize(s));
      YH_HIP(hipStreamBeginCapture(sr->cap_stream, hipStreamCaptureModeThreadLocal));
      rc = sr_enqueue(pl, false, sr->cap_stream);
      hipGraph_t gph = nullptr;
      hipError_t e = hipStreamEndCapture(sr->cap_stream, &gph);
      if (rc) {
        if (gph) (void)hipGraphDestroy(gph);
        return rc;
      }
      YH_HIP(e);
      sr->graph = gph;
      YH_HIP(hipGraphInstantiate(&sr->exec, sr->graph, nullptr, nullptr, 0));
      sr->graph_cur = pl->cur;
    }
    YH_HIP(hipGraphLaunch(sr->exec, s));
  }
  pl->count += nsteps;
  pl->have_prev = false;
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
  YH_HIP(hipMemcpy(&h, pl->sr->st, sizeof h, hipMemcpyDeviceToHost));
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
