// adjoint.hip -- gradient of a sensor-mismatch loss through the FK stepper.
//
// Forward (no reference counterpart; the reference has no gradient path):
//   u' = m*(u + tc*(L_D u - dt*I(u,v,w))),  v' = m*(v - tc*dt*G_v),  w' = ...
//   L_D u = dt/hx^2 sum_{W,E} c*0.5*(D_c + D_n)*(u_n - u) + (same in y)
// with the mask coefficients c in {0,1,2} of reactionDiffusion.cu:162-169,
// the mirror neighbours of coord_i/coord_j (helper_functions.cu:69-79), and
// the Fenton-Karma currents with the Heaviside smoothed to
// 0.5*(1 + tanh(x/w)) so the loss is differentiable.
// Loss = sum over recorded steps n (n % rec == 0) and sensors s of
// (u^n(s) - target)^2.
//
// Reverse: lambda^n = J(X^n)^T lambda^{n+1} + dL/dX^n, gD += (dF/dD)^T
// lambda^{n+1}, gtau += (dF/dtau)^T lambda^{n+1}.  The transpose of the
// mirrored, masked operator is applied in gather form: cell j collects the
// coefficient of every (cell c, direction) whose neighbour is j, including the
// two mirror images at each edge.  The forward keeps a checkpoint every
// `ckpt` steps; each segment is recomputed into a buffer before its reverse
// sweep.  gD and gtau accumulate in FP64 (gtau per workgroup, in a fixed
// order, so results are deterministic).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "plan.h"

namespace yh {

struct AdjK {
  int nx, ny;
  float rxh, ryh;  // dt/hx^2, dt/hy^2
  float dt, tc;
  float hw;        // Heaviside width
  float td, tr, tsi, t0, tvp, tv1m, tv2m, twp, twm, uc, uv, ucsi, kk;
  // reciprocals, rounded once on the host from FP64: 1/hw, 2/hw, 1/tau_d,
  // 1/tau_r, 0.5/tau_si, 1/tau_0, 1/tau_v+, 1/tau_w-, 1/tau_w+
  float ihw, tihw, itd, itr, ihtsi, it0, itvp, itwm, itwp;
};

struct FkEval {
  float I, Gv, Gw;
  float dIdu, dIdv, dIdw, dGvdu, dGvdv, dGwdu, dGwdw;
  float dI_dtau[4];  // tau_d, tau_r, tau_si, tau_0
  float dGv_dtvp, dGw_dtwp;
};

// s = sigmoid(2y) = 0.5*(1 + tanh y) and its complement sc = 1 - s, both
// from a = exp(-2|y|) in (0, 1] with the hardware exp2 and reciprocal:
// the larger is 1/(1 + a), the smaller a/(1 + a), so each keeps a relative
// error of a few ulp (1 - 2/(exp(2y) + 1) loses the smaller one, and
// 1 - tanh^2, to cancellation: 2e-3 relative at |y| = 5).
__device__ __forceinline__ void sigmoid2(float y, float* s, float* sc) {
#pragma clang fp contract(off)
  const float a = __builtin_amdgcn_exp2f(-2.8853900817779268f * __builtin_fabsf(y));  // 2/ln 2
  const float big = __builtin_amdgcn_rcpf(1.0f + a);
  const float small = a * big;
  *s = y >= 0.0f ? big : small;
  *sc = y >= 0.0f ? small : big;
}

// Smoothed Heaviside h = 0.5*(1 + tanh(x/w)), its complement hc = 1 - h and
// its derivative dh = (1 - tanh^2)/(2w) = (2/w) h (1 - h).
__device__ __forceinline__ void heav(float x, float iw, float tiw, float* h, float* hc,
                                     float* dh) {
#pragma clang fp contract(off)
  sigmoid2(x * iw, h, hc);
  *dh = tiw * (*h * *hc);
}

// Fenton & Karma 1998 currents and derivatives (smoothed Heaviside).  No
// FMA contraction here or in adj_fwd_cell: every operation rounds on its
// own, so the kernels that inline them agree bit for bit.  pc = 1 - p and
// qc = 1 - q come from heav(); 1 + tanh(k(u - u_csi)) = 2 ss and
// 1 - tanh^2 = 4 ss ssc with (ss, ssc) = sigmoid2(k(u - u_csi)).
__device__ FkEval fk_eval(const AdjK& k, float u, float v, float w) {
#pragma clang fp contract(off)
  FkEval r;
  float p, pc, dp, q, qc, dq, ss, ssc;
  heav(u - k.uc, k.ihw, k.tihw, &p, &pc, &dp);
  heav(u - k.uv, k.ihw, k.tihw, &q, &qc, &dq);
  const float f = (1.0f - u) * (u - k.uc);
  const float df = 1.0f + k.uc - 2.0f * u;
  const float Jfi = -v * p * f * k.itd;
  const float Jso = u * pc * k.it0 + p * k.itr;
  sigmoid2(k.kk * (u - k.ucsi), &ss, &ssc);
  const float opth = 2.0f * ss;           // 1 + tanh
  const float sech2 = 4.0f * (ss * ssc);  // 1 - tanh^2
  const float Jsi = -w * opth * k.ihtsi;
  r.I = Jfi + Jso + Jsi;
  r.dIdu = -v * (dp * f + p * df) * k.itd + pc * k.it0 - u * dp * k.it0 + dp * k.itr -
           w * k.kk * sech2 * k.ihtsi;
  r.dIdv = -p * f * k.itd;
  r.dIdw = -opth * k.ihtsi;
  const float tvm = qc * k.tv1m + q * k.tv2m;
  const float itvm = 1.0f / tvm;
  const float dtvm = dq * (k.tv2m - k.tv1m);
  const float A = pc * (1.0f - v) * itvm - p * v * k.itvp;
  const float dAdu = -dp * (1.0f - v) * itvm - pc * (1.0f - v) * dtvm * (itvm * itvm) -
                     dp * v * k.itvp;
  const float dAdv = -pc * itvm - p * k.itvp;
  const float B = pc * (1.0f - w) * k.itwm - p * w * k.itwp;
  const float dBdu = -dp * (1.0f - w) * k.itwm - dp * w * k.itwp;
  const float dBdw = -pc * k.itwm - p * k.itwp;
  r.Gv = -A;
  r.Gw = -B;
  r.dGvdu = -dAdu;
  r.dGvdv = -dAdv;
  r.dGwdu = -dBdu;
  r.dGwdw = -dBdw;
  r.dI_dtau[0] = -Jfi * k.itd;
  r.dI_dtau[1] = -p * (k.itr * k.itr);
  r.dI_dtau[2] = -2.0f * Jsi * k.ihtsi;
  r.dI_dtau[3] = -u * pc * (k.it0 * k.it0);
  r.dGv_dtvp = -p * v * (k.itvp * k.itvp);
  r.dGw_dtwp = -p * w * (k.itwp * k.itwp);
  return r;
}

struct Coef4 {
  float W, E, N, S;
  bool sc;
};

__device__ __forceinline__ Coef4 coef_of(const uint8_t* codes, int c) {
  Coef4 r{1.f, 1.f, 1.f, 1.f, true};
  if (codes) {
    const uint32_t b = codes[c];
    r.sc = b != kHole;
    const uint32_t q = r.sc ? b : 0u;
    r.W = (float)(q & 3u);
    r.E = (float)((q >> 2) & 3u);
    r.N = (float)((q >> 4) & 3u);
    r.S = (float)((q >> 6) & 3u);
  }
  return r;
}

// Mask coefficients of a stencil code (0xFF: hole).
__device__ __forceinline__ Coef4 coef_from_code(uint32_t b) {
  Coef4 r;
  r.sc = b != kHole;
  const uint32_t q = r.sc ? b : 0u;
  r.W = (float)(q & 3u);
  r.E = (float)((q >> 2) & 3u);
  r.N = (float)((q >> 4) & 3u);
  r.S = (float)((q >> 6) & 3u);
  return r;
}
constexpr uint8_t kCodeAllOnes = 0x55;  // cW = cE = cN = cS = 1: no mask

// ---------------------------------------------------------------- forward
// One forward step of one cell from its state, its D, its four (mirror)
// neighbours' u and D and its mask coefficients; shared by both forward
// kernels, so they agree bit for bit.
__device__ __forceinline__ void adj_fwd_cell(const AdjK& k, float u, float v, float w, float Dc,
                                             float uW, float uE, float uN, float uS, float DW,
                                             float DE, float DN, float DS, const Coef4& cf,
                                             float& un, float& vn, float& wn) {
#pragma clang fp contract(off)
  const float lap = k.rxh * (cf.W * 0.5f * (Dc + DW) * (uW - u) +
                             cf.E * 0.5f * (Dc + DE) * (uE - u)) +
                    k.ryh * (cf.N * 0.5f * (Dc + DN) * (uN - u) +
                             cf.S * 0.5f * (Dc + DS) * (uS - u));
  const FkEval r = fk_eval(k, u, v, w);
  un = cf.sc ? u + k.tc * (lap - k.dt * r.I) : 0.0f;
  vn = cf.sc ? v + k.tc * (-k.dt * r.Gv) : 0.0f;
  wn = cf.sc ? w + k.tc * (-k.dt * r.Gw) : 0.0f;
}

__global__ __launch_bounds__(256) void k_adj_fwd(const AdjK k, CFields in, Fields out,
                                                 const float* __restrict__ D,
                                                 const uint8_t* __restrict__ codes,
                                                 const int* __restrict__ smap,
                                                 float* __restrict__ trace_row) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= k.nx || j >= k.ny) return;
  const int nx = k.nx, ny = k.ny;
  const int c = i + nx * j;
  const int W = mirror(i - 1, nx) + nx * j, E = mirror(i + 1, nx) + nx * j;
  const int N = i + nx * mirror(j + 1, ny), S = i + nx * mirror(j - 1, ny);
  const Coef4 cf = coef_of(codes, c);
  const float* U = in.f[0];
  float un, vn, wn;
  adj_fwd_cell(k, U[c], in.f[1][c], in.f[2][c], D[c], U[W], U[E], U[N], U[S], D[W], D[E], D[N],
               D[S], cf, un, vn, wn);
  out.f[0][c] = un;
  out.f[1][c] = vn;
  out.f[2][c] = wn;
  if (trace_row) {
    const int s = smap[c];
    if (s >= 0) trace_row[s] = un;
  }
}

// K forward steps (K <= kAfKmax) of one kAfTX x kAfTY output tile through
// LDS.  The tile's region is the output grown by K cells on each side and
// clipped to the domain; mirror neighbours at the domain edge lie inside the
// region, so in substep s only the cells s+1 away from region sides interior
// to the domain lose an exact input and are skipped.  u is double-buffered
// (neighbours read it); v and w are read by their own cell only and are
// updated in place.  Outputs: with seg, the state after every substep into
// seg slots slot0+1 .. slot0+K (the segment recompute); otherwise the final
// state into out.  Sensor traces for every step with step % rec == 0.
constexpr int kAfTX = 16, kAfTY = 16, kAfKmax = 8, kAfThreads = 1024;
constexpr int kAfRN = (kAfTX + 2 * kAfKmax) * (kAfTY + 2 * kAfKmax);

__global__ __launch_bounds__(kAfThreads) void k_adj_fwd_tiled(
    const AdjK k, CFields in, Fields out, const float* __restrict__ D,
    const uint8_t* __restrict__ codes, const int* __restrict__ smap, int K, int tiles_x,
    int step0, int rec, float* __restrict__ traces, int ns, float* __restrict__ seg, size_t n,
    int slot0) {
  __shared__ float s_u[2][kAfRN];
  __shared__ float s_v[kAfRN], s_w[kAfRN], s_D[kAfRN];
  __shared__ uint8_t s_c[kAfRN];
  __shared__ int s_sm[kAfRN];  // sensor slot of each region cell (-1: none), read once
  const int nx = k.nx, ny = k.ny;
  const int tx = blockIdx.x % tiles_x, ty = blockIdx.x / tiles_x;
  const int ox0 = tx * kAfTX, oy0 = ty * kAfTY;
  const int ox1 = min(ox0 + kAfTX, nx), oy1 = min(oy0 + kAfTY, ny);
  const int gx0 = max(ox0 - K, 0), gy0 = max(oy0 - K, 0);
  const int gx1 = min(ox1 + K, nx), gy1 = min(oy1 + K, ny);
  const int rw = gx1 - gx0, rn = rw * (gy1 - gy0);
  for (int l = threadIdx.x; l < rn; l += kAfThreads) {
    const size_t c = (size_t)(gx0 + l % rw) + (size_t)nx * (gy0 + l / rw);
    s_u[0][l] = in.f[0][c];
    s_v[l] = in.f[1][c];
    s_w[l] = in.f[2][c];
    s_D[l] = D[c];
    s_c[l] = codes ? codes[c] : kCodeAllOnes;
    s_sm[l] = traces ? smap[c] : -1;
  }
  __syncthreads();
  for (int s = 0; s < K; ++s) {
    const int b = s & 1;
    const int cx0 = gx0 > 0 ? gx0 + s + 1 : 0, cx1 = gx1 < nx ? gx1 - s - 1 : nx;
    const int cy0 = gy0 > 0 ? gy0 + s + 1 : 0, cy1 = gy1 < ny ? gy1 - s - 1 : ny;
    const int cw = cx1 - cx0, cn = cw * (cy1 - cy0);
    const int step = step0 + s + 1;
    float* row = (traces && step % rec == 0) ? traces + (size_t)(step / rec - 1) * ns : nullptr;
    const float* U = s_u[b];
    for (int q = threadIdx.x; q < cn; q += kAfThreads) {
      const int gi = cx0 + q % cw, gj = cy0 + q / cw;
      const int li = gi - gx0, lj = gj - gy0;
      const int l = li + rw * lj;
      const int lW = (mirror(gi - 1, nx) - gx0) + rw * lj;
      const int lE = (mirror(gi + 1, nx) - gx0) + rw * lj;
      const int lN = li + rw * (mirror(gj + 1, ny) - gy0);
      const int lS = li + rw * (mirror(gj - 1, ny) - gy0);
      float un, vn, wn;
      adj_fwd_cell(k, U[l], s_v[l], s_w[l], s_D[l], U[lW], U[lE], U[lN], U[lS], s_D[lW],
                   s_D[lE], s_D[lN], s_D[lS], coef_from_code(s_c[l]), un, vn, wn);
      s_u[b ^ 1][l] = un;
      s_v[l] = vn;
      s_w[l] = wn;
      if (gi >= ox0 && gi < ox1 && gj >= oy0 && gj < oy1) {
        const size_t c = (size_t)gi + (size_t)nx * gj;
        if (seg) {
          float* base = seg + (size_t)(slot0 + s + 1) * 3 * n;
          base[c] = un;
          base[n + c] = vn;
          base[2 * n + c] = wn;
        } else if (s == K - 1) {
          out.f[0][c] = un;
          out.f[1][c] = vn;
          out.f[2][c] = wn;
        }
        if (row) {
          const int sm = s_sm[l];
          if (sm >= 0) row[sm] = un;
        }
      }
    }
    __syncthreads();
  }
}

static hipError_t launch_adj_fwd_tiled(const AdjK& k, CFields in, Fields out, const float* D,
                                       const uint8_t* codes, const int* smap, int K, int step0,
                                       int rec, float* traces, int ns, float* seg, size_t n,
                                       int slot0, hipStream_t s) {
  if (K < 1 || K > kAfKmax) return hipErrorInvalidValue;
  const int tiles_x = (k.nx + kAfTX - 1) / kAfTX, tiles_y = (k.ny + kAfTY - 1) / kAfTY;
  hipLaunchKernelGGL(k_adj_fwd_tiled, dim3(tiles_x * tiles_y), dim3(kAfThreads), 0, s, k, in,
                     out, D, codes, smap, K, tiles_x, step0, rec, traces, ns, seg, n, slot0);
  return hipGetLastError();
}

// ---------------------------------------------------------------- reverse
// One reverse step of one cell (i, j): lam_out = J(X)^T lam_in (+ source),
// its dL/dD term gd and its dL/dtau terms gt.  X is the state the forward
// step started from.  The accessor gives u of X, D, lambda_u of lam_in and
// the stencil code at offsets (di, dj) from the cell; every offset used is a
// cell of the domain.  Shared by both reverse kernels (no contraction), so
// they agree bit for bit on lambda and on each step's terms.
template <class Acc>
__device__ __forceinline__ void adj_rev_cell(const AdjK& k, const Acc& A, int i, int j, float v,
                                             float w, float lv_in, float lw_in, float src,
                                             float& lu, float& lv, float& lw, double& gd,
                                             float gt[6]) {
#pragma clang fp contract(off)
  const int nx = k.nx, ny = k.ny;
  const Coef4 cf = coef_from_code(A.code(0, 0));
  const float u = A.U(0, 0);
  const float m = cf.sc ? 1.0f : 0.0f;
  const float mu = m * A.LU(0, 0), mv = m * lv_in, mw = m * lw_in;
  const float Dc = A.D(0, 0);
  const float tc = k.tc;
  // own diffusion terms, neighbours W, E, N, S by mirror (helper_functions.cu:69-79)
  const int ox[4] = {i > 0 ? -1 : 1, i < nx - 1 ? 1 : -1, 0, 0};
  const int oy[4] = {0, 0, j < ny - 1 ? 1 : -1, j > 0 ? -1 : 1};
  const float co[4] = {cf.W, cf.E, cf.N, cf.S};
  const float rr[4] = {k.rxh, k.rxh, k.ryh, k.ryh};
  float asum = 0.0f;
  gd = 0.0;
#pragma unroll
  for (int d = 0; d < 4; ++d) {
    asum += rr[d] * co[d] * 0.5f * (Dc + A.D(ox[d], oy[d]));
    gd += (double)(tc * mu * rr[d] * co[d] * 0.5f * (A.U(ox[d], oy[d]) - u));
  }
  // coefficients pointing at this cell from its neighbours (gather)
  float lin = 0.0f;
  auto from = [&](int di, int dj, float coef_c, float r) {
    const float mc = A.LU(di, dj) * (coef_c != 0.0f ? 1.0f : 0.0f);  // coef 0 also covers holes
    const float a = r * coef_c * 0.5f * (A.D(di, dj) + Dc);
    lin += tc * mc * a;
    gd += (double)(tc * mc * r * coef_c * 0.5f * (u - A.U(di, dj)));
  };
  if (i + 1 < nx) from(1, 0, coef_from_code(A.code(1, 0)).W, k.rxh);
  if (i - 1 >= 0) from(-1, 0, coef_from_code(A.code(-1, 0)).E, k.rxh);
  if (i == 1) from(-1, 0, coef_from_code(A.code(-1, 0)).W, k.rxh);
  if (i == nx - 2) from(1, 0, coef_from_code(A.code(1, 0)).E, k.rxh);
  if (j + 1 < ny) from(0, 1, coef_from_code(A.code(0, 1)).S, k.ryh);
  if (j - 1 >= 0) from(0, -1, coef_from_code(A.code(0, -1)).N, k.ryh);
  if (j == 1) from(0, -1, coef_from_code(A.code(0, -1)).S, k.ryh);
  if (j == ny - 2) from(0, 1, coef_from_code(A.code(0, 1)).N, k.ryh);
  const FkEval r = fk_eval(k, u, v, w);
  const float tdt = tc * k.dt;
  lu = mu * (1.0f - tc * asum - tdt * r.dIdu) + lin - mv * tdt * r.dGvdu - mw * tdt * r.dGwdu;
  lv = -mu * tdt * r.dIdv + mv * (1.0f - tdt * r.dGvdv);
  lw = -mu * tdt * r.dIdw + mw * (1.0f - tdt * r.dGwdw);
  lu += src;
#pragma unroll
  for (int q = 0; q < 4; ++q) gt[q] = mu * (-tdt * r.dI_dtau[q]);
  gt[4] = mv * (-tdt * r.dGv_dtvp);
  gt[5] = mw * (-tdt * r.dGw_dtwp);
}

// Accessor over device fields (row stride nx, cell index c).
struct GlobalAcc {
  const float *Uf, *Df, *LUf;
  const uint8_t* codes;
  int c, nx;
  __device__ float U(int di, int dj) const { return Uf[c + di + nx * dj]; }
  __device__ float D(int di, int dj) const { return Df[c + di + nx * dj]; }
  __device__ float LU(int di, int dj) const { return LUf[c + di + nx * dj]; }
  __device__ uint32_t code(int di, int dj) const {
    return codes ? codes[c + di + nx * dj] : kCodeAllOnes;
  }
};

// The sum of gt over a workgroup of NT threads (FP64, fixed order), added to
// the workgroup's slot.
template <int NT>
__device__ __forceinline__ void adj_gtau_block(const double gt[6], double* slot) {
  __shared__ double red[NT / 64][6];
  const int tid = threadIdx.y * blockDim.x + threadIdx.x;
  const int lane = tid & 63, wv = tid >> 6;
  // the six shuffle trees interleaved level by level, so their latencies
  // overlap; each value's adds keep their order
  double v[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) v[q] = gt[q];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int q = 0; q < 6; ++q) v[q] += __shfl_down(v[q], off, 64);
  }
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < 6; ++q) red[wv][q] = v[q];
  }
  __syncthreads();
  if (tid < 6) {
    double s = 0.0;
    for (int w = 0; w < NT / 64; ++w) s += red[w][tid];
    slot[tid] += s;
  }
}

__global__ __launch_bounds__(256) void k_adj_rev(const AdjK k, CFields X, CFields lam_in,
                                                 Fields lam_out, const float* __restrict__ D,
                                                 const uint8_t* __restrict__ codes,
                                                 const int* __restrict__ smap,
                                                 const float* __restrict__ src_row,
                                                 double* __restrict__ gD,
                                                 double* __restrict__ gtau_part) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  double gtd[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (i < k.nx && j < k.ny) {
    const int c = i + k.nx * j;
    GlobalAcc A{X.f[0], D, lam_in.f[0], codes, c, k.nx};
    float src = 0.0f;
    if (src_row) {
      const int s = smap[c];
      if (s >= 0) src = src_row[s];
    }
    float lu, lv, lw, gt[6];
    double gd;
    adj_rev_cell(k, A, i, j, X.f[1][c], X.f[2][c], lam_in.f[1][c], lam_in.f[2][c], src, lu, lv,
                 lw, gd, gt);
    lam_out.f[0][c] = lu;
    lam_out.f[1][c] = lv;
    lam_out.f[2][c] = lw;
    gD[c] += gd;
#pragma unroll
    for (int q = 0; q < 6; ++q) gtd[q] = (double)gt[q];
  }
  adj_gtau_block<256>(gtd, gtau_part + (size_t)(blockIdx.y * gridDim.x + blockIdx.x) * 6);
}

// Accessor over an LDS region of row stride rw (cell index l).
struct LdsAcc {
  const float *Uf, *Df, *LUf;
  const uint8_t* codes;
  int l, rw;
  __device__ float U(int di, int dj) const { return Uf[l + di + rw * dj]; }
  __device__ float D(int di, int dj) const { return Df[l + di + rw * dj]; }
  __device__ float LU(int di, int dj) const { return LUf[l + di + rw * dj]; }
  __device__ uint32_t code(int di, int dj) const { return codes[l + di + rw * dj]; }
};

// K reverse steps (K <= kArKmax), n_top, n_top-1, ..., of one kArTX x kArTY
// output tile through LDS; the region is the output grown by K and clipped
// to the domain, as in the tiled forward (lambda has the same one-cell
// dependence).  u of each step's state is staged in LDS (neighbours read
// it); v and w of the cell come from the segment buffer.  lambda_u is
// double-buffered, lambda_v and lambda_w are read by their own cell only and
// updated in place.  dL/dD of the output cells accumulates over the K steps
// in LDS (FP64) and is added once; dL/dtau per thread, then per workgroup.
constexpr int kArTX = 16, kArTY = 16, kArKmax = 8, kArThreads = 1024;
constexpr int kArRN = (kArTX + 2 * kArKmax) * (kArTY + 2 * kArKmax);

__global__ __launch_bounds__(kArThreads) void k_adj_rev_tiled(
    const AdjK k, const float* __restrict__ seg, const float* __restrict__ x0, size_t n,
    int slot_top, int n_top, CFields lam_in,
    Fields lam_out, const float* __restrict__ D, const uint8_t* __restrict__ codes,
    const int* __restrict__ smap, const float* __restrict__ src, int rec, int ns,
    double* __restrict__ gD, double* __restrict__ gtau_part, int K, int tiles_x) {
  __shared__ float s_U[kArRN], s_D[kArRN], s_lu[2][kArRN], s_lv[kArRN], s_lw[kArRN];
  __shared__ uint8_t s_c[kArRN];
  __shared__ double s_gd[kArTX * kArTY];
  const int nx = k.nx, ny = k.ny;
  const int tx = blockIdx.x % tiles_x, ty = blockIdx.x / tiles_x;
  const int ox0 = tx * kArTX, oy0 = ty * kArTY;
  const int ox1 = min(ox0 + kArTX, nx), oy1 = min(oy0 + kArTY, ny);
  const int gx0 = max(ox0 - K, 0), gy0 = max(oy0 - K, 0);
  const int gx1 = min(ox1 + K, nx), gy1 = min(oy1 + K, ny);
  const int rw = gx1 - gx0, rn = rw * (gy1 - gy0);
  for (int l = threadIdx.x; l < rn; l += kArThreads) {
    const size_t c = (size_t)(gx0 + l % rw) + (size_t)nx * (gy0 + l / rw);
    s_D[l] = D[c];
    s_c[l] = codes ? codes[c] : kCodeAllOnes;
    s_lu[0][l] = lam_in.f[0][c];
    s_lv[l] = lam_in.f[1][c];
    s_lw[l] = lam_in.f[2][c];
  }
  for (int o = threadIdx.x; o < kArTX * kArTY; o += kArThreads) s_gd[o] = 0.0;
  double gtd[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  // A region has at most kArRN <= kArThreads cells, so thread t holds region
  // cell t (u of the step's state) and computes the t-th cell of the step's
  // shrinking region.  The values a step needs from global memory (u, the
  // computed cell's v and w, its sensor source) are loaded one step ahead,
  // during the previous step's arithmetic.
  static_assert(kArRN <= kArThreads, "one region cell per thread");
  const int tid = threadIdx.x;
  struct Ahead {
    float u = 0.0f, v = 0.0f, w = 0.0f, src = 0.0f;
  };
  auto fetch = [&](int s) {
    Ahead a;
    const float* X = slot_top - s == 0 ? x0 : seg + (size_t)(slot_top - s) * 3 * n;
    if (tid < rn) a.u = X[(size_t)(gx0 + tid % rw) + (size_t)nx * (gy0 + tid / rw)];
    const int cx0 = gx0 > 0 ? gx0 + s + 1 : 0, cx1 = gx1 < nx ? gx1 - s - 1 : nx;
    const int cy0 = gy0 > 0 ? gy0 + s + 1 : 0, cy1 = gy1 < ny ? gy1 - s - 1 : ny;
    const int cw = cx1 - cx0, cn = cw * (cy1 - cy0);
    if (tid < cn) {
      const size_t c = (size_t)(cx0 + tid % cw) + (size_t)nx * (cy0 + tid / cw);
      a.v = X[n + c];
      a.w = X[2 * n + c];
      const int step = n_top - s;
      if (src && step >= 1 && step % rec == 0) {
        const int sm = smap[c];
        if (sm >= 0) a.src = src[(size_t)(step / rec - 1) * ns + sm];
      }
    }
    return a;
  };
  Ahead next = fetch(0);
  for (int s = 0; s < K; ++s) {
    const int b = s & 1;
    const Ahead cur = next;
    if (tid < rn) s_U[tid] = cur.u;
    if (s + 1 < K) next = fetch(s + 1);
    __syncthreads();
    const int cx0 = gx0 > 0 ? gx0 + s + 1 : 0, cx1 = gx1 < nx ? gx1 - s - 1 : nx;
    const int cy0 = gy0 > 0 ? gy0 + s + 1 : 0, cy1 = gy1 < ny ? gy1 - s - 1 : ny;
    const int cw = cx1 - cx0, cn = cw * (cy1 - cy0);
    if (tid < cn) {
      const int gi = cx0 + tid % cw, gj = cy0 + tid / cw;
      const int l = (gi - gx0) + rw * (gj - gy0);
      LdsAcc A{s_U, s_D, s_lu[b], s_c, l, rw};
      float lu, lv, lw, gt[6];
      double gd;
      adj_rev_cell(k, A, gi, gj, cur.v, cur.w, s_lv[l], s_lw[l], cur.src, lu, lv, lw, gd, gt);
      s_lu[b ^ 1][l] = lu;
      s_lv[l] = lv;
      s_lw[l] = lw;
      if (gi >= ox0 && gi < ox1 && gj >= oy0 && gj < oy1) {
        s_gd[(gi - ox0) + kArTX * (gj - oy0)] += gd;
#pragma unroll
        for (int t = 0; t < 6; ++t) gtd[t] += (double)gt[t];
      }
    }
    __syncthreads();
  }
  const float* LU = s_lu[K & 1];
  const int ow = ox1 - ox0, on = ow * (oy1 - oy0);
  for (int q = threadIdx.x; q < on; q += kArThreads) {
    const int gi = ox0 + q % ow, gj = oy0 + q / ow;
    const int l = (gi - gx0) + rw * (gj - gy0);
    const size_t c = (size_t)gi + (size_t)nx * gj;
    lam_out.f[0][c] = LU[l];
    lam_out.f[1][c] = s_lv[l];
    lam_out.f[2][c] = s_lw[l];
    gD[c] += s_gd[(gi - ox0) + kArTX * (gj - oy0)];
  }
  adj_gtau_block<kArThreads>(gtd, gtau_part + (size_t)blockIdx.x * 6);
}

static hipError_t launch_adj_rev_tiled(const AdjK& k, const float* seg, const float* x0,
                                       size_t n, int slot_top, int n_top, CFields li, Fields lo,
                                       const float* D,
                                       const uint8_t* codes, const int* smap, const float* src,
                                       int rec, int ns, double* gD, double* gtau_part, int K,
                                       hipStream_t s) {
  if (K < 1 || K > kArKmax) return hipErrorInvalidValue;
  const int tiles_x = (k.nx + kArTX - 1) / kArTX, tiles_y = (k.ny + kArTY - 1) / kArTY;
  hipLaunchKernelGGL(k_adj_rev_tiled, dim3(tiles_x * tiles_y), dim3(kArThreads), 0, s, k, seg, x0,
                     n, slot_top, n_top, li, lo, D, codes, smap, src, rec, ns, gD, gtau_part, K,
                     tiles_x);
  return hipGetLastError();
}

// lambda^N: zero, plus the sensor source of the last recorded step.
__global__ void k_adj_init(Fields lam, size_t n, const int* __restrict__ smap,
                           const float* __restrict__ src_row) {
  for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n;
       c += (size_t)gridDim.x * blockDim.x) {
    float s = 0.0f;
    if (src_row) {
      const int q = smap[c];
      if (q >= 0) s = src_row[q];
    }
    lam.f[0][c] = s;
    lam.f[1][c] = 0.0f;
    lam.f[2][c] = 0.0f;
  }
}

__global__ void k_fill_d(double* p, size_t n, double v) {
  for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n;
       c += (size_t)gridDim.x * blockDim.x)
    p[c] = v;
}

}  // namespace yh

struct yh_adjoint {
  yh_params p;
  yh_derived d;
  int nsteps = 0, rec = 1, ckpt = 100, ns = 0, nrec = 0, nckpt = 0;
  size_t n = 0;
  std::vector<int> sensors;
  float* X[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
  float* X0[3] = {nullptr, nullptr, nullptr};
  float* ck = nullptr;   // nckpt x 3 x n
  float* seg = nullptr;  // (ckpt + 1) x 3 x n
  float* lam[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
  int* smap = nullptr;
  uint8_t* solid = nullptr;
  uint8_t* codes = nullptr;
  float* traces = nullptr;   // nrec x ns (device)
  float* targets = nullptr;  // nrec x ns (device)
  float* src = nullptr;      // nrec x ns: 2 (trace - target)
  double* gD = nullptr;
  double* gtau_part = nullptr;
  int nblocks = 0;
  std::vector<float> h_traces, h_targets;
  bool have_targets = false, forward_done = false;
  const float* D = nullptr;
  yh::AdjK k{};
  int fwdK = 8;  // steps per launch, forward and reverse: 1 = one-step kernels, else LDS-tiled
  int nparts = 0;  // dL/dtau workgroup slots: the larger of the two reverse grids
};

namespace yh {

static void adj_free(yh_adjoint* a) {
  if (!a) return;
  for (int b = 0; b < 2; ++b)
    for (int q = 0; q < 3; ++q) {
      if (a->X[b][q]) (void)hipFree(a->X[b][q]);
      if (a->lam[b][q]) (void)hipFree(a->lam[b][q]);
    }
  for (int q = 0; q < 3; ++q)
    if (a->X0[q]) (void)hipFree(a->X0[q]);
  void* ptrs[] = {a->ck, a->seg, a->smap, a->solid, a->codes, a->traces, a->targets, a->src,
                  a->gD, a->gtau_part};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  delete a;
}

static void adj_k(yh_adjoint* a, const double tau6[6]) {
  AdjK& k = a->k;
  const yh_params& p = a->p;
  k.nx = p.nx;
  k.ny = p.ny;
  k.rxh = (float)(a->d.dt / (a->d.hx * a->d.hx));
  k.ryh = (float)(a->d.dt / (a->d.hy * a->d.hy));
  k.dt = (float)a->d.dt;
  k.tc = (float)p.tc;
  k.hw = (float)p.heaviside_width;
  const yh_fk_params& f = p.fk;
  k.td = (float)tau6[0];
  k.tr = (float)tau6[1];
  k.tsi = (float)tau6[2];
  k.t0 = (float)tau6[3];
  k.tvp = (float)tau6[4];
  k.twp = (float)tau6[5];
  k.tv1m = (float)f.tau_v1m;
  k.tv2m = (float)f.tau_v2m;
  k.twm = (float)f.tau_wm;
  k.uc = (float)f.u_c;
  k.uv = (float)f.u_v;
  k.ucsi = (float)f.u_csi;
  k.kk = (float)f.k;
  k.ihw = (float)(1.0 / p.heaviside_width);
  k.tihw = (float)(2.0 / p.heaviside_width);
  k.itd = (float)(1.0 / tau6[0]);
  k.itr = (float)(1.0 / tau6[1]);
  k.ihtsi = (float)(0.5 / tau6[2]);
  k.it0 = (float)(1.0 / tau6[3]);
  k.itvp = (float)(1.0 / tau6[4]);
  k.itwp = (float)(1.0 / tau6[5]);
  k.itwm = (float)(1.0 / f.tau_wm);
}

static inline float* slot(float* base, size_t n, int s, int var) {
  return base + ((size_t)s * 3 + var) * n;
}

}  // namespace yh

using namespace yh;

extern "C" {

int yh_adjoint_create(const yh_params* p, int nsteps, int rec_every, int ckpt_every,
                      const int* sensors, int nsensors, yh_adjoint** out) {
  if (!p || !out || nsteps < 1 || rec_every < 1 || ckpt_every < 1 || nsensors < 1 || !sensors)
    return YH_EINVAL;
  if (p->model != YH_MODEL_3V_FK || !(p->heaviside_width > 0) || p->reduceSym)
    return YH_EINVAL;
  if (!standard_branch(*p)) return YH_EUNSUPPORTED;
  *out = nullptr;
  yh_derived d;
  int rc = yh_params_derive(p, &d);
  if (rc) return rc;
  const size_t n = (size_t)p->nx * p->ny;
  for (int s = 0; s < nsensors; ++s)
    if (sensors[s] < 0 || (size_t)sensors[s] >= n) return YH_EINVAL;
  yh_adjoint* a = new (std::nothrow) yh_adjoint();
  if (!a) return YH_ENOMEM;
  a->p = *p;
  a->d = d;
  a->nsteps = nsteps;
  a->rec = rec_every;
  a->ckpt = std::min(ckpt_every, nsteps);
  a->ns = nsensors;
  a->nrec = nsteps / rec_every;
  a->nckpt = (nsteps + a->ckpt - 1) / a->ckpt;
  a->n = n;
  a->sensors.assign(sensors, sensors + nsensors);
  std::vector<int> smap(n, -1);
  for (int s = 0; s < nsensors; ++s) {
    if (smap[sensors[s]] >= 0) {  // duplicate sensor cell
      adj_free(a);
      return YH_EINVAL;
    }
    smap[sensors[s]] = s;
  }
  const double tau6[6] = {p->fk.tau_d, p->fk.tau_r, p->fk.tau_si,
                          p->fk.tau_0, p->fk.tau_vp, p->fk.tau_wp};
  adj_k(a, tau6);
  const dim3 bg((p->nx + 63) / 64, (p->ny + 3) / 4);  // the 64 x 4 one-step blocks
  a->nblocks = bg.x * bg.y;
  a->nparts = std::max(a->nblocks, ((p->nx + kArTX - 1) / kArTX) * ((p->ny + kArTY - 1) / kArTY));
  hipError_t e = hipSuccess;
  auto M = [&](void** ptr, size_t bytes) {
    if (e == hipSuccess) e = hipMalloc(ptr, std::max<size_t>(bytes, 4));
  };
  for (int b = 0; b < 2; ++b)
    for (int q = 0; q < 3; ++q) {
      M((void**)&a->X[b][q], n * 4);
      M((void**)&a->lam[b][q], n * 4);
    }
  for (int q = 0; q < 3; ++q) M((void**)&a->X0[q], n * 4);
  M((void**)&a->ck, (size_t)a->nckpt * 3 * n * 4);
  M((void**)&a->seg, (size_t)(a->ckpt + 1) * 3 * n * 4);
  M((void**)&a->smap, n * 4);
  M((void**)&a->traces, (size_t)a->nrec * a->ns * 4);
  M((void**)&a->targets, (size_t)a->nrec * a->ns * 4);
  M((void**)&a->src, (size_t)a->nrec * a->ns * 4);
  M((void**)&a->gD, n * 8);
  M((void**)&a->gtau_part, (size_t)a->nparts * 6 * 8);
  if (e == hipSuccess) e = hipMemcpy(a->smap, smap.data(), n * 4, hipMemcpyHostToDevice);
  for (int q = 0; q < 3 && e == hipSuccess; ++q) e = hipMemset(a->X0[q], 0, n * 4);
  if (e != hipSuccess) {
    rc = hip_fail(e, "yh_adjoint_create");
    adj_free(a);
    return rc;
  }
  *out = a;
  return YH_OK;
}

void yh_adjoint_destroy(yh_adjoint* a) {
  if (a) (void)hipDeviceSynchronize();
  adj_free(a);
}

int yh_adjoint_set_substeps(yh_adjoint* a, int k) {
  if (!a || k < 1 || k > kAfKmax || k > kArKmax) return YH_EINVAL;
  a->fwdK = k;
  return YH_OK;
}

int yh_adjoint_set_mask(yh_adjoint* a, const uint8_t* solid) {
  if (!a || !solid) return YH_EINVAL;
  if (!a->p.solidSwitch) return YH_ESTATE;
  if (!a->solid) YH_HIP(hipMalloc(&a->solid, a->n));
  if (!a->codes) YH_HIP(hipMalloc(&a->codes, a->n));
  std::vector<uint8_t> m(a->n);
  for (size_t c = 0; c < a->n; ++c) m[c] = solid[c] ? 1 : 0;
  YH_HIP(hipMemcpy(a->solid, m.data(), a->n, hipMemcpyHostToDevice));
  YH_HIP(launch_codes(a->solid, a->codes, a->p.nx, a->p.ny, nullptr));
  YH_HIP(hipDeviceSynchronize());
  return YH_OK;
}

int yh_adjoint_set_initial(yh_adjoint* a, const float* u, const float* v, const float* w) {
  if (!a || !u || !v || !w) return YH_EINVAL;
  const float* s[3] = {u, v, w};
  for (int q = 0; q < 3; ++q) YH_HIP(hipMemcpy(a->X0[q], s[q], a->n * 4, hipMemcpyHostToDevice));
  a->forward_done = false;
  return YH_OK;
}

int yh_adjoint_set_targets(yh_adjoint* a, const float* targets) {
  if (!a || !targets) return YH_EINVAL;
  const size_t m = (size_t)a->nrec * a->ns;
  a->h_targets.assign(targets, targets + m);
  YH_HIP(hipMemcpy(a->targets, targets, m * 4, hipMemcpyHostToDevice));
  a->have_targets = true;
  return YH_OK;
}

int yh_adjoint_forward(yh_adjoint* a, const float* d_D, const double tau6[6], void* stream,
                       double* loss, float* traces) {
  if (!a || !d_D || !tau6) return YH_EINVAL;
  for (int q = 0; q < 6; ++q)
    if (!(tau6[q] > 0)) return YH_EINVAL;
  if (a->p.solidSwitch && !a->codes) return YH_ESTATE;
  hipStream_t s = (hipStream_t)stream;
  adj_k(a, tau6);
  a->D = d_D;
  const size_t n = a->n;
  dim3 bb(64, 4), bg((a->p.nx + 63) / 64, (a->p.ny + 3) / 4);
  // checkpoint 0 = X^0; a launch that ends on a later checkpoint step writes
  // its state into that checkpoint, which the next launch reads
  for (int q = 0; q < 3; ++q)
    YH_HIP(hipMemcpyAsync(slot(a->ck, n, 0, q), a->X0[q], n * 4, hipMemcpyDeviceToDevice, s));
  CFields in{{slot(a->ck, n, 0, 0), slot(a->ck, n, 0, 1), slot(a->ck, n, 0, 2)}};
  int nb = 0;  // the work buffer the next launch writes (never the one it reads)
  for (int done = 0; done < a->nsteps;) {
    // launches end on every checkpoint step
    const int K = std::min({a->fwdK, a->nsteps - done, (done / a->ckpt + 1) * a->ckpt - done});
    const int end = done + K;
    const bool to_ck = end % a->ckpt == 0 && end < a->nsteps;
    const int ci = end / a->ckpt;
    Fields out = to_ck ? Fields{{slot(a->ck, n, ci, 0), slot(a->ck, n, ci, 1), slot(a->ck, n, ci, 2)}}
                       : Fields{{a->X[nb][0], a->X[nb][1], a->X[nb][2]}};
    if (a->fwdK == 1) {
      const int step = done + 1;
      float* row = (step % a->rec == 0) ? a->traces + (size_t)(step / a->rec - 1) * a->ns : nullptr;
      hipLaunchKernelGGL(k_adj_fwd, bg, bb, 0, s, a->k, in, out, d_D, a->codes, a->smap, row);
      YH_HIP(hipGetLastError());
    } else {
      YH_HIP(launch_adj_fwd_tiled(a->k, in, out, d_D, a->codes, a->smap, K, done, a->rec,
                                  a->traces, a->ns, nullptr, n, 0, s));
    }
    in = CFields{{out.f[0], out.f[1], out.f[2]}};
    if (!to_ck) nb ^= 1;
    done = end;
  }
  const size_t m = (size_t)a->nrec * a->ns;
  a->h_traces.resize(m);
  YH_HIP(hipMemcpyAsync(a->h_traces.data(), a->traces, m * 4, hipMemcpyDeviceToHost, s));
  YH_HIP(hipStreamSynchronize(s));
  if (traces) std::memcpy(traces, a->h_traces.data(), m * 4);
  double L = 0.0;
  if (a->have_targets) {
    std::vector<float> src(m);
    for (size_t q = 0; q < m; ++q) {
      const double dlt = (double)a->h_traces[q] - (double)a->h_targets[q];
      L += dlt * dlt;
      src[q] = (float)(2.0 * dlt);
    }
    YH_HIP(hipMemcpyAsync(a->src, src.data(), m * 4, hipMemcpyHostToDevice, s));
    YH_HIP(hipStreamSynchronize(s));
  }
  if (loss) *loss = L;
  a->forward_done = true;
  return YH_OK;
}

int yh_adjoint_backward(yh_adjoint* a, void* stream, double* d_gD, double gtau[6]) {
  if (!a || !d_gD || !gtau) return YH_EINVAL;
  if (!a->forward_done || !a->have_targets) return YH_ESTATE;
  hipStream_t s = (hipStream_t)stream;
  const size_t n = a->n;
  const int N = a->nsteps;
  dim3 bb(64, 4), bg((a->p.nx + 63) / 64, (a->p.ny + 3) / 4);
  hipLaunchKernelGGL(k_fill_d, dim3(256), dim3(256), 0, s, a->gD, n, 0.0);
  hipLaunchKernelGGL(k_fill_d, dim3(64), dim3(256), 0, s, a->gtau_part, (size_t)a->nparts * 6, 0.0);
  // lambda^N
  const float* srcN = (N % a->rec == 0) ? a->src + (size_t)(N / a->rec - 1) * a->ns : nullptr;
  int lc = 0;
  Fields l0{{a->lam[0][0], a->lam[0][1], a->lam[0][2]}};
  hipLaunchKernelGGL(k_adj_init, dim3(256), dim3(256), 0, s, l0, n, (const int*)a->smap, srcN);
  YH_HIP(hipGetLastError());
  for (int seg = a->nckpt - 1; seg >= 0; --seg) {
    const int s0 = seg * a->ckpt;
    const int s1 = std::min(N, s0 + a->ckpt);  // reverse steps n = s1-1 .. s0
    // recompute X^{s0+1} .. X^{s1-1} into seg slots 1 .. s1-s0-1; X^{s0} is
    // read from its checkpoint (slot 0 stands for it)
    float* x0 = slot(a->ck, n, seg, 0);  // the checkpoint's three fields, n apart
    for (int t = 1; t < s1 - s0;) {
      CFields in = t == 1 ? CFields{{x0, x0 + n, x0 + 2 * n}}
                          : CFields{{slot(a->seg, n, t - 1, 0), slot(a->seg, n, t - 1, 1),
                                     slot(a->seg, n, t - 1, 2)}};
      Fields out{{slot(a->seg, n, t, 0), slot(a->seg, n, t, 1), slot(a->seg, n, t, 2)}};
      const int K = std::min(a->fwdK, s1 - s0 - t);
      if (a->fwdK == 1) {
        hipLaunchKernelGGL(k_adj_fwd, bg, bb, 0, s, a->k, in, out, a->D, a->codes, a->smap,
                           (float*)nullptr);
        YH_HIP(hipGetLastError());
      } else {  // slots t .. t+K-1
        YH_HIP(launch_adj_fwd_tiled(a->k, in, out, a->D, a->codes, a->smap, K, s0 + t - 1,
                                    a->rec, nullptr, a->ns, a->seg, n, t - 1, s));
      }
      t += K;
    }
    for (int step = s1 - 1; step >= s0;) {
      const int t = step - s0;
      CFields li{{a->lam[lc][0], a->lam[lc][1], a->lam[lc][2]}};
      Fields lo{{a->lam[lc ^ 1][0], a->lam[lc ^ 1][1], a->lam[lc ^ 1][2]}};
      if (a->fwdK == 1) {
        CFields X = t == 0 ? CFields{{x0, x0 + n, x0 + 2 * n}}
                           : CFields{{slot(a->seg, n, t, 0), slot(a->seg, n, t, 1),
                                      slot(a->seg, n, t, 2)}};
        const float* src = (step >= 1 && step % a->rec == 0)
                               ? a->src + (size_t)(step / a->rec - 1) * a->ns
                               : nullptr;
        hipLaunchKernelGGL(k_adj_rev, bg, bb, 0, s, a->k, X, li, lo, a->D, a->codes, a->smap,
                           src, a->gD, a->gtau_part);
        YH_HIP(hipGetLastError());
        step -= 1;
      } else {  // steps step, step-1, ..., step-K+1 of this segment
        const int K = std::min(a->fwdK, t + 1);
        YH_HIP(launch_adj_rev_tiled(a->k, a->seg, x0, n, t, step, li, lo, a->D, a->codes,
                                    a->smap, a->src, a->rec, a->ns, a->gD, a->gtau_part, K, s));
        step -= K;
      }
      lc ^= 1;
    }
  }
  YH_HIP(hipMemcpyAsync(d_gD, a->gD, n * 8, hipMemcpyDeviceToDevice, s));
  std::vector<double> part((size_t)a->nparts * 6);
  YH_HIP(hipMemcpyAsync(part.data(), a->gtau_part, part.size() * 8, hipMemcpyDeviceToHost, s));
  YH_HIP(hipStreamSynchronize(s));
  for (int q = 0; q < 6; ++q) gtau[q] = 0.0;
  for (int b = 0; b < a->nparts; ++b)
    for (int q = 0; q < 6; ++q) gtau[q] += part[(size_t)b * 6 + q];
  return YH_OK;
}

}  // extern "C"
