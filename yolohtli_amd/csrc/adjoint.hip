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
};

struct FkEval {
  float I, Gv, Gw;
  float dIdu, dIdv, dIdw, dGvdu, dGvdv, dGwdu, dGwdw;
  float dI_dtau[4];  // tau_d, tau_r, tau_si, tau_0
  float dGv_dtvp, dGw_dtwp;
};

__device__ __forceinline__ void heav(float x, float w, float* h, float* dh) {
  const float t = tanhf(x / w);
  *h = 0.5f * (1.0f + t);
  *dh = 0.5f / w * (1.0f - t * t);
}

// Fenton & Karma 1998 currents and derivatives (smoothed Heaviside).
__device__ FkEval fk_eval(const AdjK& k, float u, float v, float w) {
  FkEval r;
  float p, dp, q, dq;
  heav(u - k.uc, k.hw, &p, &dp);
  heav(u - k.uv, k.hw, &q, &dq);
  const float f = (1.0f - u) * (u - k.uc);
  const float df = 1.0f + k.uc - 2.0f * u;
  const float Jfi = -v * p * f / k.td;
  const float Jso = u * (1.0f - p) / k.t0 + p / k.tr;
  const float th = tanhf(k.kk * (u - k.ucsi));
  const float Jsi = -w * (1.0f + th) / (2.0f * k.tsi);
  r.I = Jfi + Jso + Jsi;
  r.dIdu = -v * (dp * f + p * df) / k.td + (1.0f - p) / k.t0 - u * dp / k.t0 + dp / k.tr -
           w * k.kk * (1.0f - th * th) / (2.0f * k.tsi);
  r.dIdv = -p * f / k.td;
  r.dIdw = -(1.0f + th) / (2.0f * k.tsi);
  const float tvm = (1.0f - q) * k.tv1m + q * k.tv2m;
  const float dtvm = dq * (k.tv2m - k.tv1m);
  const float A = (1.0f - p) * (1.0f - v) / tvm - p * v / k.tvp;
  const float dAdu = -dp * (1.0f - v) / tvm - (1.0f - p) * (1.0f - v) * dtvm / (tvm * tvm) -
                     dp * v / k.tvp;
  const float dAdv = -(1.0f - p) / tvm - p / k.tvp;
  const float B = (1.0f - p) * (1.0f - w) / k.twm - p * w / k.twp;
  const float dBdu = -dp * (1.0f - w) / k.twm - dp * w / k.twp;
  const float dBdw = -(1.0f - p) / k.twm - p / k.twp;
  r.Gv = -A;
  r.Gw = -B;
  r.dGvdu = -dAdu;
  r.dGvdv = -dAdv;
  r.dGwdu = -dBdu;
  r.dGwdw = -dBdw;
  r.dI_dtau[0] = -Jfi / k.td;
  r.dI_dtau[1] = -p / (k.tr * k.tr);
  r.dI_dtau[2] = -Jsi / k.tsi;
  r.dI_dtau[3] = -u * (1.0f - p) / (k.t0 * k.t0);
  r.dGv_dtvp = -p * v / (k.tvp * k.tvp);
  r.dGw_dtwp = -p * w / (k.twp * k.twp);
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

// ---------------------------------------------------------------- forward
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
  const float u = U[c], v = in.f[1][c], w = in.f[2][c];
  const float Dc = D[c];
  const float lap = k.rxh * (cf.W * 0.5f * (Dc + D[W]) * (U[W] - u) +
                             cf.E * 0.5f * (Dc + D[E]) * (U[E] - u)) +
                    k.ryh * (cf.N * 0.5f * (Dc + D[N]) * (U[N] - u) +
                             cf.S * 0.5f * (Dc + D[S]) * (U[S] - u));
  const FkEval r = fk_eval(k, u, v, w);
  const float un = cf.sc ? u + k.tc * (lap - k.dt * r.I) : 0.0f;
  out.f[0][c] = un;
  out.f[1][c] = cf.sc ? v + k.tc * (-k.dt * r.Gv) : 0.0f;
  out.f[2][c] = cf.sc ? w + k.tc * (-k.dt * r.Gw) : 0.0f;
  if (trace_row) {
    const int s = smap[c];
    if (s >= 0) trace_row[s] = un;
  }
}

// ---------------------------------------------------------------- reverse
// One reverse step: lam_out = J(X)^T lam_in (+ sensor source), gD and gtau
// accumulated.  X is the state the forward step started from.
__global__ __launch_bounds__(256) void k_adj_rev(const AdjK k, CFields X, CFields lam_in,
                                                 Fields lam_out, const float* __restrict__ D,
                                                 const uint8_t* __restrict__ codes,
                                                 const int* __restrict__ smap,
                                                 const float* __restrict__ src_row,
                                                 double* __restrict__ gD,
                                                 double* __restrict__ gtau_part) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  float gt[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  if (i < k.nx && j < k.ny) {
    const int nx = k.nx, ny = k.ny;
    const int c = i + nx * j;
    const float* U = X.f[0];
    const float* LU = lam_in.f[0];
    const Coef4 cf = coef_of(codes, c);
    const float u = U[c], v = X.f[1][c], w = X.f[2][c];
    const float m = cf.sc ? 1.0f : 0.0f;
    const float mu = m * LU[c], mv = m * lam_in.f[1][c], mw = m * lam_in.f[2][c];
    const float Dc = D[c];
    const float tc = k.tc;
    // own diffusion terms
    const int nb[4] = {mirror(i - 1, nx) + nx * j, mirror(i + 1, nx) + nx * j,
                       i + nx * mirror(j + 1, ny), i + nx * mirror(j - 1, ny)};
    const float co[4] = {cf.W, cf.E, cf.N, cf.S};
    const float rr[4] = {k.rxh, k.rxh, k.ryh, k.ryh};
    float asum = 0.0f;
    double gd = 0.0;
#pragma unroll
    for (int d = 0; d < 4; ++d) {
      asum += rr[d] * co[d] * 0.5f * (Dc + D[nb[d]]);
      gd += (double)(tc * mu * rr[d] * co[d] * 0.5f * (U[nb[d]] - u));
    }
    // coefficients pointing at this cell from its neighbours (gather)
    float lin = 0.0f;
    auto from = [&](int cc, float coef_c, float r) {
      const float mc = LU[cc] * (coef_c != 0.0f ? 1.0f : 0.0f);  // coef 0 also covers holes
      const float a = r * coef_c * 0.5f * (D[cc] + Dc);
      lin += tc * mc * a;
      gd += (double)(tc * mc * r * coef_c * 0.5f * (u - U[cc]));
    };
    if (i + 1 < nx) { const int cc = c + 1; from(cc, coef_of(codes, cc).W, k.rxh); }
    if (i - 1 >= 0) { const int cc = c - 1; from(cc, coef_of(codes, cc).E, k.rxh); }
    if (i == 1) { const int cc = c - 1; from(cc, coef_of(codes, cc).W, k.rxh); }
    if (i == nx - 2) { const int cc = c + 1; from(cc, coef_of(codes, cc).E, k.rxh); }
    if (j + 1 < ny) { const int cc = c + nx; from(cc, coef_of(codes, cc).S, k.ryh); }
    if (j - 1 >= 0) { const int cc = c - nx; from(cc, coef_of(codes, cc).N, k.ryh); }
    if (j == 1) { const int cc = c - nx; from(cc, coef_of(codes, cc).S, k.ryh); }
    if (j == ny - 2) { const int cc = c + nx; from(cc, coef_of(codes, cc).N, k.ryh); }
    const FkEval r = fk_eval(k, u, v, w);
    const float tdt = tc * k.dt;
    float lu = mu * (1.0f - tc * asum - tdt * r.dIdu) + lin - mv * tdt * r.dGvdu -
               mw * tdt * r.dGwdu;
    const float lv = -mu * tdt * r.dIdv + mv * (1.0f - tdt * r.dGvdv);
    const float lw = -mu * tdt * r.dIdw + mw * (1.0f - tdt * r.dGwdw);
    if (src_row) {
      const int s = smap[c];
      if (s >= 0) lu += src_row[s];
    }
    lam_out.f[0][c] = lu;
    lam_out.f[1][c] = lv;
    lam_out.f[2][c] = lw;
    gD[c] += gd;
#pragma unroll
    for (int q = 0; q < 4; ++q) gt[q] = mu * (-tdt * r.dI_dtau[q]);
    gt[4] = mv * (-tdt * r.dGv_dtvp);
    gt[5] = mw * (-tdt * r.dGw_dtwp);
  }
  // per-workgroup sum of the tau gradients, added to this workgroup's slot
  __shared__ double red[4][6];
  const int tid = threadIdx.y * blockDim.x + threadIdx.x;
  const int lane = tid & 63, wv = tid >> 6;
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    double s = (double)gt[q];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if (lane == 0) red[wv][q] = s;
  }
  __syncthreads();
  if (tid < 6) {
    double s = red[0][tid] + red[1][tid] + red[2][tid] + red[3][tid];
    gtau_part[(size_t)(blockIdx.y * gridDim.x + blockIdx.x) * 6 + tid] += s;
  }
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
  dim3 bb(64, 4), bg((p->nx + 63) / 64, (p->ny + 3) / 4);
  a->nblocks = bg.x * bg.y;
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
  M((void**)&a->gtau_part, (size_t)a->nblocks * 6 * 8);
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

int yh_adjoint_set_mask(yh_adjoint* a, const uint8_t* solid) {
  if (!a || !solid) return YH_EINVAL;
  if (!a->p.solidSwitch) return YH_ESTATE;
  if (!a->solid) {
    YH_HIP(hipMalloc(&a->solid, a->n));
    YH_HIP(hipMalloc(&a->codes, a->n));
  }
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
  for (int q = 0; q < 3; ++q)
    YH_HIP(hipMemcpyAsync(a->X[0][q], a->X0[q], n * 4, hipMemcpyDeviceToDevice, s));
  int cur = 0;
  for (int step = 1; step <= a->nsteps; ++step) {
    if ((step - 1) % a->ckpt == 0) {  // checkpoint X^{step-1}
      const int ci = (step - 1) / a->ckpt;
      for (int q = 0; q < 3; ++q)
        YH_HIP(hipMemcpyAsync(slot(a->ck, n, ci, q), a->X[cur][q], n * 4,
                              hipMemcpyDeviceToDevice, s));
    }
    float* row = (step % a->rec == 0) ? a->traces + (size_t)(step / a->rec - 1) * a->ns : nullptr;
    CFields in{{a->X[cur][0], a->X[cur][1], a->X[cur][2]}};
    Fields out{{a->X[cur ^ 1][0], a->X[cur ^ 1][1], a->X[cur ^ 1][2]}};
    hipLaunchKernelGGL(k_adj_fwd, bg, bb, 0, s, a->k, in, out, d_D, a->codes, a->smap, row);
    YH_HIP(hipGetLastError());
    cur ^= 1;
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
  hipLaunchKernelGGL(k_fill_d, dim3(64), dim3(256), 0, s, a->gtau_part, (size_t)a->nblocks * 6, 0.0);
  // lambda^N
  const float* srcN = (N % a->rec == 0) ? a->src + (size_t)(N / a->rec - 1) * a->ns : nullptr;
  int lc = 0;
  Fields l0{{a->lam[0][0], a->lam[0][1], a->lam[0][2]}};
  hipLaunchKernelGGL(k_adj_init, dim3(256), dim3(256), 0, s, l0, n, (const int*)a->smap, srcN);
  YH_HIP(hipGetLastError());
  for (int seg = a->nckpt - 1; seg >= 0; --seg) {
    const int s0 = seg * a->ckpt;
    const int s1 = std::min(N, s0 + a->ckpt);  // reverse steps n = s1-1 .. s0
    // recompute X^{s0} .. X^{s1-1} into seg slots 0 .. s1-s0-1
    for (int q = 0; q < 3; ++q)
      YH_HIP(hipMemcpyAsync(slot(a->seg, n, 0, q), slot(a->ck, n, seg, q), n * 4,
                            hipMemcpyDeviceToDevice, s));
    for (int t = 1; t < s1 - s0; ++t) {
      CFields in{{slot(a->seg, n, t - 1, 0), slot(a->seg, n, t - 1, 1), slot(a->seg, n, t - 1, 2)}};
      Fields out{{slot(a->seg, n, t, 0), slot(a->seg, n, t, 1), slot(a->seg, n, t, 2)}};
      hipLaunchKernelGGL(k_adj_fwd, bg, bb, 0, s, a->k, in, out, a->D, a->codes, a->smap,
                         (float*)nullptr);
      YH_HIP(hipGetLastError());
    }
    for (int step = s1 - 1; step >= s0; --step) {
      const int t = step - s0;
      CFields X{{slot(a->seg, n, t, 0), slot(a->seg, n, t, 1), slot(a->seg, n, t, 2)}};
      CFields li{{a->lam[lc][0], a->lam[lc][1], a->lam[lc][2]}};
      Fields lo{{a->lam[lc ^ 1][0], a->lam[lc ^ 1][1], a->lam[lc ^ 1][2]}};
      const float* src = (step >= 1 && step % a->rec == 0)
                             ? a->src + (size_t)(step / a->rec - 1) * a->ns
                             : nullptr;
      hipLaunchKernelGGL(k_adj_rev, bg, bb, 0, s, a->k, X, li, lo, a->D, a->codes, a->smap, src,
                         a->gD, a->gtau_part);
      YH_HIP(hipGetLastError());
      lc ^= 1;
    }
  }
  YH_HIP(hipMemcpyAsync(d_gD, a->gD, n * 8, hipMemcpyDeviceToDevice, s));
  std::vector<double> part((size_t)a->nblocks * 6);
  YH_HIP(hipMemcpyAsync(part.data(), a->gtau_part, part.size() * 8, hipMemcpyDeviceToHost, s));
  YH_HIP(hipStreamSynchronize(s));
  for (int q = 0; q < 6; ++q) gtau[q] = 0.0;
  for (int b = 0; b < a->nblocks; ++b)
    for (int q = 0; q < 6; ++q) gtau[q] += part[(size_t)b * 6 + q];
  return YH_OK;
}

}  // extern "C"
