// yh_internal.h -- shared definitions of the HIP kernels and the C-ABI layer.
//
// Kernel constants travel as one by-value struct (KParams) instead of the
// reference's ~45 __constant__ symbols (main.cu:309-402), so one process can
// hold many plans and a batched launch can carry per-tissue parameters.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/yolohtli.h"

namespace yh {

// ------------------------------------------------------------ kernel params
struct KParams {
  int nx, ny;
  float rx, ry;       // dt*Dxx/hx^2, dt*Dyy/hy^2 (main.cu:149-150)
  float dt, tc, tcdt; // tcdt = tc*dt
  float rs;           // gate diffusion scale: rscale when gateDiff, else 0
  // 2V reference reaction (reactionDiffusion.cu:131,141)
  float mu, alpha, beta, gamma, delta, eps, theta;
  // Fenton-Karma (reciprocal time constants)
  float i_td, i_tr, i_2tsi, i_t0, i_tvp, i_tv1m, i_tv2m, i_twp, i_twm;
  float u_c, u_v, u_csi, kk;
  float hw;           // tanh-smoothed Heaviside width (0 = sharp)
};

// Fields of the state (u, v, w); unused entries are null.
struct Fields {
  float* f[3];
};
struct CFields {
  const float* f[3];
};

void make_kparams(const yh_params& p, const yh_derived& d, KParams* k);
int nvars_of(int model);

// ----------------------------------------------------------------- models
// Reaction part of one explicit-Euler step.  Each functor returns the
// increments -tc*dt*I for every variable; diffusion is added by the caller.
enum { M2V = 0, MFK = 1 };

__device__ __forceinline__ float fast_tanh(float x) {
  // tanh(x) = 1 - 2/(exp(2x)+1); saturates correctly for |x| large.
  const float e = __expf(2.0f * x);
  return 1.0f - 2.0f / (e + 1.0f);
}

template <int MODEL>
struct Model;

template <>
struct Model<M2V> {
  static constexpr int NV = 2;  // u, v
  static constexpr int ND = 2;  // both diffuse (v scaled by rs)
  __device__ __forceinline__ static void react(const KParams& k, float u, float v, float /*w*/,
                                               float& du, float& dv, float& /*dw*/) {
    // I_sum = -(mu u (1-u)(u-alpha) - u v);  I_v = -(eps (delta (u-gamma)(beta-u) - v - theta))
    const float Is = -(k.mu * u * (1.0f - u) * (u - k.alpha) - u * v);
    const float Iv = -(k.eps * (k.delta * (u - k.gamma) * (k.beta - u) - v - k.theta));
    du = -k.tcdt * Is;
    dv = -k.tcdt * Iv;
  }
};

template <>
struct Model<MFK> {
  static constexpr int NV = 3;  // u, v, w
  static constexpr int ND = 1;  // only u diffuses
  __device__ __forceinline__ static void react(const KParams& k, float u, float v, float w,
                                               float& du, float& dv, float& dw) {
    // Fenton & Karma 1998 with the sharp Heaviside STEP(a,b) = a >= b.
    const bool p = u >= k.u_c;
    const bool q = u >= k.u_v;
    const float Jfi = p ? -v * (1.0f - u) * (u - k.u_c) * k.i_td : 0.0f;
    const float Jso = p ? k.i_tr : u * k.i_t0;
    const float th = fast_tanh(k.kk * (u - k.u_csi));
    const float Jsi = -w * (1.0f + th) * k.i_2tsi;
    const float Gv = p ? v * k.i_tvp : -(1.0f - v) * (q ? k.i_tv2m : k.i_tv1m);
    const float Gw = p ? w * k.i_twp : -(1.0f - w) * k.i_twm;
    du = -k.tcdt * (Jfi + Jso + Jsi);
    dv = -k.tcdt * Gv;
    dw = -k.tcdt * Gw;
  }
};

// ------------------------------------------------------------ index helpers
// coord_i / coord_j (helper_functions.cu:69-79): mirror about the edge cells.
__host__ __device__ __forceinline__ int mirror(int k, int n) {
  return k < 0 ? -k : (k >= n ? 2 * (n - 1) - k : k);
}
// Mirror, then clamp: used for loads of halo cells that may lie further out
// than one reflection (those cells never reach an output).
__device__ __forceinline__ int reflect_clamp(int k, int n) {
  int m = mirror(k, n);
  return m < 0 ? 0 : (m >= n ? n - 1 : m);
}

// Stencil code of one cell, built once per mask: 2-bit coefficients
// cW | cE<<2 | cN<<4 | cS<<6 of reactionDiffusion.cu:162-169 written as
// sum_dir c_dir*(u_dir - u); 0xFF marks a hole.
constexpr uint8_t kHole = 0xFF;

// ------------------------------------------------------------- launchers
hipError_t launch_codes(const uint8_t* solid, uint8_t* codes, int nx, int ny, hipStream_t s);
hipError_t launch_direct(int model, const KParams& k, CFields in, Fields out, Fields vt,
                         const uint8_t* solid, hipStream_t s);
hipError_t launch_direct_codes(int model, const KParams& k, CFields in, Fields out,
                               const uint8_t* codes, hipStream_t s);
hipError_t launch_tiled(int model, bool masked, int rows, int waves, int K, const KParams& k,
                        CFields in, Fields out, const uint8_t* codes, hipStream_t s);
bool tiled_supported(int rows, int waves);
hipError_t launch_tips(const KParams& k, int masked_disk, int tipOffsetX, int tipOffsetY,
                       double Uth, const float* past, const float* present, double time,
                       yh_tip* tips, int* keys, int* count, int cap, hipStream_t s);
hipError_t launch_finite(const float* const* f, int nf, size_t n, int* flag, hipStream_t s);

// Map a HIP error to a yh code, remembering the message.
int hip_fail(hipError_t e, const char* what);
void set_last_error(const char* msg);

}  // namespace yh

#define YH_HIP(call)                                        \
  do {                                                      \
    hipError_t yh_e_ = (call);                              \
    if (yh_e_ != hipSuccess) return yh::hip_fail(yh_e_, #call); \
  } while (0)
