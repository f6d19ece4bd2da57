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
  float trx, try_;    // tc*rx, tc*ry
  float dt, tc, tcdt; // tcdt = tc*dt
  float rs;           // gate diffusion scale: rscale when gateDiff, else 0
  // 2V reference reaction (reactionDiffusion.cu:131,141), folded with tc*dt:
  //   -tcdt*I_sum = u*(a2_A*(1-u)*(u-alpha) - a2_B*v)
  //   -tcdt*I_v   = a2_E*(u-gamma)*(beta-u) - a2_F*v - a2_G
  float alpha, beta, gamma;
  float a2_A, a2_B, a2_E, a2_F, a2_G;
  // Fenton-Karma, folded with tc*dt (tau's as reciprocals)
  float u_c, u_v;
  float fk_A1, fk_A2, fk_A3, fk_A4;  // tcdt/tau_d, tcdt/tau_r, tcdt/tau_0, tcdt/tau_si
  float fk_b, fk_bu;                 // 2k*log2(e), 2k*log2(e)*u_csi
  // J_si with tcdt/tau_si folded into the sigmoid: A4/(1 + 2^(bu - b u)) =
  // 1/(2^(bs - b u) + 1/A4) with bs = bu - log2(A4)
  float fk_bs, fk_iA4;
  // A1 (1-u)(u-u_c) = u (P1 - A1 u) + P0: P1 = A1 (1 + u_c), P0 = -A1 u_c
  float fk_P1, fk_P0;
  float fk_B1, fk_B2, fk_B3;         // tcdt/tau_vp, tcdt/tau_v1m, tcdt/tau_v2m
  float fk_C1, fk_C2;                // tcdt/tau_wp, tcdt/tau_wm
  float hw;                          // tanh-smoothed Heaviside width (gradient path)
  int iso;                           // trx == try: one stencil weight for both axes
  int wt;                            // host side: the launcher picks the write-through kernel
};

// A store of one float of field `base` (element idx < 2^30) written through
// to memory (the sc1 policy: coherent at device scope) instead of left dirty
// in this XCD's L2, where it would be written back at the end of the launch,
// on the critical path before the next launch.  The descriptor covers 4 GiB
// from base (raw buffer, 32-bit elements: the third-dword format gfx9
// descriptors use), so the unsigned byte offset of any element of a
// 2^30-cell grid is in range.
__device__ __forceinline__ void store_through(float* base, uint32_t idx, float v) {
  __builtin_amdgcn_raw_buffer_store_b32(
      __float_as_uint(v), __builtin_amdgcn_make_buffer_rsrc(base, 0, -1, 0x00020000),
      (int)(idx * 4u), 0, 16 /* sc1 */);
}

// Fields of the state (u, v, w); unused entries are null.
struct Fields {
  float* f[3];
};
struct CFields {
  const float* f[3];
};

void make_kparams(const yh_params& p, const yh_derived& d, KParams* k);
int nvars_of(int model);

// Per-cell action-potential-duration state of sAPD_kernel (spaceAPD.cu:278-374):
// threshold-crossing times front/back and the two alternating durations.
struct ApdState {
  float* front = nullptr;
  float* back = nullptr;
  float* apd1 = nullptr;
  float* apd2 = nullptr;
  uint8_t* first = nullptr;
  float th = 0.15f;  // apdTh, spaceAPD.cu:288
  float dt = 0.0f;
  int count0 = 0;    // step count before the launch
};

// One sAPD update (spaceAPD.cu:328-358, stimulus off) for a cell going from
// uo to un at step `count`.
__device__ __forceinline__ void apd_cell(float uo, float un, int count, float dt, float th,
                                         float& front, float& back, float& apd1, float& apd2,
                                         uint32_t& first) {
  if ((un > th) && (uo < th)) front = dt * ((float)count - (un - th) / (un - uo));
  if ((un < th) && (uo > th)) back = dt * ((float)count - (un - th) / (un - uo));
  if ((back > 0.0f) && (front > 0.0f) && (first == 0u)) {
    apd1 = back - front;
    front = 0.0f;
    back = 0.0f;
    first = 1u;
  }
  if ((back > 0.0f) && (front > 0.0f) && (first != 0u)) {
    apd2 = back - front;
    front = 0.0f;
    back = 0.0f;
    first = 0u;
  }
}

// ----------------------------------------------------------------- models
// Reaction part of one explicit-Euler step: each functor returns the
// increments -tc*dt*I of every variable; diffusion is added by the caller.
// Selects are computed on both branches so the compiler emits v_cndmask, not
// divergent branches.
enum { M2V = 0, MFK = 1 };

template <int MODEL>
struct Model;

template <>
struct Model<M2V> {
  static constexpr int NV = 2;  // u, v
  static constexpr int ND = 2;  // both diffuse (v scaled by rs)
  __device__ __forceinline__ static void react(const KParams& k, float u, float v, float /*w*/,
                                               float& du, float& dv, float& /*dw*/) {
    // I_sum = -(mu u (1-u)(u-alpha) - u v);  I_v = -(eps (delta (u-gamma)(beta-u) - v - theta))
    const float s2 = __builtin_fmaf(u, -k.a2_A, k.a2_A);  // A (1-u)
    const float s3 = (u - k.alpha) * s2;
    du = u * __builtin_fmaf(-k.a2_B, v, s3);
    const float t3 = (u - k.gamma) * (k.beta - u);
    dv = __builtin_fmaf(-k.a2_F, v, __builtin_fmaf(k.a2_E, t3, -k.a2_G));
  }
  // The Euler update of one cell in the tiled kernels: ud = u plus its
  // diffusion term, vd = v's diffusion term.  u' = ud + u (s3 - B v) is one
  // written-out fma, so no product is left for the compiler to contract with
  // the caller's code: every inlined copy (the first substep is peeled, and
  // each tile shape is its own instantiation) rounds a cell identically.
  __device__ __forceinline__ static void step(const KParams& k, float ud, float vd, float& u,
                                              float& v) {
    const float s2 = __builtin_fmaf(u, -k.a2_A, k.a2_A);
    const float s3 = (u - k.alpha) * s2;
    const float t3 = (u - k.gamma) * (k.beta - u);
    const float dv = __builtin_fmaf(-k.a2_F, v, __builtin_fmaf(k.a2_E, t3, -k.a2_G));
    const float un = __builtin_fmaf(u, __builtin_fmaf(-k.a2_B, v, s3), ud);
    v = v + __builtin_fmaf(k.rs, vd, dv);
    u = un;
  }
  struct Pinned {};
  __device__ __forceinline__ static Pinned pin(const KParams&) { return {}; }
};

template <>
struct Model<MFK> {
  static constexpr int NV = 3;  // u, v, w
  static constexpr int ND = 1;  // only u diffuses
  __device__ __forceinline__ static void react(const KParams& k, float u, float v, float w,
                                               float& du, float& dv, float& dw) {
    // Fenton & Karma 1998 with the sharp Heaviside STEP(a,b) = a >= b:
    //   J_fi = -p v (1-u)(u-u_c)/tau_d,  J_so = (1-p) u/tau_0 + p/tau_r,
    //   J_si = -w (1 + tanh(k(u-u_csi)))/(2 tau_si) = -w sigmoid(2k(u-u_csi))/tau_si
    const bool p = u >= k.u_c;
    const bool q = u >= k.u_v;
    const float pa = __builtin_fmaf(u, __builtin_fmaf(u, -k.fk_A1, k.fk_P1), k.fk_P0);
    const float fi = __builtin_fmaf(pa, v, -k.fk_A2);                // -tcdt (J_fi + 1/tau_r)
    const float so = -k.fk_A3 * u;                                   // -tcdt u/tau_0
    const float r1 = p ? fi : so;
    const float ex = __builtin_amdgcn_exp2f(__builtin_fmaf(u, -k.fk_b, k.fk_bs));
    const float sg = __builtin_amdgcn_rcpf(ex + k.fk_iA4);           // A4 sigmoid(2k(u-u_csi))
    du = __builtin_fmaf(w, sg, r1);
    const float bq = q ? k.fk_B3 : k.fk_B2;
    const float vo = __builtin_fmaf(-bq, v, bq);                     // (1-v) tcdt/tau_v-
    const float vc = -k.fk_B1 * v;
    dv = p ? vc : vo;
    const float wo = __builtin_fmaf(-k.fk_C2, w, k.fk_C2);
    const float wc = -k.fk_C1 * w;
    dw = p ? wc : wo;
  }

  // Uniform constants held in vector registers for a whole tile: each meets
  // another uniform operand in one instruction (a select between two
  // constants, an fma with two constant operands), and a VALU instruction
  // reads at most one scalar register.  The empty asm makes the values opaque
  // so they stay in VGPRs instead of being re-materialised per cell.
  // C1, C2 also stop a select between two fields of a KParams held in
  // memory (the batch table) from becoming a per-cell load of the selected
  // field.
  struct Pinned {
    float bs, B2, B3, P1, C1, C2;
  };
  __device__ __forceinline__ static Pinned pin(const KParams& k) {
    Pinned q{k.fk_bs, k.fk_B2, k.fk_B3, k.fk_P1, k.fk_C1, k.fk_C2};
    asm volatile("" : "+v"(q.bs), "+v"(q.B2), "+v"(q.B3), "+v"(q.P1), "+v"(q.C1), "+v"(q.C2));
    return q;
  }
  // One whole explicit-Euler step of a cell, the same update as react():
  // ud = u plus its diffusion term; u <- ud - tcdt*I_ion, and v, w take their
  // gate steps as v + b*(1-v) = fma(b, 1-v, v) (opening) and
  // v - b*v = fma(-b, v, v) (closing), likewise w.  These keep the fixed
  // points 1 and 0 exact and carry b at its own precision; forms with the
  // folded constant 1-b round it to half an ulp of 1, which for b ~ 1e-3
  // is a relative error of 1e-5 in the rate, and moves the opening fixed
  // point of fma(v, 1-b, b) by (ulp/2)/b.
  __device__ __forceinline__ static void step(const KParams& k, const Pinned& q, float ud,
                                              float& u, float& v, float& w) {
    const bool p = u >= k.u_c;
    const bool qv = u >= k.u_v;
    const float pa = __builtin_fmaf(u, __builtin_fmaf(u, -k.fk_A1, q.P1), k.fk_P0);
    const float fi = __builtin_fmaf(pa, v, -k.fk_A2);
    const float so = -k.fk_A3 * u;
    const float r1 = p ? fi : so;
    const float ex = __builtin_amdgcn_exp2f(__builtin_fmaf(u, -k.fk_b, q.bs));
    const float sg = __builtin_amdgcn_rcpf(ex + k.fk_iA4);  // A4 sigmoid(2k(u-u_csi))
    u = ud + __builtin_fmaf(w, sg, r1);
    // opening fma(b, 1 - v, v), closing fma(b, 0 - v, v) = fma(-b, v, v):
    // one fma per gate with the rate and the target (1 or 0) selected
    const float bq = qv ? q.B3 : q.B2;
    const float tgt = p ? 0.0f : 1.0f;
    const float gv = p ? k.fk_B1 : bq;
    v = __builtin_fmaf(gv, tgt - v, v);
    const float gw = p ? q.C1 : q.C2;
    w = __builtin_fmaf(gw, tgt - w, w);
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

// ------------------------------------------------------------------- tips
// spiralTip_kernel + tipRecordPlane (tipTracker.cu:20-241) in FP64 from the
// FP32 fields, appending {x, y, 0, 0, t} and the key 2*cell+root.
__device__ __forceinline__ void tip_record(int i, int j, float tx, float ty, float t, int key,
                                           yh_tip* tips, int* keys, int* count, int cap) {
  if ((tx > 0.0f) && (tx < 1.0f) && (ty > 0.0f) && (ty < 1.0f)) {
    const int slot = atomicAdd(count, 1);
    if (slot < cap) {
      yh_tip d;
      d.x = (float)i + tx;
      d.y = (float)j + ty;
      d.vx = 0.0f;
      d.vy = 0.0f;
      d.t = t;
      tips[slot] = d;
      keys[slot] = key;
    }
  }
}

// Whether cell (i, j) takes part in the tip search.  masked_disk (the
// solidSwitch branch, tipTracker.cu:47-51): only cells of the disk
// (i-nx/2)^2 + (j-ny/2)^2 < r2; otherwise the interior 1 <= i < nx-2,
// 1 <= j < ny-2 (:126).
__device__ __forceinline__ bool tip_inside(int i, int j, int nx, int ny, bool masked_disk, int r2) {
  if (masked_disk) {
    const int ic = i - nx / 2, jc = j - ny / 2;
    return (ic * ic + jc * jc) < r2;
  }
  return (i >= 1) && (i < nx - 2) && (j >= 1) && (j < ny - 2);
}

// Tips of cell (i, j) between the isolines u = Uth of past and present from
// the corner values (1 = (i,j), 2 = east, 3 = north-east, 4 = north, each
// replaced by the cell itself past the last column/row).  masked_disk: both
// roots recorded without a discriminant test (tipTracker.cu:92-117);
// otherwise roots recorded when disc >= 0 (:126-197).
__device__ __forceinline__ void tip_solve(int i, int j, int nx, bool masked_disk, double Uth,
                                          double x1, double x2, double x3, double x4, double y1,
                                          double y2, double y3, double y4, float t, yh_tip* tips,
                                          int* keys, int* count, int cap) {
#pragma clang fp contract(off)
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
  if (!masked_disk && !(disc >= 0.0)) return;
  const int s0 = i + nx * j;
  tip_record(i, j, (float)((px + disc) / den1), (float)((py + disc) / den2), t, 2 * s0, tips,
             keys, count, cap);
  tip_record(i, j, (float)((px - disc) / den1), (float)((py - disc) / den2), t, 2 * s0 + 1, tips,
             keys, count, cap);
}

// Tips of cell (i, j) between the isolines u = Uth of past and present.
__device__ __forceinline__ void tip_cell(int i, int j, int nx, int ny, bool masked_disk, int r2,
                                         double Uth, const float* __restrict__ past,
                                         const float* __restrict__ present, float t,
                                         yh_tip* tips, int* keys, int* count, int cap) {
  if (!tip_inside(i, j, nx, ny, masked_disk, r2)) return;
  const int s0 = i + nx * j;
  const int sx = (i < nx - 1) ? s0 + 1 : s0;
  const int sy = (j < ny - 1) ? s0 + nx : s0;
  const int sxy = ((j < ny - 1) && (i < nx - 1)) ? s0 + nx + 1 : s0;
  tip_solve(i, j, nx, masked_disk, Uth, present[s0], present[sx], present[sxy], present[sy],
            past[s0], past[sx], past[sxy], past[sy], t, tips, keys, count, cap);
}

// Stencil code of one cell, built once per mask: 2-bit coefficients
// cW | cE<<2 | cN<<4 | cS<<6 of reactionDiffusion.cu:162-169 written as
// sum_dir c_dir*(u_dir - u); 0xFF marks a hole.
constexpr uint8_t kHole = 0xFF;

// ------------------------------------------------- other stepper branches
// Constants of the general path (general.hip): Dirichlet, lap4, anisotropy,
// RK2/RK4, mouse stimulus, unfolded as the reference writes them.
struct GParams {
  int nx, ny, masked, neumann, lap4, aniso, vdiff;
  float rx, ry, rxy, rbx, rby, q4, fx4, fy4, dt, tc, rscale, bval;  // q4 = qx4 + qy4
  float mu, alpha, beta, gamma, delta, eps, theta;
  float u_c, u_v, u_csi, kk, tau_d, tau_r, tau_si, tau_0, tau_vp, tau_v1m, tau_v2m, tau_wp,
      tau_wm;
  int stim_on, stim_x, stim_y;
};
GParams make_gparams(const yh_params& p, int stim_on, int stim_x, int stim_y);
// The branches the fast kernels implement: Neumann, Euler, and the 5-point
// Laplacian (with a mask, lap4 and the anisotropic term have no effect in the
// reference, reactionDiffusion.cu:154-185).
inline bool standard_branch(const yh_params& p) {
  return p.neumannBC && p.timeIntOrder == 1 && (p.solidSwitch || (!p.lap4 && !p.anisotropy));
}
// One step of the general path: `order` stage launches (1, 2 or 4); scratch
// holds 9 device arrays of n floats (stage increments x2, rhs) when order > 1.
hipError_t launch_general_step(int model, const GParams& g, int order, CFields in, Fields out,
                               Fields vt, const uint8_t* solid, float* const* scratch,
                               hipStream_t s, const ApdState* apd);

// ------------------------------------------------------------- launchers
hipError_t launch_codes(const uint8_t* solid, uint8_t* codes, int nx, int ny, hipStream_t s);
hipError_t launch_direct(int model, const KParams& k, CFields in, Fields out, Fields vt,
                         const uint8_t* solid, hipStream_t s);
hipError_t launch_direct_codes(int model, const KParams& k, CFields in, Fields out,
                               const uint8_t* codes, hipStream_t s,
                               const ApdState* apd = nullptr);
// prev (optional): u of the output cells after K-1 substeps, for the tip test.
// occ (optional): instead of launching, the workgroups of the instantiation
// this call would launch that one CU holds at once.
hipError_t launch_tiled(int model, bool masked, int rows, int waves, int cols, int K,
                        const KParams& k, CFields in, Fields out, const uint8_t* codes,
                        hipStream_t s, const ApdState* apd = nullptr, float* prev = nullptr,
                        int* occ = nullptr);
bool tiled_supported(int rows, int waves, int cols);
bool tiled_apd_supported(int rows, int waves, int cols);
bool tiled_prev_supported(int rows, int waves, int cols);
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
