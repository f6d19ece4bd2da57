// general.hip -- the reference kernel's other branches on the device:
// Dirichlet boundaries, the 4th-order compact Laplacian (lap4), the
// anisotropic cross term, RK2 / RK4 and the circular mouse stimulus
// (reactionDiffusion.cu:36-561).  One launch per Runge-Kutta stage, one
// thread per cell.
//
// The reference writes each stage state into its input while neighbours read
// it (:114-119) and reads lap4's neighbour currents J from the same launch
// (:221-236).  Here a stage reads only the step's input y0 and the previous
// stage's increments d: the stage field is y0 + ki*d_{s-1} (computed where it
// is read), a neighbour's current is evaluated from that neighbour's stage
// state, and the increments of all cells are complete before the next stage
// starts.  Reads outside the domain that the reference makes in the masked
// Dirichlet branch (:364-373 at the domain edge) count as non-tissue.
#include <hip/hip_runtime.h>

#include <cmath>

#include "plan.h"

namespace yh {

GParams make_gparams(const yh_params& p, int stim_on, int stim_x, int stim_y) {
  // parameterSetup (saveFiles.cu:144-170) and the startup recomputation after
  // the symmetry-reduction halving (main.cu:148-155)
  const double pi = 3.14159265359;  // globalVariables.cuh:27
  const double hx = p.Lx / (p.nx - 1.0), hy = p.Ly / (p.ny - 1.0);
  const double th = p.degrad * pi / 180.0;
  const double Dxx = p.diff_par * std::cos(th) * std::cos(th) + p.diff_per * std::sin(th) * std::sin(th);
  const double Dyy = p.diff_par * std::sin(th) * std::sin(th) + p.diff_per * std::cos(th) * std::cos(th);
  const double Dxy = (p.diff_par - p.diff_per) * std::sin(th) * std::cos(th);
  const double dt = p.reduceSym ? 0.5 * p.dt : p.dt;
  GParams g;
  g.nx = p.nx;
  g.ny = p.ny;
  g.masked = p.solidSwitch;
  g.neumann = p.neumannBC;
  g.lap4 = p.lap4 ? 1 : 0;
  g.aniso = p.anisotropy ? 1 : 0;
  g.vdiff = (p.model == YH_MODEL_2V_REF && p.gateDiff) ? 1 : 0;
  g.rx = (float)(dt * Dxx / (hx * hx));
  g.ry = (float)(dt * Dyy / (hy * hy));
  g.rxy = (float)(2.0 * Dxy * dt / (4.0 * hx * hy));
  g.rbx = (float)(hx * Dxy / (Dxx * hy));
  g.rby = (float)(hy * Dxy / (Dyy * hx));
  g.q4 = (float)(dt * Dyy / (hy * hy * 12.0) + dt * Dxx / (hx * hx * 12.0));
  g.fx4 = (float)(dt / 12.0);
  g.fy4 = (float)(dt / 12.0);
  g.dt = (float)dt;
  g.tc = (float)p.tc;
  g.rscale = (float)p.rscale;
  g.bval = (float)p.boundaryVal;
  g.mu = (float)p.mu;
  g.alpha = (float)p.alpha;
  g.beta = (float)p.beta;
  g.gamma = (float)p.gamma;
  g.delta = (float)p.delta;
  g.eps = (float)p.eps;
  g.theta = (float)p.theta;
  const yh_fk_params& f = p.fk;
  g.u_c = (float)f.u_c;
  g.u_v = (float)f.u_v;
  g.u_csi = (float)f.u_csi;
  g.kk = (float)f.k;
  g.tau_d = (float)f.tau_d;
  g.tau_r = (float)f.tau_r;
  g.tau_si = (float)f.tau_si;
  g.tau_0 = (float)f.tau_0;
  g.tau_vp = (float)f.tau_vp;
  g.tau_v1m = (float)f.tau_v1m;
  g.tau_v2m = (float)f.tau_v2m;
  g.tau_wp = (float)f.tau_wp;
  g.tau_wm = (float)f.tau_wm;
  g.stim_on = stim_on;
  g.stim_x = stim_x;
  g.stim_y = stim_y;
  return g;
}

// Ionic currents (:131-141 for 2V, Fenton & Karma 1998 for FK) at a cell of
// the stage field, with the mouse stimulus when the cell is in its disk.
template <int MODEL>
__device__ __forceinline__ void g_currents(const GParams& g, float u, float v, float w, int i,
                                           int j, float& Iu, float& Iv, float& Iw) {
  // :54-61: (ic-cx)^2 + (jc-cy)^2 < 400 with the point in cells
  const int di = i - g.stim_x, dj = j - g.stim_y;
  const float st = (g.stim_on && di * di + dj * dj < 400) ? 24.7f : 0.0f;
  if (MODEL == M2V) {
    Iu = -(g.mu * u * (1.0f - u) * (u - g.alpha) - u * v) - st;
    Iv = -(g.eps * (g.delta * (u - g.gamma) * (g.beta - u) - v - g.theta));
    Iw = 0.0f;
  } else {
    const float p = u >= g.u_c ? 1.0f : 0.0f, q = u >= g.u_v ? 1.0f : 0.0f;
    const float J_fi = -v * p * (1.0f - u) * (u - g.u_c) / g.tau_d;
    const float J_so = u * (1.0f - p) / g.tau_0 + p / g.tau_r;
    const float J_si = -w * (1.0f + tanhf(g.kk * (u - g.u_csi))) / (2.0f * g.tau_si);
    const float tvm = (1.0f - q) * g.tau_v1m + q * g.tau_v2m;
    Iu = J_fi + J_so + J_si - st;
    Iv = -((1.0f - p) * (1.0f - v) / tvm - p * v / g.tau_vp);
    Iw = -((1.0f - p) * (1.0f - w) / g.tau_wm - p * w / g.tau_wp);
  }
}

// Stage field of variable a at cell index c.
struct StageField {
  CFields y0, dp;
  float ki;
  __device__ __forceinline__ float operator()(int a, int c) const {
    const float x = y0.f[a][c];
    return ki != 0.0f ? __builtin_fmaf(ki, dp.f[a][c], x) : x;
  }
};

// Diagonal neighbours of the Neumann square branch (:207-221, :250-264).
__device__ __forceinline__ void g_diag(int nx, int ny, int i, int j, int& SW, int& SE, int& NW,
                                       int& NE) {
  SW = (i > 0 && j > 0) ? i - 1 + nx * (j - 1)
       : ((i == 0 && j > 0) ? i + 1 + nx * (j - 1)
                            : ((i > 0 && j == 0) ? i - 1 + nx * (j + 1) : i + 1 + nx * (j + 1)));
  SE = (i < nx - 1 && j > 0) ? i + 1 + nx * (j - 1)
       : ((i == nx - 1 && j > 0) ? i - 1 + nx * (j - 1)
                                 : ((i < nx - 1 && j == 0) ? i + 1 + nx * (j + 1) : i - 1 + nx * (j + 1)));
  NW = (i > 0 && j < ny - 1) ? i - 1 + nx * (j + 1)
       : ((i == 0 && j < ny - 1) ? i + 1 + nx * (j + 1)
                                 : ((i > 0 && j == ny - 1) ? i - 1 + nx * (j - 1) : i + 1 + nx * (j - 1)));
  NE = (i < nx - 1 && j < ny - 1) ? i + 1 + nx * (j + 1)
       : ((i == nx - 1 && j < ny - 1) ? i - 1 + nx * (j + 1)
                                      : ((i < nx - 1 && j == ny - 1) ? i + 1 + nx * (j - 1) : i - 1 + nx * (j - 1)));
}

// Anisotropic term of variable a on the Neumann square branch (:266-313),
// scaled by s (1 for u, rscale for v).
__device__ __forceinline__ float g_aniso_neumann(const GParams& g, const StageField& Y, int a,
                                                 int i, int j, float s) {
  const int nx = g.nx, ny = g.ny, c = i + nx * j;
  int SW, SE, NW, NE;
  g_diag(nx, ny, i, j, SW, SE, NW, NE);
  const float b_S = (j > 0) ? 0.0f : ((i == 0 || i == nx - 1) ? 0.0f
                    : g.rby * (Y(a, c + 1) - Y(a, c - 1)));
  const float b_N = (j < ny - 1) ? 0.0f : ((i == 0 || i == nx - 1) ? 0.0f
                    : -g.rby * (Y(a, c + 1) - Y(a, c - 1)));
  const float b_W = (i > 0) ? 0.0f : ((j == 0 || j == ny - 1) ? 0.0f
                    : g.rbx * (Y(a, c + nx) - Y(a, c - nx)));
  const float b_E = (i < nx - 1) ? 0.0f : ((j == 0 || j == ny - 1) ? 0.0f
                    : -g.rbx * (Y(a, c + nx) - Y(a, c - nx)));
  float d = ((b_S + b_N) * g.ry + (b_W + b_E) * g.rx) * s;
  const float b_SW = (i > 0 && j > 0) ? 0.0f
                     : ((i == 0 && j > 1) ? g.rbx * (Y(a, c) - Y(a, c - 2 * nx))
                        : ((i > 1 && j == 0) ? g.rby * (Y(a, c) - Y(a, c - 2)) : 0.0f));
  const float b_SE = (i < nx - 1 && j > 0) ? 0.0f
                     : ((i == nx - 1 && j > 1) ? -g.rbx * (Y(a, c) - Y(a, c - 2 * nx))
                        : ((i < nx - 2 && j == 0) ? g.rby * (Y(a, c + 2) - Y(a, c)) : 0.0f));
  const float b_NW = (i > 0 && j < ny - 1) ? 0.0f
                     : ((i == 0 && j < ny - 2) ? g.rbx * (Y(a, c + 2 * nx) - Y(a, c))
                        : ((i > 1 && j == ny - 1) ? -g.rby * (Y(a, c) - Y(a, c - 2)) : 0.0f));
  const float b_NE = (i < nx - 1 && j < ny - 1) ? 0.0f
                     : ((i == nx - 1 && j < ny - 2) ? -g.rbx * (Y(a, c + 2 * nx) - Y(a, c))
                        : ((i < nx - 2 && j == ny - 1) ? -g.rby * (Y(a, c + 2) - Y(a, c)) : 0.0f));
  d += g.rxy * ((Y(a, SW) + b_SW) + (Y(a, NE) + b_NE) - (Y(a, SE) + b_SE) - (Y(a, NW) + b_NW)) * s;
  return d;
}

// Dirichlet neighbour (:362-373, :399-408, :434-437): the field inside the
// domain where (with a mask) the cell and the neighbour are tissue, else the
// boundary value.
__device__ __forceinline__ float g_dir(const GParams& g, const StageField& Y, int a,
                                       const uint8_t* solid, bool sc, int ni, int nj) {
  if (ni < 0 || ni >= g.nx || nj < 0 || nj >= g.ny) return g.bval;
  const int n = ni + g.nx * nj;
  if (g.masked && !(sc && solid[n])) return g.bval;
  return Y(a, n);
}

// One Runge-Kutta stage (:114-503 for one rk4idx): increments d = F(y_s) of
// every cell; rhs accumulates wsum*d (:501-502); the last stage writes
// y0 + tc*rhs, zero in holes, and velTan = rhs/dt (:506-561).
template <int MODEL>
__global__ __launch_bounds__(256) void k_stage(GParams g, CFields y0, CFields dp, float ki,
                                               float wsum, int first, int last, Fields dcur,
                                               Fields rhs, Fields out, Fields vt,
                                               const uint8_t* __restrict__ solid, ApdState apd,
                                               int apd_on) {
  constexpr int NV = Model<MODEL>::NV;
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y * 4 + threadIdx.y;
  if (i >= g.nx || j >= g.ny) return;
  const int nx = g.nx, ny = g.ny, c = i + nx * j;
  const StageField Y{y0, dp, ki};
  const float u = Y(0, c), v = Y(1, c), w = NV > 2 ? Y(NV - 1, c) : 0.0f;
  float Iu, Iv, Iw;
  g_currents<MODEL>(g, u, v, w, i, j, Iu, Iv, Iw);
  const bool vdiff = g.vdiff != 0;
  float du = 0.0f, dv = 0.0f;
  bool sc = true;
  if (g.neumann) {
    const int S = i + nx * mirror(j - 1, ny), N = i + nx * mirror(j + 1, ny);
    const int W = mirror(i - 1, nx) + nx * j, E = mirror(i + 1, nx) + nx * j;
    if (g.masked) {
      // :154-185 (lap4 and the anisotropic term are not applied, as in the reference)
      sc = solid[c] != 0;
      const bool sw = solid[W] != 0, se = solid[E] != 0, sn = solid[N] != 0, ss = solid[S] != 0;
      const float cxx = (sw && se) && (sw && sc) ? 1.0f : ((sw && sc) ? 2.0f : 0.0f);
      const float cxy = sc ? ((sw || se) ? 2.0f : 0.0f) : 0.0f;
      const float cxz = (sw && se) && (sc && se) ? 1.0f : ((sc && se) ? 2.0f : 0.0f);
      const float cyx = (sn && ss) && (sn && sc) ? 1.0f : ((sn && sc) ? 2.0f : 0.0f);
      const float cyy = sc ? ((sn || ss) ? 2.0f : 0.0f) : 0.0f;
      const float cyz = (sn && ss) && (sc && ss) ? 1.0f : ((sc && ss) ? 2.0f : 0.0f);
      du = (cxx * Y(0, W) - cxy * u + cxz * Y(0, E)) * g.rx + (cyx * Y(0, N) - cyy * u + cyz * Y(0, S)) * g.ry;
      if (vdiff)
        dv = (cxx * Y(1, W) - cxy * v + cxz * Y(1, E)) * g.rx * g.rscale
           + (cyx * Y(1, N) - cyy * v + cyz * Y(1, S)) * g.ry * g.rscale;
    } else {
      // :188-199
      du = (Y(0, W) - 2.0f * u + Y(0, E)) * g.rx + (Y(0, N) - 2.0f * u + Y(0, S)) * g.ry;
      if (vdiff)
        dv = (Y(1, W) - 2.0f * v + Y(1, E)) * g.rx * g.rscale + (Y(1, N) - 2.0f * v + Y(1, S)) * g.ry * g.rscale;
      if (g.lap4) {
        // :201-247; J of a neighbour = its current in this stage
        int SW, SE, NW, NE;
        g_diag(nx, ny, i, j, SW, SE, NW, NE);
        float JW[3], JE[3], JN[3], JS[3];
        const int iw = mirror(i - 1, nx), ie = mirror(i + 1, nx);
        const int jn = mirror(j + 1, ny), js = mirror(j - 1, ny);
        g_currents<MODEL>(g, Y(0, W), Y(1, W), NV > 2 ? Y(NV - 1, W) : 0.0f, iw, j, JW[0], JW[1], JW[2]);
        g_currents<MODEL>(g, Y(0, E), Y(1, E), NV > 2 ? Y(NV - 1, E) : 0.0f, ie, j, JE[0], JE[1], JE[2]);
        g_currents<MODEL>(g, Y(0, N), Y(1, N), NV > 2 ? Y(NV - 1, N) : 0.0f, i, jn, JN[0], JN[1], JN[2]);
        g_currents<MODEL>(g, Y(0, S), Y(1, S), NV > 2 ? Y(NV - 1, S) : 0.0f, i, js, JS[0], JS[1], JS[2]);
        du += -2.0f * g.q4 * ((Y(0, W) - u + Y(0, E)) + (Y(0, N) - u + Y(0, S)));
        du += g.q4 * (Y(0, SW) + Y(0, SE) + Y(0, NW) + Y(0, NE));
        du -= ((JW[0] - 2.0f * Iu + JE[0]) * g.fx4 + (JN[0] - 2.0f * Iu + JS[0]) * g.fy4);
        if (vdiff) {
          dv += -g.rscale * 2.0f * g.q4 * ((Y(1, W) - v + Y(1, E)) + (Y(1, N) - v + Y(1, S)));
          dv += g.rscale * g.q4 * (Y(1, SW) + Y(1, SE) + Y(1, NW) + Y(1, NE));
          dv -= ((JW[1] - 2.0f * Iv + JE[1]) * g.fx4 + (JN[1] - 2.0f * Iv + JS[1]) * g.fy4);
        }
      }
      if (g.aniso) {
        // :249-356
        du += g_aniso_neumann(g, Y, 0, i, j, 1.0f);
        if (vdiff) dv += g_aniso_neumann(g, Y, 1, i, j, g.rscale);
      }
    }
  } else {
    // Dirichlet (:360-490)
    sc = g.masked ? solid[c] != 0 : true;
    du = (g_dir(g, Y, 0, solid, sc, i - 1, j) - 2.0f * u + g_dir(g, Y, 0, solid, sc, i + 1, j)) * g.rx
       + (g_dir(g, Y, 0, solid, sc, i, j + 1) - 2.0f * u + g_dir(g, Y, 0, solid, sc, i, j - 1)) * g.ry;
    if (vdiff)
      dv = (g_dir(g, Y, 1, solid, sc, i - 1, j) - 2.0f * v + g_dir(g, Y, 1, solid, sc, i + 1, j)) * g.rx * g.rscale
         + (g_dir(g, Y, 1, solid, sc, i, j + 1) - 2.0f * v + g_dir(g, Y, 1, solid, sc, i, j - 1)) * g.ry * g.rscale;
    if (g.aniso) {
      du += g.rxy * (g_dir(g, Y, 0, solid, sc, i - 1, j - 1) + g_dir(g, Y, 0, solid, sc, i + 1, j + 1)
                     - g_dir(g, Y, 0, solid, sc, i + 1, j - 1) - g_dir(g, Y, 0, solid, sc, i - 1, j + 1));
      if (vdiff)
        dv += g.rxy * (g_dir(g, Y, 1, solid, sc, i - 1, j - 1) + g_dir(g, Y, 1, solid, sc, i + 1, j + 1)
                       - g_dir(g, Y, 1, solid, sc, i + 1, j - 1) - g_dir(g, Y, 1, solid, sc, i - 1, j + 1)) * g.rscale;
    }
  }
  // :498-499
  float d[3];
  d[0] = du - g.dt * Iu;
  d[1] = dv - g.dt * Iv;
  d[2] = -g.dt * Iw;
#pragma unroll
  for (int a = 0; a < NV; ++a) {
    const float r = (first ? 0.0f : rhs.f[a][c]) + wsum * d[a];
    if (last) {
      const float x = y0.f[a][c] + g.tc * r;
      out.f[a][c] = sc ? x : 0.0f;
      if (vt.f[a]) vt.f[a][c] = sc ? r / g.dt : 0.0f;
    } else {
      rhs.f[a][c] = r;
      dcur.f[a][c] = d[a];
    }
  }
  if (last && apd_on && sc) {
    const float u0 = y0.f[0][c];
    const float un = out.f[0][c];
    float fr = apd.front[c], bk = apd.back[c], a1 = apd.apd1[c], a2 = apd.apd2[c];
    uint32_t fi = apd.first[c];
    apd_cell(u0, un, apd.count0 + 1, apd.dt, apd.th, fr, bk, a1, a2, fi);
    apd.front[c] = fr;
    apd.back[c] = bk;
    apd.apd1[c] = a1;
    apd.apd2[c] = a2;
    apd.first[c] = (uint8_t)fi;
  }
}

hipError_t launch_general_step(int model, const GParams& g, int order, CFields in, Fields out,
                               Fields vt, const uint8_t* solid, float* const* scratch,
                               hipStream_t s, const ApdState* apd) {
  // :70-95, RK4 weights as the reference writes them
  float ki[4] = {0.0f, 0.0f, 0.0f, 0.0f}, sw[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  if (order == 1) {
    sw[0] = 1.0f;
  } else if (order == 2) {
    ki[1] = 0.5f;
    sw[1] = 1.0f;
  } else if (order == 4) {
    ki[1] = 0.5f;
    ki[2] = 0.5f;
    ki[3] = 1.0f;
    sw[0] = 0.166666666666667f;
    sw[1] = 0.333333333333333f;
    sw[2] = 0.333333333333333f;
    sw[3] = 0.166666666666667f;
  } else {
    return hipErrorInvalidValue;
  }
  if (order > 1 && !scratch) return hipErrorInvalidValue;
  // scratch: d ping-pong [0..2], [3..5]; rhs [6..8]
  Fields dbuf[2] = {Fields{{nullptr, nullptr, nullptr}}, Fields{{nullptr, nullptr, nullptr}}};
  Fields rhs{{nullptr, nullptr, nullptr}};
  if (scratch) {
    dbuf[0] = Fields{{scratch[0], scratch[1], scratch[2]}};
    dbuf[1] = Fields{{scratch[3], scratch[4], scratch[5]}};
    rhs = Fields{{scratch[6], scratch[7], scratch[8]}};
  }
  const ApdState a = apd ? *apd : ApdState();
  dim3 b(64, 4), gr((g.nx + 63) / 64, (g.ny + 3) / 4);
  for (int st = 0; st < order; ++st) {
    const Fields& dcur = dbuf[st & 1];
    const Fields& dprev = dbuf[(st + 1) & 1];
    CFields dp{{dprev.f[0], dprev.f[1], dprev.f[2]}};
    if (st == 0) dp = in;  // ki[0] = 0: never read
    const int first = st == 0, last = st == order - 1;
    if (model == M2V)
      hipLaunchKernelGGL(k_stage<M2V>, gr, b, 0, s, g, in, dp, ki[st], sw[st], first, last, dcur,
                         rhs, out, vt, solid, a, apd ? 1 : 0);
    else
      hipLaunchKernelGGL(k_stage<MFK>, gr, b, 0, s, g, in, dp, ki[st], sw[st], first, last, dcur,
                         rhs, out, vt, solid, a, apd ? 1 : 0);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace yh
