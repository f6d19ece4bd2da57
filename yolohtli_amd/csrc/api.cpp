// api.cpp -- C ABI: parameters, workload generators, single-step device
// wrappers and the plan object of the plain stepper.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "plan.h"
#include "yh_internal.h"

namespace yh {

static thread_local std::string g_last_error;

void set_last_error(const char* msg) { g_last_error = msg ? msg : ""; }

int hip_fail(hipError_t e, const char* what) {
  char buf[512];
  std::snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
  set_last_error(buf);
  if (e == hipErrorOutOfMemory) return YH_ENOMEM;
  return YH_EHIP;
}

int nvars_of(int model) { return model == YH_MODEL_3V_FK ? 3 : 2; }

void make_kparams(const yh_params& p, const yh_derived& d, KParams* k) {
  const double tcdt = p.tc * d.dt;
  k->nx = p.nx;
  k->ny = p.ny;
  k->rx = (float)d.rx;
  k->ry = (float)d.ry;
  k->trx = (float)(p.tc * d.rx);
  k->try_ = (float)(p.tc * d.ry);
  k->dt = (float)d.dt;
  k->tc = (float)p.tc;
  k->tcdt = (float)tcdt;
  k->rs = p.gateDiff ? (float)p.rscale : 0.0f;
  // 2V: -tcdt*I_sum = u*(tcdt*mu*(1-u)*(u-alpha) - tcdt*v)
  //     -tcdt*I_v   = tcdt*eps*delta*(u-gamma)*(beta-u) - tcdt*eps*v - tcdt*eps*theta
  k->alpha = (float)p.alpha;
  k->beta = (float)p.beta;
  k->gamma = (float)p.gamma;
  k->a2_A = (float)(tcdt * p.mu);
  k->a2_B = (float)tcdt;
  k->a2_E = (float)(tcdt * p.eps * p.delta);
  k->a2_F = (float)(tcdt * p.eps);
  k->a2_G = (float)(tcdt * p.eps * p.theta);
  const yh_fk_params& f = p.fk;
  k->u_c = (float)f.u_c;
  k->u_v = (float)f.u_v;
  k->fk_A1 = (float)(tcdt / f.tau_d);
  k->fk_A2 = (float)(tcdt / f.tau_r);
  k->fk_A3 = (float)(tcdt / f.tau_0);
  k->fk_A4 = (float)(tcdt / f.tau_si);
  const double log2e = 1.4426950408889634;
  k->fk_b = (float)(2.0 * f.k * log2e);
  k->fk_bu = (float)(2.0 * f.k * log2e * f.u_csi);
  k->fk_bs = (float)(2.0 * f.k * log2e * f.u_csi - std::log2(tcdt / f.tau_si));
  k->fk_iA4 = (float)(f.tau_si / tcdt);
  k->fk_P1 = (float)(tcdt / f.tau_d * (1.0 + f.u_c));
  k->fk_P0 = (float)(-tcdt / f.tau_d * f.u_c);
  k->fk_B1 = (float)(tcdt / f.tau_vp);
  k->fk_B2 = (float)(tcdt / f.tau_v1m);
  k->fk_B3 = (float)(tcdt / f.tau_v2m);
  k->fk_C1 = (float)(tcdt / f.tau_wp);
  k->fk_C2 = (float)(tcdt / f.tau_wm);
  k->hw = (float)p.heaviside_width;
  k->iso = k->trx == k->try_ ? 1 : 0;
  k->wt = 0;
}

// splitmix64 (Steele, Lea & Flood 2014).
struct SplitMix {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
};

}  // namespace yh

using namespace yh;

// ====================================================================== C ABI
extern "C" {

const char* yh_strerror(int code) {
  switch (code) {
    case YH_OK: return "ok";
    case YH_EINVAL: return "invalid argument";
    case YH_EUNSUPPORTED:
      return "this path runs the standard branch only (Neumann, Euler, 5-point)";
    case YH_EHIP: return g_last_error.empty() ? "HIP runtime error" : g_last_error.c_str();
    case YH_ENOMEM: return "out of memory";
    case YH_EUNSTABLE: return "explicit Euler diffusion number rx+ry exceeds 0.5";
    case YH_ESINGULAR: return "symmetry-reduction matrix is singular";
    case YH_ENOTIP: return "no spiral tip found";
    case YH_EOVERFLOW: return "tip log overflow";
    case YH_ENONFINITE: return "non-finite value in the state";
    case YH_ESTATE: return "call out of order";
    case YH_ESCRATCH: return "grid exceeds the 2^21 cells of yh_step_noalloc's fixed scratch";
    default: return "unknown error";
  }
}

void yh_fk_params_default(yh_fk_params* f) {
  if (!f) return;
  // Fenton & Karma (1998), Table 1, Beeler-Reuter fit.  The paper writes
  // tau_v-(u) = H(u - u_v) tau_v1- + H(u_v - u) tau_v2- with tau_v1- = 1250,
  // tau_v2- = 19.6; this library (as Fenton et al., Chaos 12:852, 2002) names
  // the constant used below u_v tau_v1m, so rest (u < u_v) recovers v with
  // 19.6 ms and u_v <= u < u_c with 1250 ms, as in the paper.
  f->tau_d = 0.25;  // C_m/g_fi with g_fi = 4
  f->tau_r = 33.33;
  f->tau_si = 29.0;
  f->tau_0 = 12.5;
  f->tau_vp = 3.33;
  f->tau_v1m = 19.6;
  f->tau_v2m = 1250.0;
  f->tau_wp = 870.0;
  f->tau_wm = 41.0;
  f->u_c = 0.13;
  f->u_v = 0.04;
  f->u_csi = 0.85;
  f->k = 10.0;
}

void yh_params_default(yh_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  // parameterSetup, saveFiles.cu:105-231 (solver subset).
  p->nx = 512;
  p->ny = 512;
  p->Lx = 12.0;
  p->Ly = 12.0;
  p->dt = 0.02;
  p->diff_par = 0.001;
  p->diff_per = 0.001;
  p->degrad = 0.0;
  p->rscale = 0.01;
  p->tc = 1.0;
  p->alpha = 0.2;
  p->beta = 1.1;
  p->gamma = 0.0;
  p->delta = 1.0;
  p->eps = 0.005;
  p->mu = 1.0;
  p->theta = 0.0;
  p->model = YH_MODEL_2V_REF;
  p->neumannBC = 1;
  p->solidSwitch = 0;
  p->gateDiff = 1;
  p->reduceSym = 0;
  // The reference defaults are RK4 + lap4 (saveFiles.cu:131-132); these
  // defaults select Euler + 5-point, the branch of the temporally blocked
  // kernels.  RK2/RK4, lap4, anisotropy and Dirichlet run on the general path.
  p->timeIntOrder = 1;
  p->lap4 = 0;
  p->anisotropy = 0;
  p->boundaryVal = 0.0;
  p->Uth = 0.7;
  p->tipOffsetX = 160;
  p->tipOffsetY = 160;
  yh_fk_params_default(&p->fk);
  p->heaviside_width = 0.0;
}

int yh_params_derive(const yh_params* p, yh_derived* d) {
  if (!p || !d) return YH_EINVAL;
  if (p->nx < 4 || p->ny < 4) return YH_EINVAL;
  // cell indices are int32 and tip/contour keys 2*cell + root must fit too
  if ((int64_t)p->nx * p->ny > (int64_t)1 << 30) return YH_EINVAL;
  if (!(p->Lx > 0) || !(p->Ly > 0) || !(p->dt > 0)) return YH_EINVAL;
  if (p->model != YH_MODEL_2V_REF && p->model != YH_MODEL_3V_FK) return YH_EINVAL;
  if (p->timeIntOrder != 1 && p->timeIntOrder != 2 && p->timeIntOrder != 4) return YH_EINVAL;
  if (p->model == YH_MODEL_3V_FK) {
    const yh_fk_params& f = p->fk;
    if (!(f.tau_d > 0 && f.tau_r > 0 && f.tau_si > 0 && f.tau_0 > 0 && f.tau_vp > 0 &&
          f.tau_v1m > 0 && f.tau_v2m > 0 && f.tau_wp > 0 && f.tau_wm > 0))
      return YH_EINVAL;
  }
  const double pi = 3.14159265359;  // the reference's pi (globalVariables.cuh:27)
  d->hx = p->Lx / (p->nx - 1.0);
  d->hy = p->Ly / (p->ny - 1.0);
  const double th = p->degrad * pi / 180.0;
  d->Dxx = p->diff_par * std::cos(th) * std::cos(th) + p->diff_per * std::sin(th) * std::sin(th);
  d->Dyy = p->diff_par * std::sin(th) * std::sin(th) + p->diff_per * std::cos(th) * std::cos(th);
  d->dt = p->reduceSym ? 0.5 * p->dt : p->dt;
  d->rx = d->dt * d->Dxx / (d->hx * d->hx);
  d->ry = d->dt * d->Dyy / (d->hy * d->hy);
  d->invdx = 0.5 / d->hx;
  d->invdy = 0.5 / d->hy;
  if (d->rx + d->ry > 0.5) return YH_EUNSTABLE;
  return YH_OK;
}

uint64_t yh_seed(uint64_t base, uint32_t config, uint32_t tissue, uint32_t purpose) {
  SplitMix m{base};
  uint64_t h = m.next();
  const uint32_t parts[3] = {config, tissue, purpose};
  for (uint32_t v : parts) {
    SplitMix q{h ^ ((uint64_t)v * 0xD1B54A32D192ED03ull)};
    h = q.next();
  }
  return h;
}

int yh_init_cross_field(int nx, int ny, int u_cols, float rest, float refr, float* u, float* v,
                        float* w) {
  if (nx <= 0 || ny <= 0 || !u || !v) return YH_EINVAL;
  // initGates, main.cu:593-618: u = 1 for i < nx/8 (here u_cols), v = 1 for j >= ny/2.
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const size_t c = (size_t)i + (size_t)nx * j;
      u[c] = i < u_cols ? 1.0f : 0.0f;
      const float g = j >= ny / 2 ? refr : rest;
      v[c] = g;
      if (w) w[c] = g;
    }
  return YH_OK;
}

int yh_init_random(int nx, int ny, uint64_t seed, float gate, float* u, float* v, float* w) {
  if (nx <= 0 || ny <= 0 || !u || !v) return YH_EINVAL;
  SplitMix m{seed};
  const size_t n = (size_t)nx * ny;
  for (size_t c = 0; c < n; ++c) {
    u[c] = (float)m.uniform();
    v[c] = gate;
    if (w) w[c] = gate;
  }
  return YH_OK;
}

int yh_make_holes(int nx, int ny, int nholes, double rmin, double rmax, uint64_t seed,
                  uint8_t* mask) {
  if (nx <= 0 || ny <= 0 || nholes < 0 || !mask || rmin < 0 || rmax < rmin) return YH_EINVAL;
  std::memset(mask, 1, (size_t)nx * ny);
  SplitMix m{seed};
  for (int h = 0; h < nholes; ++h) {
    const double cx = m.uniform() * nx;
    const double cy = m.uniform() * ny;
    const double r = rmin + m.uniform() * (rmax - rmin);
    const int i0 = std::max(0, (int)std::floor(cx - r)), i1 = std::min(nx - 1, (int)std::ceil(cx + r));
    const int j0 = std::max(0, (int)std::floor(cy - r)), j1 = std::min(ny - 1, (int)std::ceil(cy + r));
    for (int j = j0; j <= j1; ++j)
      for (int i = i0; i <= i1; ++i) {
        const double dx = i - cx, dy = j - cy;
        if (dx * dx + dy * dy <= r * r) mask[(size_t)i + (size_t)nx * j] = 0;
      }
  }
  return YH_OK;
}

int yh_rd_step_dev(const yh_params* p, float* const out[3], const float* const in[3],
                   float* const veltan[3], const uint8_t* solid, void* stream) {
  return yh_rd_step_dev_ex(p, out, in, veltan, solid, 0, 0, 0, stream);
}

int yh_rd_step_dev_ex(const yh_params* p, float* const out[3], const float* const in[3],
                      float* const veltan[3], const uint8_t* solid, int stim_on, int stim_x,
                      int stim_y, void* stream) {
  yh_derived d;
  int rc = yh_params_derive(p, &d);
  if (rc) return rc;
  const int nv = nvars_of(p->model);
  if (!out || !in) return YH_EINVAL;
  for (int a = 0; a < nv; ++a) {
    if (!out[a] || !in[a]) return YH_EINVAL;
    if (veltan && !veltan[a]) return YH_EINVAL;
    for (int b = 0; b < nv; ++b)
      if ((const float*)out[a] == in[b]) return YH_EINVAL;
  }
  if (p->solidSwitch && !solid) return YH_EINVAL;
  CFields ci{{in[0], in[1], nv > 2 ? in[2] : nullptr}};
  Fields co{{out[0], out[1], nv > 2 ? out[2] : nullptr}};
  Fields cv{{nullptr, nullptr, nullptr}};
  if (veltan) cv = Fields{{veltan[0], veltan[1], nv > 2 ? veltan[2] : nullptr}};
  if (standard_branch(*p) && !stim_on) {
    KParams k;
    make_kparams(*p, d, &k);
    YH_HIP(launch_direct(p->model, k, ci, co, cv, p->solidSwitch ? solid : nullptr,
                         (hipStream_t)stream));
    return YH_OK;
  }
  // the stages of RK2/RK4 need scratch fields: a plan owns them
  if (p->timeIntOrder != 1) return YH_EUNSUPPORTED;
  const GParams g = make_gparams(*p, stim_on, stim_x, stim_y);
  YH_HIP(launch_general_step(p->model, g, 1, ci, co, cv, p->solidSwitch ? solid : nullptr,
                             nullptr, (hipStream_t)stream, nullptr));
  return YH_OK;
}

int yh_tip_detect_dev(const yh_params* p, const float* g_past, const float* g_present,
                      double time, yh_tip* d_tips, int* d_keys, int* d_count, int cap,
                      void* stream) {
  yh_derived d;
  int rc = yh_params_derive(p, &d);
  if (rc) return rc;
  if (!g_past || !g_present || !d_tips || !d_keys || !d_count || cap < 0) return YH_EINVAL;
  KParams k;
  make_kparams(*p, d, &k);
  YH_HIP(launch_tips(k, p->solidSwitch, p->tipOffsetX, p->tipOffsetY, p->Uth, g_past, g_present,
                     time, d_tips, d_keys, d_count, cap, (hipStream_t)stream));
  return YH_OK;
}

}  // extern "C"
