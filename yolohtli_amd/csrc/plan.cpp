// plan.cpp -- yh_plan: owns the ping-pong state, mask codes and tip log of
// one tissue; drives the plain stepper (the display() loop, main.cu:876-885).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <cstdint>
#include <cstring>
#include <new>
#include <vector>

#include "plan.h"

namespace yh {

int64_t tiled_tile_count(int nx, int ny, int rows, int waves, int cols, int K) {
  const int wout = 64 * cols - 2 * K, hout = rows * waves - 2 * K;
  if (wout <= 0 || hout <= 0) return INT64_MAX;
  return (int64_t)((nx + wout - 1) / wout) * ((ny + hout - 1) / hout);
}

static int64_t tile_count(int nx, int ny, int rows, int waves, int cols, int K) {
  return tiled_tile_count(nx, ny, rows, waves, cols, K);
}

// YH_KERNEL_AUTO without timing anything (yh_autotune times the candidates).
// When 64 x 112 regions at K=9 would fill between half and all of the CUs,
// one each: 64 x 64 regions in 16 waves of 4 rows at K=8 if two of them fit
// on a CU at once (`pairs`) and their tiles fit in one such round (1024^2:
// 484 tiles, the best measured there, with remainders merged into 5 x 16),
// else the 64 x 112 regions.  64 x 64 regions in 8 waves at K=8 when there
// are more tiles than that, so two workgroups per CU overlap one's loads and
// stores with the other's substeps (the best measured at 4096^2); else 64 x
// 64 regions in 16 waves of 4 rows, as deep as one round allows.
yh_kernel_cfg auto_cfg(int nx, int ny, int cus, bool pairs) {
  const int64_t t716 = tile_count(nx, ny, 7, 16, 1, 9);
  const int64_t ncu = cus > 0 ? cus : 256;
  yh_kernel_cfg c;
  if (t716 > ncu)
    c = yh_kernel_cfg{YH_KERNEL_TILED, 8, 8, 8, 1};
  else if (2 * t716 >= ncu)
    c = pairs && tile_count(nx, ny, 4, 16, 1, 8) <= 2 * ncu
            ? yh_kernel_cfg{YH_KERNEL_TILED, 8, 4, 16, 1}
            : yh_kernel_cfg{YH_KERNEL_TILED, 9, 7, 16, 1};
  else {
    // a grid too small to fill half the CUs with 64 x 112 regions: 64 x 64
    // regions at the deepest fusion whose tiles still fit on the CUs at one
    // workgroup each (512^2: K=16, 256 tiles; 0.98 against 1.24 us per
    // step at K=8), since a launch's fixed part then weighs more than the
    // recomputed halo
    int K = 8;
    for (int k2 = 16; k2 > 8; --k2)
      if (tile_count(nx, ny, 4, 16, 1, k2) <= ncu) {
        K = k2;
        break;
      }
    c = yh_kernel_cfg{YH_KERNEL_TILED, K, 4, 16, 1};
  }
  // the reflected halo is exact for at most min(nx, ny) - 1 substeps
  c.substeps = std::max(1, std::min(c.substeps, std::min(nx, ny) - 1));
  return c;
}

// Workgroups of the tiled kernel of this shape (model, mask, the isotropy
// of k, with or without the previous-u output) that one CU holds at once
// (its occupancy), cached per instantiation in a fixed table, so neither
// the query nor the cache allocates (the no-allocation entry point uses it).
int tiled_per_cu(int model, bool masked, const KParams& k, int rows, int waves, int cols,
                 bool prev) {
  struct Entry {
    int key[7];
    int n;
  };
  static std::mutex mu;
  static Entry table[256];
  static int used = 0;
  const int key[7] = {model, masked ? 1 : 0, k.iso, rows, waves, cols, prev ? 1 : 0};
  {
    std::lock_guard<std::mutex> g(mu);
    for (int e = 0; e < used; ++e)
      if (std::equal(key, key + 7, table[e].key)) return table[e].n;
  }
  static float dummy;  // selects the previous-u instantiation; never written in a query
  int n = 0;
  if (launch_tiled(model, masked, rows, waves, cols, 1, k, CFields{{nullptr, nullptr, nullptr}},
                   Fields{{nullptr, nullptr, nullptr}}, nullptr, nullptr, nullptr,
                   prev ? &dummy : nullptr, &n) != hipSuccess ||
      n < 1)
    n = 1;
  std::lock_guard<std::mutex> g(mu);
  if (used < 256) {  // every instantiation fits; past that, query again
    std::copy(key, key + 7, table[used].key);
    table[used].n = n;
    ++used;
  }
  return n;
}

static int tiled_per_cu(const yh_plan* pl, int rows, int waves, int cols, bool prev) {
  return tiled_per_cu(pl->p.model, pl->p.solidSwitch != 0, pl->k, rows, waves, cols, prev);
}

yh_kernel_cfg resolve_cfg(const yh_plan* pl) {
  yh_kernel_cfg c = pl->cfg;
  if (c.variant == YH_KERNEL_AUTO)
    c = auto_cfg(pl->p.nx, pl->p.ny, pl->cus, tiled_per_cu(pl, 4, 16, 1, false) >= 2);
  if (c.variant == YH_KERNEL_TILED) {
    const int kmax = std::min(pl->p.nx, pl->p.ny) - 1;
    c.substeps = std::max(1, std::min(c.substeps, kmax));
    if (pl->apd_on && !tiled_apd_supported(c.rows, c.waves, c.cols)) {
      c.rows = 4;  // the fused APD update is built for these shapes only
      c.waves = 16;
      c.cols = 1;
    }
  }
  return c;
}

// Whether a remainder of r < K steps can be merged into the last K-step
// launch of a call: the shape one row taller per wave (same waves, one
// column per lane, where a tiled update does not depend on where launches
// split) is built, with the previous-u output when `prev`, and runs K + r
// substeps in no more rounds of workgroups than a K-step launch of the
// plan's shape (tiles over the CUs times the workgroups a CU holds at
// once), so the call takes one launch fewer for no longer a launch (a
// launch's fixed part is about a third of a K=9 launch at 1024^2).
static bool merge_ok(const yh_plan* pl, const yh_kernel_cfg& c, int r, bool prev) {
  const int K = c.substeps, rows = c.rows + 1, Km = K + r;
  if (c.cols != 1 || !tiled_supported(rows, c.waves, c.cols)) return false;
  if (prev && !tiled_prev_supported(rows, c.waves, c.cols)) return false;
  const int nx = pl->p.nx, ny = pl->p.ny;
  // the reflected halo is exact for at most min(nx, ny) - 1 substeps
  if (Km > 16 || Km > std::min(nx, ny) - 1) return false;
  if (2 * Km >= 64 * c.cols || 2 * Km >= rows * c.waves) return false;
  const int64_t cus = std::max(1, pl->cus);
  const int64_t slots_k = cus * tiled_per_cu(pl, c.rows, c.waves, c.cols, prev);
  const int64_t slots_m = cus * tiled_per_cu(pl, rows, c.waves, c.cols, prev);
  const int64_t tk = tile_count(nx, ny, c.rows, c.waves, c.cols, K);
  const int64_t tm = tile_count(nx, ny, rows, c.waves, c.cols, Km);
  return (tm + slots_m - 1) / slots_m <= (tk + slots_k - 1) / slots_k;
}

// Whether a tiled launch of `tiles` workgroups, `per_cu` of which fit on a
// CU, writes its outputs through to memory (the isotropic kernels come in
// both forms): when all its workgroups are resident at once and most CUs
// hold two (1.5 x the CUs or more).  Dirty lines left in each XCD's L2 are
// written back at the end of the launch, when the CUs are idle; written
// through, that traffic overlaps the second workgroup's loads (4x16 K=8 at
// 1024^2: 16.9 -> 15.6 us a launch).  A launch of one workgroup per CU ran
// 3% slower written through (512^2), one of many rounds of workgroups no
// faster (4096^2) and the batched kernel over 50 M cells 5% slower, so
// those keep plain stores.
bool tiled_write_through(int64_t tiles, int cus, int per_cu) {
  const int64_t c = std::max(1, cus);
  return 2 * tiles >= 3 * c && tiles <= c * per_cu;
}

static int write_through(const yh_plan* pl, int rows, int waves, int cols, int K, bool prev) {
  return tiled_write_through(tile_count(pl->p.nx, pl->p.ny, rows, waves, cols, K), pl->cus,
                             tiled_per_cu(pl, rows, waves, cols, prev));
}

int plan_advance(yh_plan* pl, int nsteps, hipStream_t s) {
  if (nsteps < 0) return YH_EINVAL;
  if (nsteps == 0) return YH_OK;
  if (pl->p.solidSwitch && !pl->mask_set) return YH_ESTATE;
  pl->state_epoch += 1;
  const yh_kernel_cfg c = resolve_cfg(pl);
  const bool masked = pl->p.solidSwitch != 0;
  int remaining = nsteps;
  auto fields = [&](int b, CFields* ci, Fields* co) {
    *ci = CFields{{pl->buf[b][0], pl->buf[b][1], pl->buf[b][2]}};
    *co = Fields{{pl->buf[b ^ 1][0], pl->buf[b ^ 1][1], pl->buf[b ^ 1][2]}};
  };
  ApdState apd = pl->apd;
  const ApdState* ap = pl->apd_on ? &apd : nullptr;
  int64_t done = pl->count;
  if (plan_general(pl)) {
    const int order = pl->p.timeIntOrder;
    if (order > 1) {  // completes an allocation an earlier failure left partial
      for (int q = 0; q < 9; ++q)
        if (!pl->gscratch[q]) YH_HIP(hipMalloc(&pl->gscratch[q], pl->n * sizeof(float)));
    }
    const GParams g = make_gparams(pl->p, pl->stim_on, pl->stim_x, pl->stim_y);
    for (int q = 0; q < nsteps; ++q) {
      CFields ci;
      Fields co;
      fields(pl->cur, &ci, &co);
      apd.count0 = (int)done;
      YH_HIP(launch_general_step(pl->p.model, g, order, ci, co, Fields{{nullptr, nullptr, nullptr}},
                                 masked ? pl->solid_d : nullptr, order > 1 ? pl->gscratch : nullptr,
                                 s, ap));
      pl->cur ^= 1;
      done += 1;
    }
    pl->have_prev = true;
    pl->past_prev_u = false;
    pl->count += nsteps;
    return YH_OK;
  }
  // keep_prev: the last tiled launch writes u one substep before its end
  // (shapes built with that output, no APD), else a single-step tail
  const bool tiled = c.variant == YH_KERNEL_TILED;
  const bool prev_fused =
      tiled && pl->keep_prev && !ap && tiled_prev_supported(c.rows, c.waves, c.cols);
  const int tail = (pl->keep_prev && !prev_fused) ? 1 : 0;
  if (prev_fused && !pl->prev_u) YH_HIP(hipMalloc(&pl->prev_u, pl->n * sizeof(float)));
  if (tiled) {
    const int body = remaining - tail, r = body % c.substeps;
    const bool merge = !ap && r > 0 && body > c.substeps && merge_ok(pl, c, r, prev_fused);
    while (remaining > tail) {
      int K = std::min(c.substeps, remaining - tail), rows = c.rows;
      if (merge && remaining - tail == c.substeps + r) {  // the last full launch + the rest
        K = c.substeps + r;
        rows = c.rows + 1;
      }
      CFields ci;
      Fields co;
      fields(pl->cur, &ci, &co);
      apd.count0 = (int)done;
      float* prev = (prev_fused && K == remaining) ? pl->prev_u : nullptr;
      KParams kl = pl->k;
      kl.wt = !ap && write_through(pl, rows, c.waves, c.cols, K, prev != nullptr);
      YH_HIP(launch_tiled(pl->p.model, masked, rows, c.waves, c.cols, K, kl, ci, co,
                          masked ? pl->codes_d : nullptr, s, ap, prev));
      pl->cur ^= 1;
      remaining -= K;
      done += K;
    }
  }
  while (remaining > 0) {
    CFields ci;
    Fields co;
    fields(pl->cur, &ci, &co);
    apd.count0 = (int)done;
    YH_HIP(launch_direct_codes(pl->p.model, pl->k, ci, co, masked ? pl->codes_d : nullptr, s,
                               ap));
    pl->cur ^= 1;
    remaining -= 1;
    done += 1;
  }
  pl->have_prev = !tiled || pl->keep_prev || c.substeps == 1;
  pl->past_prev_u = prev_fused;
  pl->count += nsteps;
  return YH_OK;
}

static void plan_free(yh_plan* pl) {
  if (!pl) return;
  for (int b = 0; b < 2; ++b)
    for (int a = 0; a < 3; ++a)
      if (pl->buf[b][a]) (void)hipFree(pl->buf[b][a]);
  if (pl->solid_d) (void)hipFree(pl->solid_d);
  if (pl->codes_d) (void)hipFree(pl->codes_d);
  if (pl->tips_d) (void)hipFree(pl->tips_d);
  if (pl->keys_d) (void)hipFree(pl->keys_d);
  if (pl->tipcount_d) (void)hipFree(pl->tipcount_d);
  if (pl->flag_d) (void)hipFree(pl->flag_d);
  if (pl->prev_u) (void)hipFree(pl->prev_u);
  void* apd[] = {pl->apd.front, pl->apd.back, pl->apd.apd1, pl->apd.apd2, pl->apd.first};
  for (void* q : apd)
    if (q) (void)hipFree(q);
  if (pl->sr) sr_free(pl->sr);
  for (float* q : pl->gscratch)
    if (q) (void)hipFree(q);
  void* cont[] = {pl->cont_pts, pl->cont_keys, pl->cont_count, pl->cont_area, pl->sapd};
  for (void* q : cont)
    if (q) (void)hipFree(q);
  delete pl;
}

}  // namespace yh

using namespace yh;

extern "C" {

int yh_plan_create(const yh_params* p, yh_plan** out) {
  if (!p || !out) return YH_EINVAL;
  *out = nullptr;
  yh_derived d;
  int rc = yh_params_derive(p, &d);
  if (rc) return rc;
  yh_plan* pl = new (std::nothrow) yh_plan();
  if (!pl) return YH_ENOMEM;
  pl->p = *p;
  pl->d = d;
  make_kparams(*p, d, &pl->k);
  pl->nv = nvars_of(p->model);
  pl->n = (size_t)p->nx * (size_t)p->ny;
  pl->tipcap = 500000;  // TIPVECSIZE, globalVariables.cuh:12
  hipError_t e = hipGetDevice(&pl->device);
  if (e == hipSuccess)
    e = hipDeviceGetAttribute(&pl->cus, hipDeviceAttributeMultiprocessorCount, pl->device);
  const size_t bytes = pl->n * sizeof(float);
  for (int b = 0; b < 2 && e == hipSuccess; ++b)
    for (int a = 0; a < pl->nv && e == hipSuccess; ++a) e = hipMalloc(&pl->buf[b][a], bytes);
  if (e == hipSuccess) e = hipMalloc(&pl->tips_d, sizeof(yh_tip) * pl->tipcap);
  if (e == hipSuccess) e = hipMalloc(&pl->keys_d, sizeof(int) * pl->tipcap);
  if (e == hipSuccess) e = hipMalloc(&pl->tipcount_d, sizeof(int));
  if (e == hipSuccess) e = hipMalloc(&pl->flag_d, sizeof(int));
  if (e == hipSuccess) e = hipMemset(pl->tipcount_d, 0, sizeof(int));
  for (int b = 0; b < 2 && e == hipSuccess; ++b)
    for (int a = 0; a < pl->nv && e == hipSuccess; ++a) e = hipMemset(pl->buf[b][a], 0, bytes);
  if (e != hipSuccess) {
    rc = hip_fail(e, "yh_plan_create");
    plan_free(pl);
    return rc;
  }
  *out = pl;
  return YH_OK;
}

void yh_plan_destroy(yh_plan* plan) {
  if (plan) (void)hipDeviceSynchronize();
  plan_free(plan);
}

int yh_plan_params(const yh_plan* pl, yh_params* p, yh_derived* d) {
  if (!pl) return YH_EINVAL;
  if (p) *p = pl->p;
  if (d) *d = pl->d;
  return YH_OK;
}

int yh_set_mask(yh_plan* pl, const uint8_t* solid) {
  if (!pl || !solid) return YH_EINVAL;
  if (!pl->solid_d) YH_HIP(hipMalloc(&pl->solid_d, pl->n));
  if (!pl->codes_d) YH_HIP(hipMalloc(&pl->codes_d, pl->n));
  std::vector<uint8_t> m(pl->n);
  for (size_t c = 0; c < pl->n; ++c) m[c] = solid[c] ? 1 : 0;
  YH_HIP(hipMemcpy(pl->solid_d, m.data(), pl->n, hipMemcpyHostToDevice));
  YH_HIP(launch_codes(pl->solid_d, pl->codes_d, pl->p.nx, pl->p.ny, nullptr));
  YH_HIP(hipDeviceSynchronize());
  pl->mask_set = true;
  pl->state_epoch += 1;
  return YH_OK;
}

int yh_set_state(yh_plan* pl, const float* u, const float* v, const float* w) {
  if (!pl || !u || !v || (pl->nv == 3 && !w)) return YH_EINVAL;
  const float* src[3] = {u, v, w};
  const size_t bytes = pl->n * sizeof(float);
  YH_HIP(hipDeviceSynchronize());
  for (int a = 0; a < pl->nv; ++a)
    YH_HIP(hipMemcpy(pl->buf[pl->cur][a], src[a], bytes, hipMemcpyHostToDevice));
  pl->have_prev = false;
  pl->count = 0;
  pl->state_epoch += 1;
  return YH_OK;
}

int yh_get_state(yh_plan* pl, float* u, float* v, float* w) {
  if (!pl || !u || !v || (pl->nv == 3 && !w)) return YH_EINVAL;
  float* dst[3] = {u, v, w};
  const size_t bytes = pl->n * sizeof(float);
  YH_HIP(hipDeviceSynchronize());
  for (int a = 0; a < pl->nv; ++a)
    YH_HIP(hipMemcpy(dst[a], pl->buf[pl->cur][a], bytes, hipMemcpyDeviceToHost));
  return YH_OK;
}

int yh_state_dev(yh_plan* pl, float** u, float** v, float** w) {
  if (!pl) return YH_EINVAL;
  pl->state_epoch += 1;  // the caller may write the state through these
  if (u) *u = pl->buf[pl->cur][0];
  if (v) *v = pl->buf[pl->cur][1];
  if (w) *w = pl->nv == 3 ? pl->buf[pl->cur][2] : nullptr;
  return YH_OK;
}

int yh_set_kernel(yh_plan* pl, const yh_kernel_cfg* cfg) {
  if (!pl || !cfg) return YH_EINVAL;
  if (cfg->variant < YH_KERNEL_AUTO || cfg->variant > YH_KERNEL_TILED) return YH_EINVAL;
  if (cfg->variant == YH_KERNEL_TILED) {
    if (!tiled_supported(cfg->rows, cfg->waves, cfg->cols)) return YH_EINVAL;
    if (cfg->substeps < 1 || cfg->substeps > 16) return YH_EINVAL;
    if (2 * cfg->substeps >= 64 * cfg->cols || 2 * cfg->substeps >= cfg->rows * cfg->waves)
      return YH_EINVAL;
  }
  pl->cfg = *cfg;
  return YH_OK;
}

int yh_get_kernel(const yh_plan* pl, yh_kernel_cfg* cfg) {
  if (!pl || !cfg) return YH_EINVAL;
  *cfg = resolve_cfg(pl);
  return YH_OK;
}

int yh_step(yh_plan* pl, int nsteps, void* stream) {
  if (!pl) return YH_EINVAL;
  return plan_advance(pl, nsteps, (hipStream_t)stream);
}

int64_t yh_step_count(const yh_plan* pl) { return pl ? pl->count : -1; }

int yh_track_tips(yh_plan* pl, void* stream) {
  if (!pl) return YH_EINVAL;
  if (!pl->have_prev) return YH_ESTATE;
  // main.cu:961-964 after swapSoA: g_past = previous state, g_present = newest.
  const double t = pl->d.dt * (double)pl->count;
  const float* past = pl->past_prev_u ? pl->prev_u : pl->buf[pl->cur ^ 1][0];
  YH_HIP(launch_tips(pl->k, pl->p.solidSwitch, pl->p.tipOffsetX, pl->p.tipOffsetY, pl->p.Uth,
                     past, pl->buf[pl->cur][0], t, pl->tips_d, pl->keys_d,
                     pl->tipcount_d, pl->tipcap, (hipStream_t)stream));
  return YH_OK;
}

int yh_read_tips(yh_plan* pl, yh_tip* out, int cap, int* n) {
  if (!pl || !n || cap < 0 || (cap > 0 && !out)) return YH_EINVAL;
  YH_HIP(hipDeviceSynchronize());
  int cnt = 0;
  YH_HIP(hipMemcpy(&cnt, pl->tipcount_d, sizeof(int), hipMemcpyDeviceToHost));
  const int have = std::min(cnt, pl->tipcap);
  std::vector<yh_tip> t(have);
  std::vector<int> k(have);
  if (have > 0) {
    YH_HIP(hipMemcpy(t.data(), pl->tips_d, sizeof(yh_tip) * have, hipMemcpyDeviceToHost));
    YH_HIP(hipMemcpy(k.data(), pl->keys_d, sizeof(int) * have, hipMemcpyDeviceToHost));
  }
  std::vector<int> order(have);
  for (int q = 0; q < have; ++q) order[q] = q;
  std::sort(order.begin(), order.end(), [&](int a, int b) {
    if (t[a].t != t[b].t) return t[a].t < t[b].t;
    return k[a] < k[b];
  });
  const int m = std::min(have, cap);
  for (int q = 0; q < m; ++q) out[q] = t[order[q]];
  *n = cnt;
  return (cnt > pl->tipcap || cnt > cap) ? YH_EOVERFLOW : YH_OK;
}

int yh_clear_tips(yh_plan* pl) {
  if (!pl) return YH_EINVAL;
  YH_HIP(hipDeviceSynchronize());
  YH_HIP(hipMemset(pl->tipcount_d, 0, sizeof(int)));
  return YH_OK;
}

int yh_set_stimulus(yh_plan* pl, int on, int x, int y) {
  if (!pl) return YH_EINVAL;
  pl->stim_on = on != 0;
  pl->stim_x = x;
  pl->stim_y = y;
  return YH_OK;
}

int yh_set_option(yh_plan* pl, int option, int64_t value) {
  if (!pl) return YH_EINVAL;
  switch (option) {
    case YH_OPT_KEEP_PREV:
      pl->keep_prev = value != 0;
      return YH_OK;
    case YH_OPT_SR_LAUNCHES:
      if (value != 2 && value != 3) return YH_EINVAL;
      pl->sr_launches = (int)value;
      return YH_OK;
    default:
      return YH_EINVAL;
  }
}

int yh_apd_enable(yh_plan* pl, int on, double threshold) {
  if (!pl) return YH_EINVAL;
  if (!on) {
    pl->apd_on = false;
    return YH_OK;
  }
  // each buffer on its own, so a call after a failed allocation completes it
  if (!pl->apd.front) YH_HIP(hipMalloc(&pl->apd.front, pl->n * sizeof(float)));
  if (!pl->apd.back) YH_HIP(hipMalloc(&pl->apd.back, pl->n * sizeof(float)));
  if (!pl->apd.apd1) YH_HIP(hipMalloc(&pl->apd.apd1, pl->n * sizeof(float)));
  if (!pl->apd.apd2) YH_HIP(hipMalloc(&pl->apd.apd2, pl->n * sizeof(float)));
  if (!pl->apd.first) YH_HIP(hipMalloc(&pl->apd.first, pl->n));
  YH_HIP(hipDeviceSynchronize());
  YH_HIP(hipMemset(pl->apd.front, 0, pl->n * sizeof(float)));
  YH_HIP(hipMemset(pl->apd.back, 0, pl->n * sizeof(float)));
  YH_HIP(hipMemset(pl->apd.apd1, 0, pl->n * sizeof(float)));
  YH_HIP(hipMemset(pl->apd.apd2, 0, pl->n * sizeof(float)));
  YH_HIP(hipMemset(pl->apd.first, 0, pl->n));
  pl->apd.th = (float)threshold;
  pl->apd.dt = (float)pl->d.dt;
  pl->apd_on = true;
  return YH_OK;
}

int yh_apd_get(yh_plan* pl, float* apd1, float* apd2, float* front, float* back) {
  if (!pl) return YH_EINVAL;
  if (!pl->apd.front) return YH_ESTATE;
  YH_HIP(hipDeviceSynchronize());
  const size_t b = pl->n * sizeof(float);
  if (apd1) YH_HIP(hipMemcpy(apd1, pl->apd.apd1, b, hipMemcpyDeviceToHost));
  if (apd2) YH_HIP(hipMemcpy(apd2, pl->apd.apd2, b, hipMemcpyDeviceToHost));
  if (front) YH_HIP(hipMemcpy(front, pl->apd.front, b, hipMemcpyDeviceToHost));
  if (back) YH_HIP(hipMemcpy(back, pl->apd.back, b, hipMemcpyDeviceToHost));
  return YH_OK;
}

int yh_check_finite(yh_plan* pl, void* stream) {
  if (!pl) return YH_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  YH_HIP(hipMemsetAsync(pl->flag_d, 0, sizeof(int), s));
  const float* f[3] = {pl->buf[pl->cur][0], pl->buf[pl->cur][1], pl->buf[pl->cur][2]};
  YH_HIP(launch_finite(f, pl->nv, pl->n, pl->flag_d, s));
  int flag = 0;
  YH_HIP(hipMemcpyAsync(&flag, pl->flag_d, sizeof(int), hipMemcpyDeviceToHost, s));
  YH_HIP(hipStreamSynchronize(s));
  return flag ? YH_ENONFINITE : YH_OK;
}

}  // extern "C"
