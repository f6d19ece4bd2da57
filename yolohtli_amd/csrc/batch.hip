// batch.hip -- many tissues of uneven size and parameters in one launch.
//
// All tissues share one model and one mask mode; sizes, D and the ionic time
// constants vary per tissue.  The state of every tissue lives in one
// concatenated SoA allocation per variable (tissue t at offset off[t],
// 256-byte aligned).  A work list holds one item per (tissue, tile); the
// kernel k_batch walks it with a persistent grid-stride loop in which every
// XCD owns a contiguous range of items, and reads each item's KParams from a
// per-tissue table (uniform per workgroup, so they land in scalar registers).
// One launch advances every tissue by K steps; a whole run of many launches is
// captured once into a hipGraph and replayed.  yh_batch_step_each launches the
// single-tissue kernel once per tissue per chunk instead (the comparison).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <new>
#include <tuple>
#include <vector>

#include "plan.h"
#include "tile.h"

namespace yh {

struct BatchTissue {
  KParams k;
  unsigned long long off;
  int tiles_x;
  int pad;
};

// ISO = 1 when every tissue of the batch has isotropic diffusion (fixed at
// compile time, as in k_tiled), -1 to take each tissue's flag at run time.
template <int MODEL, bool MASKED, int R, int NW, int CW, int ISO>
__global__ __launch_bounds__(64 * NW) void k_batch(const BatchTissue* __restrict__ tissues,
                                                   const int2* __restrict__ items, int nitems,
                                                   CFields in, Fields out,
                                                   const uint8_t* __restrict__ codes, int K) {
  __shared__ TileLds<MODEL, NW, CW> lds;
  const int G = gridDim.x;
  int x = blockIdx.x & 7, local = blockIdx.x >> 3;
  // XCD x handles items [lo, hi) with its workgroups b = x, x+8, x+16, ...
  int per_x = (G + 7 - x) >> 3;
  int lo = (int)(((long long)nitems * x) >> 3);
  int hi = (int)(((long long)nitems * (x + 1)) >> 3);
  if (G < 8) {  // too few workgroups to give every XCD one: plain striding
    local = blockIdx.x;
    per_x = G;
    lo = 0;
    hi = nitems;
  }
  for (int it = lo + local; it < hi; it += per_x) {
    const int2 item = items[it];
    const BatchTissue& T = tissues[item.x];
    const unsigned long long off = T.off;
    CFields ti{{in.f[0] + off, in.f[1] + off, in.f[2] ? in.f[2] + off : nullptr}};
    Fields to{{out.f[0] + off, out.f[1] + off, out.f[2] ? out.f[2] + off : nullptr}};
    const int ty = item.y / T.tiles_x;
    __syncthreads();  // the previous item's last reads of lds are done
    tile_advance<MODEL, MASKED, R, NW, CW, false, false, ISO>(
        T.k, ti, to, MASKED ? codes + off : nullptr, K, item.y - ty * T.tiles_x, ty, lds);
  }
}

template <int MODEL, bool MASKED, int R, int NW, int CW, int ISO>
static hipError_t launch_batch_t(int grid, const BatchTissue* t, const int2* items, int nitems,
                                 CFields in, Fields out, const uint8_t* codes, int K,
                                 hipStream_t s) {
  hipLaunchKernelGGL((k_batch<MODEL, MASKED, R, NW, CW, ISO>), dim3(grid), dim3(64 * NW), 0, s,
                     t, items, nitems, in, out, codes, K);
  return hipGetLastError();
}

template <int MODEL, bool MASKED, int R, int NW, int CW, int ISO>
static int max_resident(int* out) {
  int per_cu = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &per_cu, (const void*)k_batch<MODEL, MASKED, R, NW, CW, ISO>, 64 * NW, 0);
  if (e != hipSuccess) return hip_fail(e, "hipOccupancyMaxActiveBlocksPerMultiprocessor");
  int dev = 0, cus = 0;
  YH_HIP(hipGetDevice(&dev));
  YH_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  *out = std::max(1, per_cu) * cus;
  return YH_OK;
}

#define YH_BATCH_DISPATCH(FN, ...)                                                         \
  do {                                                                                     \
    if (model == M2V) {                                                                    \
      if (masked) { YH_BATCH_SHAPES(M2V, true, FN, __VA_ARGS__) }                          \
      else { YH_BATCH_SHAPES(M2V, false, FN, __VA_ARGS__) }                                \
    } else {                                                                               \
      if (masked) { YH_BATCH_SHAPES(MFK, true, FN, __VA_ARGS__) }                          \
      else { YH_BATCH_SHAPES(MFK, false, FN, __VA_ARGS__) }                                \
    }                                                                                      \
  } while (0)

#define YH_BATCH_ONE(M, MK, RR, WW, CC, FN, ...) \
  if (rows == RR && waves == WW && cols == CC) return FN<M, MK, RR, WW, CC, -1>(__VA_ARGS__);

// the default shape also comes with isotropic diffusion fixed at compile time
#define YH_BATCH_ISO(M, MK, RR, WW, CC, FN, ...)                  \
  if (rows == RR && waves == WW && cols == CC)                    \
    return iso ? FN<M, MK, RR, WW, CC, 1>(__VA_ARGS__)            \
               : FN<M, MK, RR, WW, CC, -1>(__VA_ARGS__);

#define YH_BATCH_SHAPES(M, MK, FN, ...)            \
  YH_BATCH_ONE(M, MK, 4, 8, 1, FN, __VA_ARGS__)   \
  YH_BATCH_ISO(M, MK, 8, 8, 1, FN, __VA_ARGS__)   \
  YH_BATCH_ONE(M, MK, 4, 16, 1, FN, __VA_ARGS__)  \
  YH_BATCH_ONE(M, MK, 8, 16, 1, FN, __VA_ARGS__)  \
  YH_BATCH_ONE(M, MK, 16, 8, 1, FN, __VA_ARGS__)  \
  YH_BATCH_ONE(M, MK, 7, 16, 1, FN, __VA_ARGS__)  \
  YH_BATCH_ONE(M, MK, 4, 8, 2, FN, __VA_ARGS__)   \
  YH_BATCH_ONE(M, MK, 8, 8, 2, FN, __VA_ARGS__)   \
  YH_BATCH_ONE(M, MK, 4, 16, 2, FN, __VA_ARGS__)

static bool batch_shape_ok(int rows, int waves, int cols) {
  return (cols == 1 && ((rows == 4 && (waves == 8 || waves == 16)) ||
                        (rows == 8 && (waves == 8 || waves == 16)) || (rows == 16 && waves == 8) ||
                        (rows == 7 && waves == 16))) ||
         (cols == 2 && ((rows == 4 && (waves == 8 || waves == 16)) || (rows == 8 && waves == 8)));
}

static hipError_t launch_batch(int model, bool masked, bool iso, int rows, int waves, int cols,
                               int grid, const BatchTissue* t, const int2* items, int nitems, CFields in,
                               Fields out, const uint8_t* codes, int K, hipStream_t s) {
  if (K < 1) return hipErrorInvalidValue;
  YH_BATCH_DISPATCH(launch_batch_t, grid, t, items, nitems, in, out, codes, K, s);
  return hipErrorInvalidValue;
}

static int batch_resident(int model, bool masked, bool iso, int rows, int waves, int cols,
                          int* out) {
  YH_BATCH_DISPATCH(max_resident, out);
  return YH_EINVAL;
}

}  // namespace yh

struct yh_batch {
  int nt = 0, model = 0, nv = 2;
  bool masked = false;
  bool iso = true;  // every tissue has isotropic diffusion
  std::vector<yh_params> p;
  std::vector<yh_derived> d;
  std::vector<yh::KParams> k;
  std::vector<size_t> off, ncell;
  size_t total = 0;
  float* buf[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
  int cur = 0;
  uint8_t* solid_d = nullptr;
  uint8_t* codes_d = nullptr;
  std::vector<char> mask_set;
  // the configuration the autotuner picks for the 256-tissue workload; K is
  // capped at creation by the smallest tissue (reflected halos are exact for
  // at most min(nx, ny) - 1 substeps)
  yh_kernel_cfg cfg{YH_KERNEL_TILED, 6, 8, 8, 1};
  int persistent = 0;
  bool use_graph = true;
  int64_t count = 0;
  // per-K work lists: K -> (tissue table, items, nitems)
  struct WL {
    yh::BatchTissue* tissues = nullptr;
    int2* items = nullptr;
    int nitems = 0;
    yh_kernel_cfg cfg{};
  };
  std::map<int, WL> wl;
  // graph cache: (nsteps, cur, K, rows, waves, cols, persistent) -> exec
  std::map<std::tuple<int, int, int, int, int, int, int>, hipGraphExec_t> graphs;
  std::vector<hipGraph_t> graph_objs;
  hipStream_t cap_stream = nullptr;
};

namespace yh {

static void batch_clear_wl(yh_batch* b) {
  for (auto& kv : b->wl) {
    if (kv.second.tissues) (void)hipFree(kv.second.tissues);
    if (kv.second.items) (void)hipFree(kv.second.items);
  }
  b->wl.clear();
}

static void batch_clear_graphs(yh_batch* b) {
  for (auto& kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
  for (auto g : b->graph_objs) (void)hipGraphDestroy(g);
  b->graphs.clear();
  b->graph_objs.clear();
}

static int batch_worklist(yh_batch* b, int K, yh_batch::WL** out) {
  auto it = b->wl.find(K);
  if (it != b->wl.end() && it->second.cfg.rows == b->cfg.rows &&
      it->second.cfg.waves == b->cfg.waves && it->second.cfg.cols == b->cfg.cols) {
    *out = &it->second;
    return YH_OK;
  }
  if (it != b->wl.end()) {
    batch_clear_graphs(b);  // captured launches point at the list being replaced
    (void)hipDeviceSynchronize();
    (void)hipFree(it->second.tissues);
    (void)hipFree(it->second.items);
    b->wl.erase(it);
  }
  const int wout = 64 * b->cfg.cols - 2 * K, hout = b->cfg.rows * b->cfg.waves - 2 * K;
  if (wout <= 0 || hout <= 0) return YH_EINVAL;
  std::vector<BatchTissue> tab(b->nt);
  std::vector<int2> items;
  for (int t = 0; t < b->nt; ++t) {
    tab[t].k = b->k[t];
    tab[t].off = b->off[t];
    tab[t].tiles_x = (b->p[t].nx + wout - 1) / wout;
    tab[t].pad = 0;
    const int tiles_y = (b->p[t].ny + hout - 1) / hout;
    for (int q = 0; q < tab[t].tiles_x * tiles_y; ++q) items.push_back(make_int2(t, q));
  }
  yh_batch::WL w;
  w.cfg = b->cfg;
  w.nitems = (int)items.size();
  hipError_t e = hipMalloc(&w.tissues, sizeof(BatchTissue) * b->nt);
  if (e == hipSuccess) e = hipMalloc(&w.items, sizeof(int2) * items.size());
  if (e == hipSuccess)
    e = hipMemcpy(w.tissues, tab.data(), sizeof(BatchTissue) * b->nt, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(w.items, items.data(), sizeof(int2) * items.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {  // nothing half-built is kept
    if (w.tissues) (void)hipFree(w.tissues);
    if (w.items) (void)hipFree(w.items);
    return hip_fail(e, "batch work list");
  }
  b->wl[K] = w;
  *out = &b->wl[K];
  return YH_OK;
}

// Enqueue nsteps as launches of K = cfg.substeps (the last one shorter).
static int batch_enqueue(yh_batch* b, int nsteps, hipStream_t s) {
  const yh_kernel_cfg c = b->cfg;
  int remaining = nsteps;
  while (remaining > 0) {
    const int K = std::min(c.substeps, remaining);
    yh_batch::WL* w = nullptr;
    int rc = batch_worklist(b, K, &w);
    if (rc) return rc;
    int grid = w->nitems;
    if (b->persistent) {
      int res = 0;
      rc = batch_resident(b->model, b->masked, b->iso, c.rows, c.waves, c.cols, &res);
      if (rc) return rc;
      grid = std::min(grid, (res + 7) / 8 * 8);
    }
    const int bi = b->cur;
    CFields in{{b->buf[bi][0], b->buf[bi][1], b->buf[bi][2]}};
    Fields out{{b->buf[bi ^ 1][0], b->buf[bi ^ 1][1], b->buf[bi ^ 1][2]}};
    YH_HIP(launch_batch(b->model, b->masked, b->iso, c.rows, c.waves, c.cols, grid, w->tissues,
                        w->items, w->nitems, in, out, b->masked ? b->codes_d : nullptr, K, s));
    b->cur ^= 1;
    remaining -= K;
  }
  return YH_OK;
}

static void batch_free(yh_batch* b) {
  if (!b) return;
  batch_clear_graphs(b);
  batch_clear_wl(b);
  for (int q = 0; q < 2; ++q)
    for (int a = 0; a < 3; ++a)
      if (b->buf[q][a]) (void)hipFree(b->buf[q][a]);
  if (b->solid_d) (void)hipFree(b->solid_d);
  if (b->codes_d) (void)hipFree(b->codes_d);
  if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
  delete b;
}

}  // namespace yh

using namespace yh;

extern "C" {

int yh_batch_create(const yh_params* per_tissue, int ntissues, yh_batch** out) {
  if (!per_tissue || ntissues < 1 || !out) return YH_EINVAL;
  *out = nullptr;
  yh_batch* b = new (std::nothrow) yh_batch();
  if (!b) return YH_ENOMEM;
  b->nt = ntissues;
  b->model = per_tissue[0].model;
  b->masked = per_tissue[0].solidSwitch != 0;
  b->nv = nvars_of(b->model);
  size_t off = 0;
  for (int t = 0; t < ntissues; ++t) {
    const yh_params& p = per_tissue[t];
    if (p.model != b->model || (p.solidSwitch != 0) != b->masked || p.reduceSym) {
      batch_free(b);
      return YH_EINVAL;
    }
    if (!standard_branch(p)) {
      batch_free(b);
      return YH_EUNSUPPORTED;
    }
    yh_derived d;
    int rc = yh_params_derive(&p, &d);
    if (rc) {
      batch_free(b);
      return rc;
    }
    KParams k;
    make_kparams(p, d, &k);
    b->p.push_back(p);
    b->d.push_back(d);
    b->k.push_back(k);
    b->iso = b->iso && k.iso;
    b->off.push_back(off);
    const size_t n = (size_t)p.nx * p.ny;
    b->ncell.push_back(n);
    off += (n + 63) / 64 * 64;
  }
  b->total = off;
  b->mask_set.assign(ntissues, 0);
  for (int t = 0; t < ntissues; ++t)
    b->cfg.substeps = std::min(b->cfg.substeps, std::min(b->p[t].nx, b->p[t].ny) - 1);
  hipError_t e = hipSuccess;
  for (int q = 0; q < 2 && e == hipSuccess; ++q)
    for (int a = 0; a < b->nv && e == hipSuccess; ++a) {
      e = hipMalloc(&b->buf[q][a], sizeof(float) * b->total);
      if (e == hipSuccess) e = hipMemset(b->buf[q][a], 0, sizeof(float) * b->total);
    }
  if (e == hipSuccess && b->masked) {
    e = hipMalloc(&b->solid_d, b->total);
    if (e == hipSuccess) e = hipMalloc(&b->codes_d, b->total);
    if (e == hipSuccess) e = hipMemset(b->codes_d, 0x55, b->total);
  }
  if (e != hipSuccess) {
    int rc = hip_fail(e, "yh_batch_create");
    batch_free(b);
    return rc;
  }
  *out = b;
  return YH_OK;
}

void yh_batch_destroy(yh_batch* b) {
  if (b) (void)hipDeviceSynchronize();
  batch_free(b);
}

int yh_batch_size(const yh_batch* b, int* ntissues, int64_t* total_cells) {
  if (!b) return YH_EINVAL;
  if (ntissues) *ntissues = b->nt;
  if (total_cells) {
    int64_t s = 0;
    for (size_t n : b->ncell) s += (int64_t)n;
    *total_cells = s;
  }
  return YH_OK;
}

int yh_batch_set_mask(yh_batch* b, int t, const uint8_t* solid) {
  if (!b || t < 0 || t >= b->nt || !solid) return YH_EINVAL;
  if (!b->masked) return YH_ESTATE;
  const size_t n = b->ncell[t];
  std::vector<uint8_t> m(n);
  for (size_t c = 0; c < n; ++c) m[c] = solid[c] ? 1 : 0;
  YH_HIP(hipDeviceSynchronize());
  YH_HIP(hipMemcpy(b->solid_d + b->off[t], m.data(), n, hipMemcpyHostToDevice));
  YH_HIP(launch_codes(b->solid_d + b->off[t], b->codes_d + b->off[t], b->p[t].nx, b->p[t].ny,
                      nullptr));
  YH_HIP(hipDeviceSynchronize());
  b->mask_set[t] = 1;
  return YH_OK;
}

int yh_batch_set_state(yh_batch* b, int t, const float* u, const float* v, const float* w) {
  if (!b || t < 0 || t >= b->nt || !u || !v || (b->nv == 3 && !w)) return YH_EINVAL;
  const float* src[3] = {u, v, w};
  YH_HIP(hipDeviceSynchronize());
  for (int a = 0; a < b->nv; ++a)
    YH_HIP(hipMemcpy(b->buf[b->cur][a] + b->off[t], src[a], sizeof(float) * b->ncell[t],
                     hipMemcpyHostToDevice));
  return YH_OK;
}

int yh_batch_get_state(yh_batch* b, int t, float* u, float* v, float* w) {
  if (!b || t < 0 || t >= b->nt || !u || !v || (b->nv == 3 && !w)) return YH_EINVAL;
  float* dst[3] = {u, v, w};
  YH_HIP(hipDeviceSynchronize());
  for (int a = 0; a < b->nv; ++a)
    YH_HIP(hipMemcpy(dst[a], b->buf[b->cur][a] + b->off[t], sizeof(float) * b->ncell[t],
                     hipMemcpyDeviceToHost));
  return YH_OK;
}

int yh_batch_set_kernel(yh_batch* b, const yh_kernel_cfg* cfg, int persistent) {
  if (!b || !cfg) return YH_EINVAL;
  if (cfg->variant != YH_KERNEL_TILED || !batch_shape_ok(cfg->rows, cfg->waves, cfg->cols))
    return YH_EINVAL;
  if (cfg->substeps < 1 || cfg->substeps > 16 || 2 * cfg->substeps >= 64 * cfg->cols ||
      2 * cfg->substeps >= cfg->rows * cfg->waves)
    return YH_EINVAL;
  for (int t = 0; t < b->nt; ++t)
    if (cfg->substeps > std::min(b->p[t].nx, b->p[t].ny) - 1) return YH_EINVAL;
  b->cfg = *cfg;
  b->persistent = persistent ? 1 : 0;
  return YH_OK;
}

int yh_batch_set_graph(yh_batch* b, int on) {
  if (!b) return YH_EINVAL;
  b->use_graph = on != 0;
  return YH_OK;
}

int yh_batch_step(yh_batch* b, int nsteps, void* stream) {
  if (!b || nsteps < 0) return YH_EINVAL;
  if (nsteps == 0) return YH_OK;
  if (b->masked)
    for (int t = 0; t < b->nt; ++t)
      if (!b->mask_set[t]) return YH_ESTATE;
  hipStream_t s = (hipStream_t)stream;
  if (!b->use_graph) {
    int rc = batch_enqueue(b, nsteps, s);
    if (rc) return rc;
    b->count += nsteps;
    return YH_OK;
  }
  const auto key = std::make_tuple(nsteps, b->cur, b->cfg.substeps, b->cfg.rows, b->cfg.waves,
                                   b->cfg.cols, b->persistent);
  auto it = b->graphs.find(key);
  const int nlaunch = (nsteps + b->cfg.substeps - 1) / b->cfg.substeps;
  if (it == b->graphs.end()) {
    // build the work lists outside the capture (they allocate)
    for (int rem = nsteps; rem > 0; rem -= b->cfg.substeps) {
      yh_batch::WL* w = nullptr;
      int rc = batch_worklist(b, std::min(b->cfg.substeps, rem), &w);
      if (rc) return rc;
    }
    int res = 0;
    int rc = batch_resident(b->model, b->masked, b->iso, b->cfg.rows, b->cfg.waves, b->cfg.cols,
                            &res);
    if (rc) return rc;
    if (!b->cap_stream) YH_HIP(hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking));
    const int cur0 = b->cur;
    YH_HIP(hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal));
    rc = batch_enqueue(b, nsteps, b->cap_stream);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->cap_stream, &g);
    b->cur = cur0;  // capture only recorded the launches
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    YH_HIP(e);
    hipGraphExec_t ex = nullptr;
    YH_HIP(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
    b->graph_objs.push_back(g);
    it = b->graphs.emplace(key, ex).first;
  }
  YH_HIP(hipGraphLaunch(it->second, s));
  if (nlaunch & 1) b->cur ^= 1;
  b->count += nsteps;
  return YH_OK;
}

int yh_batch_step_each(yh_batch* b, int nsteps, void* stream) {
  if (!b || nsteps < 0) return YH_EINVAL;
  if (b->masked)
    for (int t = 0; t < b->nt; ++t)
      if (!b->mask_set[t]) return YH_ESTATE;
  hipStream_t s = (hipStream_t)stream;
  const yh_kernel_cfg c = b->cfg;
  int remaining = nsteps;
  while (remaining > 0) {
    const int K = std::min(c.substeps, remaining);
    const int bi = b->cur;
    for (int t = 0; t < b->nt; ++t) {
      const size_t o = b->off[t];
      CFields in{{b->buf[bi][0] + o, b->buf[bi][1] + o, b->nv > 2 ? b->buf[bi][2] + o : nullptr}};
      Fields out{{b->buf[bi ^ 1][0] + o, b->buf[bi ^ 1][1] + o,
                  b->nv > 2 ? b->buf[bi ^ 1][2] + o : nullptr}};
      YH_HIP(launch_tiled(b->model, b->masked, c.rows, c.waves, c.cols, K, b->k[t], in, out,
                          b->masked ? b->codes_d + o : nullptr, s));
    }
    b->cur ^= 1;
    remaining -= K;
  }
  b->count += nsteps;
  return YH_OK;
}

int yh_batch_autotune(yh_batch* b, int steps, void* stream, yh_kernel_cfg* best, float* best_ms) {
  if (!b || steps < 1) return YH_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  // save the state
  std::vector<float*> save(b->nv, nullptr);
  for (int a = 0; a < b->nv; ++a) {
    YH_HIP(hipMalloc(&save[a], sizeof(float) * b->total));
    YH_HIP(hipMemcpyAsync(save[a], b->buf[b->cur][a], sizeof(float) * b->total,
                          hipMemcpyDeviceToDevice, s));
  }
  const yh_kernel_cfg cfg0 = b->cfg;
  const int pers0 = b->persistent;
  const bool graph0 = b->use_graph;
  const int64_t count0 = b->count;
  b->use_graph = false;
  hipEvent_t e0, e1;
  YH_HIP(hipEventCreate(&e0));
  YH_HIP(hipEventCreate(&e1));
  int minside = 1 << 30;
  for (int t = 0; t < b->nt; ++t) minside = std::min(minside, std::min(b->p[t].nx, b->p[t].ny));
  const int shapes[][3] = {{4, 8, 1}, {8, 8, 1}, {4, 16, 1}, {8, 16, 1}, {7, 16, 1},
                           {4, 8, 2}, {4, 16, 2}};
  const int Ks[] = {2, 4, 6, 8, 9, 10, 12};
  yh_kernel_cfg bestc = cfg0;
  int bestp = pers0;
  float bestt = 1e30f;
  int rc = YH_OK;
  for (int K : Ks) {
    if (K > minside - 1) continue;
    for (auto& sh : shapes) {
      if (2 * K >= 64 * sh[2] || 2 * K >= sh[0] * sh[1]) continue;
      for (int pers = 0; pers < 2; ++pers) {
        b->cfg = yh_kernel_cfg{YH_KERNEL_TILED, K, sh[0], sh[1], sh[2]};
        b->persistent = pers;
        const int n = std::max(K, (steps / K) * K);
        rc = batch_enqueue(b, K, s);
        if (rc) break;
        if (hipEventRecord(e0, s) != hipSuccess) { rc = YH_EHIP; break; }
        rc = batch_enqueue(b, n, s);
        if (rc) break;
        if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
          rc = YH_EHIP;
          break;
        }
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (ms / n < bestt) {
          bestt = ms / n;
          bestc = b->cfg;
          bestp = pers;
        }
      }
      if (rc) break;
    }
    if (rc) break;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  for (int a = 0; a < b->nv; ++a)
    (void)hipMemcpyAsync(b->buf[b->cur][a], save[a], sizeof(float) * b->total,
                         hipMemcpyDeviceToDevice, s);
  (void)hipStreamSynchronize(s);
  for (int a = 0; a < b->nv; ++a) (void)hipFree(save[a]);
  b->use_graph = graph0;
  b->count = count0;
  if (rc) {
    b->cfg = cfg0;
    b->persistent = pers0;
    return rc;
  }
  b->cfg = bestc;
  b->persistent = bestp;
  if (best) *best = bestc;
  if (best_ms) *best_ms = bestt;
  return YH_OK;
}

int yh_batch_get_kernel(const yh_batch* b, yh_kernel_cfg* cfg, int* persistent) {
  if (!b) return YH_EINVAL;
  if (cfg) *cfg = b->cfg;
  if (persistent) *persistent = b->persistent;
  return YH_OK;
}

}  // extern "C"
