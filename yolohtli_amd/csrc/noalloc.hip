// noalloc.hip -- yh_step_noalloc: the plain stepper on caller-owned device
// fields without any device allocation.
//
// yh_step keeps its state, stencil codes and previous-u output in buffers the
// plan allocates (some lazily, on first use).  Callers that cannot allocate
// device memory -- code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc -- get the same kernels here with those
// buffers in one file-scope device array, `scratch`, of a fixed size: the
// stencil codes (1 B per cell) and one intermediate state (4 B per variable
// and cell) for grids of at most kNoAllocMaxCells cells.  Each device holds
// its own copy of the array until the process ends.  Calls share that copy:
// two calls that overlap on one device race (serialise them, e.g. on one
// stream).  Every byte a call reads from it is written by the same call
// first (codes by k_codes, the state by the launch before), so nothing needs
// clearing between calls.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "plan.h"

namespace yh {

constexpr size_t kNoAllocMaxCells = size_t(1) << 21;  // e.g. 1448 x 1448, 2048 x 1024
constexpr size_t kFieldBytes = kNoAllocMaxCells * sizeof(float);  // a multiple of 256
constexpr size_t kScratchBytes = 3 * kFieldBytes + kNoAllocMaxCells;
constexpr int kMaxDevices = 64;

// Layout: field 0, 1, 2 of the intermediate state, then the stencil codes.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

static std::mutex g_mu;
static unsigned char* g_scratch[kMaxDevices] = {};  // device address of `scratch` per device
static int g_cus[kMaxDevices] = {};

// The device address of this device's copy of `scratch` (the host-side name
// is a stand-in, not a device address) and the device's CU count, looked up
// once per device with that device current.
static int scratch_of_current_device(unsigned char** ptr, int* cus) {
  int dev = 0;
  YH_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return YH_EINVAL;
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    YH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    int n = 0;
    YH_HIP(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev));
    g_cus[dev] = n;
    g_scratch[dev] = static_cast<unsigned char*>(p);
  }
  *ptr = g_scratch[dev];
  *cus = g_cus[dev];
  return YH_OK;
}

}  // namespace yh

using namespace yh;

extern "C" {

int yh_step_noalloc(const yh_params* p, float* const out[3], const float* const in[3],
                    const uint8_t* solid, int nsteps, void* stream) {
  if (!p || !out || !in || nsteps < 0) return YH_EINVAL;
  yh_derived d;
  int rc = yh_params_derive(p, &d);
  if (rc) return rc;
  if (!standard_branch(*p)) return YH_EUNSUPPORTED;
  const int nv = nvars_of(p->model);
  for (int a = 0; a < nv; ++a)
    if (!in[a] || !out[a] || in[a] == out[a]) return YH_EINVAL;
  if (p->solidSwitch && !solid) return YH_EINVAL;
  const size_t n = (size_t)p->nx * (size_t)p->ny;
  if (n > kNoAllocMaxCells) return YH_ESCRATCH;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (nsteps == 0) {
    for (int a = 0; a < nv; ++a)
      YH_HIP(hipMemcpyAsync(out[a], in[a], n * sizeof(float), hipMemcpyDeviceToDevice, s));
    return YH_OK;
  }
  unsigned char* base = nullptr;
  int cus = 0;
  rc = scratch_of_current_device(&base, &cus);
  if (rc) return rc;
  float* mid[3] = {nullptr, nullptr, nullptr};
  for (int a = 0; a < nv; ++a) mid[a] = reinterpret_cast<float*>(base + a * kFieldBytes);
  uint8_t* codes = base + 3 * kFieldBytes;
  const bool masked = p->solidSwitch != 0;
  if (masked) YH_HIP(launch_codes(solid, codes, p->nx, p->ny, s));
  KParams k;
  make_kparams(*p, d, &k);
  // the kernel a plan left on YH_KERNEL_AUTO runs, with the plan's store
  // policy per launch (neither the remainder merge nor the store policy
  // changes a value: the result does not depend on where launches split,
  // so it equals yh_step's bit for bit)
  const bool pairs = tiled_per_cu(p->model, masked, k, 4, 16, 1, false) >= 2;
  const yh_kernel_cfg c = auto_cfg(p->nx, p->ny, cus, pairs);
  const int per_cu = tiled_per_cu(p->model, masked, k, c.rows, c.waves, c.cols, false);
  const int K = c.substeps;
  const int launches = (nsteps + K - 1) / K;
  // launch i writes `out` when launches - 1 - i is even, else the scratch
  // state, so the last one writes `out` and no launch reads what it writes
  const float* src[3] = {in[0], in[1], nv == 3 ? in[2] : nullptr};
  int remaining = nsteps;
  for (int i = 0; i < launches; ++i) {
    float* const* dst = ((launches - 1 - i) % 2 == 0) ? out : mid;
    const int steps = std::min(K, remaining);
    CFields ci{{src[0], src[1], src[2]}};
    Fields co{{dst[0], dst[1], nv == 3 ? dst[2] : nullptr}};
    KParams kl = k;
    kl.wt = tiled_write_through(
        tiled_tile_count(p->nx, p->ny, c.rows, c.waves, c.cols, steps), cus, per_cu);
    YH_HIP(launch_tiled(p->model, masked, c.rows, c.waves, c.cols, steps, kl, ci, co,
                        masked ? codes : nullptr, s));
    for (int a = 0; a < 3; ++a) src[a] = co.f[a];
    remaining -= steps;
  }
  return YH_OK;
}

}  // extern "C"
