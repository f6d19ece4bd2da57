// contour.hip -- contour points of countour_kernel (spaceAPD.cu:18-153) on
// the device: mode 1 zero-crossings of a field with linear interpolation
// (the space-APD nodal lines of sAPD), mode 2 a single and mode 3 a double
// threshold contour of v where u < conTh1.  Points are appended with an
// atomic counter and a key 2*cell + sub, so a sort by key restores the cell
// order whatever the scheduling.
//
// At the last column (row) the reference reads field2 at (i+1, j) (i, j+1)
// anyway, a wrapped or out-of-bounds element; here the cell's own value
// stands in, so that direction has no crossing, as its zpm test assumes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "plan.h"

namespace yh {

__device__ __forceinline__ void contour_push(float x, float y, float t, int key, float* pts,
                                             int* keys, int* count, int cap) {
  const int slot = atomicAdd(count, 1);
  if (slot < cap) {
    pts[3 * slot] = x;
    pts[3 * slot + 1] = y;
    pts[3 * slot + 2] = t;
    keys[slot] = key;
  }
}

__global__ __launch_bounds__(256) void k_contour(int nx, int ny, int mode, float th1, float th2,
                                                 float th3, const float* __restrict__ f1,
                                                 const float* __restrict__ f2,
                                                 const uint8_t* __restrict__ area, float t,
                                                 float* pts, int* keys, int* count, int cap) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y * 4 + threadIdx.y;
  if (i >= nx || j >= ny) return;
  const int c = i + nx * j;
  const bool in_area = area ? area[c] != 0 : true;  // stimArea: avoid the pacing site
  const float here = f2[c];
  const float east = i < nx - 1 ? f2[c + 1] : here;
  const float north = j < ny - 1 ? f2[c + nx] : here;
  // a negative product of a value with its east (north) neighbour's is a
  // sign change across that edge; past the last column (row) the cell's
  // own square stands in, which is never negative
  auto crosses = [&](float a, float ae, float an, bool& xe, bool& xn) {
    xe = (i < nx - 1 ? a * ae : a * a) < 0.0f;
    xn = (j < ny - 1 ? a * an : a * a) < 0.0f;
  };
  bool xe, xn;
  if (mode == 1) {  // zero lines of field2 (:55-69), interpolated along the edge
    crosses(here, east, north, xe, xn);
    if (xe || (xn && in_area)) {  // area masks the north test only, as the reference's precedence
      const float tiny = 2.220446049250313e-16f;  // MACHINE_EPS, globalVariables.cuh:26
      const float x = fabsf(here - east) > tiny ? (float)i + here / (here - east) : (float)i;
      const float y = fabsf(here - north) > tiny ? (float)j + here / (here - north) : (float)j;
      contour_push(x, y, t, 2 * c, pts, keys, count, cap);
    }
  } else if (mode == 2) {  // field2 = th2 where field1 < th1 (:73-93), at the cell
    if ((f1[c] < th1) && in_area) {
      crosses(here - th2, east - th2, north - th2, xe, xn);
      if (xe || xn) contour_push((float)i, (float)j, t, 2 * c, pts, keys, count, cap);
    }
  } else {  // two lines around th2 inside a band of th3 (:97-138)
    const float dh = here - th2, de = east - th2, dn = north - th2;
    const float band = th3 + 0.1f;
    if ((f1[c] < th1) && (fabsf(dh) < band) && (fabsf(de) < band) && (fabsf(dn) < band) &&
        in_area) {
      // the lower line sits at 2*th2, as the reference subtracts th2 twice
      crosses(dh - th2, de - th2, dn - th2, xe, xn);
      if (xe || xn) contour_push((float)i, (float)j, t, 2 * c, pts, keys, count, cap);
      crosses(dh + th3, de + th3, dn + th3, xe, xn);
      if (xe || xn) contour_push((float)i, (float)j, t, 2 * c + 1, pts, keys, count, cap);
    }
  }
}

// sAPD of sAPD_kernel without stimulus (spaceAPD.cu:364): +1 where APD1 > APD2, else -1.
__global__ void k_sapd(const float* __restrict__ apd1, const float* __restrict__ apd2, float* sapd,
                       size_t n) {
  for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n;
       q += (size_t)gridDim.x * blockDim.x)
    sapd[q] = (apd1[q] - apd2[q] > 0.0f) ? 1.0f : -1.0f;
}

// Zeroes the hit count on the stream with a kernel, not a memset, so the
// entry point can be captured into a graph.
__global__ void k_zero_count(int* count) { *count = 0; }

static hipError_t launch_contour(int nx, int ny, int mode, double th1, double th2, double th3,
                                 const float* f1, const float* f2, const uint8_t* area, float t,
                                 float* pts, int* keys, int* count, int cap, hipStream_t s) {
  dim3 b(64, 4), g((nx + 63) / 64, (ny + 3) / 4);
  hipLaunchKernelGGL(k_contour, g, b, 0, s, nx, ny, mode, (float)th1, (float)th2, (float)th3, f1,
                     f2, area, t, pts, keys, count, cap);
  return hipGetLastError();
}

}  // namespace yh

using namespace yh;

extern "C" {

int yh_contour_dev(int nx, int ny, int mode, double th1, double th2, double th3,
                   const float* field1, const float* field2, const uint8_t* area, float time,
                   float* d_pts, int* d_keys, int* d_count, int cap, void* stream) {
  if (nx < 1 || ny < 1 || mode < 1 || mode > 3 || !field2 || !d_pts || !d_keys || !d_count ||
      cap < 0)
    return YH_EINVAL;
  if (mode != 1 && !field1) return YH_EINVAL;
  if ((int64_t)nx * ny > (int64_t)1 << 30) return YH_EINVAL;  // keys 2*cell + sub in int32
  hipStream_t s = (hipStream_t)stream;
  // countour_wrapper zeroes the count before every search (spaceAPD.cu:261)
  hipLaunchKernelGGL(k_zero_count, dim3(1), dim3(1), 0, s, d_count);
  YH_HIP(hipGetLastError());
  YH_HIP(launch_contour(nx, ny, mode, th1, th2, th3, field1, field2, area, time, d_pts, d_keys,
                        d_count, cap, s));
  return YH_OK;
}

int yh_contour(yh_plan* pl, int mode, double th1, double th2, double th3, const uint8_t* area,
               float* pts, int cap, int* n) {
  if (!pl || !n || mode < 1 || mode > 3 || cap < 0 || (cap > 0 && !pts)) return YH_EINVAL;
  if (mode == 1 && !pl->apd_on) return YH_ESTATE;  // mode 1 traces the APD alternans sign
  const size_t cells = pl->n;
  const int pcap = (int)std::min<size_t>(2 * cells, (size_t)1 << 30);
  // each buffer on its own, so a call after a failed allocation completes it
  if (!pl->cont_pts) YH_HIP(hipMalloc(&pl->cont_pts, sizeof(float) * 3 * pcap));
  if (!pl->cont_keys) YH_HIP(hipMalloc(&pl->cont_keys, sizeof(int) * pcap));
  if (!pl->cont_count) YH_HIP(hipMalloc(&pl->cont_count, sizeof(int)));
  if (!pl->cont_area) YH_HIP(hipMalloc(&pl->cont_area, cells));
  const uint8_t* darea = nullptr;
  if (area) {
    YH_HIP(hipMemcpy(pl->cont_area, area, cells, hipMemcpyHostToDevice));
    darea = pl->cont_area;
  }
  const float* f1 = pl->buf[pl->cur][0];
  const float* f2 = pl->buf[pl->cur][1];
  if (mode == 1) {
    if (!pl->sapd) YH_HIP(hipMalloc(&pl->sapd, sizeof(float) * cells));
    hipLaunchKernelGGL(k_sapd, dim3(1024), dim3(256), 0, 0, pl->apd.apd1, pl->apd.apd2, pl->sapd,
                       cells);
    YH_HIP(hipGetLastError());
    f2 = pl->sapd;
  }
  const float t = (float)(pl->d.dt * (double)pl->count);  // physicalTime
  YH_HIP(hipMemsetAsync(pl->cont_count, 0, sizeof(int), 0));
  YH_HIP(launch_contour(pl->p.nx, pl->p.ny, mode, th1, th2, th3, f1, f2, darea, t, pl->cont_pts,
                        pl->cont_keys, pl->cont_count, pcap, 0));
  YH_HIP(hipDeviceSynchronize());
  int cnt = 0;
  YH_HIP(hipMemcpy(&cnt, pl->cont_count, sizeof(int), hipMemcpyDeviceToHost));
  const int have = std::min(cnt, pcap);
  std::vector<float> p(3 * (size_t)have);
  std::vector<int> k(have);
  if (have > 0) {
    YH_HIP(hipMemcpy(p.data(), pl->cont_pts, sizeof(float) * 3 * have, hipMemcpyDeviceToHost));
    YH_HIP(hipMemcpy(k.data(), pl->cont_keys, sizeof(int) * have, hipMemcpyDeviceToHost));
  }
  std::vector<int> order(have);
  for (int q = 0; q < have; ++q) order[q] = q;
  std::sort(order.begin(), order.end(), [&](int a, int b) { return k[a] < k[b]; });
  const int m = std::min(have, cap);
  for (int q = 0; q < m; ++q)
    for (int c = 0; c < 3; ++c) pts[3 * q + c] = p[3 * (size_t)order[q] + c];
  *n = cnt;
  return cnt > cap ? YH_EOVERFLOW : YH_OK;
}

}  // extern "C"
