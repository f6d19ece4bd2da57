// tile.h -- one temporally blocked tile of the plain stepper (device code),
// shared by the single-tissue kernel (stepper.hip) and the batched work-list
// kernel (batch.hip).
#pragma once

#include <type_traits>

#include "yh_internal.h"

namespace yh {

__device__ __forceinline__ float lane_from(float v, int addr) {
  return __int_as_float(__builtin_amdgcn_ds_bpermute(addr, __float_as_int(v)));
}

// Workgroup -> tile with consecutive tiles on the same XCD: workgroups are
// dispatched round-robin over the 8 XCDs, so XCD x receives a contiguous band
// of tiles and neighbouring halo rows hit the same L2.
__device__ __forceinline__ int xcd_tile(int bid, int ntiles) {
  const int q = ntiles >> 3, rem = ntiles & 7;
  const int x = bid & 7, idx = bid >> 3;
  return x * q + (x < rem ? x : rem) + idx;
}

// Element idx of an array of at most 2^30 elements (the entry points reject
// larger grids) as a 32-bit byte offset from the array base, so loads and
// stores take the scalar-base + 32-bit vector-offset form instead of a
// 64-bit address per lane.
template <typename T>
__device__ __forceinline__ T* at32(T* base, uint32_t idx) {
  using B = typename std::conditional<std::is_const<T>::value, const char, char>::type;
  return reinterpret_cast<T*>(reinterpret_cast<B*>(base) + (size_t)(idx * (uint32_t)sizeof(T)));
}

// Region of one workgroup: (64*CW) columns x (NW*R) rows; lane L owns columns
// CW*L .. CW*L+CW-1, wave w owns rows w*R .. w*R+R-1.  After K substeps the
// inner (64*CW-2K) x (NW*R-2K) cells are exact; the outer ring of width K is
// recomputed by the neighbouring workgroup.  In substep s only rows that stay
// exact after it, s+1 <= row < NW*R-s-1, are computed (the test is uniform
// over a wave).  With a mask every cell carries its five stencil weights in
// registers, so masked and regular cells cost the same.
template <int MODEL, int NW, int CW>
struct TileLds {
  // [buffer][first/last row][wave][var][column]
  float xch[2][2][NW][Model<MODEL>::ND][64 * CW];
};

// With APD, every substep also updates the action-potential-duration state of
// the cells this tile outputs (their value is exact at every substep), so
// the measurement costs no extra launch or state traffic.  With PREV, u of
// the output cells one substep before the end is written to prev_u (0 in
// holes): the previous state the tip test needs, without a single-step tail.
// K >= 1 (the launchers reject K < 1).
// ISO: -1 takes k.iso at run time; 1 (0) fixes an isotropic (anisotropic)
// diffusion at compile time, which drops that branch from the row loop.
// WT: the new state (and prev_u) is written through to memory
// (store_through) rather than left dirty in the XCD's L2, for launches that
// hold two workgroups on a CU: the write-back then overlaps the other
// workgroup's loads instead of delaying the end of the launch.
template <int MODEL, bool MASKED, int R, int NW, int CW, bool APD = false, bool PREV = false,
          int ISO = -1, bool WT = false>
__device__ __forceinline__ void tile_advance(const KParams& k, CFields in, Fields out,
                                             const uint8_t* __restrict__ codes, int K, int tx,
                                             int ty, TileLds<MODEL, NW, CW>& lds,
                                             const ApdState& apd = ApdState(),
                                             float* __restrict__ prev_u = nullptr) {
  using M = Model<MODEL>;
  constexpr int NV = M::NV, ND = M::ND;
  constexpr int WR = 64 * CW;
  constexpr int H = NW * R;
  auto& xch = lds.xch;

  const int lane = threadIdx.x & 63;
  // wave-uniform: row tests below become scalar branches
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int wout = WR - 2 * K;
  const int hout = H - 2 * K;
  const int nx = k.nx, ny = k.ny;
  const int gx0 = tx * wout - K + lane * CW;
  const int gy0 = ty * hout - K + wave * R;

  int sx[CW];
  bool flip_x[CW];
#pragma unroll
  for (int c = 0; c < CW; ++c) {
    sx[c] = reflect_clamp(gx0 + c, nx);
    flip_x[c] = (gx0 + c < 0) || (gx0 + c >= nx);
  }

  float X[NV][R][CW];
  // Masked stencil: the coefficients c in {0,1,2} of reactionDiffusion.cu:162-169
  // per cell and direction, and their sum, as floats.  Sums of c*nb over
  // integers c are exact for a constant field, so the update below leaves
  // constant fields unchanged, as the reference's c*(nb-u) form does.
  // Holes get c = 0.
  constexpr int RM = MASKED ? R : 1;
  float cw[RM][CW], ce[RM][CW], cn[RM][CW], cs[RM][CW], csum[RM][CW];
  uint32_t hole[CW];    // bit r: cell (r, c) is a hole
#pragma unroll
  for (int c = 0; c < CW; ++c) hole[c] = 0;
  // Rows are loaded first, last, then 1 .. R-2: the row exchange of the
  // first substep needs only the first and the last row, and row r of that
  // substep only rows up to r+1, so (with the first substep peeled out of
  // the loop below) each wave starts computing while its later rows are
  // still arriving.
#pragma unroll
  for (int ri = 0; ri < R; ++ri) {
    const int r = ri == 0 ? 0 : ri == 1 ? R - 1 : ri - 1;
    const int gy = gy0 + r;
    const int sy = reflect_clamp(gy, ny);
    const bool flip_y = (gy < 0) || (gy >= ny);
#pragma unroll
    for (int c = 0; c < CW; ++c) {
      const uint32_t idx = (uint32_t)sy * (uint32_t)nx + (uint32_t)sx[c];
#pragma unroll
      for (int a = 0; a < NV; ++a) X[a][r][c] = *at32(in.f[a], idx);
      if (MASKED) {
        const uint32_t b = *at32(codes, idx);
        const bool h = b == kHole;
        const uint32_t q = h ? 0u : b;
        uint32_t w2 = q & 3u, e2 = (q >> 2) & 3u, n2 = (q >> 4) & 3u, s2 = (q >> 6) & 3u;
        // a reflected column keeps the weights of the cell it mirrors when
        // its lane swaps its neighbour sources below (one column per lane)
        if (CW != 1 && flip_x[c]) { const uint32_t t = w2; w2 = e2; e2 = t; }
        if (flip_y) { const uint32_t t = n2; n2 = s2; s2 = t; }
        const int rm = MASKED ? r : 0;
        cw[rm][c] = (float)w2;
        ce[rm][c] = (float)e2;
        cn[rm][c] = (float)n2;
        cs[rm][c] = (float)s2;
        csum[rm][c] = (float)(w2 + e2 + n2 + s2);
        hole[c] |= (h ? 1u : 0u) << r;
        if (h) {
#pragma unroll
          for (int a = 0; a < NV; ++a) X[a][r][c] = 0.0f;  // hole cells stay 0 (:521-522)
        }
      }
    }
  }

  // APD state of this lane's cells (only output cells are written back)
  float a_front[APD ? R : 1][CW], a_back[APD ? R : 1][CW];
  float a_d1[APD ? R : 1][CW], a_d2[APD ? R : 1][CW];
  uint32_t a_first[APD ? R : 1][CW];
  if (APD) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const size_t row = (size_t)reflect_clamp(gy0 + r, ny) * nx;
#pragma unroll
      for (int c = 0; c < CW; ++c) {
        const size_t idx = row + sx[c];
        a_front[APD ? r : 0][c] = apd.front[idx];
        a_back[APD ? r : 0][c] = apd.back[idx];
        a_d1[APD ? r : 0][c] = apd.apd1[idx];
        a_d2[APD ? r : 0][c] = apd.apd2[idx];
        a_first[APD ? r : 0][c] = apd.first[idx];
      }
    }
  }

  // ds_bpermute sources of the west and east neighbours.  A reflected halo
  // column (one column per lane) takes them the other way round and keeps
  // its mirror cell's weights, so it adds the same terms in the same order
  // as that cell and stays its bitwise mirror through the substeps: results
  // then do not depend on where launches split, with or without a mask.
  // (The N/S pair needs nothing: its first product is exact for integer
  // weights, so the fma that adds the second rounds the exact pair sum.)
  int addr_w = ((lane + 63) & 63) << 2;
  int addr_e = ((lane + 1) & 63) << 2;
  if (CW == 1 && flip_x[0]) {
    const int t = addr_w;
    addr_w = addr_e;
    addr_e = t;
  }
  [[maybe_unused]] const typename M::Pinned pinned = M::pin(k);
  const bool iso = ISO < 0 ? k.iso : ISO == 1;
  const float dscale = iso ? k.trx : 1.0f;  // the laplacian terms below carry trx otherwise

  // One substep; called once with s = 0 (peeled, so the loads above are
  // waited for row by row) and then in a loop.
  auto substep = [&](const int s) __attribute__((always_inline)) {
    const int b = s & 1;
#pragma unroll
    for (int a = 0; a < ND; ++a)
#pragma unroll
      for (int c = 0; c < CW; ++c) {
        xch[b][0][wave][a][lane * CW + c] = X[a][0][c];
        xch[b][1][wave][a][lane * CW + c] = X[a][R - 1][c];
      }
    __syncthreads();
    float south[ND][CW], north[ND][CW];
#pragma unroll
    for (int a = 0; a < ND; ++a)
#pragma unroll
      for (int c = 0; c < CW; ++c) {
        south[a][c] = wave > 0 ? xch[b][1][wave - 1][a][lane * CW + c] : X[a][0][c];
        north[a][c] = wave < NW - 1 ? xch[b][0][wave + 1][a][lane * CW + c] : X[a][R - 1][c];
      }
    float prev[ND][CW];
#pragma unroll
    for (int a = 0; a < ND; ++a)
#pragma unroll
      for (int c = 0; c < CW; ++c) prev[a][c] = south[a][c];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int lr = wave * R + r;
      const bool active = (lr >= s + 1) && (lr < H - s - 1);
      float old[ND][CW];
#pragma unroll
      for (int a = 0; a < ND; ++a)
#pragma unroll
        for (int c = 0; c < CW; ++c) old[a][c] = X[a][r][c];
      if (active) {
        // dif[0]: u plus its diffusion term; dif[1] (2V): v's diffusion term
        float dif[ND][CW];
#pragma unroll
        for (int a = 0; a < ND; ++a) {
          const float wl = lane_from(old[a][CW - 1], addr_w);  // west of column 0
          const float er = lane_from(old[a][0], addr_e);       // east of column CW-1
#pragma unroll
          for (int c = 0; c < CW; ++c) {
            const float ctr = old[a][c];
            const float wv = c == 0 ? wl : old[a][c - 1];
            const float ev = c == CW - 1 ? er : old[a][c + 1];
            const float sv = prev[a][c];
            const float nv = (r < R - 1) ? X[a][r + 1][c] : north[a][c];
            float lap_or_t;  // iso: sum c*(nb - ctr); otherwise the laplacian term
            if (MASKED) {
              const int rm = MASKED ? r : 0;
              if (iso) {
                float t = cs[rm][c] * sv;
                t = __builtin_fmaf(cn[rm][c], nv, t);
                t = __builtin_fmaf(ce[rm][c], ev, t);
                t = __builtin_fmaf(cw[rm][c], wv, t);
                lap_or_t = __builtin_fmaf(-csum[rm][c], ctr, t);
              } else {
                float tx = __builtin_fmaf(ce[rm][c], ev, cw[rm][c] * wv);
                tx = __builtin_fmaf(-(cw[rm][c] + ce[rm][c]), ctr, tx);
                float ty = __builtin_fmaf(cs[rm][c], sv, cn[rm][c] * nv);
                ty = __builtin_fmaf(-(cn[rm][c] + cs[rm][c]), ctr, ty);
                lap_or_t = __builtin_fmaf(k.try_, ty, k.trx * tx);
              }
            } else {
              if (iso) {
                lap_or_t = __builtin_fmaf(-4.0f, ctr, (wv + ev) + (nv + sv));
              } else {
                const float lx = __builtin_fmaf(-2.0f, ctr, wv + ev);
                const float ly = __builtin_fmaf(-2.0f, ctr, nv + sv);
                lap_or_t = __builtin_fmaf(k.try_, ly, k.trx * lx);
              }
            }
            dif[a][c] = a == 0 ? __builtin_fmaf(dscale, lap_or_t, ctr) : dscale * lap_or_t;
          }
        }
#pragma unroll
        for (int c = 0; c < CW; ++c) {
          const float u_old = X[0][r][c];
          if constexpr (MODEL == MFK) {
            M::step(k, pinned, dif[0][c], X[0][r][c], X[1][r][c], X[NV - 1][r][c]);
          } else {
            M::step(k, dif[0][c], dif[ND - 1][c], X[0][r][c], X[1][r][c]);
          }
          if (APD && !(MASKED && ((hole[c] >> r) & 1u))) {
            apd_cell(u_old, X[0][r][c], apd.count0 + s + 1, apd.dt, apd.th,
                     a_front[APD ? r : 0][c], a_back[APD ? r : 0][c], a_d1[APD ? r : 0][c],
                     a_d2[APD ? r : 0][c], a_first[APD ? r : 0][c]);
          }
          if ((PREV || !APD) && s == K - 1) {
            // The rows active in the last substep are exactly the output
            // rows: store each as soon as it is final (overlapping the
            // stores with the rest of the substep), and with PREV also u
            // one substep before the end.
            int px = gx0 + c, py = gy0 + r;
            asm volatile("" : "+v"(px), "+s"(py));  // addresses computed here, not hoisted
            const int lc = lane * CW + c;
            if (lc >= K && lc < WR - K && px < nx && py < ny) {
              const bool h = MASKED && ((hole[c] >> r) & 1u);
              const uint32_t idx = (uint32_t)py * (uint32_t)nx + (uint32_t)px;
              if (PREV) {
                if (WT) store_through(prev_u, idx, h ? 0.0f : u_old);
                else *at32(prev_u, idx) = h ? 0.0f : u_old;
              }
              if (!APD) {
#pragma unroll
                for (int a = 0; a < NV; ++a) {
                  if (WT) store_through(out.f[a], idx, h ? 0.0f : X[a][r][c]);
                  else *at32(out.f[a], idx) = h ? 0.0f : X[a][r][c];
                }
              }
            }
          }
        }
      }
#pragma unroll
      for (int a = 0; a < ND; ++a)
#pragma unroll
        for (int c = 0; c < CW; ++c) prev[a][c] = old[a][c];
    }
  };
  substep(0);
  for (int s = 1; s < K; ++s) substep(s);
  if (!APD) return;

#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int lr = wave * R + r;
    const int gy = gy0 + r;
    if (lr >= K && lr < H - K && gy < ny) {
#pragma unroll
      for (int c = 0; c < CW; ++c) {
        const int lc = lane * CW + c;
        const int gx = gx0 + c;
        if (lc >= K && lc < WR - K && gx < nx) {
          const size_t idx = (size_t)gy * nx + gx;
          const bool h = MASKED && ((hole[c] >> r) & 1u);
#pragma unroll
          for (int a = 0; a < NV; ++a) out.f[a][idx] = h ? 0.0f : X[a][r][c];
          if (APD) {
            apd.front[idx] = a_front[APD ? r : 0][c];
            apd.back[idx] = a_back[APD ? r : 0][c];
            apd.apd1[idx] = a_d1[APD ? r : 0][c];
            apd.apd2[idx] = a_d2[APD ? r : 0][c];
            apd.first[idx] = (uint8_t)a_first[APD ? r : 0][c];
          }
        }
      }
    }
  }
}

}  // namespace yh
