// plan.h -- the plan object behind yh_plan (internal).
#pragma once

#include <hip/hip_runtime.h>

#include "yh_internal.h"

struct yh_sr;  // symmetry-reduction session state (sr.hip)

struct yh_plan {
  yh_params p;
  yh_derived d;
  yh::KParams k;
  int nv = 2;
  size_t n = 0;
  int device = 0;
  int cus = 0;  // compute units of that device (the AUTO kernel choice)
  float* buf[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
  int cur = 0;             // buf[cur] holds the current state
  bool have_prev = false;  // u one step earlier is known: in prev_u or buf[cur^1][0]
  bool past_prev_u = false;  // it is in prev_u (written by the last tiled launch)
  float* prev_u = nullptr;   // lazily allocated, n floats
  uint8_t* solid_d = nullptr;
  uint8_t* codes_d = nullptr;
  bool mask_set = false;
  yh_tip* tips_d = nullptr;
  int* keys_d = nullptr;
  int* tipcount_d = nullptr;
  int tipcap = 0;
  int* flag_d = nullptr;
  yh_kernel_cfg cfg{YH_KERNEL_AUTO, 4, 8, 8, 1};
  int64_t count = 0;
  bool keep_prev = true;
  bool apd_on = false;
  yh::ApdState apd;  // device arrays (front, back, apd1, apd2, first) when apd_on
  yh_sr* sr = nullptr;
  int stim_on = 0, stim_x = 0, stim_y = 0;  // mouse stimulus (yh_set_stimulus)
  int sr_launches = 2;      // launches per symmetry-reduction step (YH_OPT_SR_LAUNCHES)
  long long state_epoch = 0;  // bumped whenever the state changes outside yh_sr_step
  float* gscratch[9] = {};  // stage fields of the general path (RK2/RK4), lazily allocated
  // contour search (yh_contour), lazily allocated
  float* cont_pts = nullptr;
  int* cont_keys = nullptr;
  int* cont_count = nullptr;
  uint8_t* cont_area = nullptr;
  float* sapd = nullptr;
};

namespace yh {
// Steps of this plan run on the general per-stage path (another branch, or
// the mouse stimulus on).
inline bool plan_general(const yh_plan* pl) { return !standard_branch(pl->p) || pl->stim_on; }
// The YH_KERNEL_AUTO rule for an nx x ny grid on a device of `cus` compute
// units (a tiled shape and its substeps, capped at min(nx, ny) - 1); pairs:
// a CU holds two workgroups of the 4 x 16 shape at once.
yh_kernel_cfg auto_cfg(int nx, int ny, int cus, bool pairs = false);
// Workgroups of the tiled kernel of this shape one CU holds at once
// (cached; allocates nothing).
int tiled_per_cu(int model, bool masked, const KParams& k, int rows, int waves, int cols,
                 bool prev);
// Tiles of one launch of K substeps on a rows x waves x cols shape.
int64_t tiled_tile_count(int nx, int ny, int rows, int waves, int cols, int K);
// Whether a tiled launch of `tiles` workgroups, `per_cu` per CU at most,
// writes its outputs through to memory (KParams::wt).
bool tiled_write_through(int64_t tiles, int cus, int per_cu);
// Resolve YH_KERNEL_AUTO to a concrete configuration for this plan.
yh_kernel_cfg resolve_cfg(const yh_plan* pl);
// Advance `nsteps` on the plan's buffers, keeping u of the step before the
// last (for the tip test) when keep_prev: the last tiled launch writes it to
// prev_u, or a single-step tail leaves it in buf[cur^1].
int plan_advance(yh_plan* pl, int nsteps, hipStream_t s);
void sr_free(yh_sr* sr);
}  // namespace yh
