/*
 * yolohtli.h -- C ABI of the MI355X-native tissue time-stepper.
 *
 * Pure C: no HIP or torch types appear in any signature.  Streams are passed
 * as `void*` (a hipStream_t, or NULL for the null stream).  Every entry point
 * returns an int error code (YH_OK == 0) and never calls exit().
 *
 * Ownership: the caller owns host arrays; a plan owns the device buffers it
 * allocates.  Functions named *_dev operate on caller-owned device pointers,
 * exactly as the reference's host wrappers do.  One plan is used by one host
 * thread at a time; every call is ordered on the stream it is given.
 *
 * Memory layout of every field: structure of arrays, float32, row-major with
 * index i + nx*j (reference globalVariables.cuh:4 I2D, reactionDiffusion.cu:47).
 * Masks are uint8, 1 = tissue, 0 = hole (main.cu:679 `mesh > 0.5`).
 */
#ifndef YOLOHTLI_H
#define YOLOHTLI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- errors */
enum {
  YH_OK = 0,
  YH_EINVAL = 1,       /* bad argument (NULL pointer, size, range) */
  YH_EUNSUPPORTED = 2, /* the entry point runs only the standard branch (Neumann,
                          Euler, 5-point) and the parameters select another */
  YH_EHIP = 3,         /* the HIP runtime reported an error */
  YH_ENOMEM = 4,       /* device or host allocation failed */
  YH_EUNSTABLE = 5,    /* explicit Euler diffusion number rx+ry > 0.5 */
  YH_ESINGULAR = 6,    /* symmetry-reduction 3x3 solve hit a zero pivot */
  YH_ENOTIP = 7,       /* no spiral tip found (informational) */
  YH_EOVERFLOW = 8,    /* tip log saturated; later tips were dropped */
  YH_ENONFINITE = 9,   /* NaN/Inf found in the state */
  YH_ESTATE = 10,      /* call made in the wrong order */
  YH_ESCRATCH = 11     /* grid larger than yh_step_noalloc's fixed scratch */
};

/* Human-readable text for an error code. */
const char* yh_strerror(int code);

/* ------------------------------------------------------------ parameters */
enum { YH_MODEL_2V_REF = 0, YH_MODEL_3V_FK = 1 };

/* Fenton-Karma three-variable model (Fenton & Karma, Chaos 8:20, 1998).
 * Not present in the reference; defaults are the paper's Beeler-Reuter fit.
 * v opens with tau_v1m while u < u_v and with tau_v2m while u_v <= u < u_c
 * (the naming of Fenton et al., Chaos 12:852, 2002; the 1998 paper calls the
 * same two constants tau_v2- = 19.6 and tau_v1- = 1250). */
typedef struct yh_fk_params {
  double tau_d, tau_r, tau_si, tau_0; /* ms: fast-in, slow-out, slow-in, outward */
  double tau_vp, tau_v1m, tau_v2m;    /* ms: v gate closing / opening (u<u_v, u>=u_v) */
  double tau_wp, tau_wm;              /* ms: w gate closing / opening */
  double u_c, u_v, u_csi, k;          /* thresholds and tanh steepness */
} yh_fk_params;

/* Solver subset of the reference paramVar (typeDefinition.cuh:35-125), same
 * field names and meaning.  Derived quantities follow saveFiles.cu:144-170 and
 * main.cu:148-155: hx = Lx/(nx-1); dt is halved when reduceSym is set.
 * Branches (reactionDiffusion.cu:36-561): neumannBC 1 (mirror) or 0
 * (Dirichlet, boundaryVal); timeIntOrder 1 (Euler), 2 (RK2) or 4 (RK4); lap4
 * (4th-order compact Laplacian) and anisotropy (cross term from degrad,
 * diff_par != diff_per) act on the square domain, and with Dirichlet the
 * anisotropic term also with a mask; with a mask and Neumann both have no
 * effect, as in the reference.  The standard branch (Neumann, Euler, 5-point)
 * runs on the temporally blocked kernels; the others on a per-stage path. */
typedef struct yh_params {
  int nx, ny;
  double Lx, Ly;                 /* cm */
  double dt;                     /* ms, before the symmetry-reduction halving */
  double diff_par, diff_per;     /* cm^2/ms */
  double degrad;                 /* fibre angle in degrees */
  double rscale;                 /* gate diffusion scale (2V model) */
  double tc, alpha, beta, gamma, delta, eps, mu, theta; /* 2V reaction */
  int model;                     /* YH_MODEL_2V_REF or YH_MODEL_3V_FK */
  int neumannBC, solidSwitch, gateDiff, reduceSym;
  int timeIntOrder, lap4, anisotropy; /* see above */
  double boundaryVal, Uth;
  int tipOffsetX, tipOffsetY;
  yh_fk_params fk;
  double heaviside_width;        /* 0: sharp STEP(a,b)=(a>=b); >0: 0.5(1+tanh(x/w)) */
} yh_params;

/* Quantities derived once on the host (saveFiles.cu:144-170, main.cu:148-155). */
typedef struct yh_derived {
  double hx, hy, dt, Dxx, Dyy, rx, ry, invdx, invdy;
} yh_derived;

/* Defaults of the reference parameterSetup (saveFiles.cu:105-231), with the
 * standard branch selected: Euler (timeIntOrder=1), 5-point (lap4=0). */
void yh_params_default(yh_params* p);
/* Fenton-Karma 1998 Beeler-Reuter parameter set. */
void yh_fk_params_default(yh_fk_params* fk);
/* Validate and derive (YH_EINVAL for a timeIntOrder other than 1, 2, 4, for
 * nx or ny below 4, or for more than 2^30 cells: cell indices are int32 and
 * the tip and contour keys 2*cell + root must fit in one). */
int yh_params_derive(const yh_params* p, yh_derived* d);

/* ------------------------------------------------------------- workloads */
/* splitmix64 sub-stream seed for (base, config, tissue, purpose). */
uint64_t yh_seed(uint64_t base, uint32_t config, uint32_t tissue, uint32_t purpose);
/* Cross-field initial condition (main.cu:577-621): u=1 for i<u_cols, gate
 * variables = refr for j>=ny/2 and = rest otherwise.  w may be NULL. */
int yh_init_cross_field(int nx, int ny, int u_cols, float rest, float refr,
                        float* u, float* v, float* w);
/* u ~ U[0,1) i.i.d. from splitmix64(seed); v (and w) = gate. */
int yh_init_random(int nx, int ny, uint64_t seed, float gate,
                   float* u, float* v, float* w);
/* Seeded circular holes (replaces the missing common/holes*.dat files read at
 * main.cu:634-684): centres U[0,nx)xU[0,ny), radii U[rmin,rmax] cells; cell is
 * a hole when (i-cx)^2+(j-cy)^2 <= r^2.  mask: 1 = tissue, 0 = hole. */
int yh_make_holes(int nx, int ny, int nholes, double rmin, double rmax,
                  uint64_t seed, uint8_t* mask);

/* ---------------------------------------------- single-step device wrapper */
/* One time step on caller-owned device arrays.  Replaces
 * reactionDiffusion_wrapper (reactionDiffusion.cu:568-577): in[] are read,
 * out[] written (never aliasing in[]), veltan[] (may be NULL) receives
 * rhs/dt (reactionDiffusion.cu:529-530,551-552).  Arrays: u,v for the 2V
 * model; u,v,w for the FK model.  solid may be NULL unless solidSwitch. */
int yh_rd_step_dev(const yh_params* p, float* const out[3], const float* const in[3],
                   float* const veltan[3], const uint8_t* solid, void* stream);
/* The same with the wrapper's mouse-stimulus arguments (stimLockMouse, point:
 * reactionDiffusion.cu:54-61,134): while stim_on, cells within radius^2 < 400
 * of (stim_x, stim_y) get the stimulus current 24.7.  Every branch runs with
 * timeIntOrder 1; RK2/RK4 need stage fields and run through a plan
 * (YH_EUNSUPPORTED here). */
int yh_rd_step_dev_ex(const yh_params* p, float* const out[3], const float* const in[3],
                      float* const veltan[3], const uint8_t* solid, int stim_on, int stim_x,
                      int stim_y, void* stream);

/* Spiral-tip record: the reference vec5dyn (typeDefinition.cuh:19-20). */
typedef struct yh_tip { float x, y, vx, vy, t; } yh_tip;

/* Spiral-tip detection between two u fields on the device.  Replaces
 * tip_wrapper with tipAlgorithm 1 (tipTracker.cu:20-241,569-611): g_past and
 * g_present keep the reference's roles.  Found tips are appended to
 * d_tips/d_keys at positions [*d_count, ...) (d_count is NOT reset), keys are
 * 2*(i+nx*j)+root for deterministic ordering; entries past cap are dropped
 * while *d_count still counts them. */
int yh_tip_detect_dev(const yh_params* p, const float* g_past, const float* g_present,
                      double time, yh_tip* d_tips, int* d_keys, int* d_count, int cap,
                      void* stream);

/* Contour points of countour_kernel (spaceAPD.cu:18-153) between two
 * device fields, replacing countour_wrapper (:257-275), which zeroes the
 * count first (so does this call).  mode 1: zero-crossings of field2 toward
 * (i+1, j) / (i, j+1), interpolated linearly (the space-APD nodal lines, field2
 * = sAPD); mode 2: cells where field1 < th1 and field2 crosses th2 (single
 * contour); mode 3: the double contour of :97-138 (up to two points per cell).
 * th1..th3 are the reference's conTh1..conTh3.  area (may be NULL =
 * everywhere) is the reference's stimArea.  Each point is (x, y, time) in
 * d_pts, its key 2*cell + sub in d_keys; entries past cap are dropped while
 * *d_count still counts them.  At the last column / row the reference reads a
 * wrapped or out-of-bounds neighbour; the cell's own value stands in here. */
int yh_contour_dev(int nx, int ny, int mode, double th1, double th2, double th3,
                   const float* field1, const float* field2, const uint8_t* area, float time,
                   float* d_pts, int* d_keys, int* d_count, int cap, void* stream);

/* ------------------------------------------------------------------ plans */
typedef struct yh_plan yh_plan;

/* Kernel choice for the plain stepper.  AUTO picks a tiled shape and K from
 * the grid size and the device's CU count without timing anything;
 * yh_autotune times the candidates and sets the fastest. */
enum { YH_KERNEL_AUTO = 0, YH_KERNEL_DIRECT = 1, YH_KERNEL_TILED = 2 };
typedef struct yh_kernel_cfg {
  int variant;   /* YH_KERNEL_* */
  int substeps;  /* k fused time steps per launch (TILED), 1..16, 2k < rows*waves */
  int rows;      /* rows held in registers per thread (TILED): 4, 7, 8 or 16 */
  int waves;     /* wavefronts per workgroup (TILED): 4, 8 or 16 */
  int cols;      /* columns per lane (TILED): 1 or 2.  Built shapes (rows x
                  * waves x cols): 4x4, 8x4, 16x4, 4x8, 8x8, 16x8, 4x16, 8x16,
                  * 7x16 with cols 1; 4x4, 8x4, 4x8, 8x8, 4x16 with cols 2 */
} yh_kernel_cfg;

/* Creates a plan: checks parameters, allocates ping-pong state on the current
 * device.  Mirrors main.cu:213-402 (allocation + constant upload). */
int yh_plan_create(const yh_params* p, yh_plan** out);
void yh_plan_destroy(yh_plan* plan);
int yh_plan_params(const yh_plan* plan, yh_params* p, yh_derived* d);
/* Mask (1 = tissue).  Required when solidSwitch; stencil codes are built here. */
int yh_set_mask(yh_plan* plan, const uint8_t* solid);
/* Host state in/out; w is ignored (may be NULL) for the 2V model. */
int yh_set_state(yh_plan* plan, const float* u, const float* v, const float* w);
int yh_get_state(yh_plan* plan, float* u, float* v, float* w);
/* Device pointers of the current state (valid until the next step call). */
int yh_state_dev(yh_plan* plan, float** u, float** v, float** w);
int yh_set_kernel(yh_plan* plan, const yh_kernel_cfg* cfg);
int yh_get_kernel(const yh_plan* plan, yh_kernel_cfg* cfg);
/* Advance nsteps explicit-Euler steps (the display() plain loop,
 * main.cu:876-885).  After return the plan also holds the state one step
 * earlier, so yh_track_tips sees (u^{n-1}, u^n) as main.cu:961-964 does.
 * With a tiled kernel of one column per lane, a remainder of nsteps % K
 * steps may run inside the call's last launch on the shape one row taller
 * per wave; the result does not depend on how steps are split into
 * launches. */
int yh_step(yh_plan* plan, int nsteps, void* stream);
/* yh_step without any device allocation, on caller-owned device fields (no
 * counterpart in the reference: for callers that cannot allocate device
 * memory, such as code under HIP graph capture, tight serving loops and
 * harnesses that forbid hipMalloc).  Advances nsteps explicit-Euler steps
 * (standard branch only, else YH_EUNSUPPORTED) from in[] into out[] (never
 * aliasing; in[] is not written) with the kernel a plan on YH_KERNEL_AUTO
 * runs, so out[] equals yh_step's state bit for bit.  solid: device mask,
 * required when solidSwitch.  The stencil codes and the intermediate state
 * live in a fixed device array of the library, one per device for the
 * process's lifetime: grids above 2^21 cells return YH_ESCRATCH, and calls
 * that overlap on one device race (order them, e.g. on one stream).  No tip
 * or APD output; nsteps = 0 copies in[] to out[]. */
int yh_step_noalloc(const yh_params* p, float* const out[3], const float* const in[3],
                    const uint8_t* solid, int nsteps, void* stream);
/* Steps taken since the last yh_set_state (param.count, main.cu:884). */
int64_t yh_step_count(const yh_plan* plan);
/* Detect tips between the previous and the current state and append them to
 * the plan's device tip log, stamped t = dt*count (main.cu:961-964). */
int yh_track_tips(yh_plan* plan, void* stream);
/* Synchronise and copy the tip log, ordered by (t, cell, root).  *n receives
 * the number of tips logged (may exceed cap: YH_EOVERFLOW). */
int yh_read_tips(yh_plan* plan, yh_tip* out, int cap, int* n);
int yh_clear_tips(yh_plan* plan);
/* Device flag check: YH_ENONFINITE when any state value is NaN/Inf. */
int yh_check_finite(yh_plan* plan, void* stream);

/* Mouse stimulus of the plan's steps (stimLockMouse and point,
 * reactionDiffusion.cu:54-61,134): on != 0 applies the current 24.7 within
 * radius^2 < 400 cells of (x, y) in every following step until switched off.
 * While on, the plan steps on the per-stage path. */
int yh_set_stimulus(yh_plan* plan, int on, int x, int y);

/* Plan options.  YH_OPT_SR_LAUNCHES (default 2): launches per symmetry-
 * reduction step, 2 (the step's BFECC launch also takes the RD step and the
 * tip test of the next step) or 3 (RD + tips in a launch of their own); the
 * results are the same bit for bit.  YH_OPT_KEEP_PREV (default 1): every yh_step keeps u of the
 * step before its last for yh_track_tips (the last k-substep launch writes
 * it, or, for tile shapes built without that output and with APD on, a
 * single-step tail leaves it in the other state buffer); 0 drops it (no tip
 * tracking until the next yh_step with 1). */
enum { YH_OPT_KEEP_PREV = 1, YH_OPT_SR_LAUNCHES = 2 };
int yh_set_option(yh_plan* plan, int option, int64_t value);

/* Action-potential durations measured while stepping (sAPD_kernel,
 * spaceAPD.cu:278-374, stimulus off; threshold apdTh = 0.15 at :288).  Every
 * step updates, per cell, the up-crossing time `front`, the down-crossing
 * time `back` (both linearly interpolated, :298-306) and, once both are set,
 * the alternating durations apd1/apd2 = back - front.  The update is fused
 * into the stepping kernels; yh_apd_enable zeroes the state.  The crossing
 * arguments are passed in time order (old state, new state): display() passes
 * them swapped after its pointer swap (main.cu:882,1035-1038), which makes
 * its durations negative. */
int yh_apd_enable(yh_plan* plan, int on, double threshold);
/* Host copies of the per-cell fields; any pointer may be NULL. */
int yh_apd_get(yh_plan* plan, float* apd1, float* apd2, float* front, float* back);

/* Contours of the plan's current state (main.cu:968-977): modes 2 and 3
 * between u and v, mode 1 on sAPD = sign(APD1 - APD2) of the plan's APD
 * state (spaceAPD.cu:364; YH_ESTATE unless APD is enabled).  area: host mask
 * (may be NULL).  pts receives up to cap points (x, y, t) sorted by (cell,
 * sub); *n the number found (YH_EOVERFLOW when above cap).  Synchronous. */
int yh_contour(yh_plan* plan, int mode, double th1, double th2, double th3, const uint8_t* area,
               float* pts, int cap, int* n);

/* Autotuner (no reference counterpart; the reference launches fixed 16x16
 * blocks, globalVariables.cuh:15-16): times the direct kernel and the tiled
 * kernel over tile shapes x substep depths on the plan's own state (restored
 * afterwards), installs the fastest, and caches it per (nx, ny, model,
 * masked, call size) for the process.  steps: the number of steps per
 * yh_step call the plan will be driven with (capped at 2000); each
 * candidate is timed over whole calls of that size, at least 400 steps in
 * all, so a call's remainder launch counts.  best / best_ms (ms per step)
 * may be NULL. */
int yh_autotune(yh_plan* plan, int steps, void* stream, yh_kernel_cfg* best, float* best_ms);

/* ----------------------------------------------------- symmetry reduction */
/* State of a co-moving-frame run (main.cu:156-161 c, phi; integrals of
 * integralTrapz.cu:97-177; disk centre of symmetryReduction.cu:95-107). */
typedef struct yh_sr_info {
  double c[3];      /* translation (cx, cy) and rotation c_theta of the last solve */
  double phi[3];    /* accumulated frame shift and angle */
  double I[12];     /* the 12 disk inner products of the last step */
  int cx, cy;       /* disk centre offset from (nx/2, ny/2) */
  int64_t count;    /* SR steps taken */
  int status;       /* bit 0: a solve was singular */
  int ntips;        /* tips found in the last step */
} yh_sr_info;

/* Starts a symmetry-reduction run on the plan's current state (requires a
 * plan created with reduceSym = 1, so dt is the halved SR step, main.cu:148).
 * tip0 is the disk centre used at count 0 (tipx0_d, main.cu:373-376). */
int yh_sr_begin(yh_plan* plan, float tip0x, float tip0y, void* stream);
/* nsteps SR steps (display(), main.cu:894-953): RD, tips, disk inner
 * products, 3x3 solve, BFECC advection, phi += dt*c; all on the device.
 * c_phi_out (may be NULL) receives 6 doubles per step: c and phi before that
 * step's solve (clist/philist, main.cu:902-903).  Tips of every step go to the
 * plan's tip log. */
int yh_sr_step(yh_plan* plan, int nsteps, void* stream, double* c_phi_out);
/* Replay steady SR steps from one captured hipGraph (default on). */
int yh_sr_set_graph(yh_plan* plan, int on);
int yh_sr_info_get(yh_plan* plan, yh_sr_info* out);
/* solve_matrix (symmetryReduction.cu:329-420) on the host: c from phi and the
 * 12 integrals by the reference's unpivoted elimination; YH_ESINGULAR when the
 * result is not finite. */
int yh_solve_matrix(const double phi[3], const double I[12], double c[3]);

/* --------------------------------------------------------------- batches */
/* Many tissues advanced together (no reference counterpart: the reference
 * steps one tissue, main.cu:876-885).  All tissues share model and mask mode;
 * nx, ny, D and the ionic constants may differ.  One launch advances every
 * tissue by k steps from a (tissue, tile) work list; yh_batch_step replays a
 * captured hipGraph of the launches when graphs are on (default).  The
 * kernel starts as k = 6 on 8 rows x 8 waves, one workgroup per item, with k
 * capped at the smallest tissue's min(nx, ny) - 1; yh_batch_autotune times
 * the alternatives. */
typedef struct yh_batch yh_batch;
int yh_batch_create(const yh_params* per_tissue, int ntissues, yh_batch** out);
void yh_batch_destroy(yh_batch* batch);
int yh_batch_size(const yh_batch* batch, int* ntissues, int64_t* total_cells);
int yh_batch_set_mask(yh_batch* batch, int tissue, const uint8_t* solid);
int yh_batch_set_state(yh_batch* batch, int tissue, const float* u, const float* v,
                       const float* w);
int yh_batch_get_state(yh_batch* batch, int tissue, float* u, float* v, float* w);
/* cfg: YH_KERNEL_TILED shape and substeps; persistent = 1 runs a grid of the
 * resident workgroup count that loops over the work list, 0 one workgroup per
 * item. */
int yh_batch_set_kernel(yh_batch* batch, const yh_kernel_cfg* cfg, int persistent);
int yh_batch_get_kernel(const yh_batch* batch, yh_kernel_cfg* cfg, int* persistent);
int yh_batch_set_graph(yh_batch* batch, int on);
int yh_batch_step(yh_batch* batch, int nsteps, void* stream);
/* The same steps as per-tissue launches of the single-tissue kernel, one per
 * tissue per k-step chunk (the comparison the batched launch is measured
 * against). */
int yh_batch_step_each(yh_batch* batch, int nsteps, void* stream);
/* Times tile shapes x substep depths x {persistent, one-per-item} on the
 * batch's own state (restored afterwards) and installs the fastest. */
int yh_batch_autotune(yh_batch* batch, int steps, void* stream, yh_kernel_cfg* best,
                      float* best_ms);

/* --------------------------------------------------------------- adjoint */
/* Gradient of loss = sum_{n % rec == 0} sum_s (u^n(sensor_s) - target)^2
 * through nsteps of the FK stepper with a spatial diffusion field D and the
 * tanh-smoothed Heaviside (params.heaviside_width > 0, model FK).  No
 * reference counterpart.  Checkpoints every ckpt_every steps; the reverse
 * sweep recomputes each segment.  Sensors are distinct cell indices. */
typedef struct yh_adjoint yh_adjoint;
int yh_adjoint_create(const yh_params* p, int nsteps, int rec_every, int ckpt_every,
                      const int* sensors, int nsensors, yh_adjoint** out);
void yh_adjoint_destroy(yh_adjoint* adj);
/* Steps per launch of the forward and the reverse sweep, 1..8 (default 8):
 * 1 runs one step per launch, more run that many steps of a tile through LDS
 * (the same per-cell arithmetic: forward and lambda bit for bit, the
 * gradients up to the grouping of their FP64 sums). */
int yh_adjoint_set_substeps(yh_adjoint* adj, int k);
int yh_adjoint_set_mask(yh_adjoint* adj, const uint8_t* solid);
int yh_adjoint_set_initial(yh_adjoint* adj, const float* u, const float* v, const float* w);
/* targets: (nsteps/rec_every) x nsensors, row = recorded step. */
int yh_adjoint_set_targets(yh_adjoint* adj, const float* targets);
/* d_D: device float32 field (nx*ny); it must stay valid until the matching
 * yh_adjoint_backward.  tau6 = (tau_d, tau_r, tau_si, tau_0, tau_vp, tau_wp).
 * loss is 0 until targets are set; traces (host, may be NULL) receives u at
 * the sensors, one row per recorded step. */
int yh_adjoint_forward(yh_adjoint* adj, const float* d_D, const double tau6[6], void* stream,
                       double* loss, float* traces);
/* d_gD: device float64 (nx*ny) receives dL/dD; gtau receives dL/dtau6. */
int yh_adjoint_backward(yh_adjoint* adj, void* stream, double* d_gD, double gtau[6]);

/* ---------------------------------------------------------- file formats */
/* The reference's text files, written with its printf formats so its MATLAB
 * tooling and its own loaders read them unchanged.  Host memory only. */

/* dataparamcsv.csv (printParameters, printFunctions.cu:266-403): 91 lines
 * "label,value" in the reference's positional order.  Solver fields come from
 * p, derived terms as saveFiles.cu:140-170 and main.cu:148-155 compute them;
 * the reference's display and recording fields carry their parameterSetup
 * defaults (saveFiles.cu:111-225).  ic_path / results_path may be NULL ("NA");
 * tipx, tipy: the last tip point. */
int yh_params_write_csv(const char* path, const yh_params* p, const char* ic_path,
                        const char* results_path, float tipx, float tipy);
/* loadParamValues (saveFiles.cu:540-711): the second comma-separated field of
 * each line, parsed with strtof, assigned positionally to the solver fields of
 * p.  Fields the file does not carry (model, Fenton-Karma set, Heaviside
 * width) keep their values.  YH_EINVAL for a file of fewer than 91 lines. */
int yh_params_read_csv(const char* path, yh_params* p);
/* raw_data.dat (print2D2column, printFunctions.cu:58-78): "%f %f\n" (u v) per
 * cell, i fastest; with w non-NULL (FK) a third column.  The reader parses it
 * back as loadData does (saveFiles.cu:512-536). */
int yh_state_write_dat(const char* path, int nx, int ny, const float* u, const float* v,
                       const float* w);
int yh_state_read_dat(const char* path, int nx, int ny, float* u, float* v, float* w);
/* dataTip.dat + dataTipSize.dat (printTip, printFunctions.cu:149-195): each
 * tip as "%f %f %f %f %f\n" (x y vx vy t) and, when n > 0, "%d\n" with n in
 * the size file; append != 0 appends, as every call after the first does. */
int yh_tips_write_dat(const char* tip_path, const char* size_path, const yh_tip* tips, int n,
                      int append);
/* c_phi_list_sym.dat (printSym, printFunctions.cu:249-264): one line
 * "%f %f %f %f %f %f\n" per step, c then phi: the layout yh_sr_step's
 * c_phi_out fills. */
int yh_sym_write_dat(const char* path, const double* c_phi, int nsteps);
/* Obstacle mask file read at main.cu:663-681: nx*ny numbers separated by white
 * space; cell i + nx*j is tissue when its number is > 0.5.  The writer uses
 * MATLAB's save -ascii layout ("%16.7e" per line) of common/Hole_generator.m. */
int yh_mask_read_dat(const char* path, int nx, int ny, uint8_t* mask);
int yh_mask_write_dat(const char* path, int nx, int ny, const uint8_t* mask);

#ifdef __cplusplus
}
#endif
#endif /* YOLOHTLI_H */
