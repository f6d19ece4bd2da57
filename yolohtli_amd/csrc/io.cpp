// io.cpp -- the reference's text files: parameter CSV, state, tip log,
// symmetry history and obstacle mask.  Host code only; every writer uses the
// reference's printf formats so its MATLAB tooling and its own loaders read
// the files unchanged.
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/yolohtli.h"

namespace {

struct File {
  FILE* f = nullptr;
  File(const char* path, const char* mode) : f(path ? std::fopen(path, mode) : nullptr) {}
  ~File() {
    if (f) std::fclose(f);
  }
  // fclose result decides whether buffered writes reached the file
  int close() {
    const int rc = f ? std::fclose(f) : EOF;
    f = nullptr;
    return rc == 0 ? YH_OK : YH_EINVAL;
  }
};

// Quantities printParameters writes that the solver parameters determine,
// computed as parameterSetup (saveFiles.cu:140-201) and the startup
// recomputation after the symmetry-reduction halving (main.cu:148-155).
struct CsvDerived {
  double hx, hy, dt, Dxx, Dyy, Dxy, rxy, rbx, rby, rx, ry, invdx, invdy, qx4, qy4, fx4, fy4;
};

CsvDerived csv_derived(const yh_params& p) {
  const double pi = 3.14159265359;  // globalVariables.cuh:27
  CsvDerived d;
  d.hx = p.Lx / (p.nx - 1.0);
  d.hy = p.Ly / (p.ny - 1.0);
  const double th = p.degrad * pi / 180.0;
  d.Dxx = p.diff_par * std::cos(th) * std::cos(th) + p.diff_per * std::sin(th) * std::sin(th);
  d.Dyy = p.diff_par * std::sin(th) * std::sin(th) + p.diff_per * std::cos(th) * std::cos(th);
  d.Dxy = (p.diff_par - p.diff_per) * std::sin(th) * std::cos(th);
  d.dt = p.reduceSym ? 0.5 * p.dt : p.dt;
  d.rxy = 2.0 * d.Dxy * d.dt / (4.0 * d.hx * d.hy);
  d.rbx = d.hx * d.Dxy / (d.Dxx * d.hy);
  d.rby = d.hy * d.Dxy / (d.Dyy * d.hx);
  d.rx = d.dt * d.Dxx / (d.hx * d.hx);
  d.ry = d.dt * d.Dyy / (d.hy * d.hy);
  d.invdx = 0.5 / d.hx;
  d.invdy = 0.5 / d.hy;
  d.qx4 = d.dt * d.Dyy / (d.hy * d.hy * 12.0);
  d.qy4 = d.dt * d.Dxx / (d.hx * d.hx * 12.0);
  d.fx4 = d.dt / 12.0;
  d.fy4 = d.dt / 12.0;
  return d;
}

}  // namespace

extern "C" int yh_params_write_csv(const char* path, const yh_params* p, const char* ic_path,
                                   const char* results_path, float tipx, float tipy) {
  if (!path || !p || p->nx < 2 || p->ny < 2) return YH_EINVAL;
  File fh(path, "w");
  if (!fh.f) return YH_EINVAL;
  FILE* f = fh.f;
  const CsvDerived d = csv_derived(*p);
  const auto F = [](double x) { return (double)(float)x; };  // the reference's (float) casts
  // Display and recording fields: parameterSetup defaults (saveFiles.cu:111-137,174-225).
  const int saveEveryIt = 0, plotTip = 1, recordTip = 0, plotContour = 0, recordContour = 0;
  const int stimulate = 0, plotTimeSeries = 1, recordTimeSeries = 0, tipAlgorithm = 1;
  const int tipGrad = 0, contourMode = 3, clock_ = 0, counterclock = 1;
  const double physicalTimeLim = 100000.0, startRecTime = 0.0;
  const int eSize = 2;
  // fibThreshold is declared bool (typeDefinition.cuh:83): its 0.1 is stored as true
  const double stimPeriod = 600.0, stimMag = 2.0, stimDuration = 10.0, fibThreshold = 1.0;
  const int fibTerminated = 0, leapShocks = 0, nc = 100, itPerFrame = 50;
  const double sample = 2.0;
  const float minVarColor = -0.1f, maxVarColor = 1.1f, uMin = -0.1f, uMax = 1.1f;
  const float vMin = -0.1f, vMax = 0.5f;
  const double c1 = 0.8, c2 = 0.85, c3 = 0.7;

  // printFunctions.cu:290-399, line by line.
  std::fprintf(f, "Initial condition path:,%s\n", ic_path ? ic_path : "NA");
  std::fprintf(f, "Results file path:,%s\n", results_path ? results_path : "NA");
  std::fprintf(f, "Save every sampling period:,%d\n", saveEveryIt);
  std::fprintf(f, "plot tip on screen:,%d\n", plotTip);
  std::fprintf(f, "Save tip in to data file:,%d\n", recordTip);
  std::fprintf(f, "Plot contours on screen:,%d\n", plotContour);
  std::fprintf(f, "Save contours to data file:,%d\n", recordContour);
  std::fprintf(f, "Pacing stimulus:,%d\n", stimulate);
  std::fprintf(f, "Plot time series:,%d\n", plotTimeSeries);
  std::fprintf(f, "Record time series:,%d\n", recordTimeSeries);
  std::fprintf(f, "Reduce symmetry:,%d\n", p->reduceSym ? 1 : 0);
  std::fprintf(f, "Solid boundary:,%d\n", p->solidSwitch ? 1 : 0);
  std::fprintf(f, "Neumann BCs:,%d\n", p->neumannBC ? 1 : 0);
  std::fprintf(f, "Gate diffusion:,%d\n", p->gateDiff ? 1 : 0);
  std::fprintf(f, "Tip trajectory algorithm:, %d\n", tipAlgorithm);
  std::fprintf(f, "Anisotropic tisue:,%d\n", p->anisotropy ? 1 : 0);
  std::fprintf(f, "Tip gradient:,%d\n", tipGrad);
  std::fprintf(f, "Contour mode (space APD or refractory):,%d\n", contourMode);
  std::fprintf(f, "Laplacian order:,%d\n", p->lap4 ? 1 : 0);
  std::fprintf(f, "Scheme order for time integration (Euler or RK4):,%d\n", p->timeIntOrder);
  std::fprintf(f, "Conduction block clock:,%d\n", clock_);
  std::fprintf(f, "Conduction block counterclock:,%d\n", counterclock);
  std::fprintf(f, "# grid points X =,%d\n", p->nx);
  std::fprintf(f, "# grid points Y =,%d\n", p->ny);
  std::fprintf(f, "Physical Lx length,%f\n", F(p->Lx));
  std::fprintf(f, "Physical Ly length,%f\n", F(p->Ly));
  std::fprintf(f, "Physical dx,%f\n", F(d.hx));
  std::fprintf(f, "Physical dy,%f\n", F(d.hy));
  std::fprintf(f, "Time step:,%f\n", p->reduceSym ? 2.0 * (double)(float)d.dt : d.dt);
  std::fprintf(f, "Diffusion parallel component:,%f\n", F(p->diff_par));
  std::fprintf(f, "Diffusion perpendicular component:,%f\n", F(p->diff_per));
  std::fprintf(f, "Initial fiber angle:,%f\n", F(p->degrad));
  std::fprintf(f, "Diffusion Dxx:,%f\n", F(d.Dxx));
  std::fprintf(f, "Diffusion Dyy:,%f\n", F(d.Dyy));
  std::fprintf(f, "Diffusion Dxy:,%f\n", F(d.Dxy));
  std::fprintf(f, "rxy (2*Dxy*dt/(4*dx*dy)):,%f\n", F(d.rxy));
  std::fprintf(f, "rbx (hx*Dxy/(Dxx*dy)):,%f\n", F(d.rbx));
  std::fprintf(f, "rby (hy*Dxy/(Dyy*dx)):,%f\n", F(d.rby));
  std::fprintf(f, "rx (Dxx*dt/(dx*dx)):,%f\n", F(d.rx));
  std::fprintf(f, "ry (Dyy*dt/(dy*dy)):,%f\n", F(d.ry));
  std::fprintf(f, "Gate r-scale:,%f\n", F(p->rscale));
  std::fprintf(f, "invdx (1/(2*hx)),%f\n", F(d.invdx));
  std::fprintf(f, "invdy (1/(2*hy)),%f\n", F(d.invdy));
  std::fprintf(f, "qx4 (dt*hx*hx*Dyy/12):,%f\n", F(d.qx4));
  std::fprintf(f, "qy4 (dt*hy*hx*Dxx/12):,%f\n", F(d.qy4));
  std::fprintf(f, "fx4 (dt*hx*hx/12):,%f\n", F(d.fx4));
  std::fprintf(f, "fy4 (dt*hy*hy/12):,%f\n", F(d.fy4));
  std::fprintf(f, "Physical time limit:,%f\n", F(physicalTimeLim));
  std::fprintf(f, "Start recording time:,%f\n", F(startRecTime));
  std::fprintf(f, "Number of electrodes,%d\n", eSize);
  std::fprintf(f, "Electrode position x:,%f\n", F(p->nx / 2));
  std::fprintf(f, "Electrode position y:,%f\n", F(p->ny / 2));
  std::fprintf(f, "Stimulus period (ms):,%f\n", stimPeriod);
  std::fprintf(f, "Stimulus magnitude:,%f\n", F(stimMag));
  std::fprintf(f, "Stimulus duration (ms):,%f\n", F(stimDuration));
  std::fprintf(f, "Voltage threshold for fibrillation:,%f\n", F(fibThreshold));
  std::fprintf(f, "Fibrillation terminated by pacing:,%d\n", fibTerminated);
  std::fprintf(f, "Number of LEAP shocks applied:,%d\n", leapShocks);
  std::fprintf(f, "Number of points in circles:,%d\n", nc);
  std::fprintf(f, "Stimulus position x:,%f\n", F(0.25 * p->Lx));
  std::fprintf(f, "Stimulus position y:,%f\n", F(0.25 * p->Ly));
  std::fprintf(f, "Stimulus area radius:,%f\n", F(0.03 * p->Lx));
  std::fprintf(f, "APD area radius:,%f\n", F(0.15 * p->Lx));
  std::fprintf(f, "Tip offset x,%d\n", p->tipOffsetX);
  std::fprintf(f, "Tip offset y,%d\n", p->tipOffsetY);
  std::fprintf(f, "Integral area radius:,%f\n", F(0.5 * ((p->tipOffsetX + p->tipOffsetY) * d.hx)));
  std::fprintf(f, "Dirichlet BC value:,%f\n", F(p->boundaryVal));
  std::fprintf(f, "Iterations per frame:,%d\n", itPerFrame);
  std::fprintf(f, "Sampling period (ms):,%f\n", sample);
  std::fprintf(f, "Min signal range,%f\n", (double)minVarColor);
  std::fprintf(f, "Max signal range,%f\n", (double)maxVarColor);
  std::fprintf(f, "Secondary window size x,%d\n", p->nx);
  std::fprintf(f, "Secondary window size y,%d\n", p->ny);
  std::fprintf(f, "Secondary window min signal 1,%f\n", (double)uMin);
  std::fprintf(f, "Secondary window max signal 1,%f\n", (double)uMax);
  std::fprintf(f, "Secondary window min signal 2,%f\n", (double)vMin);
  std::fprintf(f, "Secondary window max signal 2,%f\n", (double)vMax);
  std::fprintf(f, "Last tip point X,%f\n", (double)tipx);
  std::fprintf(f, "Last tip point Y,%f\n", (double)tipy);
  std::fprintf(f, "Contour threshold 1:,%f\n", F(c1));
  std::fprintf(f, "Contour threshold 2:,%f\n", F(c2));
  std::fprintf(f, "Contour threshold 3:,%f\n", F(c3));
  std::fprintf(f, "Filament voltage threshold:,%f\n", F(p->Uth));
  std::fprintf(f, "time scale (tc):,%f\n", F(p->tc));
  std::fprintf(f, "alpha:,%f\n", F(p->alpha));
  std::fprintf(f, "beta:,%f\n", F(p->beta));
  std::fprintf(f, "gamma:,%f\n", F(p->gamma));
  std::fprintf(f, "delta:,%f\n", F(p->delta));
  std::fprintf(f, "epsilon:,%f\n", F(p->eps));
  std::fprintf(f, "mu:,%f\n", F(p->mu));
  std::fprintf(f, "theta:,%f\n", F(p->theta));
  return fh.close();
}

// Line numbers (0-based) of the solver fields in the positional layout that
// loadParamValues reads (saveFiles.cu:596-708; its index starts at 1 and is
// pre-incremented, so the first value read is line 2).
enum CsvLine {
  L_REDUCESYM = 10, L_SOLID = 11, L_NEUMANN = 12, L_GATEDIFF = 13, L_ANISO = 15, L_LAP4 = 18,
  L_TIMEINT = 19, L_NX = 22, L_NY = 23, L_LX = 24, L_LY = 25, L_DT = 28, L_DPAR = 29,
  L_DPER = 30, L_DEGRAD = 31, L_RSCALE = 40, L_TIPOFFX = 63, L_TIPOFFY = 64, L_BVAL = 66,
  L_UTH = 82, L_TC = 83, L_ALPHA = 84, L_BETA = 85, L_GAMMA = 86, L_DELTA = 87, L_EPS = 88,
  L_MU = 89, L_THETA = 90, L_COUNT = 91
};

extern "C" int yh_params_read_csv(const char* path, yh_params* p) {
  if (!path || !p) return YH_EINVAL;
  File fh(path, "r");
  if (!fh.f) return YH_EINVAL;
  // second comma-separated field of each line, strtof (saveFiles.cu:566-582)
  std::vector<float> v;
  char line[512];
  while (std::fgets(line, sizeof(line), fh.f)) {
    const char* comma = std::strchr(line, ',');
    v.push_back(comma ? std::strtof(comma + 1, nullptr) : 0.0f);
  }
  if ((int)v.size() < L_COUNT) return YH_EINVAL;
  // integer fields must be finite and in int range (the conversion of any
  // other float is undefined); such a file is rejected
  bool bad = false;
  const auto I = [&](int l) {
    const float x = v[l];
    if (!(x > -2147483520.0f && x < 2147483520.0f)) {
      bad = true;
      return 0;
    }
    return (int)x;
  };
  const auto D = [&](int l) { return (double)v[l]; };
  yh_params q = *p;  // *p changes only when the whole file is valid
  q.reduceSym = I(L_REDUCESYM);
  q.solidSwitch = I(L_SOLID);
  q.neumannBC = I(L_NEUMANN);
  q.gateDiff = I(L_GATEDIFF);
  q.anisotropy = I(L_ANISO);
  q.lap4 = I(L_LAP4);
  q.timeIntOrder = I(L_TIMEINT);
  q.nx = I(L_NX);
  q.ny = I(L_NY);
  q.Lx = D(L_LX);
  q.Ly = D(L_LY);
  q.dt = D(L_DT);
  q.diff_par = D(L_DPAR);
  q.diff_per = D(L_DPER);
  q.degrad = D(L_DEGRAD);
  q.rscale = D(L_RSCALE);
  q.tipOffsetX = I(L_TIPOFFX);
  q.tipOffsetY = I(L_TIPOFFY);
  q.boundaryVal = D(L_BVAL);
  q.Uth = D(L_UTH);
  q.tc = D(L_TC);
  q.alpha = D(L_ALPHA);
  q.beta = D(L_BETA);
  q.gamma = D(L_GAMMA);
  q.delta = D(L_DELTA);
  q.eps = D(L_EPS);
  q.mu = D(L_MU);
  q.theta = D(L_THETA);
  if (bad) return YH_EINVAL;
  *p = q;
  return YH_OK;
}

extern "C" int yh_state_write_dat(const char* path, int nx, int ny, const float* u,
                                  const float* v, const float* w) {
  if (!path || nx < 1 || ny < 1 || !u || !v) return YH_EINVAL;
  File fh(path, "w");
  if (!fh.f) return YH_EINVAL;
  const size_t n = (size_t)nx * ny;
  for (size_t k = 0; k < n; ++k) {
    if (w)
      std::fprintf(fh.f, "%f %f %f\n", (double)u[k], (double)v[k], (double)w[k]);
    else
      std::fprintf(fh.f, "%f %f\n", (double)u[k], (double)v[k]);
  }
  return fh.close();
}

extern "C" int yh_state_read_dat(const char* path, int nx, int ny, float* u, float* v, float* w) {
  if (!path || nx < 1 || ny < 1 || !u || !v) return YH_EINVAL;
  File fh(path, "r");
  if (!fh.f) return YH_EINVAL;
  const size_t n = (size_t)nx * ny;
  for (size_t k = 0; k < n; ++k) {
    const int got = w ? std::fscanf(fh.f, "%f %f %f", &u[k], &v[k], &w[k])
                      : std::fscanf(fh.f, "%f %f", &u[k], &v[k]);
    if (got != (w ? 3 : 2)) return YH_EINVAL;
  }
  return YH_OK;
}

extern "C" int yh_tips_write_dat(const char* tip_path, const char* size_path, const yh_tip* tips,
                                 int n, int append) {
  if (!tip_path || !size_path || n < 0 || (n > 0 && !tips)) return YH_EINVAL;
  File f1(tip_path, append ? "a" : "w");
  File f2(size_path, append ? "a" : "w");
  if (!f1.f || !f2.f) return YH_EINVAL;
  if (n > 0) {
    for (int i = 0; i < n; ++i)
      std::fprintf(f1.f, "%f %f %f %f %f\n", (double)tips[i].x, (double)tips[i].y,
                   (double)tips[i].vx, (double)tips[i].vy, (double)tips[i].t);
    std::fprintf(f2.f, "%d\n", n);
  }
  const int r1 = f1.close(), r2 = f2.close();
  return r1 ? r1 : r2;
}

extern "C" int yh_sym_write_dat(const char* path, const double* c_phi, int nsteps) {
  if (!path || nsteps < 0 || (nsteps > 0 && !c_phi)) return YH_EINVAL;
  File fh(path, "w");
  if (!fh.f) return YH_EINVAL;
  for (int i = 0; i < nsteps; ++i) {
    const double* r = c_phi + 6 * (size_t)i;
    std::fprintf(fh.f, "%f %f %f %f %f %f\n", r[0], r[1], r[2], r[3], r[4], r[5]);
  }
  return fh.close();
}

extern "C" int yh_mask_read_dat(const char* path, int nx, int ny, uint8_t* mask) {
  if (!path || nx < 1 || ny < 1 || !mask) return YH_EINVAL;
  File fh(path, "r");
  if (!fh.f) return YH_EINVAL;
  const size_t n = (size_t)nx * ny;
  for (size_t k = 0; k < n; ++k) {
    float m;
    if (std::fscanf(fh.f, "%f", &m) != 1) return YH_EINVAL;
    mask[k] = m > 0.5f ? 1 : 0;
  }
  return YH_OK;
}

extern "C" int yh_mask_write_dat(const char* path, int nx, int ny, const uint8_t* mask) {
  if (!path || nx < 1 || ny < 1 || !mask) return YH_EINVAL;
  File fh(path, "w");
  if (!fh.f) return YH_EINVAL;
  const size_t n = (size_t)nx * ny;
  for (size_t k = 0; k < n; ++k) std::fprintf(fh.f, "%16.7e\n", mask[k] ? 1.0 : 0.0);
  return fh.close();
}
