// Ensemble: many independent runs (own grid, Ma, Kn, CFL, tmax; each its own CFL dt)
// advanced together by one HIP-graph-captured step.
//
// Layout: the runs' states are contiguous parts of shared SoA planes ([35][Ntot]; run r
// owns cells [off_r, off_r + N_r), each in its own Julia order).  Kernels run over blocks
// of one run each: a block table maps blockIdx -> (run, first cell), so a block reads its
// run's parameters and step scalars once and a finished run costs one load per block.
// One step = five launches (begin, pass 1, dt, update, pass 2) captured once into a graph
// and replayed; the per-run dt reduction is an atomic max into the run's scalars.
//
// The block size of each of the three cell kernels is a tunable: hyq_ensemble_tune times
// every candidate on the ensemble's own state (saved and restored around the trial steps)
// and keeps the fastest per kernel.
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>
#include <vector>

#include "../../include/hyqmom.h"
#include "hyq_ensemble_kernels.h"

using namespace hyq;
using namespace hyq_ens;

namespace {

thread_local std::string g_eerr;
int efail(int code, const char* what, hipError_t e = hipSuccess) {
  g_eerr = what;
  if (e != hipSuccess) {
    g_eerr += ": ";
    g_eerr += hipGetErrorString(e);
  }
  return code;
}
#define ECHK(x, msg)                                       \
  do {                                                     \
    hipError_t _e = (x);                                   \
    if (_e != hipSuccess) return efail(HYQ_EHIP, msg, _e); \
  } while (0)

constexpr int kCand[3] = {64, 128, 256};  // block-size candidates

__global__ void ke_begin(RunDev* R, int nr) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nr) return;
  RunDev& q = R[r];
  // a run whose z wave-speed solve met a non-finite Jacobian stops there (the reference throws)
  const bool act = q.t < q.tmax && q.nn < q.nnmax && q.err == 0ull;
  q.active = act ? 1.0 : 0.0;
  if (act) q.nn += 1;
  q.vmax_bits = 0ull;
}

__global__ void ke_dt(RunDev* R, int nr) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nr) return;
  RunDev& q = R[r];
  if (q.active == 0.0) return;
  const double vmax = __longlong_as_double((long long)q.vmax_bits);
  const double hmin = jmin(jmin(q.dx, q.dy), q.dz);
  double dt = jmin(q.CFL * hmin / vmax, jmin(q.dtmax, q.Kn));
  dt = jmin(dt, q.tmax - q.t);
  q.dt = dt;
  q.t += dt;
}

int cand_index(int b) {
  for (int i = 0; i < 3; ++i)
    if (kCand[i] == b) return i;
  return -1;
}

}  // namespace

struct hyq_ensemble {
  int nr = 0, prec = 0;
  long long P = 0;  // total cells = plane stride
  std::vector<RunDev> runs;
  RunDev* R = nullptr;
  double *M = nullptr, *Mr = nullptr, *Mbak = nullptr;
  void *F = nullptr, *vp = nullptr;
  uint32_t* flags = nullptr;
  Blk* T[3] = {nullptr, nullptr, nullptr};
  int nblk[3] = {0, 0, 0};
  int blocks[3] = {128, 128, 128};  // pass 1, update, pass 2
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
};

namespace {

void launch_kernel(hyq_ensemble* e, int which, int b, hipStream_t s) {
  const int ci = cand_index(b);
  const EnsPtrs p{e->M, e->Mr, e->F, e->vp, e->flags, e->R, e->P};
  if (e->prec == HYQ_PREC_FP64) launch_fp64(which, b, (unsigned)e->nblk[ci], e->T[ci], p, s);
  else if (e->prec == HYQ_PREC_MIXED) launch_mixed(which, b, (unsigned)e->nblk[ci], e->T[ci], p, s);
  else launch_newton(which, b, (unsigned)e->nblk[ci], e->T[ci], p, s);
}

// One step; with ev != null, events around the three cell kernels (ev[0..3]).
void enqueue(hyq_ensemble* e, hipStream_t s, hipEvent_t* ev) {
  const unsigned gr = (unsigned)((e->nr + 63) / 64);
  hipLaunchKernelGGL(ke_begin, dim3(gr), dim3(64), 0, s, e->R, e->nr);
  if (ev) hipEventRecord(ev[0], s);
  launch_kernel(e, 0, e->blocks[0], s);
  if (ev) hipEventRecord(ev[1], s);
  hipLaunchKernelGGL(ke_dt, dim3(gr), dim3(64), 0, s, e->R, e->nr);
  launch_kernel(e, 1, e->blocks[1], s);
  if (ev) hipEventRecord(ev[2], s);
  launch_kernel(e, 2, e->blocks[2], s);
  if (ev) hipEventRecord(ev[3], s);
}

void drop_graph(hyq_ensemble* e) {
  if (e->exec) (void)hipGraphExecDestroy(e->exec);
  if (e->graph) (void)hipGraphDestroy(e->graph);
  e->exec = nullptr;
  e->graph = nullptr;
}

int build_graph(hyq_ensemble* e) {
  drop_graph(e);
  ECHK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal), "begin capture");
  enqueue(e, e->stream, nullptr);
  ECHK(hipStreamEndCapture(e->stream, &e->graph), "end capture");
  ECHK(hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0), "graph instantiate");
  return HYQ_OK;
}

}  // namespace

extern "C" {

const char* hyq_ensemble_last_error(void) { return g_eerr.c_str(); }

int hyq_ensemble_create(const hyq_params* runs, int32_t nruns, hyq_ensemble** out) {
  if (!runs || nruns <= 0 || !out) return efail(HYQ_EINVAL, "bad argument");
  const int prec = runs[0].precision;
  if (prec != HYQ_PREC_FP64 && prec != HYQ_PREC_MIXED && prec != HYQ_PREC_MIXED_NEWTON)
    return efail(HYQ_EINVAL, "ensemble precision must be FP64, MIXED or MIXED_NEWTON");
  hyq_ensemble* e = new hyq_ensemble();
  e->nr = nruns;
  e->prec = prec;
  long long off = 0;
  for (int r = 0; r < nruns; ++r) {
    const hyq_params& p = runs[r];
    if (p.Nmom != 35 || p.Nx < 1 || p.Ny < 1 || p.Nz < 1 || p.precision != prec) {
      delete e;
      return efail(HYQ_EINVAL, "run parameters (Nmom 35, sizes >= 1, one precision for all)");
    }
    RunDev q;
    std::memset(&q, 0, sizeof(q));
    q.off = off;
    q.N = (long long)p.Nx * p.Ny * p.Nz;
    q.Nx = p.Nx; q.Ny = p.Ny; q.Nz = p.Nz;
    q.flag2D = p.flag2D; q.positivity = p.positivity;
    q.Ma = p.Ma; q.Kn = p.Kn; q.CFL = p.CFL; q.dtmax = p.dtmax;
    q.dx = (p.xmax - p.xmin) / p.Nx; q.dy = (p.ymax - p.ymin) / p.Ny; q.dz = (p.zmax - p.zmin) / p.Nz;
    q.tmax = p.tmax;
    q.nnmax = (1ll << 62);
    e->runs.push_back(q);
    off += q.N;
    if (off >= hyq::kMaxUpdateCells) {
      delete e;
      return efail(HYQ_EINVAL, "more than 2^29 - 1 cells in all runs");
    }
  }
  e->P = off;
  const size_t fsz = prec == HYQ_PREC_FP64 ? sizeof(double) : sizeof(float);
  bool ok = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) == hipSuccess &&
            hipMalloc(&e->R, sizeof(RunDev) * nruns) == hipSuccess &&
            hipMalloc(&e->M, sizeof(double) * 35 * e->P) == hipSuccess &&
            hipMalloc(&e->Mr, sizeof(double) * 35 * e->P) == hipSuccess &&
            hipMalloc(&e->F, fsz * 105 * e->P) == hipSuccess &&
            hipMalloc(&e->vp, fsz * 6 * e->P) == hipSuccess &&
            hipMalloc(&e->flags, sizeof(uint32_t) * e->P) == hipSuccess;
  for (int i = 0; i < 4 && ok; ++i) ok = hipEventCreate(&e->ev[i]) == hipSuccess;
  // block tables, one per candidate block size
  for (int ci = 0; ci < 3 && ok; ++ci) {
    std::vector<Blk> tab;
    for (int r = 0; r < nruns; ++r)
      for (long long f = 0; f < e->runs[r].N; f += kCand[ci]) tab.push_back(Blk{r, 0, e->runs[r].off + f});
    e->nblk[ci] = (int)tab.size();
    ok = hipMalloc(&e->T[ci], sizeof(Blk) * tab.size()) == hipSuccess &&
         hipMemcpy(e->T[ci], tab.data(), sizeof(Blk) * tab.size(), hipMemcpyHostToDevice) == hipSuccess;
  }
  if (ok) ok = hipMemcpy(e->R, e->runs.data(), sizeof(RunDev) * nruns, hipMemcpyHostToDevice) == hipSuccess;
  if (ok) ok = hipMemset(e->M, 0, sizeof(double) * 35 * e->P) == hipSuccess;
  if (!ok) {
    hyq_ensemble_destroy(e);
    return efail(HYQ_ENOMEM, "ensemble allocation");
  }
  *out = e;
  return HYQ_OK;
}

int hyq_ensemble_destroy(hyq_ensemble* e) {
  if (!e) return HYQ_OK;
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  drop_graph(e);
  (void)hipFree(e->R);
  (void)hipFree(e->M);
  (void)hipFree(e->Mr);
  (void)hipFree(e->Mbak);
  (void)hipFree(e->F);
  (void)hipFree(e->vp);
  (void)hipFree(e->flags);
  for (int i = 0; i < 3; ++i) (void)hipFree(e->T[i]);
  for (int i = 0; i < 4; ++i)
    if (e->ev[i]) (void)hipEventDestroy(e->ev[i]);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
  return HYQ_OK;
}

// State of run r (host, Julia layout (Nx,Ny,Nz,35)); resets the run's t and nn to 0.
int hyq_ensemble_set_state(hyq_ensemble* e, int32_t r, const double* M) {
  if (!e || !M || r < 0 || r >= e->nr) return efail(HYQ_EINVAL, "bad argument");
  const RunDev& q = e->runs[r];
  ECHK(hipStreamSynchronize(e->stream), "sync");
  for (int n = 0; n < 35; ++n)
    ECHK(hipMemcpy(e->M + (size_t)n * e->P + q.off, M + (size_t)n * q.N, sizeof(double) * q.N,
                   hipMemcpyHostToDevice), "state upload");
  RunDev h;
  ECHK(hipMemcpy(&h, e->R + r, sizeof(RunDev), hipMemcpyDeviceToHost), "run scalars");
  h.t = 0.0;
  h.nn = 0;
  h.active = 0.0;
  h.err = 0ull;
  ECHK(hipMemcpy(e->R + r, &h, sizeof(RunDev), hipMemcpyHostToDevice), "run scalars");
  return HYQ_OK;
}

int hyq_ensemble_get_state(hyq_ensemble* e, int32_t r, double* M) {
  if (!e || !M || r < 0 || r >= e->nr) return efail(HYQ_EINVAL, "bad argument");
  const RunDev& q = e->runs[r];
  ECHK(hipStreamSynchronize(e->stream), "sync");
  for (int n = 0; n < 35; ++n)
    ECHK(hipMemcpy(M + (size_t)n * q.N, e->M + (size_t)n * e->P + q.off, sizeof(double) * q.N,
                   hipMemcpyDeviceToHost), "state download");
  return HYQ_OK;
}

// Per-run time, last dt and step count (arrays of nruns; any may be NULL).
int hyq_ensemble_status(hyq_ensemble* e, double* t, double* dt, int64_t* nn) {
  if (!e) return efail(HYQ_EINVAL, "bad argument");
  std::vector<RunDev> h(e->nr);
  ECHK(hipStreamSynchronize(e->stream), "sync");
  ECHK(hipMemcpy(h.data(), e->R, sizeof(RunDev) * e->nr, hipMemcpyDeviceToHost), "run scalars");
  for (int r = 0; r < e->nr; ++r) {
    if (t) t[r] = h[r].t;
    if (dt) dt[r] = h[r].dt;
    if (nn) nn[r] = h[r].nn;
  }
  return HYQ_OK;
}

// nsteps steps of every run still below its tmax (runs past it stay put): graph replays
// when use_graph != 0, else direct launches.  Synchronizes before returning.
int hyq_ensemble_run(hyq_ensemble* e, int64_t nsteps, int32_t use_graph) {
  if (!e || nsteps < 0) return efail(HYQ_EINVAL, "bad argument");
  if (use_graph && !e->exec) {
    const int rc = build_graph(e);
    if (rc) return rc;
  }
  for (int64_t s = 0; s < nsteps; ++s) {
    if (use_graph) {
      ECHK(hipGraphLaunch(e->exec, e->stream), "graph launch");
    } else {
      enqueue(e, e->stream, nullptr);
    }
  }
  ECHK(hipGetLastError(), "ensemble launch");
  ECHK(hipStreamSynchronize(e->stream), "ensemble sync");
  std::vector<RunDev> h(e->nr);
  ECHK(hipMemcpy(h.data(), e->R, sizeof(RunDev) * e->nr, hipMemcpyDeviceToHost), "run scalars");
  std::string bad;
  for (int r = 0; r < e->nr; ++r)
    if (h[r].err) bad += (bad.empty() ? "" : ", ") + std::to_string(r);
  if (!bad.empty())
    return efail(HYQ_ENONFINITE, ("non-finite Jacobian in the z wave-speed solve of run(s) " + bad +
                                  " (eigenvalues6z_hyperbolic_3D.jl:108,121 throws); they stopped there")
                                     .c_str());
  return HYQ_OK;
}

// Block sizes (pass 1, update, pass 2), each one of 64, 128, 256.
int hyq_ensemble_set_blocks(hyq_ensemble* e, const int32_t* blocks) {
  if (!e || !blocks) return efail(HYQ_EINVAL, "bad argument");
  for (int i = 0; i < 3; ++i)
    if (cand_index(blocks[i]) < 0) return efail(HYQ_EINVAL, "block size must be 64, 128 or 256");
  for (int i = 0; i < 3; ++i) e->blocks[i] = blocks[i];
  drop_graph(e);
  return HYQ_OK;
}

// Autotuner: for each candidate block size, `trial_steps` steps from the current state
// (saved and restored; run scalars too) timed per kernel with HIP events; each kernel
// keeps its fastest candidate.  ms_out [3][3] (candidate x kernel) may be NULL.
int hyq_ensemble_tune(hyq_ensemble* e, int32_t trial_steps, int32_t* blocks_out, double* ms_out) {
  if (!e || trial_steps < 1) return efail(HYQ_EINVAL, "bad argument");
  if (!e->Mbak) ECHK(hipMalloc(&e->Mbak, sizeof(double) * 35 * e->P), "tuner backup");
  std::vector<RunDev> rbak(e->nr);
  ECHK(hipStreamSynchronize(e->stream), "sync");
  ECHK(hipMemcpy(rbak.data(), e->R, sizeof(RunDev) * e->nr, hipMemcpyDeviceToHost), "run scalars");
  ECHK(hipMemcpyAsync(e->Mbak, e->M, sizeof(double) * 35 * e->P, hipMemcpyDeviceToDevice, e->stream),
       "tuner backup");
  double ms[3][3];
  for (int ci = 0; ci < 3; ++ci) {
    for (int k = 0; k < 3; ++k) e->blocks[k] = kCand[ci];
    for (int k = 0; k < 3; ++k) ms[ci][k] = 0.0;
    for (int s = 0; s < trial_steps; ++s) {
      enqueue(e, e->stream, e->ev);
      ECHK(hipStreamSynchronize(e->stream), "tuner step");
      float a, b, c;
      ECHK(hipEventElapsedTime(&a, e->ev[0], e->ev[1]), "event");
      ECHK(hipEventElapsedTime(&b, e->ev[1], e->ev[2]), "event");
      ECHK(hipEventElapsedTime(&c, e->ev[2], e->ev[3]), "event");
      ms[ci][0] += a;
      ms[ci][1] += b;
      ms[ci][2] += c;
    }
    ECHK(hipMemcpyAsync(e->M, e->Mbak, sizeof(double) * 35 * e->P, hipMemcpyDeviceToDevice, e->stream),
         "tuner restore");
    ECHK(hipMemcpyAsync(e->R, rbak.data(), sizeof(RunDev) * e->nr, hipMemcpyHostToDevice, e->stream),
         "tuner restore");
    ECHK(hipStreamSynchronize(e->stream), "tuner restore");
  }
  for (int k = 0; k < 3; ++k) {
    int best = 0;
    for (int ci = 1; ci < 3; ++ci)
      if (ms[ci][k] < ms[best][k]) best = ci;
    e->blocks[k] = kCand[best];
    if (blocks_out) blocks_out[k] = kCand[best];
  }
  if (ms_out)
    for (int ci = 0; ci < 3; ++ci)
      for (int k = 0; k < 3; ++k) ms_out[ci * 3 + k] = ms[ci][k] / trial_steps;
  drop_graph(e);
  return HYQ_OK;
}

}  // extern "C"
