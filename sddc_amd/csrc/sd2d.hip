// sd2d.hip - C ABI (include/sd2d.h) of the MI355X SD2D-MHD residual and
// RK path: context set-up, device layouts, and the launch sequence.
//
// Device layout ("tile-major SoA"): cells are grouped in tiles of CPW
// cells (the cells of one workgroup of the stage kernel); inside a tile
// every (variable, point) pair is a contiguous run of CPW doubles, so the
// threads of one grid line read and write whole runs, and a tile of state
// is one contiguous block.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/sd2d.h"
#include "sd2d_host.hpp"
#include "sd2d_stage.hpp"
#include "sd2d_stage_mx.hpp"
#include "sd2d_visc.hpp"
#include "sd2d_adj.hpp"
#include "sd2d_rkseq.hpp"

using namespace sd2d;

struct AdjState;                      // the adjoint's buffers (sd2d_adjhost.hpp)
static void adj_free(AdjState *A);

struct sd2d_ctx {
  int N = 0, ncell = 0, ntile = 0, cpw = 0;
  sd2d_params p{};
  int device = 0;
  hipStream_t stream = nullptr;
  Ops ops{};
  OpsF opsf{};                                  // FP32 operators (SD2D_MIXED)
  float *metf = nullptr;                        // FP32 metrics (SD2D_MIXED)
  Phys ph{};
  double *q[3] = {nullptr, nullptr, nullptr};  // RK state buffers
  int cur = 0;                                  // buffer holding the state
  double *fsr[4] = {nullptr, nullptr, nullptr, nullptr};  // FiveStageRk registers
  double *met = nullptr;
  int *ftab = nullptr;
  int *c2s = nullptr, *s2c = nullptr;          // cell <-> slot maps of the tile shape
  double avf[NMAX * NMAX] = {};               // AV smoothing filter (kernel arguments)
  double *gmax = nullptr;                       // eigvmax_g, umax_g
  double *lastch = nullptr;                     // c_h of the last evaluation (per member)
  unsigned long long *slots = nullptr;
  double *sm_ind = nullptr;
  double *normbuf = nullptr, *normout = nullptr;
  double *stagef = nullptr;                     // Fortran-layout staging (Q or resid)
  double *nablaf = nullptr;                     // Fortran-layout nablaQs
  // ensemble group: nmem members share the launches (per-tile member and
  // valid-cell count, per-member dt); 1 and null otherwise.  phm holds the
  // free stream of every member (one entry for one run).
  int nmem = 1;
  long long nlaunch = 0;                // stage launches of sd2d_step so far
  int next_stage = 1;                   // sd2d_rk_stage: the next stage of the step
  int check_every = 0;                  // opt-in non-finite check (sd2d_set_check_finite)
  int *nfflag = nullptr;                // its per-member flags (device)
  // the line family (0: xi, 1: eta) whose wave stages a tile into LDS: the
  // one with fewer end faces whose neighbour lies in another tile (fewer
  // neighbour gathers), from the face table
  int sfam = 1;
  int *tinfo = nullptr;
  double *dtm = nullptr;
  Phys *phm = nullptr;
  AdjState *adj = nullptr;              // sd2d_fwd / sd2d_vjp state
  std::vector<double> xs;               // solution-point positions (face quadrature)
  // viscous path (vismode 1, mu or eta0 non-zero)
  bool vis = false;
  double *gq = nullptr, *dei = nullptr, *visres = nullptr, *metw = nullptr;
  double ctime = 0.0;
  int iter = 0;
  double resid_rms = 0.0, divb_rms = 0.0;
  int norms_valid = 0;
  bool have_av = false;
  // profiling (sd2d_set_profiling)
  bool prof = false;
  std::vector<hipEvent_t> ev;     // pairs (start, stop)
  size_t nev_used = 0;
  double prof_ms = 0.0;
  int prof_n = 0;
  std::string err;
};

static std::string g_err = "no error";

// The nablaQs staging of sd2d_compresid_noalloc: a file-scope device array,
// so that entry point allocates nothing.  Each GPU holds its own copy for
// the life of the process; its address on a device comes from
// hipGetSymbolAddress with that device current (in host code the array's
// name is only a host-side stand-in and must never be passed to a kernel
// or a copy).  Calls that overlap on one GPU share it and race.
static __device__ __align__(256) double scratch[SD2D_NOALLOC_MAX_DOUBLES];
constexpr int NOALLOC_MAX_DEVICES = 64;
static double *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};
static hipError_t g_scratch_err[NOALLOC_MAX_DEVICES] = {};
static std::mutex g_scratch_mu;

// Look up (once per device, the device current) the address of `scratch`.
static void scratch_init(int dev) {
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[dev]) return;
  void *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  g_scratch_err[dev] = e;
  if (e == hipSuccess) g_scratch_ptr[dev] = static_cast<double *>(ptr);
}

// vismode 0 is refused: the reference then forms the Powell source and the
// divB norm from a nablaQs it never computes (compflux.f90:13-15 against
// compresid.f90:58, iterations.f90:43), so its result is undefined.
static const char *const vismode_msg =
    "vismode must be 1: with vismode 0 the reference reads nablaQs without computing it "
    "(compflux.f90:13-15, compresid.f90:58), so there is no defined result to reproduce";

// Calls that read or replace the state, or evaluate the residual outside
// the stage sequence, are refused between sd2d_rk_stage calls of one step:
// the state buffer then holds a stage's working values (RK_S5P3 stage 4
// writes its output in place) and a residual evaluation would raise the
// running maxima in the middle of the step.
#define SD2D_NOT_MID_STEP(c, what)                                                   \
  do {                                                                           \
    if ((c)->next_stage != 1)                                                    \
      return fail((c), SD2D_EINVAL,                                              \
                  std::string(what) + ": a step is in progress (sd2d_rk_stage): the next stage is " + \
                      std::to_string((c)->next_stage));                          \
  } while (0)

#define HIPCHK(ctx, call)                                                       \
  do {                                                                          \
    hipError_t e_ = (call);                                                     \
    if (e_ != hipSuccess) {                                                     \
      (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_);           \
      return SD2D_EHIP;                                                         \
    }                                                                           \
  } while (0)

static int fail(sd2d_ctx *c, int code, const std::string &m) {
  if (c) c->err = m; else g_err = m;
  return code;
}

// ---------------------------------------------------------------------------
// small device kernels
// ---------------------------------------------------------------------------
// Fortran Q(9,N,N,NCELL) -> tile layout (qo) through the slot -> cell map
// (padding slots get zeros; no valid lane reads them).
template <int N>
__global__ void k_f2t(const double *__restrict__ f, double *__restrict__ t,
                      const int *__restrict__ s2c, int ntile) {
  constexpr int CPW = Geo<N>::CPW, N2 = Geo<N>::N2, QT = Geo<N>::QT;
  const size_t tot = (size_t)ntile * QT;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < tot;
       i += (size_t)gridDim.x * blockDim.x) {
    const size_t tile = i / QT;
    int v, p, ci;
    qo_inv<N>(int(i - tile * QT), v, p, ci);
    const int cell = s2c[tile * CPW + ci];
    t[i] = cell >= 0 ? f[v + 9 * (p + (size_t)N2 * cell)] : 0.0;
  }
}

template <int N>
__global__ void k_t2f(const double *__restrict__ t, double *__restrict__ f,
                      const int *__restrict__ c2s, int ncell) {
  constexpr int CPW = Geo<N>::CPW, N2 = Geo<N>::N2;
  const size_t tot = (size_t)ncell * 9 * N2;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < tot;
       i += (size_t)gridDim.x * blockDim.x) {
    const int v = int(i % 9);
    const size_t r = i / 9;
    const int p = int(r % N2);
    const size_t slot = c2s[r / N2];
    const size_t tile = slot / CPW;
    f[i] = t[tile * Geo<N>::QT + qo<N>(v, p, int(slot % CPW))];
  }
}

// running maxima: g = max over slots (one block per ensemble member).
// The values g holds on entry are the ones the stage just finished formed
// its c_h from (compflux.f90:6), so that c_h is recorded first, with the
// kernel's own formula.  No NaN reaches the slots: every point adopts a
// speed only through `speed > running max`, which a NaN never passes
// (getrusanov.f90:116-123).
__global__ void k_finalize_max(const unsigned long long *__restrict__ slots,
                               double *__restrict__ gmax, double *__restrict__ lastch) {
  __shared__ unsigned long long s[2][NSLOT];
  const int t = threadIdx.x;
  slots += 2 * NSLOT * (size_t)blockIdx.x;
  gmax += 2 * blockIdx.x;
  if (t == 0) {
    const double ge = gmax[0], gu = gmax[1];
    lastch[blockIdx.x] = sqrt(ge * (ge - gu));
  }
  s[0][t] = slots[2 * t];
  s[1][t] = slots[2 * t + 1];
  __syncthreads();
  for (int o = NSLOT / 2; o > 0; o >>= 1) {
    if (t < o) {
      s[0][t] = s[0][t] > s[0][t + o] ? s[0][t] : s[0][t + o];
      s[1][t] = s[1][t] > s[1][t + o] ? s[1][t] : s[1][t + o];
    }
    __syncthreads();
  }
  if (t == 0) {
    gmax[0] = __longlong_as_double((long long)s[0][0]);
    gmax[1] = __longlong_as_double((long long)s[1][0]);
  }
}

// Non-finite scan of a state in tile layout (the opt-in check of
// sd2d_set_check_finite): flags[member] = 1 when a tile of that member holds
// a NaN or an infinity.  Padding slots hold zeros, so they never raise it.
__global__ void k_scan_finite(const double2 *__restrict__ q, const int *__restrict__ tinfo,
                              int ntile, int qt2, int *__restrict__ flags) {
  for (int t = blockIdx.x; t < ntile; t += gridDim.x) {   // uniform over the block
    const double2 *b = q + (size_t)t * qt2;
    bool bad = false;
    for (int i = threadIdx.x; i < qt2; i += blockDim.x) {
      const double2 v = b[i];
      bad |= !(isfinite(v.x) && isfinite(v.y));
    }
    if (__syncthreads_or(bad) && threadIdx.x == 0) atomicOr(&flags[tinfo[t] >> 6], 1);
  }
}

// set the running maxima (slot 0 = given local maxima, others 0) and g
__global__ void k_set_max(unsigned long long *slots, double *gmax, double e, double u,
                          double eg, double ug) {
  const int t = threadIdx.x;
  slots[2 * t] = t == 0 ? (unsigned long long)__double_as_longlong(e) : 0ull;
  slots[2 * t + 1] = t == 0 ? (unsigned long long)__double_as_longlong(u) : 0ull;
  if (t == 0) { gmax[0] = eg; gmax[1] = ug; }
}

// deterministic sum of the per-tile norm partials (single block, fixed order)
__global__ void k_norm_reduce(const double *__restrict__ nb, int ntile,
                              double *__restrict__ out) {
  __shared__ double s[2][256];
  const int t = threadIdx.x;
  double a = 0.0, b = 0.0;
  for (int i = t; i < ntile; i += 256) { a += nb[2 * i]; b += nb[2 * i + 1]; }
  s[0][t] = a; s[1][t] = b;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o) { s[0][t] += s[0][t + o]; s[1][t] += s[1][t + o]; }
    __syncthreads();
  }
  if (t == 0) { out[0] = s[0][0]; out[1] = s[1][0]; }
}

// Nodal values of the Tecplot output (tecplotter.f90:17-60), one thread
// per plot node: for every (cell, ii, jj) occurrence of the node, in
// ascending cell order, the cell's solution polynomial at the node,
// sum_js sum_is (Q(is,js) * h(ii,is)) * h(jj,js) from 0; the occurrences
// summed from 0, divided by their count, converted to rho, u, v, w, p, Bx,
// By, Bz, psi; muAV(1:2) averaged likewise.  Plain IEEE double in the
// reference's order (no contraction to FMA), so the values are bitwise the
// reference program's for the same state.
template <int N>
__global__ void k_plot_nodes(const double *__restrict__ q, const int *__restrict__ c2s, int nnode,
                             const int *__restrict__ xadj, const int *__restrict__ occ, int np1,
                             const double *__restrict__ h, const double *__restrict__ muav,
                             double *__restrict__ plotq, double *__restrict__ plotav) {
#pragma clang fp contract(off)
  constexpr int T = Geo<N>::CPW, N2 = N * N, QT = Geo<N>::QT;
  const int nd = blockIdx.x * blockDim.x + threadIdx.x;
  if (nd >= nnode) return;
  double acc[9];
#pragma unroll
  for (int v = 0; v < 9; ++v) acc[v] = 0.0;
  double a0 = 0.0, a1 = 0.0;
  const int k0 = xadj[nd], k1 = xadj[nd + 1];
  for (int k = k0; k < k1; ++k) {
    const int code = occ[k];
    const int cell = code / (np1 * np1), loc = code % (np1 * np1);
    const int ii = loc % np1, jj = loc / np1;
    const int sl = c2s[cell];
    const double *b = q + (size_t)(sl / T) * QT;
    const int sci = sl % T;
    double qs[9];
#pragma unroll
    for (int v = 0; v < 9; ++v) qs[v] = 0.0;
    for (int js = 0; js < N; ++js)
      for (int is = 0; is < N; ++is) {
        const double hi = h[ii * N + is], hj = h[jj * N + js];
#pragma unroll
        for (int v = 0; v < 9; ++v) qs[v] = qs[v] + (b[qo<N>(v, is + N * js, sci)] * hi) * hj;
      }
#pragma unroll
    for (int v = 0; v < 9; ++v) acc[v] = acc[v] + qs[v];
    if (muav) {
      a0 = a0 + muav[2 * (size_t)cell];
      a1 = a1 + muav[2 * (size_t)cell + 1];
    }
  }
  const double f = (double)(k1 - k0);
  double Qs[9];
#pragma unroll
  for (int v = 0; v < 9; ++v) Qs[v] = acc[v] / f;
  double *o = plotq + 9 * (size_t)nd;
  const double u = Qs[1] / Qs[0], vv = Qs[2] / Qs[0], w = Qs[3] / Qs[0];
  o[0] = Qs[0];
  o[1] = u;
  o[2] = vv;
  o[3] = w;
  o[5] = Qs[5];
  o[6] = Qs[6];
  o[7] = Qs[7];
  o[8] = Qs[8];
  const double nu = u * u + vv * vv + w * w;
  const double nb = Qs[5] * Qs[5] + Qs[6] * Qs[6] + Qs[7] * Qs[7];
  o[4] = (Qs[4] - 0.5 * Qs[0] * nu - 0.5 * nb) * (GAM - 1);
  if (plotav) {
    plotav[2 * (size_t)nd] = a0 / f;
    plotav[2 * (size_t)nd + 1] = a1 / f;
  }
}

// ---------------------------------------------------------------------------
// stage launch
// ---------------------------------------------------------------------------
template <int N, int MODE>
static hipError_t launch_stage_mx_t(sd2d_ctx *c, const StageArgs &a) {
  constexpr size_t lds = stage_mx_lds_bytes<N, MODE>();
  static_assert(lds <= 65536, "stage kernel LDS above 64 KiB");
  hipLaunchKernelGGL((stage_kernel_mx<N, MODE>), dim3(c->ntile), dim3(STAGE_THREADS), lds,
                     c->stream, a, c->opsf, c->metf);
  return hipGetLastError();
}

template <int N, int MODE>
static hipError_t launch_stage_t(sd2d_ctx *c, const StageArgs &a) {
  if constexpr ((MODE & (M_COMPAT | M_VIS | M_DIVB)) == 0) {
    if (c->p.precision == SD2D_MIXED) return launch_stage_mx_t<N, MODE>(c, a);
  }
  constexpr size_t lds = stage_lds_bytes<N, MODE>();
  static_assert(lds <= 65536, "stage kernel LDS above 64 KiB");
  static_assert(2 * N * Geo<N>::CPW <= STAGE_THREADS, "one thread per grid line");
  hipLaunchKernelGGL((stage_kernel<N, MODE>), dim3(c->ntile), dim3(STAGE_THREADS), lds,
                     c->stream, a, c->ops);
  return hipGetLastError();
}

// the stage kernels that the RK sequences and sd2d_compresid launch
template <int N>
static hipError_t launch_stage_n(sd2d_ctx *c, const StageArgs &a, int mode) {
  constexpr int R = M_RK, A1 = M_AUX1, A4 = M_AUX4, V = M_AV, NR = M_NORMS;
  switch (mode) {
#define SD2D_MODE(m) case (m): return launch_stage_t<N, (m)>(c, a);
    SD2D_MODE(R)
    SD2D_MODE(R | A1)
    SD2D_MODE(R | A1 | V)
    SD2D_MODE(R | A1 | NR)
    SD2D_MODE(R | A1 | V | NR)
    SD2D_MODE(R | A1 | M_RSTORE)
    SD2D_MODE(R | A4)
    SD2D_MODE(R | A4 | V)
    SD2D_MODE(R | A4 | NR)
    SD2D_MODE(R | A4 | V | NR)
    SD2D_MODE(M_COMPAT)
    SD2D_MODE(M_COMPAT | V)
    SD2D_MODE(M_DIVB)
    // the same with the viscous divergence (FP64 only)
    SD2D_MODE(M_VIS | R)
    SD2D_MODE(M_VIS | R | A1)
    SD2D_MODE(M_VIS | R | A1 | V)
    SD2D_MODE(M_VIS | R | A1 | NR)
    SD2D_MODE(M_VIS | R | A1 | V | NR)
    SD2D_MODE(M_VIS | R | A1 | M_RSTORE)
    SD2D_MODE(M_VIS | R | A4)
    SD2D_MODE(M_VIS | R | A4 | V)
    SD2D_MODE(M_VIS | R | A4 | NR)
    SD2D_MODE(M_VIS | R | A4 | V | NR)
    SD2D_MODE(M_VIS | M_COMPAT)
    SD2D_MODE(M_VIS | M_COMPAT | V)
#undef SD2D_MODE
    default: return hipErrorInvalidValue;
  }
}

static hipError_t launch_stage_raw(sd2d_ctx *c, const StageArgs &a, int mode) {
  switch (c->N) {
    case 3: return launch_stage_n<3>(c, a, mode);
    case 4: return launch_stage_n<4>(c, a, mode);
    default: return launch_stage_n<5>(c, a, mode);
  }
}

// the viscous divergence of the residual of state qin -> c->visres
template <int N>
static hipError_t launch_vis_n(sd2d_ctx *c, const double *qin) {
  VisArgs va{};
  va.qin = qin; va.ftab = c->ftab; va.met = c->met; va.metw = c->metw;
  va.gq = c->gq; va.dei = c->dei; va.visres = c->visres;
  va.mu = c->p.mu; va.eta0 = c->p.eta0;
  va.ncell = c->ncell; va.ntile = c->ntile; va.walls = c->dei != nullptr; va.s2c = c->s2c;
  constexpr size_t lds = vis_lds_bytes<N>();
  static_assert(lds <= 65536, "viscous kernel LDS above 64 KiB");
  hipLaunchKernelGGL(vis_grad_kernel<N>, dim3(c->ntile), dim3(STAGE_THREADS), lds, c->stream,
                     va, c->ops, c->ph);
  hipError_t r = hipGetLastError();
  if (r != hipSuccess) return r;
  hipLaunchKernelGGL(vis_flux_kernel<N>, dim3(c->ntile), dim3(STAGE_THREADS), lds, c->stream,
                     va, c->ops, c->ph);
  return hipGetLastError();
}

static hipError_t launch_vis(sd2d_ctx *c, const double *qin) {
  switch (c->N) {
    case 3: return launch_vis_n<3>(c, qin);
    case 4: return launch_vis_n<4>(c, qin);
    default: return launch_vis_n<5>(c, qin);
  }
}

static hipError_t launch_stage(sd2d_ctx *c, const StageArgs &a, int mode) {
  if (!c->prof) return launch_stage_raw(c, a, mode);
  while (c->ev.size() < c->nev_used + 2) {
    hipEvent_t e;
    hipError_t r = hipEventCreate(&e);
    if (r != hipSuccess) return r;
    c->ev.push_back(e);
  }
  hipError_t r = hipEventRecord(c->ev[c->nev_used], c->stream);
  if (r != hipSuccess) return r;
  r = launch_stage_raw(c, a, mode);
  if (r != hipSuccess) return r;
  r = hipEventRecord(c->ev[c->nev_used + 1], c->stream);
  c->nev_used += 2;
  return r;
}

// accumulate the recorded stage-kernel times (stream must be idle)
static int collect_profile(sd2d_ctx *c) {
  for (size_t i = 0; i + 1 < c->nev_used; i += 2) {
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev[i], c->ev[i + 1]));
    c->prof_ms += ms;
    c->prof_n += 1;
  }
  c->nev_used = 0;
  return SD2D_OK;
}

static StageArgs base_args(sd2d_ctx *c) {
  StageArgs a{};
  a.met = c->met;
  a.ftab = c->ftab;
  a.gmax = c->gmax;
  a.slots = c->slots;
  std::memcpy(a.avf, c->avf, sizeof(a.avf));
  a.ncell = c->ncell;
  a.ntile = c->ntile;
  a.sm_ind = c->sm_ind;
  a.s2c = c->s2c;
  a.tinfo = c->tinfo;
  a.dtm = c->dtm;
  a.phm = c->phm;
  a.sfam = c->sfam;
  return a;
}

// The opt-in finiteness scan of the state in q[c->cur]: flags (c->nmem
// entries, host) get 1 for every member whose state holds a NaN or an
// infinity.
static int scan_finite(sd2d_ctx *c, int *flags) {
  const int qt2 = 9 * c->N * c->N * c->cpw / 2;
  HIPCHK(c, hipMemsetAsync(c->nfflag, 0, (size_t)c->nmem * sizeof(int), c->stream));
  hipLaunchKernelGGL(k_scan_finite, dim3(std::min(c->ntile, 8192)), dim3(256), 0, c->stream,
                     reinterpret_cast<const double2 *>(c->q[c->cur]), c->tinfo, c->ntile, qt2,
                     c->nfflag);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(flags, c->nfflag, (size_t)c->nmem * sizeof(int), hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SD2D_OK;
}

static int check_finite_now(sd2d_ctx *c) {
  int bad = 0;
  const int r = scan_finite(c, &bad);
  if (r) return r;
  if (bad)
    return fail(c, SD2D_ENONFINITE,
                "non-finite value (NaN or Inf) in the state after step " + std::to_string(c->iter));
  return SD2D_OK;
}

static int finalize_stage(sd2d_ctx *c) {
  hipLaunchKernelGGL(k_finalize_max, dim3(c->nmem), dim3(NSLOT), 0, c->stream, c->slots, c->gmax,
                     c->lastch);
  HIPCHK(c, hipGetLastError());
  return SD2D_OK;
}

// Device allocations and uploads of a context from its host tables.  An
// ensemble group passes its members' per-tile info, time steps and free
// streams (nmem > 1 entries of the maxima); one run passes its per-tile
// valid-cell counts as tinfo (member 0) and no time steps.  On error
// the context keeps what was allocated (sd2d_destroy frees it).
static int ctx_upload(sd2d_ctx *c, const HostTables &tb, int nwall, int nmem,
                      const std::vector<int> *tinfo, const std::vector<double> *dtm,
                      const std::vector<Phys> *phm) {
  const int nt = c->ntile, N = c->N, CPW = c->cpw, N2 = N * N, nc = c->ncell;
  const bool vis = c->vis;
  const sd2d_params *p = &c->p;
  const std::vector<int> &ft = tb.ft;
  const std::vector<double> &met = tb.met;

#define CK(call)                                                                   \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess) {                                                        \
      c->err = std::string(#call) + ": " + hipGetErrorString(e_);                  \
      return e_ == hipErrorOutOfMemory ? SD2D_ENOMEM : SD2D_EHIP;                  \
    }                                                                              \
  } while (0)
  {
    // end faces with the neighbour in another tile, per line family:
    // xi-lines end at local faces 3 and 1, eta-lines at 0 and 2
    const int QT = 9 * N2 * CPW;
    long long cross[2] = {0, 0};
    for (int t = 0; t < nt; ++t)
      for (int k = 0; k < 4; ++k)
        for (int ci = 0; ci < CPW; ++ci) {
          const int nbase = ft[((size_t)t * 8 + 2 * k) * CPW + ci];
          const int info = ft[((size_t)t * 8 + 2 * k + 1) * CPW + ci];
          if (((info >> 3) & 7) == FT_INTERIOR && nbase / QT != t) ++cross[(k == 1 || k == 3) ? 0 : 1];
        }
    c->sfam = cross[0] < cross[1] ? 0 : 1;
  }
  CK(hipGetDevice(&c->device));
  scratch_init(c->device);
  CK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  const size_t qn = (size_t)nt * 9 * N2 * CPW;
  for (int i = 0; i < 3; ++i) CK(hipMalloc(&c->q[i], qn * sizeof(double)));
  if (p->k_stage == 5)
    for (int i = 0; i < 4; ++i) CK(hipMalloc(&c->fsr[i], qn * sizeof(double)));
  CK(hipMalloc(&c->met, met.size() * sizeof(double)));
  std::vector<float> metf;
  if (p->precision == SD2D_MIXED) {
    metf.assign(met.begin(), met.end());
    CK(hipMalloc(&c->metf, metf.size() * sizeof(float)));
  }
  CK(hipMalloc(&c->ftab, ft.size() * sizeof(int)));
  CK(hipMalloc(&c->c2s, tb.c2s.size() * sizeof(int)));
  CK(hipMalloc(&c->s2c, tb.s2c.size() * sizeof(int)));
  CK(hipMalloc(&c->gmax, 2 * (size_t)nmem * sizeof(double)));
  CK(hipMalloc(&c->lastch, (size_t)nmem * sizeof(double)));
  CK(hipMalloc(&c->nfflag, (size_t)nmem * sizeof(int)));
  CK(hipMemsetAsync(c->lastch, 0, (size_t)nmem * sizeof(double), c->stream));
  CK(hipMalloc(&c->slots, 2 * NSLOT * (size_t)nmem * sizeof(unsigned long long)));
  CK(hipMalloc(&c->sm_ind, 2 * (size_t)nc * sizeof(double)));
  CK(hipMalloc(&c->normbuf, 2 * (size_t)nt * sizeof(double)));
  CK(hipMalloc(&c->normout, 2 * sizeof(double)));
  CK(hipMalloc(&c->stagef, (size_t)nc * 9 * N2 * sizeof(double)));
  if (vis) {
    CK(hipMalloc(&c->gq, 2 * qn * sizeof(double)));
    CK(hipMalloc(&c->visres, qn * sizeof(double)));
    CK(hipMalloc(&c->metw, tb.metw.size() * sizeof(double)));
    CK(hipMemcpyAsync(c->metw, tb.metw.data(), tb.metw.size() * sizeof(double),
                      hipMemcpyHostToDevice, c->stream));
    if (nwall > 0)   // comp_grad_internal_energy runs only with walls (compflux.f90:16)
      CK(hipMalloc(&c->dei, (size_t)nt * 2 * N2 * CPW * sizeof(double)));
  }
  CK(hipMemcpyAsync(c->met, met.data(), met.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  if (c->metf)
    CK(hipMemcpyAsync(c->metf, metf.data(), metf.size() * sizeof(float), hipMemcpyHostToDevice,
                      c->stream));
  CK(hipMemcpyAsync(c->ftab, ft.data(), ft.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
  CK(hipMemcpyAsync(c->c2s, tb.c2s.data(), tb.c2s.size() * sizeof(int), hipMemcpyHostToDevice,
                    c->stream));
  CK(hipMemcpyAsync(c->s2c, tb.s2c.data(), tb.s2c.size() * sizeof(int), hipMemcpyHostToDevice,
                    c->stream));
  if (tb.av.size() != (size_t)N2) return fail(c, SD2D_EINVAL, "AV filter of the wrong size");
  std::copy(tb.av.begin(), tb.av.end(), c->avf);
  for (int i = 0; i < 3; ++i) CK(hipMemsetAsync(c->q[i], 0, qn * sizeof(double), c->stream));
  CK(hipMemsetAsync(c->sm_ind, 0, 2 * (size_t)nc * sizeof(double), c->stream));
  if (tinfo) {
    CK(hipMalloc(&c->tinfo, tinfo->size() * sizeof(int)));
    CK(hipMemcpyAsync(c->tinfo, tinfo->data(), tinfo->size() * sizeof(int), hipMemcpyHostToDevice,
                      c->stream));
  }
  if (dtm) {
    CK(hipMalloc(&c->dtm, dtm->size() * sizeof(double)));
    CK(hipMemcpyAsync(c->dtm, dtm->data(), dtm->size() * sizeof(double), hipMemcpyHostToDevice,
                      c->stream));
  }
  {
    const std::vector<Phys> one{tb.ph};
    const std::vector<Phys> &ph = phm ? *phm : one;
    CK(hipMalloc(&c->phm, ph.size() * sizeof(Phys)));
    CK(hipMemcpyAsync(c->phm, ph.data(), ph.size() * sizeof(Phys), hipMemcpyHostToDevice,
                      c->stream));
  }
  c->nmem = nmem;
  for (int k = 0; k < nmem; ++k) {
    hipLaunchKernelGGL(k_set_max, dim3(1), dim3(NSLOT), 0, c->stream, c->slots + 2 * NSLOT * (size_t)k,
                       c->gmax + 2 * k, 0.0, 0.0, 0.0, 0.0);
    CK(hipGetLastError());
  }
  CK(hipStreamSynchronize(c->stream));
#undef CK
  return SD2D_OK;
}

extern "C" {

const char *sd2d_last_error(const sd2d_ctx *ctx) {
  return ctx ? ctx->err.c_str() : g_err.c_str();
}

int sd2d_create(const sd2d_mesh *m, const sd2d_params *p, sd2d_ctx **out) {
  if (!out) return fail(nullptr, SD2D_EINVAL, "out is NULL");
  *out = nullptr;
  if (!m || !p) return fail(nullptr, SD2D_EINVAL, "mesh or params is NULL");
  if (m->N < 3 || m->N > 5)
    return fail(nullptr, SD2D_EUNSUPPORTED, "N must be 3, 4 or 5");
  if (m->ncell <= 0 || m->nface <= 0 || !m->if2c || !m->lmat || !m->mmat || !m->jac ||
      !m->s1 || !m->s2 || !m->xs)
    return fail(nullptr, SD2D_EINVAL, "incomplete mesh");
  if (p->k_stage != 3 && p->k_stage != 5)
    return fail(nullptr, SD2D_EINVAL, "k_stage can be only 3 or 5");
  if (p->vismode != 1)
    return fail(nullptr, SD2D_EUNSUPPORTED, vismode_msg);
  if (p->precision != SD2D_FP64 && p->precision != SD2D_MIXED)
    return fail(nullptr, SD2D_EINVAL, "precision must be SD2D_FP64 or SD2D_MIXED");
  const bool vis = p->vismode == 1 && (p->mu != 0.0 || p->eta0 != 0.0);
  if (vis && p->precision == SD2D_MIXED)
    return fail(nullptr, SD2D_EUNSUPPORTED,
                "the viscous/resistive path (mu, eta0 != 0) is FP64 only");
  sd2d_ctx *c = new (std::nothrow) sd2d_ctx();
  if (!c) return fail(nullptr, SD2D_ENOMEM, "host allocation failed");
  HostTables tb;
  {
    std::string e;
    const int r = build_host_tables(m, p, tb, e);
    if (r) { delete c; return fail(nullptr, r, e); }
  }
  const int N = m->N, nc = m->ncell, CPW = tb.cpw;
  c->N = N; c->ncell = nc; c->cpw = CPW; c->ntile = tb.ntile; c->p = *p;
  c->ctime = 0.0; c->iter = 0;
  c->ops = tb.ops;
  for (int i = 0; i <= NMAX; ++i)
    for (int j = 0; j < NMAX; ++j) {
      c->opsf.L[i][j] = (float)tb.ops.L[i][j];
      c->opsf.M[i][j] = (float)tb.ops.M[i][j];
    }
  c->ph = tb.ph;
  c->vis = vis;
  c->xs.assign(m->xs, m->xs + N);
  {
    // one run: member 0 and the valid cells of every tile
    const std::vector<int> tinfo(tb.tvalid.begin(), tb.tvalid.end());
    const int r = ctx_upload(c, tb, m->nwall, 1, &tinfo, nullptr, nullptr);
    if (r) { g_err = c->err; sd2d_destroy(c); return r; }
  }
  *out = c;
  return SD2D_OK;
}

void sd2d_destroy(sd2d_ctx *c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  adj_free(c->adj);
  for (double *b : c->q) if (b) (void)hipFree(b);
  for (double *b : c->fsr) if (b) (void)hipFree(b);
  void *ptrs[] = {c->met, c->metf, c->ftab, c->gmax, c->lastch, c->nfflag, c->slots, c->sm_ind, c->normbuf,
                  c->normout, c->stagef, c->nablaf, c->gq, c->dei, c->visres, c->metw,
                  c->c2s, c->s2c, c->tinfo, c->dtm, c->phm};
  for (void *b : ptrs) if (b) (void)hipFree(b);
  for (hipEvent_t e : c->ev) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

static int f2t(sd2d_ctx *c, const double *dev_f, double *dev_t) {
  const int blocks = 1024;
  switch (c->N) {
    case 3: hipLaunchKernelGGL(k_f2t<3>, dim3(blocks), dim3(256), 0, c->stream, dev_f, dev_t, c->s2c, c->ntile); break;
    case 4: hipLaunchKernelGGL(k_f2t<4>, dim3(blocks), dim3(256), 0, c->stream, dev_f, dev_t, c->s2c, c->ntile); break;
    default: hipLaunchKernelGGL(k_f2t<5>, dim3(blocks), dim3(256), 0, c->stream, dev_f, dev_t, c->s2c, c->ntile); break;
  }
  HIPCHK(c, hipGetLastError());
  return SD2D_OK;
}

static int t2f(sd2d_ctx *c, const double *dev_t, double *dev_f) {
  const int blocks = 1024;
  switch (c->N) {
    case 3: hipLaunchKernelGGL(k_t2f<3>, dim3(blocks), dim3(256), 0, c->stream, dev_t, dev_f, c->c2s, c->ncell); break;
    case 4: hipLaunchKernelGGL(k_t2f<4>, dim3(blocks), dim3(256), 0, c->stream, dev_t, dev_f, c->c2s, c->ncell); break;
    default: hipLaunchKernelGGL(k_t2f<5>, dim3(blocks), dim3(256), 0, c->stream, dev_t, dev_f, c->c2s, c->ncell); break;
  }
  HIPCHK(c, hipGetLastError());
  return SD2D_OK;
}

int sd2d_set_state(sd2d_ctx *c, const double *q, const sd2d_maxima *mx) {
  if (!c) return fail(nullptr, SD2D_EINVAL, "ctx is NULL");
  const size_t n = (size_t)c->ncell * 9 * c->N * c->N;
  if (q) {
    HIPCHK(c, hipMemcpyAsync(c->stagef, q, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    int r = f2t(c, c->stagef, c->q[c->cur]);
    if (r) return r;
    c->next_stage = 1;   // a step in progress is abandoned
  }
  if (mx) {
    hipLaunchKernelGGL(k_set_max, dim3(1), dim3(NSLOT), 0, c->stream, c->slots, c->gmax,
                       mx->eigvmax, mx->umax, mx->eigvmax_g, mx->umax_g);
    HIPCHK(c, hipGetLastError());
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SD2D_OK;
}

int sd2d_get_state(sd2d_ctx *c, double *q, sd2d_maxima *mx) {
  if (!c) return fail(nullptr, SD2D_EINVAL, "ctx is NULL");
  if (q) SD2D_NOT_MID_STEP(c, "sd2d_get_state");
  const size_t n = (size_t)c->ncell * 9 * c->N * c->N;
  if (q) {
    int r = t2f(c, c->q[c->cur], c->stagef);
    if (r) return r;
    HIPCHK(c, hipMemcpyAsync(q, c->stagef, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  }
  if (mx) {
    unsigned long long s[2 * NSLOT];
    double g[2];
    HIPCHK(c, hipMemcpyAsync(s, c->slots, sizeof(s), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(g, c->gmax, sizeof(g), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    unsigned long long e = 0, u = 0;
    for (int i = 0; i < NSLOT; ++i) { e = std::max(e, s[2 * i]); u = std::max(u, s[2 * i + 1]); }
    long long ei = (long long)e, ui = (long long)u;
    std::memcpy(&mx->eigvmax, &ei, 8);
    std::memcpy(&mx->umax, &ui, 8);
    mx->eigvmax_g = g[0]; mx->umax_g = g[1];
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SD2D_OK;
}

int sd2d_set_time(sd2d_ctx *c, double ctime, int32_t iter) {
  if (!c) return fail(nullptr, SD2D_EINVAL, "ctx is NULL");
  c->ctime = ctime; c->iter = iter;
  return SD2D_OK;
}

int sd2d_compresid(sd2d_ctx *c, double *resid_out, double *nablaqs_out) {
  if (!c) return fail(nullptr, SD2D_EINVAL, "ctx is NULL");
  SD2D_NOT_MID_STEP(c, "sd2d_compresid");
  const size_t n = (size_t)c->ncell * 9 * c->N * c->N;
  if (!c->nablaf) HIPCHK(c, hipMalloc(&c->nablaf, 2 * n * sizeof(double)));
  int r;
  StageArgs a = base_args(c);
  a.qin = c->q[c->cur];
  a.resid_f = c->stagef;
  a.nabla_f = c->nablaf;
  int mode = M_COMPAT | (c->p.artiv == 1 && c->p.vismode == 1 ? M_AV : 0);
  if (c->vis) {
    HIPCHK(c, launch_vis(c, a.qin));
    a.visres = c->visres;
    mode |= M_VIS;
  }
  HIPCHK(c, launch_stage(c, a, mode));
  c->have_av = (mode & M_AV) != 0;
  r = finalize_stage(c);
  if (r) return r;
  if (resid_out)
    HIPCHK(c, hipMemcpyAsync(resid_out, c->stagef, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (nablaqs_out)
    HIPCHK(c, hipMemcpyAsync(nablaqs_out, c->nablaf, 2 * n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SD2D_OK;
}

int sd2d_compresid_q(sd2d_ctx *c, const double *q, double *resid_out, double *nablaqs_out) {
  int r = sd2d_set_state(c, q, nullptr);
  if (r) return r;
  return sd2d_compresid(c, resid_out, nablaqs_out);
}

// sd2d_compresid without any allocation: the same kernels and per-call
// contract, with the Fortran-layout nablaQs staged in `scratch` (through the
// address looked up for this context's device) instead of a buffer the
// library allocates on the first call.  The COMPAT stage kernel writes every
// element of that range (the xi-line threads store, the eta-line threads add
// after a barrier), so nothing from an earlier call is read.
int sd2d_compresid_noalloc(sd2d_ctx *c, double *resid_out, double *nablaqs_out) {
  if (!c) return fail(nullptr, SD2D_EINVAL, "ctx is NULL");
  SD2D_NOT_MID_STEP(c, "sd2d_compresid_noalloc");
  const size_t n = (size_t)c->ncell * 9 * c->N * c->N;
  if (2 * n > (size_t)SD2D_NOALLOC_MAX_DOUBLES)
    return fail(c, SD2D_EINVAL,
                "sd2d_compresid_noalloc: nablaQs needs " + std::to_string(2 * n) +
                    " doubles of scratch; the static buffer holds " +
                    std::to_string((size_t)SD2D_NOALLOC_MAX_DOUBLES) + " (use sd2d_compresid)");
  if (c->device < 0 || c->device >= NOALLOC_MAX_DEVICES || !g_scratch_ptr[c->device])
    return fail(c, SD2D_EHIP,
                std::string("sd2d_compresid_noalloc: no address of the static scratch on this device: ") +
                    (c->device >= 0 && c->device < NOALLOC_MAX_DEVICES
                         ? hipGetErrorString(g_scratch_err[c->device])
                         : "device index out of range"));
  if (c->prof)
    return fail(c, SD2D_EINVAL,
                "sd2d_compresid_noalloc: profiling creates HIP events; switch it off first");
  double *nab = g_scratch_ptr[c->device];
  StageArgs a = base_args(c);
  a.qin = c->q[c->cur];
  a.resid_f = c->stagef;
  a.nabla_f = nab;
  int mode = M_COMPAT | (c->p.artiv == 1 && c->p.vismode == 1 ? M_AV : 0);
  if (c->vis) {
    HIPCHK(c, launch_vis(c, a.qin));
    a.visres = c->visres;
    mode |= M_VIS;
  }
  HIPCHK(c, launch_stage_raw(c, a, mode));
  c->have_av = (mode & M_AV) != 0;
  const int r = finalize_stage(c);
  if (r) return r;
  if (resid_out)
    HIPCHK(c, hipMemcpyAsync(resid_out, c->stagef, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (nablaqs_out)
    HIPCHK(c, hipMemcpyAsync(nablaqs_out, nab, 2 * n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SD2D_OK;
}

// One RK step, or (only > 0) only stage `only` of it: the sequence is walked
// as a whole and the other stages' launches are skipped; the buffers
// persist between calls, so stages run one call at a time give the same
// result.  The rotation and ctime advance with the last stage.
static int step_rk(sd2d_ctx *c, bool norms, int only = 0) {
  RkBuffers rb{{c->q[0], c->q[1], c->q[2]}, {c->fsr[0], c->fsr[1], c->fsr[2]}, c->cur};
  const int av = (c->p.artiv == 1 && c->p.vismode == 1) ? M_AV : 0;
  int idx = 0;
  auto launch = [&](StageArgs &a, int mode) -> int {
    if (only > 0 && ++idx != only) return 0;
    if (mode & M_NORMS) a.normbuf = c->normbuf;
    a.rev = c->nlaunch++ & 1;   // alternate the tile order stage by stage
    if (c->vis) {
      HIPCHK(c, launch_vis(c, a.qin));
      a.visres = c->visres;
      mode |= M_VIS;
    }
    HIPCHK(c, launch_stage(c, a, mode));
    return finalize_stage(c);
  };
  const int r = (c->p.k_stage == 5) ? fsrk_sequence(rb, c->p.dt, norms, av, base_args(c), launch)
                                    : s5p3_sequence(rb, c->p.dt, norms, av, base_args(c), launch);
  if (r) return r;
  if (only > 0 && only < 5) return SD2D_OK;
  c->cur = rb.cur;
  // ctime bookkeeping of RK_S5P3 / FiveStageRk: only the final value
  // (ctime0 + dt) survives a step
  c->ctime = c->ctime + c->p.dt;
  return SD2D_OK;
}

int sd2d_rk_stage(sd2d_ctx *c, int32_t stage) {
  if (!c) return fail(nullptr, SD2D_EINVAL, "ctx is NULL");
  if (stage != c->next_stage)
    return fail(c, SD2D_EINVAL, "stage " + std::to_string(stage) + " out of order: the next stage is " +
                                    std::to_string(c->next_stage));
  if (stage == 1) c->iter += 1;
  const int r = step_rk(c, false, stage);
  if (r) return r;
  c->next_stage = stage == 5 ? 1 : stage + 1;
  if (stage == 5) c->have_av = c->p.artiv == 1 && c->p.vismode == 1;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (c->prof) {
    const int rp = collect_profile(c);
    if (rp) return rp;
  }
  if (stage == 5 && c->check_every > 0 && c->iter % c->check_every == 0) return check_finite_now(c);
  return SD2D_OK;
}

int sd2d_step(sd2d_ctx *c, int32_t nsteps, int32_t norm_every) {
  if (!c) return fail(nullptr, SD2D_EINVAL, "ctx is NULL");
  if (nsteps < 0) return fail(c, SD2D_EINVAL, "nsteps < 0");
  if (c->next_stage != 1)
    return fail(c, SD2D_EINVAL, "a step is in progress (sd2d_rk_stage): finish it first");
  for (int s = 0; s < nsteps; ++s) {
    c->iter += 1;
    const bool norms = norm_every > 0 && (c->iter % norm_every) == 1;
    int r = step_rk(c, norms);
    if (r) return r;
    c->have_av = c->p.artiv == 1 && c->p.vismode == 1;
    if (c->prof && c->nev_used >= 64) {   // bound the event pool
      HIPCHK(c, hipStreamSynchronize(c->stream));
      r = collect_profile(c);
      if (r) return r;
    }
    if (norms) {
      hipLaunchKernelGGL(k_norm_reduce, dim3(1), dim3(256), 0, c->stream, c->normbuf, c->ntile,
                         c->normout);
      HIPCHK(c, hipGetLastError());
      double s2[2];
      HIPCHK(c, hipMemcpyAsync(s2, c->normout, sizeof(s2), hipMemcpyDeviceToHost, c->stream));
      HIPCHK(c, hipStreamSynchronize(c->stream));
      const double nn = (double)c->p.ncellglob * c->N * c->N;
      c->resid_rms = std::sqrt(s2[0] / (nn * 9.0));
      c->divb_rms = std::sqrt(s2[1] / nn);
      c->norms_valid = 1;
    }
    if (c->check_every > 0 && (c->iter % c->check_every == 0 || s == nsteps - 1)) {
      r = check_finite_now(c);
      if (r) return r;
    }
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (c->prof) return collect_profile(c);
  return SD2D_OK;
}

// The reference's restart stream (write_data.f90:44-74): int32 iter,
// float64 ctime, then per cell int32 ICG and float64 Q(9,N,N); q holds
// Q(9,N,N,NCELL) in Fortran layout.
static bool write_restart_file(const char *path, int32_t iter, double ctime, int N, int ncell,
                               const double *q, const int32_t *icg) {
  const size_t rec = (size_t)9 * N * N;
  FILE *fh = std::fopen(path, "wb");
  if (!fh) return false;
  bool ok = std::fwrite(&iter, sizeof(int32_t), 1, fh) == 1 &&
            std::fwrite(&ctime, sizeof(double), 1, fh) == 1;
  for (int ic = 0; ok && ic < ncell; ++ic) {
    const int32_t g = icg ? icg[ic] : ic + 1;
    ok = std::fwrite(&g, sizeof(int32_t), 1, fh) == 1 &&
         std::fwrite(q + rec * ic, sizeof(double), rec, fh) == rec;
  }
  return (std::fclose(fh) == 0) && ok;
}

// READ_ALL_DATA_BINARY (read_data_restart.f90:35-67) for one rank holding
// the whole mesh: every record's cell goes to local cell ICG-1.  Returns an
// empty string or what is wrong.
static std::string read_restart_file(const char *path, int N, int ncell, int32_t &iter,
                                     double &ctime, std::vector<double> &q) {
  const size_t rec = (size_t)9 * N * N;
  FILE *fh = std::fopen(path, "rb");
  if (!fh) return std::string("cannot open ") + path;
  std::fseek(fh, 0, SEEK_END);
  const long size = std::ftell(fh);
  std::fseek(fh, 0, SEEK_SET);
  const long want = (long)(sizeof(int32_t) + sizeof(double) +
                           (sizeof(int32_t) + rec * sizeof(double)) * (size_t)ncell);
  if (size != want) {
    std::fclose(fh);
    return "restart file size does not match NCELL and N";
  }
  q.assign(rec * ncell, 0.0);
  std::vector<char> seen(ncell, 0);
  bool ok = std::fread(&iter, sizeof(int32_t), 1, fh) == 1 &&
            std::fread(&ctime, sizeof(double), 1, fh) == 1;
  for (int i = 0; ok && i < ncell; ++i) {
    int32_t g = 0;
    ok = std::fread(&g, sizeof(int32_t), 1, fh) == 1;
    if (!ok) break;
    if (g < 1 || g > ncell || seen[g - 1]) {
      std::fclose(fh);
      return "restart record with a cell number outside 1..NCELL or repeated";
    }
    seen[g - 1] = 1;
    ok = std::fread(q.data() + rec * (g - 1), sizeof(double), rec, fh) == rec;
  }
  std::fclose(fh);
  if (!ok) return std::string("read failed: ") + path;
  return std::string();
}

// The running-maxima sidecar of a checkpoint, <restart path>.max: 8 bytes
// "SD2DMAX1", int32 iter, float64 ctime (those of the restart file beside
// it), float64 eigvmax, umax, eigvmax_g, umax_g; native byte order.
static std::string max_path(const char *path) { return std::string(path) + ".max"; }

static bool write_max_file(const char *path, int32_t iter, double ctime, const sd2d_maxima &m) {
  FILE *fh = std::fopen(max_path(path).c_str(), "wb");
  if (!fh) return false;
  const double v[4] = {m.eigvmax, m.umax, m.eigvmax_g, m.umax_g};
  bool ok = std::fwrite("SD2DMAX1", 1, 8, fh) == 8 && std::fwrite(&iter, sizeof(int32_t), 1, fh) == 1 &&
            std::fwrite(&ctime, sizeof(double), 1, fh) == 1 && std::fwrite(v, sizeof(double), 4, fh) == 4;
  return (std::fclose(fh) == 0) && ok;
}

static std::string read_max_file(const char *path, int32_t iter, double ctime, sd2d_maxima &m) {
  const std::string mp = max_path(path);
  FILE *fh = std::fopen(mp.c_str(), "rb");
  if (!fh) return "cannot open the maxima sidecar " + mp;
  char magic[8];
  int32_t it = 0;
  double t = 0.0, v[4];
  const bool ok = std::fread(magic, 1, 8, fh) == 8 && std::fread(&it, sizeof(int32_t), 1, fh) == 1 &&
                  std::fread(&t, sizeof(double), 1, fh) == 1 && std::fread(v, sizeof(double), 4, fh) == 4 &&
                  std::fgetc(fh) == EOF;
  std::fclose(fh);
  if (!ok || std::memcmp(magic, "SD2DMAX1", 8) != 0) return "not a maxima sidecar: " + mp;
  if (it != iter || std::memcmp(&t, &ctime, sizeof(double)) != 0)
    return "the maxima sidecar " + mp + " belongs to another step than the restart file";
  m.eigvmax = v[0]; m.umax = v[1]; m.eigvmax_g = v[2]; m.umax_g = v[3];
  return std::string();
}

int sd2d_write_restart(sd2d_ctx *c, const char *path, const int32_t *icg) {
  if (!c || !path) return fail(c, SD2D_EINVAL, "NULL argument");
  SD2D_NOT_MID_STEP(c, "sd2d_write_restart");
  std::vector<double> q((size_t)9 * c->N * c->N * c->ncell);
  int r = sd2d_get_state(c, q.data(), nullptr);
  if (r) return r;
  if (!write_restart_file(path, c->iter, c->ctime, c->N, c->ncell, q.data(), icg))
    return fail(c, SD2D_EINVAL, std::string("write failed: ") + path);
  return SD2D_OK;
}

int sd2d_read_restart(sd2d_ctx *c, const char *path) {
  if (!c || !path) return fail(c, SD2D_EINVAL, "NULL argument");
  int32_t it = 0;
  double t = 0.0;
  std::vector<double> q;
  const std::string e = read_restart_file(path, c->N, c->ncell, it, t, q);
  if (!e.empty()) return fail(c, SD2D_EINVAL, e);
  const sd2d_maxima zero{0.0, 0.0, 0.0, 0.0};
  int r = sd2d_set_state(c, q.data(), &zero);
  if (r) return r;
  return sd2d_set_time(c, t, it);
}

int sd2d_write_checkpoint(sd2d_ctx *c, const char *path, const int32_t *icg) {
  if (!c || !path) return fail(c, SD2D_EINVAL, "NULL argument");
  SD2D_NOT_MID_STEP(c, "sd2d_write_checkpoint");
  std::vector<double> q((size_t)9 * c->N * c->N * c->ncell);
  sd2d_maxima m{};
  int r = sd2d_get_state(c, q.data(), &m);
  if (r) return r;
  if (!write_restart_file(path, c->iter, c->ctime, c->N, c->ncell, q.data(), icg) ||
      !write_max_file(path, c->iter, c->ctime, m))
    return fail(c, SD2D_EINVAL, std::string("write failed: ") + path);
  return SD2D_OK;
}

int sd2d_read_checkpoint(sd2d_ctx *c, const char *path) {
  if (!c || !path) return fail(c, SD2D_EINVAL, "NULL argument");
  int32_t it = 0;
  double t = 0.0;
  std::vector<double> q;
  std::string e = read_restart_file(path, c->N, c->ncell, it, t, q);
  sd2d_maxima m{};
  if (e.empty()) e = read_max_file(path, it, t, m);
  if (!e.empty()) return fail(c, SD2D_EINVAL, e);
  int r = sd2d_set_state(c, q.data(), &m);
  if (r) return r;
  return sd2d_set_time(c, t, it);
}

int sd2d_set_check_finite(sd2d_ctx *c, int32_t every) {
  if (!c) return fail(nullptr, SD2D_EINVAL, "ctx is NULL");
  if (every < 0) return fail(c, SD2D_EINVAL, "every must be >= 0");
  c->check_every = every;
  return SD2D_OK;
}

int sd2d_set_profiling(sd2d_ctx *c, int32_t on) {
  if (!c) return fail(nullptr, SD2D_EINVAL, "ctx is NULL");
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->prof = on != 0;
  c->nev_used = 0;
  c->prof_ms = 0.0;
  c->prof_n = 0;
  return SD2D_OK;
}

int sd2d_get_profile(sd2d_ctx *c, double *stage_ms, int32_t *nlaunch) {
  if (!c) return fail(nullptr, SD2D_EINVAL, "ctx is NULL");
  HIPCHK(c, hipStreamSynchronize(c->stream));
  int r = collect_profile(c);
  if (r) return r;
  if (stage_ms) *stage_ms = c->prof_ms;
  if (nlaunch) *nlaunch = c->prof_n;
  return SD2D_OK;
}

int sd2d_get_diag(sd2d_ctx *c, sd2d_diag *d) {
  if (!c || !d) return fail(c, SD2D_EINVAL, "NULL argument");
  double g[2], ch = 0.0;
  HIPCHK(c, hipMemcpyAsync(g, c->gmax, sizeof(g), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(&ch, c->lastch, sizeof(ch), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  d->resid_rms = c->resid_rms;
  d->divb_rms = c->divb_rms;
  d->eigvmax_g = g[0];
  d->umax_g = g[1];
  d->c_h = ch;
  const double h = std::sqrt(c->p.dxcycl * c->p.dycycl / (double)c->p.ncellglob);
  d->epsilon0 = g[0] * h / (double)(c->N - 1) * 0.05;
  d->ctime = c->ctime;
  d->iter = c->iter;
  d->norms_valid = c->norms_valid;
  return SD2D_OK;
}

int sd2d_plot_nodes(sd2d_ctx *c, int32_t nre, int32_t nnode, const int32_t *xadj,
                    const int32_t *occ, const double *muav, double *plotq, double *plotav) {
  if (!c || !xadj || !occ || !plotq) return fail(c, SD2D_EINVAL, "NULL argument");
  SD2D_NOT_MID_STEP(c, "sd2d_plot_nodes");
  if (nre < 1 || nnode < 1) return fail(c, SD2D_EINVAL, "nre and nnode must be positive");
  if (plotav && !muav) return fail(c, SD2D_EINVAL, "plotav needs muav");
  const long long np1 = nre + 1, per = np1 * np1;
  if ((long long)c->ncell * per > 0x7fffffffLL)
    return fail(c, SD2D_EINVAL, "NCELL*(nre+1)^2 exceeds 32-bit occurrence codes");
  // every index the kernel will use, checked here
  if (xadj[0] != 0) return fail(c, SD2D_EINVAL, "xadj[0] must be 0");
  for (int i = 0; i < nnode; ++i)
    if (xadj[i + 1] <= xadj[i]) return fail(c, SD2D_EINVAL, "xadj must increase strictly");
  const int nocc = xadj[nnode];
  for (int k = 0; k < nocc; ++k)
    if (occ[k] < 0 || (long long)occ[k] >= (long long)c->ncell * per)
      return fail(c, SD2D_EINVAL, "occurrence code outside NCELL*(nre+1)^2");
  // Lagrange values of the solution-point basis at the plot positions
  // (ii)/nre, the product formula of tecplotter.f90:62-80
  const int N = c->N;
  std::vector<double> h((size_t)np1 * N);
  for (int k = 0; k <= nre; ++k) {
    const double x = (double)k / (double)nre;
    for (int i = 0; i < N; ++i) {
      double hn = 1.0, hd = 1.0;
      for (int s2 = 0; s2 < N; ++s2)
        if (s2 != i) {
          hn = hn * (x - c->xs[s2]);
          hd = hd * (c->xs[i] - c->xs[s2]);
        }
      h[(size_t)k * N + i] = hn / hd;
    }
  }
  int *dx = nullptr, *docc = nullptr;
  double *dh = nullptr, *dmu = nullptr, *dq = nullptr, *dav = nullptr;
  auto release = [&]() {
    void *ptrs[] = {dx, docc, dh, dmu, dq, dav};
    for (void *b : ptrs) if (b) (void)hipFree(b);
  };
#define PLOTCHK(call)                                                              \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess) {                                                        \
      release();                                                                   \
      return fail(c, SD2D_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    }                                                                              \
  } while (0)
  PLOTCHK(hipMalloc(&dx, ((size_t)nnode + 1) * sizeof(int)));
  PLOTCHK(hipMalloc(&docc, (size_t)nocc * sizeof(int)));
  PLOTCHK(hipMalloc(&dh, h.size() * sizeof(double)));
  PLOTCHK(hipMalloc(&dq, (size_t)9 * nnode * sizeof(double)));
  if (plotav) {
    PLOTCHK(hipMalloc(&dmu, (size_t)2 * c->ncell * sizeof(double)));
    PLOTCHK(hipMalloc(&dav, (size_t)2 * nnode * sizeof(double)));
    PLOTCHK(hipMemcpyAsync(dmu, muav, (size_t)2 * c->ncell * sizeof(double), hipMemcpyHostToDevice,
                           c->stream));
  }
  PLOTCHK(hipMemcpyAsync(dx, xadj, ((size_t)nnode + 1) * sizeof(int), hipMemcpyHostToDevice, c->stream));
  PLOTCHK(hipMemcpyAsync(docc, occ, (size_t)nocc * sizeof(int), hipMemcpyHostToDevice, c->stream));
  PLOTCHK(hipMemcpyAsync(dh, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  const dim3 grid((nnode + 255) / 256), block(256);
  const double *qd = c->q[c->cur];
  switch (N) {
    case 3: hipLaunchKernelGGL(k_plot_nodes<3>, grid, block, 0, c->stream, qd, c->c2s, nnode, dx, docc,
                               (int)np1, dh, dmu, dq, dav); break;
    case 4: hipLaunchKernelGGL(k_plot_nodes<4>, grid, block, 0, c->stream, qd, c->c2s, nnode, dx, docc,
                               (int)np1, dh, dmu, dq, dav); break;
    default: hipLaunchKernelGGL(k_plot_nodes<5>, grid, block, 0, c->stream, qd, c->c2s, nnode, dx, docc,
                                (int)np1, dh, dmu, dq, dav); break;
  }
  PLOTCHK(hipGetLastError());
  PLOTCHK(hipMemcpyAsync(plotq, dq, (size_t)9 * nnode * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (plotav)
    PLOTCHK(hipMemcpyAsync(plotav, dav, (size_t)2 * nnode * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
  PLOTCHK(hipStreamSynchronize(c->stream));
#undef PLOTCHK
  release();
  return SD2D_OK;
}

int sd2d_get_av(sd2d_ctx *c, double *sm_ind, double *muav, double *epsilon0) {
  if (!c) return fail(nullptr, SD2D_EINVAL, "ctx is NULL");
  if (c->p.artiv != 1 || c->p.vismode != 1)
    return fail(c, SD2D_EINVAL, "the AV sensor runs only with vismode=1 and ARTIV=1");
  std::vector<double> sm((size_t)2 * c->ncell);
  double g[2];
  HIPCHK(c, hipMemcpyAsync(sm.data(), c->sm_ind, sm.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(g, c->gmax, sizeof(g), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  // AV.f90:106-140
  const double h = std::sqrt(c->p.dxcycl * c->p.dycycl / (double)c->p.ncellglob);
  const double eps = g[0] * h / (double)(c->N - 1) * 0.05;
  if (epsilon0) *epsilon0 = eps;
  if (sm_ind) std::memcpy(sm_ind, sm.data(), sm.size() * sizeof(double));
  if (muav) {
    const double lo[2] = {-5.0, -3.0}, hi[2] = {0.0, 1.0};
    const double pi = 3.141592653589793;
    for (int ic = 0; ic < c->ncell; ++ic)
      for (int v = 0; v < 2; ++v) {
        const double S = sm[2 * (size_t)ic + v];
        const double k = (hi[v] - lo[v]) * 0.5, s0 = (hi[v] + lo[v]) * 0.5;
        double mu;
        if (S < lo[v]) mu = 0.0;
        else if (S > hi[v]) mu = eps;
        else mu = eps / 2.0 * (1.0 + std::sin(pi * (S - s0) / 2.0 / k));
        muav[2 * (size_t)ic + v] = mu;
      }
  }
  return SD2D_OK;
}

}  // extern "C"

#include "sd2d_ens.hpp"
#include "sd2d_adjhost.hpp"
