// stage_emu.cpp - the device stage kernels (sd2d_stage.hpp,
// sd2d_stage_mx.hpp), compiled from their own source with g++ against the
// CPU stand-ins of tests/emu/hip/hip_runtime.h, launched on host threads
// with the grids, LDS sizes and arguments the library uses, on buffers of
// exactly the sizes the library allocates.  Built with AddressSanitizer and
// UBSan by tests/test_emu_stage.py, so that an access outside any buffer,
// a barrier not every work-item reaches, an LDS write beyond the launch's
// dynamic size or a read of memory nothing wrote (every buffer and the LDS
// start as NaN) fails the test.  Test infrastructure only.
//
//   stage_emu IN OUT
//
// IN (little-endian, written by tests/test_emu_stage.py):
//   int32[16]   N, ncell, nface, ninter, ninlet, noutlet, nwall, nsymp, nfree,
//               npair, k_stage, artiv, job, nsteps, precision, tile_shape
//   float64[16] mu, eta0, alpha, rinf, pinf, uinf, vinf, bxinf, byinf, dxcycl,
//               dycycl, dt, eigvmax, umax, eigvmax_g, umax_g
//   int32       IF2C(NFACE,4), IBFINTER, IBFINL, IBFOUT, IBFWAL, IBFSYMP, IBFREE,
//               CYC2FPAIR(NPAIR,2)
//   float64     Lmat, Mmat, Xs, Xs_low, Jac, S1, S2, Q(9,N,N,NCELL)
// job 0: two sd2d_compresid evaluations (the second with the lagged c_h);
//        OUT: resid, nablaQs, gmax(2), SM_IND(2,NCELL) of each.
// job 1: nsteps RK steps (k_stage); OUT: Q(9,N,N,NCELL), gmax(2).
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "sd2d_host.hpp"
#include "sd2d_rkseq.hpp"
#include "sd2d_stage_mx.hpp"

// ---------------------------------------------------------------------------
// the launch model of hip_runtime.h
// ---------------------------------------------------------------------------
namespace sd2d {
// dynamic LDS of the running workgroup: 64 KiB, then a canary region
constexpr size_t LDS_DOUBLES = 65536 / 8, CANARY_DOUBLES = 1024;
double lds[LDS_DOUBLES + CANARY_DOUBLES];
}  // namespace sd2d

namespace emu {
thread_local emu_dim3 tid3{0, 0, 0}, bid3{0, 0, 0};
constexpr int THREADS = 128;
constexpr uint64_t POISON = 0x7FC07FC07FC07FC0ull;   // NaN as one double and as two floats

struct Group {
  int live = 0, arrived = 0;
  unsigned long gen = 0;
};
struct BlockState {
  std::mutex m;
  std::condition_variable cv;
  Group blk, wave[THREADS / 64];
  uint64_t xch[THREADS];
  bool exited[THREADS];
};
static BlockState *g_bs = nullptr;
static std::mutex g_atomic;

[[noreturn]] static void die(const std::string &msg) {
  std::fprintf(stderr, "emu: %s (workgroup %u, work-item %u)\n", msg.c_str(), bid3.x, tid3.x);
  std::fflush(stderr);
  std::abort();
}

static void release(Group &g) {
  g.arrived = 0;
  ++g.gen;
  g_bs->cv.notify_all();
}

static void wait_group(Group &g, std::unique_lock<std::mutex> &lk, const char *what) {
  const unsigned long my = g.gen;
  if (++g.arrived == g.live) {
    release(g);
    return;
  }
  if (!g_bs->cv.wait_for(lk, std::chrono::seconds(60), [&] { return g.gen != my; }))
    die(std::string(what) + " never completes: not every live work-item reaches it");
}

void syncthreads() {
  std::unique_lock<std::mutex> lk(g_bs->m);
  wait_group(g_bs->blk, lk, "__syncthreads");
}

void wave_sync() {
  std::unique_lock<std::mutex> lk(g_bs->m);
  wait_group(g_bs->wave[(int)tid3.x >> 6], lk, "a wave-level LDS sync");
}

uint64_t shfl_xor_bits(uint64_t b, int mask) {
  const int t = (int)tid3.x, w = t >> 6, l = t & 63;
  if (mask < 0 || mask > 63) die("__shfl_xor lane mask outside the wavefront");
  std::unique_lock<std::mutex> lk(g_bs->m);
  g_bs->xch[t] = b;
  wait_group(g_bs->wave[w], lk, "__shfl_xor");
  const int p = (w << 6) | (l ^ mask);
  if (g_bs->exited[p]) die("__shfl_xor reads the lane of a work-item that has returned");
  const uint64_t r = g_bs->xch[p];
  wait_group(g_bs->wave[w], lk, "__shfl_xor");
  return r;
}

unsigned long long atomic_max(unsigned long long *p, unsigned long long v) {
  std::lock_guard<std::mutex> lk(g_atomic);
  const unsigned long long old = *p;
  if (v > old) *p = v;
  return old;
}

static void thread_exit() {
  std::unique_lock<std::mutex> lk(g_bs->m);
  const int t = (int)tid3.x;
  g_bs->exited[t] = true;
  for (Group *g : {&g_bs->blk, &g_bs->wave[t >> 6]}) {
    --g->live;
    if (g->arrived > 0 && g->arrived == g->live) release(*g);
  }
}

// one launch: nblocks workgroups of THREADS work-items, lds_bytes of LDS
template <class F>
static void launch(int nblocks, size_t lds_bytes, F &&body) {
  if (lds_bytes > sd2d::LDS_DOUBLES * 8) die("launch asks for more than 64 KiB of LDS");
  for (int b = 0; b < nblocks; ++b) {
    BlockState bs;
    bs.blk.live = THREADS;
    for (auto &w : bs.wave) w.live = 64;
    for (int t = 0; t < THREADS; ++t) bs.exited[t] = false;
    g_bs = &bs;
    for (size_t i = 0; i < sd2d::LDS_DOUBLES + sd2d::CANARY_DOUBLES; ++i)
      std::memcpy(&sd2d::lds[i], &POISON, 8);
    std::vector<std::thread> th;
    th.reserve(THREADS);
    for (int t = 0; t < THREADS; ++t)
      th.emplace_back([&, t, b] {
        tid3 = emu_dim3{(unsigned)t, 0, 0};
        bid3 = emu_dim3{(unsigned)b, 0, 0};
        body();
        thread_exit();
      });
    for (auto &x : th) x.join();
    // nothing may have been written past the launch's dynamic LDS size
    const unsigned char *c = reinterpret_cast<const unsigned char *>(sd2d::lds);
    for (size_t i = lds_bytes; i < (sd2d::LDS_DOUBLES + sd2d::CANARY_DOUBLES) * 8; ++i)
      if (c[i] != reinterpret_cast<const unsigned char *>(&POISON)[i % 8]) {
        bid3.x = (unsigned)b;
        die("LDS written at byte " + std::to_string(i) + " beyond the launch's " +
            std::to_string(lds_bytes) + " bytes");
      }
    g_bs = nullptr;
  }
}
}  // namespace emu

// ---------------------------------------------------------------------------
// the library's buffers and launches, on the host
// ---------------------------------------------------------------------------
namespace {

using namespace sd2d;

template <class T>
std::vector<T> dev_buffer(size_t n) {   // uninitialised device memory: NaN
  std::vector<T> v(n);
  for (auto &x : v) std::memcpy(&x, &emu::POISON, sizeof(T) <= 8 ? sizeof(T) : 8);
  return v;
}

struct Ctx {
  int N = 0, nc = 0, nt = 0, cpw = 0, precision = 0;
  HostTables tb;
  std::vector<double> q[3], fsr[3], met, gmax, sm_ind, normbuf, resid_f, nabla_f;
  std::vector<int> ftab, tinfo, s2c;
  std::vector<unsigned long long> slots;
  std::vector<float> metf;
  std::vector<Phys> phm;
  OpsF opsf{};
  int sfam = 1, cur = 0;
  long nlaunch = 0;
};

template <int N, int MODE>
void run_fp64(Ctx &c, const StageArgs &a) {
  emu::launch(c.nt, stage_lds_bytes<N, MODE>(), [&] { stage_kernel<N, MODE>(a, c.tb.ops); });
}
template <int N, int MODE>
void run_mx(Ctx &c, const StageArgs &a) {
  emu::launch(c.nt, stage_mx_lds_bytes<N, MODE>(),
              [&] { stage_kernel_mx<N, MODE>(a, c.opsf, c.metf.data()); });
}

template <int N>
int run_stage_n(Ctx &c, const StageArgs &a, int mode) {
  constexpr int R = M_RK, A1 = M_AUX1, A4 = M_AUX4, V = M_AV, NR = M_NORMS;
  const bool mx = c.precision == SD2D_MIXED;
  switch (mode) {
    case M_COMPAT | V: run_fp64<N, M_COMPAT | V>(c, a); return 0;
    case M_COMPAT: run_fp64<N, M_COMPAT>(c, a); return 0;
    case R: mx ? run_mx<N, R>(c, a) : run_fp64<N, R>(c, a); return 0;
    case R | A1: mx ? run_mx<N, R | A1>(c, a) : run_fp64<N, R | A1>(c, a); return 0;
    case R | A1 | V | NR:
      mx ? run_mx<N, R | A1 | V | NR>(c, a) : run_fp64<N, R | A1 | V | NR>(c, a);
      return 0;
    case R | A1 | M_RSTORE:
      if (mx) break;
      run_fp64<N, R | A1 | M_RSTORE>(c, a);
      return 0;
    case R | A4 | V | NR:
      if (mx) break;
      run_fp64<N, R | A4 | V | NR>(c, a);
      return 0;
    default: break;
  }
  std::fprintf(stderr, "emu: stage mode %d (precision %d) is not built into the emulator\n", mode,
               c.precision);
  return 1;
}

int run_stage(Ctx &c, const StageArgs &a, int mode) {
  switch (c.N) {
    case 3: return run_stage_n<3>(c, a, mode);
    case 4: return run_stage_n<4>(c, a, mode);
    default: return run_stage_n<5>(c, a, mode);
  }
}

// k_finalize_max: the slots' maxima become the global maxima of the next
// evaluation
void finalize(Ctx &c) {
  unsigned long long e = 0, u = 0;
  for (int t = 0; t < NSLOT; ++t) {
    e = c.slots[2 * t] > e ? c.slots[2 * t] : e;
    u = c.slots[2 * t + 1] > u ? c.slots[2 * t + 1] : u;
  }
  c.gmax[0] = __longlong_as_double((long long)e);
  c.gmax[1] = __longlong_as_double((long long)u);
}

StageArgs base_args(Ctx &c) {
  StageArgs a{};
  a.met = c.met.data();
  a.ftab = c.ftab.data();
  a.gmax = c.gmax.data();
  a.slots = c.slots.data();
  if (c.tb.av.size() != (size_t)c.N * c.N) emu::die("AV filter of the wrong size");
  std::copy(c.tb.av.begin(), c.tb.av.end(), a.avf);
  a.ncell = c.nc;
  a.ntile = c.nt;
  a.sm_ind = c.sm_ind.data();
  a.s2c = c.s2c.data();
  a.tinfo = c.tinfo.data();
  a.dtm = nullptr;
  a.phm = c.phm.data();
  a.sfam = c.sfam;
  return a;
}

template <class T>
void rd(std::ifstream &f, std::vector<T> &v, size_t n) {
  v.resize(n);
  if (n) f.read(reinterpret_cast<char *>(v.data()), (std::streamsize)(n * sizeof(T)));
  if (!f) {
    std::fprintf(stderr, "emu: input file too short\n");
    std::exit(2);
  }
}

}  // namespace

int main(int argc, char **argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: stage_emu IN OUT\n");
    return 2;
  }
  std::ifstream f(argv[1], std::ios::binary);
  std::vector<int32_t> h;
  std::vector<double> pr;
  rd(f, h, 16);
  rd(f, pr, 16);
  const int N = h[0], nc = h[1], nface = h[2];
  std::vector<int32_t> if2c, lists[6], cyc;
  rd(f, if2c, (size_t)4 * nface);
  for (int i = 0; i < 6; ++i) rd(f, lists[i], (size_t)h[3 + i]);
  rd(f, cyc, (size_t)2 * h[9]);
  std::vector<double> lmat, mmat, xs, xsl, jac, s1, s2, Q;
  rd(f, lmat, (size_t)(N + 1) * N);
  rd(f, mmat, (size_t)(N + 1) * N);
  rd(f, xs, N);
  rd(f, xsl, N - 1);
  rd(f, jac, (size_t)N * N * nc);
  rd(f, s1, (size_t)4 * (N + 1) * N * nc);
  rd(f, s2, (size_t)4 * N * (N + 1) * nc);
  rd(f, Q, (size_t)9 * N * N * nc);

  sd2d_mesh m{};
  m.N = N; m.ncell = nc; m.nface = nface; m.if2c = if2c.data();
  m.ninter = h[3]; m.ibfinter = lists[0].data();
  m.ninlet = h[4]; m.ibfinl = lists[1].data();
  m.noutlet = h[5]; m.ibfout = lists[2].data();
  m.nwall = h[6]; m.ibfwal = lists[3].data();
  m.nsymp = h[7]; m.ibfsymp = lists[4].data();
  m.nfree = h[8]; m.ibfree = lists[5].data();
  m.npair = h[9]; m.cyc2fpair = cyc.data();
  m.lmat = lmat.data(); m.mmat = mmat.data(); m.xs = xs.data(); m.xs_low = xsl.data();
  m.jac = jac.data(); m.s1 = s1.data(); m.s2 = s2.data();
  sd2d_params p{};
  p.vismode = 1; p.artiv = h[11]; p.mu = pr[0]; p.eta0 = pr[1]; p.alpha = pr[2];
  p.rinf = pr[3]; p.pinf = pr[4]; p.uinf = pr[5]; p.vinf = pr[6]; p.bxinf = pr[7];
  p.byinf = pr[8]; p.dxcycl = pr[9]; p.dycycl = pr[10]; p.ncellglob = nc; p.k_stage = h[10];
  p.dt = pr[11]; p.precision = h[14]; p.tile_shape = h[15];
  if (p.mu != 0.0 || p.eta0 != 0.0) {
    std::fprintf(stderr, "emu: the viscous kernels are not part of the emulator\n");
    return 2;
  }

  Ctx c;
  std::string err;
  if (build_host_tables(&m, &p, c.tb, err)) {
    std::fprintf(stderr, "emu: host tables: %s\n", err.c_str());
    return 3;
  }
  c.N = N; c.nc = nc; c.nt = c.tb.ntile; c.cpw = c.tb.cpw; c.precision = p.precision;
  const int N2 = N * N, CPW = c.cpw, nt = c.nt;
  const size_t QT = (size_t)9 * N2 * CPW, qn = (size_t)nt * QT, n = (size_t)nc * 9 * N2;
  for (int i = 0; i < 3; ++i) c.q[i] = dev_buffer<double>(qn);
  if (p.k_stage == 5)
    for (int i = 0; i < 3; ++i) c.fsr[i] = dev_buffer<double>(qn);
  c.met = c.tb.met;
  c.ftab = c.tb.ft;
  c.tinfo = c.tb.tvalid;   // one run: member 0, the tile's valid cells
  c.s2c = c.tb.s2c;
  c.gmax = dev_buffer<double>(2);
  c.slots = dev_buffer<unsigned long long>(2 * (size_t)NSLOT);
  c.sm_ind = dev_buffer<double>(2 * (size_t)nc);
  c.normbuf = dev_buffer<double>(2 * (size_t)nt);
  c.resid_f = dev_buffer<double>(n);
  c.nabla_f = dev_buffer<double>(2 * n);
  c.phm.assign(1, c.tb.ph);
  c.metf.assign(c.met.begin(), c.met.end());
  for (int i = 0; i <= N; ++i)
    for (int j = 0; j < N; ++j) {
      c.opsf.L[i][j] = (float)c.tb.ops.L[i][j];
      c.opsf.M[i][j] = (float)c.tb.ops.M[i][j];
    }
  {  // the staging family: the one with fewer end faces in other tiles
    long long cross[2] = {0, 0};
    for (int t = 0; t < nt; ++t)
      for (int k = 0; k < 4; ++k)
        for (int ci = 0; ci < CPW; ++ci) {
          const int nbase = c.ftab[((size_t)t * 8 + 2 * k) * CPW + ci];
          const int info = c.ftab[((size_t)t * 8 + 2 * k + 1) * CPW + ci];
          if (((info >> 3) & 7) == FT_INTERIOR && (size_t)nbase / QT != (size_t)t)
            ++cross[(k == 1 || k == 3) ? 0 : 1];
        }
    c.sfam = cross[0] < cross[1] ? 0 : 1;
  }
  // sd2d_set_state: Fortran Q(9,N,N,NCELL) -> tile layout; the maxima
  for (int cell = 0; cell < nc; ++cell) {
    const int sl = c.tb.c2s[cell], t = sl / CPW, ci = sl % CPW;
    for (int pp = 0; pp < N2; ++pp)
      for (int v = 0; v < 9; ++v)
        c.q[0][(size_t)t * QT + qo_rt(N, v, pp, ci)] = Q[v + 9 * (pp + (size_t)N2 * cell)];
  }
  for (int t = 0; t < NSLOT; ++t) c.slots[2 * t] = c.slots[2 * t + 1] = 0ull;
  c.slots[0] = (unsigned long long)__double_as_longlong(pr[12]);
  c.slots[1] = (unsigned long long)__double_as_longlong(pr[13]);
  c.gmax[0] = pr[14];
  c.gmax[1] = pr[15];

  std::ofstream o(argv[2], std::ios::binary);
  auto put = [&](const double *d, size_t k) {
    o.write(reinterpret_cast<const char *>(d), (std::streamsize)(k * sizeof(double)));
  };
  const int job = h[12];
  if (job == 0) {
    for (int call = 0; call < 2; ++call) {
      StageArgs a = base_args(c);
      a.qin = c.q[c.cur].data();
      a.resid_f = c.resid_f.data();
      a.nabla_f = c.nabla_f.data();
      if (run_stage(c, a, M_COMPAT | (p.artiv == 1 ? M_AV : 0))) return 4;
      finalize(c);
      put(c.resid_f.data(), n);
      put(c.nabla_f.data(), 2 * n);
      put(c.gmax.data(), 2);
      put(c.sm_ind.data(), 2 * (size_t)nc);
    }
  } else {
    const int av = p.artiv == 1 ? M_AV : 0;
    auto launch = [&](StageArgs &a, int mode) -> int {
      if (mode & M_NORMS) a.normbuf = c.normbuf.data();
      a.rev = (int)(c.nlaunch++ & 1);
      const int r = run_stage(c, a, mode);
      if (r) return r;
      finalize(c);
      return 0;
    };
    for (int s = 0; s < h[13]; ++s) {
      RkBuffers rb{{c.q[0].data(), c.q[1].data(), c.q[2].data()},
                   {c.fsr[0].data(), c.fsr[1].data(), c.fsr[2].data()}, c.cur};
      const int r = p.k_stage == 5 ? fsrk_sequence(rb, p.dt, true, av, base_args(c), launch)
                                   : s5p3_sequence(rb, p.dt, true, av, base_args(c), launch);
      if (r) return 4;
      c.cur = rb.cur;
    }
    std::vector<double> Qo(n);
    for (int cell = 0; cell < nc; ++cell) {
      const int sl = c.tb.c2s[cell], t = sl / CPW, ci = sl % CPW;
      for (int pp = 0; pp < N2; ++pp)
        for (int v = 0; v < 9; ++v)
          Qo[v + 9 * (pp + (size_t)N2 * cell)] = c.q[c.cur][(size_t)t * QT + qo_rt(N, v, pp, ci)];
    }
    put(Qo.data(), n);
    put(c.gmax.data(), 2);
  }
  o.close();
  return o ? 0 : 5;
}
