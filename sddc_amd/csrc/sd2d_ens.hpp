// sd2d_ens.hpp - ensembles of independent runs stepped together
// (sd2d_ens_* in include/sd2d.h).  Included at the end of sd2d.hip: it
// builds on the context set-up and the stage launches defined there.
//
// Members are grouped by order N.  The tiles of all members of a group are
// concatenated into one tile array (states, metrics, face tables with the
// neighbour offsets shifted by the member's first tile), so one launch of
// the stage kernel per RK stage advances every member of the group; each
// tile carries its member's index, through which the kernel finds the
// member's GLM maxima, time step and free stream.  The groups run on their
// own streams, forked from and joined to the ensemble's stream, and one
// whole RK_S5P3 step of all groups is captured once in a HIP graph (one per
// position of the three-buffer rotation) and replayed.

struct EnsMember {
  int group = 0, local = 0;       // group (order) and index in the group
  int tile0 = 0, cell0 = 0;       // first tile / cell in the group arrays
  int ntile = 0, ncell = 0;
  int *c2s = nullptr, *s2c = nullptr;   // member-local cell <-> slot (device)
  double ctime = 0.0;
  int iter = 0;
};

struct sd2d_ens {
  int nmem = 0;
  std::vector<sd2d_mesh> meshes;      // the caller's descriptions (arrays not copied)
  std::vector<sd2d_params> params;
  std::vector<EnsMember> mem;
  std::vector<sd2d_ctx *> grp;        // one context per order present
  std::vector<int> grp_n;             // N of each group
  std::vector<int> grp_shape;         // tile shape of each group
  hipStream_t stream = nullptr;
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> join;
  hipGraphExec_t graph[3] = {nullptr, nullptr, nullptr};   // by rotation position
  int cur = 0;                        // state buffer holding every member's state
  int check_every = 0;                // opt-in non-finite check
  std::vector<int> nonfinite;         // per member, from the latest check
  std::string err;
};

static int ens_fail(sd2d_ens *e, int code, const std::string &m) {
  if (e) e->err = m; else g_err = m;
  return code;
}

#define ENSCHK(e, call)                                                         \
  do {                                                                          \
    hipError_t e_ = (call);                                                     \
    if (e_ != hipSuccess) {                                                     \
      (e)->err = std::string(#call) + ": " + hipGetErrorString(e_);             \
      return e_ == hipErrorOutOfMemory ? SD2D_ENOMEM : SD2D_EHIP;               \
    }                                                                           \
  } while (0)

static void ens_drop_graphs(sd2d_ens *e) {
  for (auto &g : e->graph)
    if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
}

static void ens_free_group(sd2d_ens *e, int g) {
  for (auto &m : e->mem)
    if (m.group == g) {
      if (m.c2s) (void)hipFree(m.c2s);
      if (m.s2c) (void)hipFree(m.s2c);
      m.c2s = m.s2c = nullptr;
    }
  if (e->grp[g]) sd2d_destroy(e->grp[g]);
  e->grp[g] = nullptr;
}

// (Re)build group g with tile shape `shape`: host tables of every member,
// concatenated, uploaded into a fresh context.  Member states are not kept.
static int ens_build_group(sd2d_ens *e, int g, int shape) {
  ens_free_group(e, g);
  ens_drop_graphs(e);
  const int N = e->grp_n[g], CPW = cpw_of(N), N2 = N * N, QT = 9 * N2 * CPW;
  HostTables gt;
  gt.N = N; gt.cpw = CPW;
  std::vector<int> tinfo;
  std::vector<double> dtm;
  std::vector<Phys> phm;
  std::vector<std::vector<int>> mc2s, ms2c;
  int nloc = 0, nwall = 0;
  long long tiles = 0, cells = 0;
  for (int i = 0; i < e->nmem; ++i) {
    EnsMember &m = e->mem[i];
    if (m.group != g) continue;
    sd2d_params p = e->params[i];
    p.tile_shape = shape;
    HostTables tb;
    std::string err;
    const int r = build_host_tables(&e->meshes[i], &p, tb, err);
    if (r) return ens_fail(e, r, "member " + std::to_string(i) + ": " + err);
    if (nloc == 0) { gt.ops = tb.ops; gt.ph = tb.ph; gt.av = tb.av; }
    else if (std::memcmp(&gt.ops, &tb.ops, sizeof(Ops)) != 0)
      return ens_fail(e, SD2D_EINVAL, "members of one order differ in Lmat/Mmat");
    m.local = nloc++;
    m.tile0 = (int)tiles; m.cell0 = (int)cells;
    m.ntile = tb.ntile; m.ncell = tb.ncell;
    if ((tiles + tb.ntile) * (long long)QT > 0x7fffffffLL)
      return ens_fail(e, SD2D_EUNSUPPORTED, "ensemble group too large for 32-bit offsets");
    if (m.local >= (1 << 25))
      return ens_fail(e, SD2D_EUNSUPPORTED, "too many members in one group");
    // face table: neighbour offsets shift by the member's first tile
    for (size_t k = 0; k < tb.ft.size(); ++k) {
      const size_t slot_k = (k / CPW) % 8;   // [tile][k][{nbase, info}][ci]
      gt.ft.push_back((slot_k % 2 == 0) ? tb.ft[k] + (int)(tiles * QT) : tb.ft[k]);
    }
    gt.met.insert(gt.met.end(), tb.met.begin(), tb.met.end());
    for (int t = 0; t < tb.ntile; ++t) {
      const int nv = tb.tvalid[t];
      tinfo.push_back((m.local << 6) | nv);
    }
    for (int sl = 0; sl < tb.ntile * CPW; ++sl)
      gt.s2c.push_back(tb.s2c[sl] < 0 ? -1 : tb.s2c[sl] + (int)cells);
    for (int c = 0; c < tb.ncell; ++c) gt.c2s.push_back(tb.c2s[c] + (int)(tiles * CPW));
    dtm.push_back(e->params[i].dt);
    phm.push_back(tb.ph);
    mc2s.push_back(tb.c2s);
    ms2c.push_back(tb.s2c);
    nwall += e->meshes[i].nwall;
    tiles += tb.ntile; cells += tb.ncell;
  }
  gt.ntile = (int)tiles; gt.ncell = (int)cells;
  sd2d_ctx *c = new (std::nothrow) sd2d_ctx();
  if (!c) return ens_fail(e, SD2D_ENOMEM, "host allocation failed");
  e->grp[g] = c;
  int first = 0;
  while (e->mem[first].group != g) ++first;
  const sd2d_params &p0 = e->params[first];
  c->N = N; c->ncell = gt.ncell; c->cpw = CPW; c->ntile = gt.ntile; c->p = p0;
  c->p.dt = 1.0;   // the kernels scale by each member's dt (cr * dtm[member])
  c->ops = gt.ops;
  for (int i = 0; i <= NMAX; ++i)
    for (int j = 0; j < NMAX; ++j) {
      c->opsf.L[i][j] = (float)gt.ops.L[i][j];
      c->opsf.M[i][j] = (float)gt.ops.M[i][j];
    }
  c->ph = gt.ph;
  int r = ctx_upload(c, gt, nwall, nloc, &tinfo, &dtm, &phm);
  if (r) { e->err = c->err; return r; }
  // member-local maps for the state conversions
  int k = 0;
  for (auto &m : e->mem) {
    if (m.group != g) continue;
    ENSCHK(e, hipMalloc(&m.c2s, mc2s[k].size() * sizeof(int)));
    ENSCHK(e, hipMalloc(&m.s2c, ms2c[k].size() * sizeof(int)));
    ENSCHK(e, hipMemcpy(m.c2s, mc2s[k].data(), mc2s[k].size() * sizeof(int), hipMemcpyHostToDevice));
    ENSCHK(e, hipMemcpy(m.s2c, ms2c[k].data(), ms2c[k].size() * sizeof(int), hipMemcpyHostToDevice));
    ++k;
  }
  c->cur = e->cur;
  e->grp_shape[g] = shape;
  return SD2D_OK;
}

// One member's state <-> Fortran layout through the group context.
static int ens_member_state(sd2d_ens *e, int i, const double *qin, double *qout,
                            const sd2d_maxima *mxin, sd2d_maxima *mxout) {
  EnsMember &m = e->mem[i];
  sd2d_ctx *c = e->grp[m.group];
  const int N = c->N, N2 = N * N, CPW = c->cpw;
  const size_t n = (size_t)m.ncell * 9 * N2, QT = (size_t)9 * N2 * CPW;
  double *tq = c->q[e->cur] + (size_t)m.tile0 * QT;
  const int blocks = 1024;
  if (qin) {
    ENSCHK(e, hipMemcpyAsync(c->stagef, qin, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    switch (N) {
      case 3: hipLaunchKernelGGL(k_f2t<3>, dim3(blocks), dim3(256), 0, c->stream, c->stagef, tq, m.s2c, m.ntile); break;
      case 4: hipLaunchKernelGGL(k_f2t<4>, dim3(blocks), dim3(256), 0, c->stream, c->stagef, tq, m.s2c, m.ntile); break;
      default: hipLaunchKernelGGL(k_f2t<5>, dim3(blocks), dim3(256), 0, c->stream, c->stagef, tq, m.s2c, m.ntile); break;
    }
    ENSCHK(e, hipGetLastError());
  }
  if (mxin) {
    hipLaunchKernelGGL(k_set_max, dim3(1), dim3(NSLOT), 0, c->stream,
                       c->slots + 2 * NSLOT * (size_t)m.local, c->gmax + 2 * m.local,
                       mxin->eigvmax, mxin->umax, mxin->eigvmax_g, mxin->umax_g);
    ENSCHK(e, hipGetLastError());
  }
  if (qout) {
    switch (N) {
      case 3: hipLaunchKernelGGL(k_t2f<3>, dim3(blocks), dim3(256), 0, c->stream, tq, c->stagef, m.c2s, m.ncell); break;
      case 4: hipLaunchKernelGGL(k_t2f<4>, dim3(blocks), dim3(256), 0, c->stream, tq, c->stagef, m.c2s, m.ncell); break;
      default: hipLaunchKernelGGL(k_t2f<5>, dim3(blocks), dim3(256), 0, c->stream, tq, c->stagef, m.c2s, m.ncell); break;
    }
    ENSCHK(e, hipGetLastError());
    ENSCHK(e, hipMemcpyAsync(qout, c->stagef, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  }
  if (mxout) {
    std::vector<unsigned long long> s(2 * NSLOT);
    double gm[2];
    ENSCHK(e, hipMemcpyAsync(s.data(), c->slots + 2 * NSLOT * (size_t)m.local,
                             s.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream));
    ENSCHK(e, hipMemcpyAsync(gm, c->gmax + 2 * m.local, sizeof(gm), hipMemcpyDeviceToHost, c->stream));
    ENSCHK(e, hipStreamSynchronize(c->stream));
    unsigned long long a = 0, b = 0;
    for (int k = 0; k < NSLOT; ++k) { a = std::max(a, s[2 * k]); b = std::max(b, s[2 * k + 1]); }
    long long ai = (long long)a, bi = (long long)b;
    std::memcpy(&mxout->eigvmax, &ai, 8);
    std::memcpy(&mxout->umax, &bi, 8);
    mxout->eigvmax_g = gm[0]; mxout->umax_g = gm[1];
  }
  ENSCHK(e, hipStreamSynchronize(c->stream));
  return SD2D_OK;
}

// The stages of one RK_S5P3 step of group g, on the group's stream.
static int ens_group_step(sd2d_ens *e, int g) {
  sd2d_ctx *c = e->grp[g];
  c->cur = e->cur;
  const bool prof = c->prof;
  c->prof = false;
  const int r = step_rk(c, false);
  c->prof = prof;
  return r;
}

// Capture one step of every group (fork/join over the group streams).
static int ens_capture(sd2d_ens *e) {
  hipGraph_t gr = nullptr;
  ENSCHK(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
  int r = SD2D_OK;
  if (hipEventRecord(e->fork, e->stream) != hipSuccess) r = SD2D_EHIP;
  for (size_t g = 0; !r && g < e->grp.size(); ++g) {
    if (hipStreamWaitEvent(e->grp[g]->stream, e->fork, 0) != hipSuccess) { r = SD2D_EHIP; break; }
    r = ens_group_step(e, (int)g);
    if (r) { e->err = e->grp[g]->err; break; }
    if (hipEventRecord(e->join[g], e->grp[g]->stream) != hipSuccess ||
        hipStreamWaitEvent(e->stream, e->join[g], 0) != hipSuccess)
      r = SD2D_EHIP;
  }
  const hipError_t ec = hipStreamEndCapture(e->stream, &gr);
  if (r) { if (gr) (void)hipGraphDestroy(gr); return r; }
  ENSCHK(e, ec);
  const hipError_t ei = hipGraphInstantiate(&e->graph[e->cur], gr, nullptr, nullptr, 0);
  (void)hipGraphDestroy(gr);
  ENSCHK(e, ei);
  return SD2D_OK;
}

extern "C" {

const char *sd2d_ens_last_error(const sd2d_ens *e) { return e ? e->err.c_str() : g_err.c_str(); }

void sd2d_ens_destroy(sd2d_ens *e) {
  if (!e) return;
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  ens_drop_graphs(e);
  for (size_t g = 0; g < e->grp.size(); ++g) ens_free_group(e, (int)g);
  for (hipEvent_t ev : e->join) if (ev) (void)hipEventDestroy(ev);
  if (e->fork) (void)hipEventDestroy(e->fork);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
}

int sd2d_ens_create(int32_t nmem, const sd2d_mesh *meshes, const sd2d_params *params,
                    sd2d_ens **out) {
  if (!out) return ens_fail(nullptr, SD2D_EINVAL, "out is NULL");
  *out = nullptr;
  if (nmem <= 0 || !meshes || !params)
    return ens_fail(nullptr, SD2D_EINVAL, "need nmem > 0 meshes and params");
  for (int i = 0; i < nmem; ++i) {
    const sd2d_mesh &m = meshes[i];
    const sd2d_params &p = params[i];
    const std::string w = "member " + std::to_string(i) + ": ";
    if (m.N < 3 || m.N > 5) return ens_fail(nullptr, SD2D_EUNSUPPORTED, w + "N must be 3, 4 or 5");
    if (m.ncell <= 0 || m.nface <= 0 || !m.if2c || !m.lmat || !m.mmat || !m.jac || !m.s1 ||
        !m.s2 || !m.xs)
      return ens_fail(nullptr, SD2D_EINVAL, w + "incomplete mesh");
    if (p.k_stage != 3)
      return ens_fail(nullptr, SD2D_EUNSUPPORTED, w + "ensembles step with RK_S5P3 (k_stage 3)");
    if (p.vismode != 1) return ens_fail(nullptr, SD2D_EUNSUPPORTED, w + vismode_msg);
    if (p.vismode == 1 && (p.mu != 0.0 || p.eta0 != 0.0))
      return ens_fail(nullptr, SD2D_EUNSUPPORTED, w + "ensembles run the inviscid path");
    if (p.precision != params[0].precision || p.artiv != params[0].artiv ||
        p.vismode != params[0].vismode)
      return ens_fail(nullptr, SD2D_EINVAL,
                      w + "precision, ARTIV and vismode must be the same for all members");
    if (p.precision != SD2D_FP64 && p.precision != SD2D_MIXED)
      return ens_fail(nullptr, SD2D_EINVAL, w + "precision must be SD2D_FP64 or SD2D_MIXED");
    if (p.tile_shape < SD2D_TILE_DEFAULT || p.tile_shape > SD2D_TILE_ROWPAIRS)
      return ens_fail(nullptr, SD2D_EINVAL, w + "bad tile_shape");
  }
  sd2d_ens *e = new (std::nothrow) sd2d_ens();
  if (!e) return ens_fail(nullptr, SD2D_ENOMEM, "host allocation failed");
  e->nmem = nmem;
  e->meshes.assign(meshes, meshes + nmem);
  e->params.assign(params, params + nmem);
  e->mem.resize(nmem);
  for (int n = 3; n <= 5; ++n)
    for (int i = 0; i < nmem; ++i)
      if (meshes[i].N == n) { e->grp_n.push_back(n); break; }
  e->grp.assign(e->grp_n.size(), nullptr);
  e->grp_shape.assign(e->grp_n.size(), SD2D_TILE_ROWPAIRS);
  for (int i = 0; i < nmem; ++i)
    for (size_t g = 0; g < e->grp_n.size(); ++g)
      if (e->grp_n[g] == meshes[i].N) e->mem[i].group = (int)g;
  auto bail = [&](int code) { g_err = e->err; sd2d_ens_destroy(e); return code; };
  if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&e->fork, hipEventDisableTiming) != hipSuccess) {
    e->err = "stream/event creation failed";
    return bail(SD2D_EHIP);
  }
  e->join.assign(e->grp_n.size(), nullptr);
  for (auto &ev : e->join)
    if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
      e->err = "event creation failed";
      return bail(SD2D_EHIP);
    }
  for (size_t g = 0; g < e->grp_n.size(); ++g) {
    int shape = SD2D_TILE_ROWPAIRS;   // the group's shape: the first member's request
    for (int i = 0; i < nmem; ++i)
      if (e->mem[i].group == (int)g) {
        shape = params[i].tile_shape == SD2D_TILE_DEFAULT ? SD2D_TILE_ROWPAIRS : params[i].tile_shape;
        break;
      }
    const int r = ens_build_group(e, (int)g, shape);
    if (r) return bail(r);
  }
  *out = e;
  return SD2D_OK;
}

int sd2d_ens_set_state(sd2d_ens *e, int32_t i, const double *q, const sd2d_maxima *mx) {
  if (!e) return ens_fail(nullptr, SD2D_EINVAL, "ens is NULL");
  if (i < 0 || i >= e->nmem) return ens_fail(e, SD2D_EINVAL, "member index out of range");
  return ens_member_state(e, i, q, nullptr, mx, nullptr);
}

int sd2d_ens_get_state(sd2d_ens *e, int32_t i, double *q, sd2d_maxima *mx) {
  if (!e) return ens_fail(nullptr, SD2D_EINVAL, "ens is NULL");
  if (i < 0 || i >= e->nmem) return ens_fail(e, SD2D_EINVAL, "member index out of range");
  return ens_member_state(e, i, nullptr, q, nullptr, mx);
}

int sd2d_ens_set_time(sd2d_ens *e, int32_t i, double ctime, int32_t iter) {
  if (!e) return ens_fail(nullptr, SD2D_EINVAL, "ens is NULL");
  if (i < 0 || i >= e->nmem) return ens_fail(e, SD2D_EINVAL, "member index out of range");
  e->mem[i].ctime = ctime;
  e->mem[i].iter = iter;
  return SD2D_OK;
}

int sd2d_ens_get_time(sd2d_ens *e, int32_t i, double *ctime, int32_t *iter) {
  if (!e) return ens_fail(nullptr, SD2D_EINVAL, "ens is NULL");
  if (i < 0 || i >= e->nmem) return ens_fail(e, SD2D_EINVAL, "member index out of range");
  if (ctime) *ctime = e->mem[i].ctime;
  if (iter) *iter = e->mem[i].iter;
  return SD2D_OK;
}

// A member's checkpoint: its state in the reference's restart format (with
// its iter and ctime) and the running-maxima sidecar, as sd2d_write_checkpoint.
int sd2d_ens_write_checkpoint(sd2d_ens *e, int32_t i, const char *path, const int32_t *icg) {
  if (!e || !path) return ens_fail(e, SD2D_EINVAL, "NULL argument");
  if (i < 0 || i >= e->nmem) return ens_fail(e, SD2D_EINVAL, "member index out of range");
  const EnsMember &m = e->mem[i];
  const int N = e->grp_n[m.group];
  std::vector<double> q((size_t)9 * N * N * m.ncell);
  sd2d_maxima mx{};
  const int r = ens_member_state(e, i, nullptr, q.data(), nullptr, &mx);
  if (r) return r;
  if (!write_restart_file(path, m.iter, m.ctime, N, m.ncell, q.data(), icg) ||
      !write_max_file(path, m.iter, m.ctime, mx))
    return ens_fail(e, SD2D_EINVAL, std::string("write failed: ") + path);
  return SD2D_OK;
}

// Loads a checkpoint into member i: state, running maxima, iter and ctime.
int sd2d_ens_read_checkpoint(sd2d_ens *e, int32_t i, const char *path) {
  if (!e || !path) return ens_fail(e, SD2D_EINVAL, "NULL argument");
  if (i < 0 || i >= e->nmem) return ens_fail(e, SD2D_EINVAL, "member index out of range");
  EnsMember &m = e->mem[i];
  const int N = e->grp_n[m.group];
  int32_t it = 0;
  double t = 0.0;
  std::vector<double> q;
  std::string err = read_restart_file(path, N, m.ncell, it, t, q);
  sd2d_maxima mx{};
  if (err.empty()) err = read_max_file(path, it, t, mx);
  if (!err.empty()) return ens_fail(e, SD2D_EINVAL, err);
  const int r = ens_member_state(e, i, q.data(), nullptr, &mx, nullptr);
  if (r) return r;
  m.iter = it;
  m.ctime = t;
  return SD2D_OK;
}

// The opt-in finiteness scan of every member's state (after a step):
// e->nonfinite[i] = 1 for the members holding a NaN or an infinity.
static int ens_check_finite(sd2d_ens *e, int steps_done) {
  bool any = false;
  e->nonfinite.assign(e->nmem, 0);
  for (size_t g = 0; g < e->grp.size(); ++g) {
    sd2d_ctx *c = e->grp[g];
    c->cur = e->cur;
    std::vector<int> fl(c->nmem, 0);
    const int r = scan_finite(c, fl.data());
    if (r) { e->err = c->err; return r; }
    for (int i = 0; i < e->nmem; ++i)
      if (e->mem[i].group == (int)g && fl[e->mem[i].local]) { e->nonfinite[i] = 1; any = true; }
  }
  if (!any) return SD2D_OK;
  std::string who;
  for (int i = 0; i < e->nmem; ++i)
    if (e->nonfinite[i]) who += (who.empty() ? "" : ", ") + std::to_string(i);
  return ens_fail(e, SD2D_ENONFINITE,
                  "non-finite value (NaN or Inf) in the state of member(s) " + who + " after " +
                      std::to_string(steps_done) + " step(s) of this call");
}

int sd2d_ens_step(sd2d_ens *e, int32_t nsteps) {
  if (!e) return ens_fail(nullptr, SD2D_EINVAL, "ens is NULL");
  if (nsteps < 0) return ens_fail(e, SD2D_EINVAL, "nsteps < 0");
  for (int s = 0; s < nsteps; ++s) {
    if (!e->graph[e->cur]) {
      const int r = ens_capture(e);
      if (r) return r;
    }
    ENSCHK(e, hipGraphLaunch(e->graph[e->cur], e->stream));
    e->cur = (e->cur + 2) % 3;   // RK_S5P3's rotation (sd2d_rkseq.hpp)
    for (auto &m : e->mem) {
      m.ctime = m.ctime + e->params[&m - e->mem.data()].dt;
      m.iter += 1;
    }
    if (e->check_every > 0 && ((s + 1) % e->check_every == 0 || s == nsteps - 1)) {
      for (sd2d_ctx *c : e->grp) c->cur = e->cur;
      const int r = ens_check_finite(e, s + 1);
      if (r) return r;
    }
  }
  for (sd2d_ctx *c : e->grp) c->cur = e->cur;
  ENSCHK(e, hipStreamSynchronize(e->stream));
  return SD2D_OK;
}

int sd2d_ens_set_check_finite(sd2d_ens *e, int32_t every) {
  if (!e) return ens_fail(nullptr, SD2D_EINVAL, "ens is NULL");
  if (every < 0) return ens_fail(e, SD2D_EINVAL, "every must be >= 0");
  e->check_every = every;
  return SD2D_OK;
}

int sd2d_ens_nonfinite(sd2d_ens *e, int32_t *flags) {
  if (!e || !flags) return ens_fail(e, SD2D_EINVAL, "NULL argument");
  for (int i = 0; i < e->nmem; ++i)
    flags[i] = i < (int)e->nonfinite.size() ? e->nonfinite[i] : 0;
  return SD2D_OK;
}

int sd2d_ens_tune(sd2d_ens *e, int32_t trial_steps, int32_t *shapes) {
  if (!e) return ens_fail(nullptr, SD2D_EINVAL, "ens is NULL");
  if (trial_steps <= 0) return ens_fail(e, SD2D_EINVAL, "trial_steps must be positive");
  // every member's state and maxima, restored after each candidate
  std::vector<std::vector<double>> q(e->nmem);
  std::vector<sd2d_maxima> mx(e->nmem);
  auto save = [&](int g) -> int {
    for (int i = 0; i < e->nmem; ++i) {
      if (e->mem[i].group != g) continue;
      q[i].resize((size_t)e->mem[i].ncell * 9 * e->grp_n[g] * e->grp_n[g]);
      const int r = ens_member_state(e, i, nullptr, q[i].data(), nullptr, &mx[i]);
      if (r) return r;
    }
    return SD2D_OK;
  };
  auto restore = [&](int g) -> int {
    for (int i = 0; i < e->nmem; ++i) {
      if (e->mem[i].group != g) continue;
      const int r = ens_member_state(e, i, q[i].data(), nullptr, &mx[i], nullptr);
      if (r) return r;
    }
    return SD2D_OK;
  };
  for (size_t g = 0; g < e->grp.size(); ++g) {
    int r = save((int)g);
    if (r) return r;
    const int cand[3] = {SD2D_TILE_ROWPAIRS, SD2D_TILE_STRIPS, SD2D_TILE_BLOCKS};
    double best = 1e300;
    int best_shape = SD2D_TILE_ROWPAIRS;
    for (int shape : cand) {
      r = ens_build_group(e, (int)g, shape);
      if (r) return r;
      r = restore((int)g);
      if (r) return r;
      sd2d_ctx *c = e->grp[g];
      hipEvent_t t0, t1;
      ENSCHK(e, hipEventCreate(&t0));
      ENSCHK(e, hipEventCreate(&t1));
      r = ens_group_step(e, (int)g);   // warm-up
      if (!r) {
        (void)hipEventRecord(t0, c->stream);
        for (int s = 0; s < trial_steps && !r; ++s) r = step_rk(c, false);
        (void)hipEventRecord(t1, c->stream);
      }
      float ms = 0.f;
      if (!r && hipEventSynchronize(t1) == hipSuccess) (void)hipEventElapsedTime(&ms, t0, t1);
      (void)hipEventDestroy(t0);
      (void)hipEventDestroy(t1);
      c->cur = e->cur;
      if (r) { e->err = c->err; return r; }
      if (ms < best) { best = ms; best_shape = shape; }
    }
    r = ens_build_group(e, (int)g, best_shape);
    if (r) return r;
    r = restore((int)g);
    if (r) return r;
  }
  if (shapes)
    for (size_t g = 0; g < e->grp.size(); ++g) shapes[e->grp_n[g] - 3] = e->grp_shape[g];
  return SD2D_OK;
}

}  // extern "C"
