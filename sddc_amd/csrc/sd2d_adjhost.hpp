// sd2d_adjhost.hpp - the adjoint entry points (sd2d_fwd, sd2d_vjp,
// sd2d_compresid_vjp in include/sd2d.h).  Included at the end of sd2d.hip.
//
// sd2d_fwd steps the context like sd2d_step (RK_S5P3, FP64, inviscid path),
// accumulates the time-averaged wall drag J = (1/nsteps) sum_n D(Q_n) and
// keeps a checkpoint (state and running maxima) every `ckpt` steps.
// sd2d_vjp runs backwards over the checkpoint segments: a segment's step
// inputs are recomputed from its checkpoint, then each step's five stage
// inputs and the GLM maxima before each stage are recomputed with the very
// kernels of the forward run (bitwise the same states), and the step is
// reversed stage by stage:
//   Q1 = Q0 + b10 dt R(Q0)            Q0bar += Q1bar + (dR/dQ0)^T b10 dt Q1bar
//   Q2 = Q1 + b21 dt R(Q1)            ...
//   Q3 = a30 Q0 + a32 Q2 + b32 dt R(Q2)
//   Q4 = a40 Q0 + a43 Q3 + b43 dt R(Q3)
//   Q5 = a52 Q2 + a54 Q4 + b54 dt R(Q4)   (iterations.f90:158-213)
// with the drag's cotangent added to each step's output.  The lagged GLM
// speed c_h = sqrt(E(E - U)) of an evaluation comes from the running maxima
// E, U of the wave speeds over all earlier evaluations: its cotangent is
// carried backwards (pools PE, PU) to the last earlier evaluation that
// raised the maximum, and there to the flux point whose speed attained it
// (adj_argmax_kernel / adj_argmax_apply).

struct AdjState {
  int nsteps = 0, ckpt = 0;
  bool have_fwd = false;
  double J = 0.0;
  double *q[6] = {};          // stage states Q0..Q5 of one step (tile layout)
  double *b[6] = {};          // cotangents
  double *g = nullptr;        // Dbar / wbar per point
  double *divb = nullptr;     // divB per point of the evaluation being reversed
  double *part = nullptr;     // per tile partial sums
  double *gm = nullptr;       // gmax before each of the 5 stages (device)
  unsigned long long *slots = nullptr;   // scratch maxima of recomputed evaluations
  double *dcell = nullptr;    // drag per slot
  double *wq = nullptr;       // face quadrature weights
  std::vector<double *> ck;   // checkpoint states (index = step / ckpt) + final state
  std::vector<std::array<double, 4>> ckm;   // their maxima
  std::vector<int32_t> cki;   // their iter and ctime
  std::vector<double> ckt;
  std::vector<double *> seg;  // step inputs of one segment
  std::vector<std::array<double, 4>> segm;
  double fin_ctime = 0.0;
  int fin_iter = 0;
  unsigned long long *amax = nullptr;   // argmax search: maxima and locations
  int *aloc = nullptr;
  double PE = 0.0, PU = 0.0;            // pending cotangents of the running maxima
  int argmax_missed = 0;                // maxima whose flux point was not found
};

static void adj_free(AdjState *A) {
  if (!A) return;
  for (double *p : A->q) if (p) (void)hipFree(p);
  for (double *p : A->b) if (p) (void)hipFree(p);
  for (void *p : {(void *)A->g, (void *)A->divb, (void *)A->part, (void *)A->gm, (void *)A->slots,
                  (void *)A->dcell, (void *)A->wq, (void *)A->amax, (void *)A->aloc})
    if (p) (void)hipFree(p);
  for (double *p : A->ck) if (p) (void)hipFree(p);
  for (double *p : A->seg) if (p) (void)hipFree(p);
  delete A;
}

static size_t state_len(const sd2d_ctx *c) { return (size_t)c->ntile * 9 * c->N * c->N * c->cpw; }

__global__ void k_axpby(size_t n, double a, const double *__restrict__ x, double b,
                        double *__restrict__ y) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    y[i] = (b == 0.0 ? 0.0 : b * y[i]) + a * x[i];
}

__global__ void k_copy2(const double *__restrict__ s, double *__restrict__ d) {
  if (threadIdx.x < 2) d[threadIdx.x] = s[threadIdx.x];
}

static int axpby(sd2d_ctx *c, double a, const double *x, double b, double *y) {
  hipLaunchKernelGGL(k_axpby, dim3(2048), dim3(256), 0, c->stream, state_len(c), a, x, b, y);
  HIPCHK(c, hipGetLastError());
  return SD2D_OK;
}

static int adj_maxima_get(sd2d_ctx *c, std::array<double, 4> &m) {
  sd2d_maxima mx;
  const int r = sd2d_get_state(c, nullptr, &mx);
  m = {mx.eigvmax, mx.umax, mx.eigvmax_g, mx.umax_g};
  return r;
}

static int adj_maxima_set(sd2d_ctx *c, const std::array<double, 4> &m) {
  hipLaunchKernelGGL(k_set_max, dim3(1), dim3(NSLOT), 0, c->stream, c->slots, c->gmax, m[0], m[1],
                     m[2], m[3]);
  HIPCHK(c, hipGetLastError());
  return SD2D_OK;
}

// interpolatory quadrature weights on the solution-point positions xs of
// [0, 1]: w_i = integral of the Lagrange basis polynomial l_i
static std::vector<double> face_weights(const double *xs, int N) {
  std::vector<double> w(N, 0.0);
  for (int i = 0; i < N; ++i) {
    std::vector<double> c(1, 1.0);   // coefficients of prod_{j != i} (x - xj)/(xi - xj)
    for (int j = 0; j < N; ++j) {
      if (j == i) continue;
      std::vector<double> n(c.size() + 1, 0.0);
      for (size_t k = 0; k < c.size(); ++k) {
        n[k + 1] += c[k] / (xs[i] - xs[j]);
        n[k] -= c[k] * xs[j] / (xs[i] - xs[j]);
      }
      c = n;
    }
    for (size_t k = 0; k < c.size(); ++k) w[i] += c[k] / double(k + 1);
  }
  return w;
}

static int adj_alloc(sd2d_ctx *c, AdjState *A, int nck) {
  const size_t n = state_len(c);
  const int N = c->N, N2 = N * N, T = c->cpw;
  for (auto &p : A->q) if (!p) HIPCHK(c, hipMalloc(&p, n * sizeof(double)));
  for (auto &p : A->b) if (!p) HIPCHK(c, hipMalloc(&p, n * sizeof(double)));
  if (!A->g) HIPCHK(c, hipMalloc(&A->g, (size_t)c->ntile * 10 * N2 * T * sizeof(double)));
  if (!A->part) HIPCHK(c, hipMalloc(&A->part, (size_t)c->ntile * ADJ_NPART * sizeof(double)));
  if (!A->gm) HIPCHK(c, hipMalloc(&A->gm, 12 * sizeof(double)));
  if (!A->amax) HIPCHK(c, hipMalloc(&A->amax, 2 * sizeof(unsigned long long)));
  if (!A->aloc) HIPCHK(c, hipMalloc(&A->aloc, 2 * sizeof(int)));
  if (!A->slots) HIPCHK(c, hipMalloc(&A->slots, 2 * NSLOT * sizeof(unsigned long long)));
  if (!A->dcell) HIPCHK(c, hipMalloc(&A->dcell, (size_t)c->ntile * T * sizeof(double)));
  if (!A->divb) HIPCHK(c, hipMalloc(&A->divb, (size_t)c->ntile * N2 * T * sizeof(double)));
  while ((int)A->ck.size() < nck) {
    double *p = nullptr;
    HIPCHK(c, hipMalloc(&p, n * sizeof(double)));
    A->ck.push_back(p);
  }
  A->ckm.resize(A->ck.size());
  while ((int)A->seg.size() < A->ckpt) {
    double *p = nullptr;
    HIPCHK(c, hipMalloc(&p, n * sizeof(double)));
    A->seg.push_back(p);
  }
  A->segm.resize(A->seg.size());
  return SD2D_OK;
}

// time-step quadrature of the wall drag at state q (and, with qbar, its
// cotangent scale * dD/dq added to qbar)
static int drag_eval(sd2d_ctx *c, const double *q, double *D, double *qbar, double scale) {
  AdjState *A = c->adj;
  const int nsl = c->ntile * c->cpw, tb = 256, nb = (nsl + tb - 1) / tb;
  switch (c->N) {
    case 3: hipLaunchKernelGGL(drag_kernel<3>, dim3(nb), dim3(tb), 0, c->stream, q, c->ftab, c->met, A->wq, c->ops, c->ncell, c->ntile, A->dcell, qbar, scale); break;
    case 4: hipLaunchKernelGGL(drag_kernel<4>, dim3(nb), dim3(tb), 0, c->stream, q, c->ftab, c->met, A->wq, c->ops, c->ncell, c->ntile, A->dcell, qbar, scale); break;
    default: hipLaunchKernelGGL(drag_kernel<5>, dim3(nb), dim3(tb), 0, c->stream, q, c->ftab, c->met, A->wq, c->ops, c->ncell, c->ntile, A->dcell, qbar, scale); break;
  }
  HIPCHK(c, hipGetLastError());
  if (D) {
    std::vector<double> d(nsl);
    HIPCHK(c, hipMemcpyAsync(d.data(), A->dcell, nsl * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double s = 0.0;
    for (double x : d) s += x;
    *D = s;
  }
  return SD2D_OK;
}

// the RK_S5P3 stages as s5p3_sequence runs them, on the six stage buffers:
// qin, aux, coefficients and the mode of the forward launch
struct AdjStage { int in, aux, out; double cin, ca, b; };
static const AdjStage kS5P3[5] = {
    {0, -1, 1, 1.0, 0.0, 0.377268915331368},
    {1, -1, 2, 1.0, 0.0, 0.377268915331368},
    {2, 0, 3, 0.644090224936674, 0.355909775063327, 0.242995220537396},
    {3, 0, 4, 0.632066208361863, 0.367933791638137, 0.238458932846290},
    {4, 2, 5, 0.762406163401431, 0.237593836598569, 0.287632146308408}};

// Recompute the five stage inputs of one step from q[0] and the maxima m,
// recording gmax before every stage in A->gm (the forward's launches).
static int adj_recompute_step(sd2d_ctx *c, const std::array<double, 4> &m) {
  AdjState *A = c->adj;
  int r = adj_maxima_set(c, m);
  if (r) return r;
  const int av = (c->p.artiv == 1 && c->p.vismode == 1) ? M_AV : 0;
  for (int s = 0; s < 5; ++s) {
    const AdjStage &st = kS5P3[s];
    hipLaunchKernelGGL(k_copy2, dim3(1), dim3(64), 0, c->stream, c->gmax, A->gm + 2 * s);
    HIPCHK(c, hipGetLastError());
    StageArgs a = base_args(c);
    a.qin = A->q[st.in];
    a.qout = A->q[st.out];
    a.cin = st.cin;
    a.cr = st.b * c->p.dt;
    a.naux = st.aux >= 0 ? 1 : 0;
    a.aux[0] = st.aux >= 0 ? A->q[st.aux] : nullptr;
    a.ca[0] = st.ca;
    const int mode = M_RK | (a.naux ? M_AUX1 : 0) | (s == 4 ? av : 0);
    HIPCHK(c, launch_stage(c, a, mode));
    r = finalize_stage(c);
    if (r) return r;
  }
  hipLaunchKernelGGL(k_copy2, dim3(1), dim3(64), 0, c->stream, c->gmax, A->gm + 10);
  HIPCHK(c, hipGetLastError());
  return SD2D_OK;
}

static AdjArgs adj_args(sd2d_ctx *c, const double *q, double *qbar) {
  AdjState *A = c->adj;
  AdjArgs a{};
  a.q = q; a.rscale = 1.0; a.divb = A->divb; a.met = c->met; a.ftab = c->ftab;
  a.s2c = c->s2c; a.phm = c->phm; a.g = A->g; a.qbar = qbar; a.part = A->part;
  a.alpha = c->ph.alpha; a.ncell = c->ncell; a.ntile = c->ntile;
  return a;
}

// The flux point where the evaluation of state q attained its largest eig
// (kind 0) or um (kind 1), and qbar += w * d(that speed)/dq.
static int adj_argmax_push(sd2d_ctx *c, const double *q, double *qbar, int kind, double w) {
  AdjState *A = c->adj;
  AdjArgs a = adj_args(c, q, qbar);
  HIPCHK(c, hipMemsetAsync(A->amax, 0, 2 * sizeof(unsigned long long), c->stream));
  HIPCHK(c, hipMemsetAsync(A->aloc, 0x7f, 2 * sizeof(int), c->stream));
  for (int pass = 0; pass < 2; ++pass) {
    switch (c->N) {
      case 3: hipLaunchKernelGGL(adj_argmax_kernel<3>, dim3(c->ntile), dim3(STAGE_THREADS), 0, c->stream, a, c->ops, A->amax, A->aloc, pass); break;
      case 4: hipLaunchKernelGGL(adj_argmax_kernel<4>, dim3(c->ntile), dim3(STAGE_THREADS), 0, c->stream, a, c->ops, A->amax, A->aloc, pass); break;
      default: hipLaunchKernelGGL(adj_argmax_kernel<5>, dim3(c->ntile), dim3(STAGE_THREADS), 0, c->stream, a, c->ops, A->amax, A->aloc, pass); break;
    }
    HIPCHK(c, hipGetLastError());
  }
  int loc[2];
  HIPCHK(c, hipMemcpyAsync(loc, A->aloc, sizeof(loc), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (loc[kind] == 0x7f7f7f7f) { A->argmax_missed += 1; return SD2D_OK; }
  switch (c->N) {
    case 3: hipLaunchKernelGGL(adj_argmax_apply<3>, dim3(1), dim3(64), 0, c->stream, a, c->ops, loc[kind], kind, w); break;
    case 4: hipLaunchKernelGGL(adj_argmax_apply<4>, dim3(1), dim3(64), 0, c->stream, a, c->ops, loc[kind], kind, w); break;
    default: hipLaunchKernelGGL(adj_argmax_apply<5>, dim3(1), dim3(64), 0, c->stream, a, c->ops, loc[kind], kind, w); break;
  }
  HIPCHK(c, hipGetLastError());
  return SD2D_OK;
}

// Qbar += (dR/dq)^T (rscale * rbar) for the residual of q whose c_h came
// from gm; acc[0] += c_hbar, acc[1..6] += free-stream cotangents.
static int adj_resid_vjp(sd2d_ctx *c, const double *q, const double *gm, const double *rbar,
                         double rscale, double *qbar, double *acc) {
  AdjState *A = c->adj;
  // divB of q, as the stepping kernel forms it for the Powell source, with
  // scratch maxima
  StageArgs s = base_args(c);
  s.qin = q;
  s.gmax = gm;
  s.slots = A->slots;
  s.divb = A->divb;
  HIPCHK(c, launch_stage(c, s, M_DIVB));
  AdjArgs a{};
  a.q = q; a.rbar = rbar; a.rscale = rscale; a.divb = A->divb; a.met = c->met;
  a.ftab = c->ftab; a.s2c = c->s2c; a.phm = c->phm; a.gm = gm; a.g = A->g; a.qbar = qbar;
  a.part = A->part; a.alpha = c->ph.alpha; a.ncell = c->ncell; a.ntile = c->ntile;
  switch (c->N) {
    case 3:
      hipLaunchKernelGGL(adj_point_kernel<3>, dim3(c->ntile), dim3(STAGE_THREADS), 0, c->stream, a);
      hipLaunchKernelGGL(adj_line_kernel<3>, dim3(c->ntile), dim3(STAGE_THREADS), adj_lds_bytes<3>(), c->stream, a, c->ops);
      break;
    case 4:
      hipLaunchKernelGGL(adj_point_kernel<4>, dim3(c->ntile), dim3(STAGE_THREADS), 0, c->stream, a);
      hipLaunchKernelGGL(adj_line_kernel<4>, dim3(c->ntile), dim3(STAGE_THREADS), adj_lds_bytes<4>(), c->stream, a, c->ops);
      break;
    default:
      hipLaunchKernelGGL(adj_point_kernel<5>, dim3(c->ntile), dim3(STAGE_THREADS), 0, c->stream, a);
      hipLaunchKernelGGL(adj_line_kernel<5>, dim3(c->ntile), dim3(STAGE_THREADS), adj_lds_bytes<5>(), c->stream, a, c->ops);
      break;
  }
  HIPCHK(c, hipGetLastError());
  std::vector<double> part((size_t)c->ntile * ADJ_NPART);
  HIPCHK(c, hipMemcpyAsync(part.data(), A->part, part.size() * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int t = 0; t < c->ntile; ++t)
    for (int j = 0; j < ADJ_NPART; ++j) acc[j] += part[(size_t)t * ADJ_NPART + j];
  return SD2D_OK;
}

// Reverse one RK_S5P3 step: A->b[5] holds the cotangent of the step's
// output (Q5), on return A->b[0] that of its input (Q0).  A->q[0..4] and
// A->gm hold the recomputed stage inputs and their maxima.
static int adj_step_vjp(sd2d_ctx *c, double *pacc, double *chbar5) {
  AdjState *A = c->adj;
  double **B = A->b;
  const double dt = c->p.dt;
  int r;
  double acc[ADJ_NPART];
  double gm[12];   // running maxima (eigvmax_g, umax_g) before each stage and after the last
  HIPCHK(c, hipMemcpyAsync(gm, A->gm, sizeof(gm), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  // zero the stage cotangents
  for (int k = 0; k < 5; ++k)
    HIPCHK(c, hipMemsetAsync(B[k], 0, state_len(c) * sizeof(double), c->stream));
  for (int s = 4; s >= 0; --s) {
    const AdjStage &st = kS5P3[s];
    // Q_out = cin Q_in + ca Q_aux + b dt R(Q_in)
    if ((r = axpby(c, st.cin, B[st.out], 1.0, B[st.in]))) return r;
    if (st.aux >= 0 && (r = axpby(c, st.ca, B[st.out], 1.0, B[st.aux]))) return r;
    // the maxima this evaluation raised take the pending cotangents of the
    // later evaluations' c_h
    const double *g0 = gm + 2 * s, *g1 = gm + 2 * (s + 1);
    if (g1[0] > g0[0]) {
      if (A->PE != 0.0 && (r = adj_argmax_push(c, A->q[st.in], B[st.in], 0, A->PE))) return r;
      A->PE = 0.0;
    }
    if (g1[1] > g0[1]) {
      if (A->PU != 0.0 && (r = adj_argmax_push(c, A->q[st.in], B[st.in], 1, A->PU))) return r;
      A->PU = 0.0;
    }
    for (double &x : acc) x = 0.0;
    r = adj_resid_vjp(c, A->q[st.in], A->gm + 2 * st.in, B[st.out], st.b * dt, B[st.in], acc);
    if (r) return r;
    if (chbar5) chbar5[s] = acc[0];
    for (int j = 0; j < 6; ++j) pacc[j] += acc[1 + j];
    // c_h = sqrt(E (E - U)) of this evaluation, from the maxima before it
    const double E = g0[0], U = g0[1], ch = std::sqrt(E * (E - U));
    if (ch > 0.0) {
      A->PE += acc[0] * (2.0 * E - U) / (2.0 * ch);
      A->PU += acc[0] * (-E) / (2.0 * ch);
    }
  }
  return SD2D_OK;
}

extern "C" {

// the wall-drag quadrature weights on the device, once per context
static int adj_weights(sd2d_ctx *c, AdjState *A) {
  if (A->wq) return SD2D_OK;
  const std::vector<double> w = face_weights(c->xs.data(), c->N);
  HIPCHK(c, hipMalloc(&A->wq, w.size() * sizeof(double)));
  HIPCHK(c, hipMemcpy(A->wq, w.data(), w.size() * sizeof(double), hipMemcpyHostToDevice));
  return SD2D_OK;
}

int sd2d_fwd(sd2d_ctx *c, int32_t nsteps, int32_t ckpt, double *objective) {
  if (!c) return fail(nullptr, SD2D_EINVAL, "ctx is NULL");
  if (nsteps <= 0 || ckpt <= 0) return fail(c, SD2D_EINVAL, "nsteps and ckpt must be positive");
  if (c->next_stage != 1)
    return fail(c, SD2D_EINVAL, "a step is in progress (sd2d_rk_stage): finish it first");
  if (c->p.precision != SD2D_FP64 || c->p.k_stage != 3 || c->vis || c->nmem > 1)
    return fail(c, SD2D_EUNSUPPORTED,
                "the adjoint covers the FP64 inviscid path with RK_S5P3 (k_stage 3)");
  if (!c->adj) c->adj = new (std::nothrow) AdjState();
  AdjState *A = c->adj;
  if (!A) return fail(c, SD2D_ENOMEM, "host allocation failed");
  A->nsteps = nsteps;
  A->ckpt = ckpt;
  A->have_fwd = false;
  const int nck = (nsteps + ckpt - 1) / ckpt + 1;   // segment starts + the final state
  int r = adj_alloc(c, A, nck);
  if (r) return r;
  if ((r = adj_weights(c, A))) return r;
  A->cki.assign(nck, 0);
  A->ckt.assign(nck, 0.0);
  const size_t nb = state_len(c) * sizeof(double);
  double J = 0.0;
  for (int n = 0; n < nsteps; ++n) {
    if (n % ckpt == 0) {
      HIPCHK(c, hipMemcpyAsync(A->ck[n / ckpt], c->q[c->cur], nb, hipMemcpyDeviceToDevice, c->stream));
      if ((r = adj_maxima_get(c, A->ckm[n / ckpt]))) return r;
      A->cki[n / ckpt] = c->iter;
      A->ckt[n / ckpt] = c->ctime;
    }
    c->iter += 1;
    if ((r = step_rk(c, false))) return r;
    c->have_av = c->p.artiv == 1 && c->p.vismode == 1;
    double D;
    if ((r = drag_eval(c, c->q[c->cur], &D, nullptr, 0.0))) return r;
    J += D;
  }
  HIPCHK(c, hipMemcpyAsync(A->ck[nck - 1], c->q[c->cur], nb, hipMemcpyDeviceToDevice, c->stream));
  if ((r = adj_maxima_get(c, A->ckm[nck - 1]))) return r;
  A->fin_ctime = c->ctime;
  A->fin_iter = c->iter;
  A->cki[nck - 1] = c->iter;
  A->ckt[nck - 1] = c->ctime;
  A->J = J / nsteps;
  A->have_fwd = true;
  if (objective) *objective = A->J;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SD2D_OK;
}

int sd2d_vjp(sd2d_ctx *c, const double *qbar_final, double jbar, double *qbar_init,
             double *pbar, double *chbar) {
  if (!c) return fail(nullptr, SD2D_EINVAL, "ctx is NULL");
  SD2D_NOT_MID_STEP(c, "sd2d_vjp");
  AdjState *A = c->adj;
  if (!A || !A->have_fwd) return fail(c, SD2D_EINVAL, "sd2d_vjp needs a preceding sd2d_fwd");
  const int nsteps = A->nsteps, ckpt = A->ckpt, nck = (nsteps + ckpt - 1) / ckpt + 1;
  const size_t nb = state_len(c) * sizeof(double);
  int r;
  // cotangent of the final state
  if (qbar_final) {
    const size_t n = (size_t)c->ncell * 9 * c->N * c->N;
    HIPCHK(c, hipMemcpyAsync(c->stagef, qbar_final, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if ((r = f2t(c, c->stagef, A->b[5]))) return r;
  } else {
    HIPCHK(c, hipMemsetAsync(A->b[5], 0, nb, c->stream));
  }
  double pacc[6] = {0, 0, 0, 0, 0, 0};
  A->PE = A->PU = 0.0;
  A->argmax_missed = 0;
  for (int sg = (nsteps - 1) / ckpt; sg >= 0; --sg) {
    const int n0 = sg * ckpt, n1 = std::min(nsteps, n0 + ckpt);
    // the segment's step inputs, recomputed from its checkpoint
    HIPCHK(c, hipMemcpyAsync(A->q[0], A->ck[sg], nb, hipMemcpyDeviceToDevice, c->stream));
    std::array<double, 4> m = A->ckm[sg];
    for (int i = 0; i < n1 - n0; ++i) {
      HIPCHK(c, hipMemcpyAsync(A->seg[i], A->q[0], nb, hipMemcpyDeviceToDevice, c->stream));
      A->segm[i] = m;
      if (i + 1 < n1 - n0) {
        if ((r = adj_recompute_step(c, m))) return r;
        if ((r = adj_maxima_get(c, m))) return r;
        HIPCHK(c, hipMemcpyAsync(A->q[0], A->q[5], nb, hipMemcpyDeviceToDevice, c->stream));
      }
    }
    for (int i = n1 - n0 - 1; i >= 0; --i) {
      const int n = n0 + i;   // step n takes Q_n to Q_{n+1}
      HIPCHK(c, hipMemcpyAsync(A->q[0], A->seg[i], nb, hipMemcpyDeviceToDevice, c->stream));
      if ((r = adj_recompute_step(c, A->segm[i]))) return r;
      // the drag of the step's output enters the objective with weight 1/nsteps
      if (jbar != 0.0 && (r = drag_eval(c, A->q[5], nullptr, A->b[5], jbar / nsteps))) return r;
      if ((r = adj_step_vjp(c, pacc, chbar ? chbar + 5 * n : nullptr))) return r;
      HIPCHK(c, hipMemcpyAsync(A->b[5], A->b[0], nb, hipMemcpyDeviceToDevice, c->stream));
    }
  }
  if (qbar_init) {
    if ((r = t2f(c, A->b[5], c->stagef))) return r;
    const size_t n = (size_t)c->ncell * 9 * c->N * c->N;
    HIPCHK(c, hipMemcpyAsync(qbar_init, c->stagef, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  }
  if (pbar) for (int j = 0; j < 6; ++j) pbar[j] = pacc[j];
  if (A->argmax_missed)
    return fail(c, SD2D_EINVAL, "adjoint: the flux point of a running maximum was not found");
  // the context returns to the final state of sd2d_fwd
  HIPCHK(c, hipMemcpyAsync(c->q[c->cur], A->ck[nck - 1], nb, hipMemcpyDeviceToDevice, c->stream));
  if ((r = adj_maxima_set(c, A->ckm[nck - 1]))) return r;
  c->ctime = A->fin_ctime;
  c->iter = A->fin_iter;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SD2D_OK;
}

// The adjoint checkpoints of the last sd2d_fwd as files: <prefix>.NNNN.dat
// (NNNN = 0..nck-1) in the reference's restart format with the maxima
// sidecar of sd2d_write_checkpoint, and <prefix>.adj: 8 bytes "SD2DADJ1",
// int32 N, NCELL, nsteps, ckpt, final iter, float64 J, final ctime.
static std::string adj_ck_path(const char *prefix, int k) {
  char tail[16];
  std::snprintf(tail, sizeof(tail), ".%04d.dat", k);
  return std::string(prefix) + tail;
}

int sd2d_adj_save(sd2d_ctx *c, const char *prefix) {
  if (!c || !prefix) return fail(c, SD2D_EINVAL, "NULL argument");
  AdjState *A = c->adj;
  if (!A || !A->have_fwd) return fail(c, SD2D_EINVAL, "sd2d_adj_save needs a preceding sd2d_fwd");
  const int nck = (A->nsteps + A->ckpt - 1) / A->ckpt + 1;
  const size_t n = (size_t)c->ncell * 9 * c->N * c->N;
  std::vector<double> q(n);
  for (int k = 0; k < nck; ++k) {
    int r = t2f(c, A->ck[k], c->stagef);
    if (r) return r;
    HIPCHK(c, hipMemcpyAsync(q.data(), c->stagef, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const std::string path = adj_ck_path(prefix, k);
    const sd2d_maxima m{A->ckm[k][0], A->ckm[k][1], A->ckm[k][2], A->ckm[k][3]};
    if (!write_restart_file(path.c_str(), A->cki[k], A->ckt[k], c->N, c->ncell, q.data(), nullptr) ||
        !write_max_file(path.c_str(), A->cki[k], A->ckt[k], m))
      return fail(c, SD2D_EINVAL, "write failed: " + path);
  }
  const std::string mp = std::string(prefix) + ".adj";
  FILE *fh = std::fopen(mp.c_str(), "wb");
  if (!fh) return fail(c, SD2D_EINVAL, "cannot write " + mp);
  const int32_t hv[5] = {c->N, c->ncell, A->nsteps, A->ckpt, A->fin_iter};
  const double dv[2] = {A->J, A->fin_ctime};
  const bool ok = std::fwrite("SD2DADJ1", 1, 8, fh) == 8 && std::fwrite(hv, sizeof(int32_t), 5, fh) == 5 &&
                  std::fwrite(dv, sizeof(double), 2, fh) == 2;
  if (std::fclose(fh) != 0 || !ok) return fail(c, SD2D_EINVAL, "write failed: " + mp);
  return SD2D_OK;
}

int sd2d_adj_load(sd2d_ctx *c, const char *prefix) {
  if (!c || !prefix) return fail(c, SD2D_EINVAL, "NULL argument");
  SD2D_NOT_MID_STEP(c, "sd2d_adj_load");
  if (c->p.precision != SD2D_FP64 || c->p.k_stage != 3 || c->vis || c->nmem > 1)
    return fail(c, SD2D_EUNSUPPORTED,
                "the adjoint covers the FP64 inviscid path with RK_S5P3 (k_stage 3)");
  const std::string mp = std::string(prefix) + ".adj";
  FILE *fh = std::fopen(mp.c_str(), "rb");
  if (!fh) return fail(c, SD2D_EINVAL, "cannot open " + mp);
  char magic[8];
  int32_t hv[5];
  double dv[2];
  const bool ok = std::fread(magic, 1, 8, fh) == 8 && std::fread(hv, sizeof(int32_t), 5, fh) == 5 &&
                  std::fread(dv, sizeof(double), 2, fh) == 2 && std::fgetc(fh) == EOF;
  std::fclose(fh);
  if (!ok || std::memcmp(magic, "SD2DADJ1", 8) != 0)
    return fail(c, SD2D_EINVAL, "not an adjoint checkpoint manifest: " + mp);
  if (hv[0] != c->N || hv[1] != c->ncell || hv[2] <= 0 || hv[3] <= 0)
    return fail(c, SD2D_EINVAL, "the adjoint checkpoints of " + mp + " belong to another mesh or order");
  if (!c->adj) c->adj = new (std::nothrow) AdjState();
  AdjState *A = c->adj;
  if (!A) return fail(c, SD2D_ENOMEM, "host allocation failed");
  A->have_fwd = false;
  A->nsteps = hv[2];
  A->ckpt = hv[3];
  const int nck = (A->nsteps + A->ckpt - 1) / A->ckpt + 1;
  int r = adj_alloc(c, A, nck);
  if (r) return r;
  if ((r = adj_weights(c, A))) return r;
  A->cki.assign(nck, 0);
  A->ckt.assign(nck, 0.0);
  const size_t n = (size_t)c->ncell * 9 * c->N * c->N;
  for (int k = 0; k < nck; ++k) {
    const std::string path = adj_ck_path(prefix, k);
    int32_t it = 0;
    double t = 0.0;
    std::vector<double> q;
    sd2d_maxima m{};
    std::string e = read_restart_file(path.c_str(), c->N, c->ncell, it, t, q);
    if (e.empty()) e = read_max_file(path.c_str(), it, t, m);
    if (!e.empty()) return fail(c, SD2D_EINVAL, e);
    HIPCHK(c, hipMemcpyAsync(c->stagef, q.data(), n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if ((r = f2t(c, c->stagef, A->ck[k]))) return r;
    HIPCHK(c, hipStreamSynchronize(c->stream));   // q is reused by the next file
    A->ckm[k] = {m.eigvmax, m.umax, m.eigvmax_g, m.umax_g};
    A->cki[k] = it;
    A->ckt[k] = t;
  }
  A->J = dv[0];
  A->fin_ctime = dv[1];
  A->fin_iter = hv[4];
  if (A->cki[nck - 1] != A->fin_iter || std::memcmp(&A->ckt[nck - 1], &A->fin_ctime, sizeof(double)) != 0)
    return fail(c, SD2D_EINVAL, "the last adjoint checkpoint is not the final state of " + mp);
  A->have_fwd = true;
  // the context holds the final state of the forward run, as after sd2d_fwd
  HIPCHK(c, hipMemcpyAsync(c->q[c->cur], A->ck[nck - 1], state_len(c) * sizeof(double),
                           hipMemcpyDeviceToDevice, c->stream));
  if ((r = adj_maxima_set(c, A->ckm[nck - 1]))) return r;
  c->ctime = A->fin_ctime;
  c->iter = A->fin_iter;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SD2D_OK;
}

int sd2d_compresid_vjp(sd2d_ctx *c, const double *rbar, double *qbar, double *chbar, double *pbar) {
  if (!c || !rbar) return fail(c, SD2D_EINVAL, "NULL argument");
  SD2D_NOT_MID_STEP(c, "sd2d_compresid_vjp");
  if (c->p.precision != SD2D_FP64 || c->vis || c->nmem > 1)
    return fail(c, SD2D_EUNSUPPORTED, "the adjoint covers the FP64 inviscid path");
  if (!c->adj) c->adj = new (std::nothrow) AdjState();
  AdjState *A = c->adj;
  if (!A) return fail(c, SD2D_ENOMEM, "host allocation failed");
  if (A->ckpt <= 0) A->ckpt = 1;
  int r = adj_alloc(c, A, 0);
  if (r) return r;
  const size_t n = (size_t)c->ncell * 9 * c->N * c->N;
  HIPCHK(c, hipMemcpyAsync(c->stagef, rbar, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  if ((r = f2t(c, c->stagef, A->b[1]))) return r;
  HIPCHK(c, hipMemsetAsync(A->b[0], 0, state_len(c) * sizeof(double), c->stream));
  double acc[ADJ_NPART] = {0, 0, 0, 0, 0, 0, 0};
  r = adj_resid_vjp(c, c->q[c->cur], c->gmax, A->b[1], 1.0, A->b[0], acc);
  if (r) return r;
  if (qbar) {
    if ((r = t2f(c, A->b[0], c->stagef))) return r;
    HIPCHK(c, hipMemcpyAsync(qbar, c->stagef, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (chbar) *chbar = acc[0];
  if (pbar) for (int j = 0; j < 6; ++j) pbar[j] = acc[1 + j];
  return SD2D_OK;
}

}  // extern "C"
