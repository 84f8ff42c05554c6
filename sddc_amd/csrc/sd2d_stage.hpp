// sd2d_stage.hpp - the fused residual + Runge-Kutta stage kernel.
//
// One 128-thread workgroup owns a tile of T cells (T = 12/16/20 for
// N = 5/4/3) and runs one thread per grid line: thread (fam, line, cell)
// takes the N solution values of its line for all 9 variables and, in
// registers,
//   - interpolates them to the N+1 flux points of the line (Lmat),
//   - forms the flux at every flux point: the ideal-MHD + GLM flux normal
//     to the point's metric vector inside the cell, the Rusanov flux with
//     the neighbour's trace, a boundary ghost or the periodic partner at
//     the two end points,
//   - differentiates the fluxes along the line (Mmat), and the normal
//     component of the averaged field (divB) and its rotated counterpart
//     (Jz, for the AV sensor) likewise.
// The two families exchange their derivatives through LDS in two rounds
// and split the solution points' work by variable: each xi-line thread
// (fam 0) finishes variables 0-3, each eta-line thread (fam 1) variables
// 4-8 at the N points of its line: residual, Powell and GLM sources, the RK
// combination, norm partials.  The running maxima and norms are reduced
// per workgroup; the AV smoothness sensor runs per cell on the last stage.
//
// Register budget: the neighbour traces of a line's two end faces are
// gathered first, while nothing else is live; the tile's state is staged
// in LDS so the line values can be re-read cheaply; the flux points are
// then taken one at a time and folded into the derivative accumulators.
//
// A face point is evaluated by both adjacent cells from the same two
// traces in the same left/right roles (the neighbour's line is read from
// memory and interpolated with the same arithmetic as the own trace), so
// both cells store bitwise the same flux and only the state, the metrics
// and the RK registers move through HBM.
#pragma once
#include <type_traits>

#include "sd2d_device.hpp"

namespace sd2d {

constexpr int STAGE_THREADS = 128;

struct StageArgs {
  const double *qin;        // state the residual is taken of (tile layout)
  const double *aux[4];     // RK registers read pointwise (tile layout)
  double ca[4];
  double cin, cr;           // qout = cin*qin + sum ca*aux + cr*resid
  int naux;
  double *qout;             // may alias aux[i] (pointwise in-place)
  double *rstore;           // resid (tile layout) for FiveStageRk, or null
  const double *met;        // metrics (tile layout)
  const int *ftab;          // per-cell face table (tile layout)
  const double *gmax;       // eigvmax_g, umax_g of the previous evaluation
  unsigned long long *slots;  // running maxima, NSLOT x 2 (bit patterns)
  double *sm_ind;           // SM_IND(2,NCELL) or null
  double *normbuf;          // per-tile sums of resid^2 and divB^2, or null
  double *resid_f;          // resid(9,N,N,NCELL) Fortran layout, or null
  double *nabla_f;          // nablaQs(9,2,N,N,NCELL) Fortran layout, or null
  // the AV sensor's smoothing filter per direction, F[i][k] at avf[i*N + k]:
  // the projection onto the N-1 low-order points followed by the
  // interpolation back to the N solution points (AV.f90:18-63) is one linear
  // map F = Hhigh Hlow, the same along xi and eta (host-built, in the
  // kernel arguments so that its entries are scalar loads)
  double avf[NMAX * NMAX];
  const double *visres;     // viscous divergence dFv + dGv (tile layout), M_VIS
  double *divb;             // divB per solution point [tile][p][ci], M_DIVB
  const int *s2c;           // slot -> mesh cell, for the cell-indexed outputs
  // per tile (member << 6 | valid cells), member 0 for one run; an
  // ensemble of runs also has per-member dt (qout then uses cr * dt, null
  // for one run), and gmax / slots hold one entry per member
  const int *tinfo;
  const double *dtm;
  int rev;                  // walk the launch order backwards (alternate stages)
  int sfam;                 // line family whose wave stages the tile (fewer gathers)
  const Phys *phm;          // free stream and GLM alpha, one per member
  int ncell, ntile;
};

// MODE bits.  The RK registers read pointwise (none, one, or the four of
// FiveStageRk's last stage) and the residual store are part of the mode so
// that the stage update is straight-line code.
// M_VIS adds the viscous divergence of the residual evaluation (formed by
// the kernels of sd2d_visc.hpp) before the Jacobian scaling.
// M_DIVB (no RK update, no Fortran-layout output) stores only divB at every
// solution point, in tile layout: what the adjoint's reverse sweep of the
// Powell source needs, with the arithmetic of the stepping kernel.
enum { M_RK = 1, M_AV = 2, M_NORMS = 4, M_COMPAT = 8, M_AUX1 = 16, M_AUX4 = 32,
       M_RSTORE = 64, M_VIS = 128, M_DIVB = 256 };

// XCD-aware block order: consecutive positions of the tile launch order go
// to the same XCD (and its L2), so the neighbour lines a tile reads were
// staged by a nearby tile on the same XCD.  With rev set, every XCD walks
// its part of the order backwards: alternate stages then start where the
// previous stage ended, on the tiles it wrote last (still in the
// memory-side cache).  Bijective for any tile count either way.
__device__ __forceinline__ int tile_of_block(int b, int nt, int rev) {
  const int q = nt >> 3, r = nt & 7;
  const int x = b & 7;
  int i = b >> 3;
  if (rev) i = q + (x < r ? 1 : 0) - 1 - i;
  return (x < r) ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

// LDS of one workgroup (in doubles): the tile's state (qo layout), the
// derivatives of the fluxes of up to XV variables (xo layout) handed from one
// family to the other (eta -> xi: variables 0-3, then xi -> eta: 4-8), of
// the divB weight [p][ci] and (AV) of the Jz weight [p][ci], then
// reduction slots.  The AV sensor's wave reads rho from the state area and
// keeps Jz^2, then its line sums, in the Jz area, which only that wave
// uses after the exchange.
template <int N, int MODE> struct StageLds {
  static constexpr int T = Geo<N>::CPW, N2 = N * N;
  static constexpr bool AV = (MODE & M_AV) != 0;
  static constexpr int XV = 5;                 // variables per exchange round
  static constexpr int Q0 = 0;
  static constexpr int XD = 9 * N2 * T;
  static constexpr int XB = XD + XV * N2 * T;
  static constexpr int XJ = XB + N2 * T;
  static constexpr int RED = XJ + (AV ? N2 * T : 0);
  static constexpr int TOTAL = RED + 4 * (STAGE_THREADS / 64);
  static_assert(2 * N <= N2, "the AV line sums fit the Jz area");
};

template <int N, int MODE>
constexpr size_t stage_lds_bytes() { return sizeof(double) * StageLds<N, MODE>::TOTAL; }

// Trace of one line of N values with interpolation row Lr:
// acc = sum_s q[s]*Lr[s], summed in order s = 0..N-1 with FMA.  The own
// trace and the neighbour trace at a face use this same formula so that
// both cells form bitwise the same face state.
template <int N, typename R>
__device__ __forceinline__ R line_trace(const R *vals, const R *Lr) {
  R acc = R(0);
#pragma unroll
  for (int s = 0; s < N; ++s) acc = fma(vals[s], Lr[s], acc);
  return acc;
}

// Wave-wide maximum / sum with a fixed butterfly order (deterministic).
__device__ __forceinline__ double wave_max(double x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    double y = __shfl_xor(x, o, 64);
    x = (y > x) ? y : x;
  }
  return x;
}
__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}

// AV smoothness sensor (AV.f90:14-83) of the cells of one tile, by one
// wave of line threads.  The reference projects rho and Jz^2 of a cell onto
// the (N-1)^2 low-order points and interpolates back; per direction that is
// the filter F = Hhigh Hlow, so the smoothed values are (F x F) applied to
// the cell's N x N values.  A thread filters the whole cell across its line
// onto the line (ROW: the xi-line of row js = line; otherwise the eta-line
// of column is = line), then along the line, and sums its points' terms
// (smoothed - x)^2 / x^2; the N line sums of a cell meet in LDS and lane
// line 0 of the cell takes log10 of each total.  rho and jz2 hold the
// cell's values in LDS at [(p T + ci) RS] and [p T + ci] (p = is + N js;
// RS = 2 for rho read from the staged tile's first variable pair), rs and js2 the thread's own
// points in line order; only lanes of this wave exchange values, so the
// caller needs no workgroup barrier but must call this from every lane of
// the wave (a wave-level LDS sync inside).  The sums run in another order
// than the reference's and rho's quotient uses the reciprocal, so values
// differ from it in the last bits of the smoothed field.
template <int N, int T, bool ROW, int RS>
__device__ __forceinline__ void av_sensor_lines(const StageArgs &a, const double *rho,
                                                const double *jz2, double *lineb,
                                                const double (&rs)[N], const double (&js2)[N],
                                                int line, int ci, int cell, bool valid) {
  wave_lds_sync();   // the cell's Jz^2 (and rho) of the other lines are in LDS
  double e1 = 0.0, e2 = 0.0;
  if (valid) {
    double y1[N], y2[N];
#pragma unroll
    for (int m = 0; m < N; ++m) {
      double u1 = 0.0, u2 = 0.0;
#pragma unroll
      for (int k = 0; k < N; ++k) {
        const double f = a.avf[line * N + k];
        const int p = ROW ? m + N * k : k + N * m;
        u1 = fma(f, rho[(p * T + ci) * RS], u1);
        u2 = fma(f, jz2[p * T + ci], u2);
      }
      y1[m] = u1;
      y2[m] = u2;
    }
#pragma unroll
    for (int m = 0; m < N; ++m) {
      double h1 = 0.0, h2 = 0.0;
#pragma unroll
      for (int k = 0; k < N; ++k) {
        h1 = fma(a.avf[m * N + k], y1[k], h1);
        h2 = fma(a.avf[m * N + k], y2[k], h2);
      }
      // rho > 0: its quotient through the reciprocal (within a few ulp);
      // Jz^2 may be 0, so its quotient keeps IEEE division (inf or NaN as
      // in the reference)
      const double d1 = (h1 - rs[m]) * rcp(rs[m]), d2 = h2 - js2[m];
      e1 = fma(d1, d1, e1);
      e2 += (d2 * d2) / (js2[m] * js2[m]);
    }
  }
  // (lineb may be the area jz2 was read from: every lane's reads first)
  wave_lds_sync();
  if (valid) {
    lineb[line * T + ci] = e1;
    lineb[(N + line) * T + ci] = e2;
  }
  wave_lds_sync();
  // lane line 0 of a cell finishes rho's sum, lane line 1 Jz's, so the
  // wave evaluates log10 once
  if (valid && line < 2) {
    const double *sb = lineb + line * N * T + ci;
    double sum = sb[0];
#pragma unroll
    for (int l = 1; l < N; ++l) sum += sb[l * T];
    a.sm_ind[2 * (size_t)a.s2c[cell] + line] = log10(sum);
  }
}

// The values of all 9 variables along a line of a state tile q (in HBM or
// staged in LDS, FP64 or FP32): f(v, vals) is called for v = 0..8 with
// vals[s] = variable v at point p0 + ps*s of slot ci.  Variables 0-7 are
// read in pairs, one 16-byte (FP32: 8-byte) access per pair and point.
template <typename S> struct PairOf;
template <> struct PairOf<double> { using type = double2; };
template <> struct PairOf<float> { using type = float2; };
template <int N, typename R, typename S, typename F>
__device__ __forceinline__ void line_vals(const S *q, int p0, int ps, int ci, F &&f) {
  using S2 = typename PairOf<S>::type;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    R va[N], vb[N];
#pragma unroll
    for (int s = 0; s < N; ++s) {
      const S2 d = *reinterpret_cast<const S2 *>(q + qo<N>(2 * k, p0 + ps * s, ci));
      va[s] = R(d.x);
      vb[s] = R(d.y);
    }
    f(2 * k, va);
    f(2 * k + 1, vb);
  }
  R vc[N];
#pragma unroll
  for (int s = 0; s < N; ++s) vc[s] = R(q[qo<N>(8, p0 + ps * s, ci)]);
  f(8, vc);
}

// Variables V0..V1-1 of point p of slot ci of the state tile q:
// load_vars puts variable v in out[v - V0], store_vars writes in[v - V0]
// to it; both pairs inside the range in one 16-byte access.
template <int N, int V0, int V1>
__device__ __forceinline__ void load_vars(const double *q, int p, int ci, double *out) {
#pragma unroll
  for (int v = V0; v < V1; ++v) {
    if (v < 8 && (v & 1) == 0 && v + 1 < V1) {
      const double2 d = *reinterpret_cast<const double2 *>(q + qo<N>(v, p, ci));
      out[v - V0] = d.x;
      out[v + 1 - V0] = d.y;
    } else if (!(v < 8 && (v & 1) == 1 && v - 1 >= V0)) {
      out[v - V0] = q[qo<N>(v, p, ci)];
    }
  }
}
template <int N, int V0, int V1>
__device__ __forceinline__ void store_vars(double *q, int p, int ci, const double *in) {
#pragma unroll
  for (int v = V0; v < V1; ++v) {
    if (v < 8 && (v & 1) == 0 && v + 1 < V1) {
      *reinterpret_cast<double2 *>(q + qo<N>(v, p, ci)) = double2{in[v - V0], in[v + 1 - V0]};
    } else if (!(v < 8 && (v & 1) == 1 && v - 1 >= V0)) {
      q[qo<N>(v, p, ci)] = in[v - V0];
    }
  }
}

// The derivative exchange area of the stage kernels holds up to five
// variables u = 0..4 of every point of the tile, pairs (0,1), (2,3) as
// [u/2][p][ci][u%2] and u = 4 as [p][ci] after them.  xput writes w[0..U-1]
// to variables 0..U-1 of point p of slot ci, xadd adds them into w, both
// with one 16-byte (FP32: 8-byte) LDS access per pair.
template <int N>
__device__ constexpr int xo(int u, int p, int ci) {
  constexpr int T = cpw_of(N), N2 = N * N;
  return u < 4 ? (((u >> 1) * N2 + p) * T + ci) * 2 + (u & 1) : 4 * N2 * T + p * T + ci;
}
template <int N, int U, typename S>
__device__ __forceinline__ void xput(S *x, int p, int ci, const S *w) {
  using S2 = typename PairOf<S>::type;
#pragma unroll
  for (int u = 0; u < U; u += 2) {
    if (u + 1 < U && u < 4) {
      S2 d;
      d.x = w[u];
      d.y = w[u + 1];
      *reinterpret_cast<S2 *>(x + xo<N>(u, p, ci)) = d;
    } else {
      x[xo<N>(u, p, ci)] = w[u];
    }
  }
}
template <int N, int U, typename S>
__device__ __forceinline__ void xadd(const S *x, int p, int ci, S *w) {
  using S2 = typename PairOf<S>::type;
#pragma unroll
  for (int u = 0; u < U; u += 2) {
    if (u + 1 < U && u < 4) {
      const S2 d = *reinterpret_cast<const S2 *>(x + xo<N>(u, p, ci));
      w[u] += d.x;
      w[u + 1] += d.y;
    } else {
      w[u] += x[xo<N>(u, p, ci)];
    }
  }
}

// End face `rfp` (0 or N) of line `line` of family `fam`: own local face k
// (0-based) and own point index nfp (1-based) from face2fpconnec
// (initsetup.f90:356-380).
template <int N>
__device__ __forceinline__ void end_face(int fam, int line, int rfp, int &k, int &nfp) {
  k = fam == 0 ? (rfp == 0 ? 3 : 1) : (rfp == 0 ? 0 : 2);
  nfp = fam == 0 ? (rfp == 0 ? N - line : line + 1) : (rfp == 0 ? line + 1 : N - line);
}

// The neighbour's line through the shared point of an end face: its first
// point p0 and stride, and whether the point is the line's end N (otherwise
// its end 0), from the neighbour's local face kn and the own point index
// nfp (the neighbour walks the face the other way: interfaceflux_all.f90:59-60).
template <int N>
__device__ __forceinline__ void neighbour_line(int kn, int nfp, int &p0, int &stride, bool &endN) {
  const int m = N - nfp + 1;  // neighbour point index on its face (1-based)
  if (kn == 0) { p0 = m - 1; stride = N; }
  else if (kn == 1) { p0 = N * (m - 1); stride = 1; }
  else if (kn == 2) { p0 = N - m; stride = N; }
  else { p0 = N * (N - m); stride = 1; }
  endN = (kn == 1 || kn == 2);
}

// Gather for an end face: the face-table info word and, for an interior or
// periodic face, the neighbour's trace Qn at the shared face point (its
// line through that point, interpolated with the same arithmetic as the
// own trace: interfaceflux_all.f90:59-76).  When the neighbour cell is in
// the same tile, nothing is read here: lci returns its index in the tile
// and the trace is formed from the staged tile by face_trace_lds.
// The face-table words (neighbour base, info) of end face `rfp` of a line.
template <int N>
__device__ __forceinline__ void face_words(const StageArgs &a, int tile, int ci, int fam, int line,
                                           int rfp, int &nbase, int &info) {
  constexpr int T = Geo<N>::CPW;
  int k, nfp;
  end_face<N>(fam, line, rfp, k, nfp);
  const int *ft = a.ftab + ((size_t)tile * 8 + k * 2) * T + ci;
  nbase = ft[0];
  info = ft[T];
}

// The gather of face_gather from already loaded face-table words.
template <int N, typename R, typename OP>
__device__ __forceinline__ void face_gather_w(const StageArgs &a, const OP &ops, int tile, int fam,
                                              int line, int rfp, int nbase, int info, R *Qn,
                                              int &lci) {
  constexpr int T = Geo<N>::CPW, N2 = Geo<N>::N2, QT = Geo<N>::QT;
  int k, nfp;
  end_face<N>(fam, line, rfp, k, nfp);
  lci = -1;
  if (((info >> 3) & 7) == FT_INTERIOR) {
    const int ntile = nbase / QT;
    if (ntile == tile) {
      lci = nbase - tile * QT;
    } else {
      const int nci = nbase - ntile * QT;
      int p0, stride;
      bool endN;
      neighbour_line<N>(info & 3, nfp, p0, stride, endN);
      R Lr[N];
#pragma unroll
      for (int s = 0; s < N; ++s) Lr[s] = endN ? ops.L[N][s] : ops.L[0][s];
      line_vals<N, R>(a.qin + (size_t)ntile * QT, p0, stride, nci,
                      [&](int v, const R *vals) { Qn[v] = line_trace<N, R>(vals, Lr); });
    }
  }
}

// The gathers of both end faces of a line from loaded face-table words, in
// as few passes as the wave needs: pass 1 gathers, in every lane, the first
// end face whose neighbour lies in another tile; pass 2 the N end of the
// lanes that need both.  In a wave whose lanes need at most one gather each
// (a line family whose neighbours are mostly in the tile: only the lanes of
// the strip's end cells gather) pass 2 is skipped, which halves its gather
// loads.  Each face's trace is formed exactly as by face_gather_w.
template <int N, typename R, typename OP>
__device__ __forceinline__ void face_gather_pair(const StageArgs &a, const OP &ops, int tile,
                                                 int fam, int line, int nb0, int info0, int nbN,
                                                 int infoN, R *qn0, R *qnN, int &lci0,
                                                 int &lciN) {
  constexpr int T = Geo<N>::CPW, N2 = Geo<N>::N2, QT = Geo<N>::QT;
  const bool i0 = ((info0 >> 3) & 7) == FT_INTERIOR, iN = ((infoN >> 3) & 7) == FT_INTERIOR;
  const bool g0 = i0 && nb0 / QT != tile, gN = iN && nbN / QT != tile;
  lci0 = (i0 && !g0) ? nb0 - tile * QT : -1;
  lciN = (iN && !gN) ? nbN - tile * QT : -1;
  auto gather = [&](int rfp, int nbase, int info, R *Qn) {
    int k, nfp;
    end_face<N>(fam, line, rfp, k, nfp);
    int p0, stride;
    bool endN;
    neighbour_line<N>(info & 3, nfp, p0, stride, endN);
    R Lr[N];
#pragma unroll
    for (int s = 0; s < N; ++s) Lr[s] = endN ? ops.L[N][s] : ops.L[0][s];
    const int nt = nbase / QT;
    line_vals<N, R>(a.qin + (size_t)nt * QT, p0, stride, nbase - nt * QT,
                    [&](int v, const R *vals) { Qn[v] = line_trace<N, R>(vals, Lr); });
  };
  if (g0 || gN) {
    R Q[9];
    gather(g0 ? 0 : N, g0 ? nb0 : nbN, g0 ? info0 : infoN, Q);
#pragma unroll
    for (int v = 0; v < 9; ++v) {
      if (g0) qn0[v] = Q[v];
      else qnN[v] = Q[v];
    }
  }
  if (g0 && gN) gather(N, nbN, infoN, qnN);
}

// Gather for an end face: the face-table info word and, for an interior or
// periodic face, the neighbour's trace Qn at the shared face point (its
// line through that point, interpolated with the same arithmetic as the
// own trace: interfaceflux_all.f90:59-76).  When the neighbour cell is in
// the same tile, nothing is read here: lci returns its index in the tile
// and the trace is formed from the staged tile by face_trace_lds.
template <int N, typename R, typename OP>
__device__ __forceinline__ int face_gather(const StageArgs &a, const OP &ops, int tile, int ci,
                                           int fam, int line, int rfp, R *Qn, int &lci) {
  int nbase, info;
  face_words<N>(a, tile, ci, fam, line, rfp, nbase, info);
  face_gather_w<N, R>(a, ops, tile, fam, line, rfp, nbase, info, Qn, lci);
  return info;
}

// The neighbour's trace for a neighbour in the same tile, from the staged
// tile (the same values and arithmetic as face_gather).
template <int N, typename R, typename OP>
__device__ __forceinline__ void face_trace_lds(const R *qs, const OP &ops, int info, int fam,
                                               int line, int rfp, int lci, R *Qn) {
  constexpr int T = Geo<N>::CPW, N2 = Geo<N>::N2;
  int k, nfp;
  end_face<N>(fam, line, rfp, k, nfp);
  int p0, stride;
  bool endN;
  neighbour_line<N>(info & 3, nfp, p0, stride, endN);
  R Lr[N];
#pragma unroll
  for (int s = 0; s < N; ++s) Lr[s] = endN ? ops.L[N][s] : ops.L[0][s];
  line_vals<N, R>(qs, p0, stride, lci,
                  [&](int v, const R *vals) { Qn[v] = line_trace<N, R>(vals, Lr); });
}

// Flux at an end face point: Rusanov in the reference's left/right roles
// with the gathered neighbour trace (interfaceflux_all.f90:31-106,
// cyclocflux_all.f90:15-110) or with the boundary ghost (bcflux.f90).
// Returns the flux stored for this cell and the averaged state Qv.
template <int N, typename R>
__device__ __forceinline__ void face_flux(int info, int fam, int rfp, const Phys &ph,
                                          R c_h, const R *Qo, const R *Qn,
                                          R sx, R sy, R *Fh, R *Qv,
                                          R &eig, R &um) {
  const int k = fam == 0 ? (rfp == 0 ? 3 : 1) : (rfp == 0 ? 0 : 2);
  const int bc = (info >> 3) & 7;
  const int own_sl = (k == 0 || k == 3) ? -1 : 1;  // sign_l if this cell is left
  if (bc == FT_INTERIOR) {
    const int kn = info & 3, left = (info >> 2) & 1, sneg = (info >> 6) & 1;
    R QL[9], QR[9];
#pragma unroll
    for (int v = 0; v < 9; ++v) {
      QL[v] = left ? Qo[v] : Qn[v];
      QR[v] = left ? Qn[v] : Qo[v];
    }
    const int sl = left ? own_sl : ((kn == 0 || kn == 3) ? -1 : 1);
    const int sr = left ? ((kn == 1 || kn == 2) ? -1 : 1) : ((k == 1 || k == 2) ? -1 : 1);
    const R nsg = (left || !sneg) ? R(1) : R(-1);
    R Fl[9];
    rusanov<R>(QL, QR, nsg * sx, nsg * sy, sl, c_h, Fl, eig, um);
    const R fac = left ? R(1) : R(sl * sr);
#pragma unroll
    for (int v = 0; v < 9; ++v) {
      Fh[v] = fac * Fl[v];
      Qv[v] = R(0.5) * QL[v] + R(0.5) * QR[v];
    }
  } else {
    R QR[9];
    bc_ghost<R>(bc, Qo, sx, sy, ph, QR);
    rusanov<R>(Qo, QR, sx, sy, own_sl, c_h, Fh, eig, um);
#pragma unroll
    for (int v = 0; v < 9; ++v) Qv[v] = R(0.5) * Qo[v] + R(0.5) * QR[v];
  }
}

template <int N, int MODE>
__global__ __launch_bounds__(STAGE_THREADS, 2) void stage_kernel(StageArgs a, Ops ops) {
  using G = Geo<N>;
  using LD = StageLds<N, MODE>;
  constexpr int T = G::CPW, N2 = G::N2, NF = N + 1;
  constexpr bool AV = (MODE & M_AV) != 0, NORMS = (MODE & M_NORMS) != 0;
  constexpr bool COMPAT = (MODE & M_COMPAT) != 0, RK = (MODE & M_RK) != 0;
  constexpr bool DIVB = (MODE & M_DIVB) != 0;
  static_assert(!DIVB || (!RK && !COMPAT), "M_DIVB is a mode of its own");
  constexpr int NAUX = (MODE & M_AUX4) ? 4 : ((MODE & M_AUX1) ? 1 : 0);
  constexpr bool RSTORE = (MODE & M_RSTORE) != 0, VIS = (MODE & M_VIS) != 0;
  static_assert(N * T <= 64 && STAGE_THREADS == 128, "one wavefront per line family");
  extern __shared__ double lds[];
  const int tid = threadIdx.x;
  const int tile = tile_of_block(blockIdx.x, a.ntile, a.rev);   // tiles are numbered in launch order
  // one family per wavefront (wave 0: xi-lines, wave 1: eta-lines), so the
  // branches that depend on the family (in-tile or gathered neighbour
  // traces, the points step) do not diverge inside a wave
  const int fam = tid >> 6;            // 0: xi-line (row js), 1: eta-line (column is)
  const int line = (tid & 63) / T;     // N lines of T cells; the rest of the wave idles
  const int ci = (tid & 63) % T;
  const int cell = tile * T + ci;

  // ---- loads that depend on no other load, issued first: the face-table
  //      words of the line's two end faces (the neighbour gathers depend on
  //      them; every slot of a tile has them, so no lane waits for the
  //      tile's valid-cell count first) and the tile's state, which one wave
  //      stages into LDS in 16-byte loads (one contiguous block; all loads
  //      are issued before the first LDS store), the wave of the line family
  //      with fewer neighbours in other tiles, so that the two waves issue
  //      about as many vector loads before the barrier ---------------------------
  int nb0, nbN, info0, infoN;
  face_words<N>(a, tile, ci, fam, line, 0, nb0, info0);
  face_words<N>(a, tile, ci, fam, line, N, nbN, infoN);
  // the tile's member (0 for one run) and valid cells (its first slots):
  // loaded now, waited for only after the staging
  const int ti = a.tinfo[tile];
  if (fam == a.sfam) {
    constexpr int NV2 = 9 * N2 * T / 2;                  // double2 per tile
    constexpr int NSW = (NV2 + 63) / 64;                 // rounds of one wave
    static_assert((G::QT * 8) % 16 == 0, "tiles start on 16-byte boundaries");
    const double2 *src = reinterpret_cast<const double2 *>(a.qin + (size_t)tile * G::QT);
    double2 *dst = reinterpret_cast<double2 *>(lds + LD::Q0);
    const int l = tid & 63;
    // rounds k < NSW-1 are full for every lane, so only the last one is
    // guarded and the copy stays one basic block
    double2 st[NSW];
#pragma unroll
    for (int k = 0; k < NSW; ++k) {
      const int i = l + k * 64;
      st[k] = src[(k < NSW - 1 || i < NV2) ? i : 0];
    }
#pragma unroll
    for (int k = 0; k < NSW - 1; ++k) dst[l + k * 64] = st[k];
    if (l + (NSW - 1) * 64 < NV2) dst[l + (NSW - 1) * 64] = st[NSW - 1];
  }
  const int mem = ti >> 6, nvt = ti & 63;
  const bool valid = line < N && ci < nvt;
  const double *mt = a.met + (size_t)tile * G::NM * T + ci;
  const int p0 = fam ? line : N * line;      // point of s = 0 on this line
  const int ps = fam ? N : 1;                // point stride along the line
  const int mx = fam ? G::S2X : G::S1X, my = fam ? G::S2Y : G::S1Y;
  const double *qs = lds + LD::Q0;           // the staged tile (qo layout)
  // ---- neighbour traces at both end faces: gathered from memory for a
  //      neighbour in another tile, formed from the staged tile after the
  //      barrier for a neighbour in this one --------------------------------------
  double qn0[9], qnN[9];
  int lci0 = -1, lciN = -1;
  if (valid) {
    face_gather_pair<N, double>(a, ops, tile, fam, line, nb0, info0, nbN, infoN, qn0, qnN, lci0,
                                lciN);
  }
  // the metric vectors of the line's N+1 flux points, loaded while the
  // gathers are in flight
  double msx[NF], msy[NF];
  if (valid) {
#pragma unroll
    for (int r = 0; r < NF; ++r) {
      msx[r] = mt[(mx + r + NF * line) * T];
      msy[r] = mt[(my + r + NF * line) * T];
    }
  }
  // the member's lagged maxima, step and free stream (needed from the face
  // fluxes on)
  const double ge = a.gmax[2 * mem], gu = a.gmax[2 * mem + 1];
  const double cr = a.dtm ? a.cr * a.dtm[mem] : a.cr;   // b * dt of the member
  const Phys &ph = a.phm[mem];                            // the member's free stream
  const double c_h = sqrt(ge * (ge - gu));   // lagged GLM speed, compflux.f90:6
  double eig = 0.0, um = 0.0, nres = 0.0, ndiv = 0.0;
  __syncthreads();
  if (valid && lci0 >= 0) face_trace_lds<N, double>(lds + LD::Q0, ops, info0, fam, line, 0, lci0, qn0);
  if (valid && lciN >= 0) face_trace_lds<N, double>(lds + LD::Q0, ops, infoN, fam, line, N, lciN, qnN);

  double D[N][9];            // d(flux)/d(line coordinate) at the line's points
  double db[N], dj[N];       // same for the divB and Jz weights
  double qe[2][9];           // averaged face states (all-variable gradient)
  if (valid) {
    // ---- end points: own traces (compflux.f90:71-72,89-90), face fluxes
    double F0[9], wb0, wj0;
    {
      double t0[9];
      line_vals<N, double>(qs, p0, ps, ci, [&](int v, const double *vals) {
        t0[v] = line_trace<N, double>(vals, ops.L[0]);
      });
      const double sx = msx[0], sy = msy[0];
      double Qv[9];
      face_flux<N, double>(info0, fam, 0, ph, c_h, t0, qn0, sx, sy, F0, Qv, eig, um);
      // gradient weights of the averaged field (nablaQsp, compflux.f90:107-157:
      // divB = dBx/dx + dBy/dy, Jz = dBy/dx - dBx/dy)
      wb0 = Qv[5] * sx + Qv[6] * sy;
      wj0 = Qv[6] * sx - Qv[5] * sy;
      if (COMPAT) {
#pragma unroll
        for (int v = 0; v < 9; ++v) qe[0][v] = Qv[v];
      }
    }
    __builtin_amdgcn_sched_barrier(0);
    {
      double tN[9];
      line_vals<N, double>(qs, p0, ps, ci, [&](int v, const double *vals) {
        tN[v] = line_trace<N, double>(vals, ops.L[N]);
      });
      const double sx = msx[N], sy = msy[N];
      double FN[9], Qv[9];
      face_flux<N, double>(infoN, fam, N, ph, c_h, tN, qnN, sx, sy, FN, Qv, eig, um);
      const double wbN = Qv[5] * sx + Qv[6] * sy;
      const double wjN = Qv[6] * sx - Qv[5] * sy;
      if (COMPAT) {
#pragma unroll
        for (int v = 0; v < 9; ++v) qe[1][v] = Qv[v];
      }
      // derivative accumulators start with the two end points (compresid.f90:24-39)
#pragma unroll
      for (int s = 0; s < N; ++s) {
#pragma unroll
        for (int v = 0; v < 9; ++v) D[s][v] = fma(FN[v], ops.M[N][s], F0[v] * ops.M[0][s]);
        db[s] = fma(wbN, ops.M[N][s], wb0 * ops.M[0][s]);
        dj[s] = fma(wjN, ops.M[N][s], wj0 * ops.M[0][s]);
      }
    }
    // ---- interior flux points, two at a time: one pass over the line's
    //      values in LDS forms both traces, then each point's flux is folded
    //      into the accumulators in order (compflux.f90:53-104)
    auto interior_point = [&](int r, const double *tr) {
      const double sx = msx[r], sy = msy[r];
      double F[9];
      iflux_n<double>(tr, c_h, sx, sy, F, um);
      const double wb = tr[5] * sx + tr[6] * sy;
      const double wj = tr[6] * sx - tr[5] * sy;
#pragma unroll
      for (int s = 0; s < N; ++s) {
#pragma unroll
        for (int v = 0; v < 9; ++v) D[s][v] = fma(F[v], ops.M[r][s], D[s][v]);
        db[s] = fma(wb, ops.M[r][s], db[s]);
        dj[s] = fma(wj, ops.M[r][s], dj[s]);
      }
    };
#pragma unroll
    for (int r = 1; r < N; r += 2) {
      __builtin_amdgcn_sched_barrier(0);
      const bool two = r + 1 < N;
      double tr[9], tr2[9];
      line_vals<N, double>(qs, p0, ps, ci, [&](int v, const double *vals) {
        tr[v] = line_trace<N, double>(vals, ops.L[r]);
        if (two) tr2[v] = line_trace<N, double>(vals, ops.L[r + 1]);
      });
      interior_point(r, tr);
      if (two) interior_point(r + 1, tr2);
    }
    __builtin_amdgcn_sched_barrier(0);
  }

  // ---- the two families share the solution points' work by variable: the
  //      xi-line threads finish variables 0..VS-1 at their N points, the
  //      eta-line threads variables VS..8 at theirs, so both waves of the
  //      workgroup carry half of the point updates.  The RK register of the
  //      variables a thread finishes is loaded now, so that its latency hides
  //      behind the exchange (the output may be the same array, so these
  //      loads cannot move past its stores) -------------------------------------
  constexpr int VS = 4;
  static_assert(VS <= LD::XV && 9 - VS <= LD::XV, "exchange area holds one round");
  // the line family whose wave runs the AV sensor (1: the eta-lines)
  constexpr int AVF = 0;
  double ax[N][VS > 9 - VS ? VS : 9 - VS];
  const double *aux0 = a.aux[0] + (size_t)tile * G::QT;
  // 1/J at the thread's points, with them
  double jis[N];
  if (valid) {
#pragma unroll
    for (int s = 0; s < N; ++s) {
      const int is = fam ? line : s, js = fam ? s : line;
      jis[s] = mt[(G::JI + js + N * is) * T];
    }
  }
  if (RK && NAUX == 1 && valid) {
    if (fam == 0) {
#pragma unroll
      for (int s = 0; s < N; ++s) load_vars<N, 0, VS>(aux0, p0 + ps * s, ci, ax[s]);
    } else {
#pragma unroll
      for (int s = 0; s < N; ++s) load_vars<N, VS, 9>(aux0, p0 + ps * s, ci, ax[s]);
    }
  }

  // ---- exchange of the derivatives through LDS.  Round A: the eta-line
  //      threads hand variables 0..VS-1 and the divB weights to the xi-line
  //      threads.  Round B: the xi-line threads hand variables VS..8 and the
  //      summed divB weight to the eta-line threads.  The Jz weights (AV)
  //      go to the family that runs the sensor, in round B to the eta-lines
  //      (round A to the xi-lines).  Each LDS location is read and then
  //      rewritten by the same xi-line lane, so one area serves both rounds.
  //      Afterwards D of each thread holds the sum of both families for the
  //      variables it finishes, db the summed divB weight, dj (the sensor's
  //      family, AV) the summed Jz weight -----------------------------------------
  if (valid && fam == 1) {
#pragma unroll
    for (int s = 0; s < N; ++s) {
      const int p = p0 + ps * s;
      xput<N, VS>(lds + LD::XD, p, ci, D[s]);
      lds[LD::XB + p * T + ci] = db[s];
      if (AV && AVF == 0) lds[LD::XJ + p * T + ci] = dj[s];
    }
  }
  __syncthreads();
  if (valid && fam == 0) {
#pragma unroll
    for (int s = 0; s < N; ++s) {
      const int p = N * line + s;
      xadd<N, VS>(lds + LD::XD, p, ci, D[s]);
      db[s] += lds[LD::XB + p * T + ci];
      if (AV && AVF == 0) dj[s] += lds[LD::XJ + p * T + ci];
      xput<N, 9 - VS>(lds + LD::XD, p, ci, D[s] + VS);
      lds[LD::XB + p * T + ci] = db[s];
      if (AV && AVF == 1) lds[LD::XJ + p * T + ci] = dj[s];
    }
  }
  __syncthreads();
  if (valid && fam == 1) {
#pragma unroll
    for (int s = 0; s < N; ++s) {
      const int p = p0 + ps * s;
      xadd<N, 9 - VS>(lds + LD::XD, p, ci, D[s] + VS);
      db[s] = lds[LD::XB + p * T + ci];
      if (AV && AVF == 1) dj[s] += lds[LD::XJ + p * T + ci];
    }
  }

  // all-variable gradient nablaQs for the compresid output: the xi-line
  // threads store their part, the eta-line threads add theirs after the
  // barrier (each point has one writer of each kind)
  auto compat_gradient = [&](bool add) {
#pragma unroll
    for (int v = 0; v < 9; ++v) {
      double vals[N], qv[NF];
#pragma unroll
      for (int s = 0; s < N; ++s) vals[s] = qs[qo<N>(v, p0 + ps * s, ci)];
      qv[0] = qe[0][v];
      qv[N] = qe[1][v];
#pragma unroll
      for (int r = 1; r < N; ++r) qv[r] = line_trace<N, double>(vals, ops.L[r]);
#pragma unroll
      for (int s = 0; s < N; ++s) {
        double ax = 0.0, ay = 0.0;
#pragma unroll
        for (int r = 0; r < NF; ++r) {
          const double t = qv[r] * ops.M[r][s];
          ax = fma(t, mt[(mx + r + NF * line) * T], ax);
          ay = fma(t, mt[(my + r + NF * line) * T], ay);
        }
        const int p = p0 + ps * s;
        const int is = p % N, js = p / N;
        const double ji = mt[(G::JI + js + N * is) * T];
        double *o = a.nabla_f + v + 9 * 2 * ((size_t)p + (size_t)N2 * a.s2c[cell]);
        if (add) { o[0] += ax * ji; o[9] += ay * ji; }
        else { o[0] = ax * ji; o[9] = ay * ji; }
      }
    }
  };
  // ---- solution points: point (is, js) = (s, line) of a xi-line, (line, s)
  //      of an eta-line; variables V0..V1-1 of it -----------------------------------
  double rho_s[N], jz2_s[N];
  auto points = [&](auto v0c, auto v1c) {
    constexpr int V0 = decltype(v0c)::value, V1 = decltype(v1c)::value;
#pragma unroll
    for (int s = 0; s < N; ++s) {
      const int p = p0 + ps * s;
      const double ji = jis[s];
      double R[9], q[9];
      // the velocity and the field enter the sources of every variable;
      // the energy and psi only their own update
      load_vars<N, 0, 8>(qs, p, ci, q);
      if (V1 == 9) q[8] = qs[qo<N>(8, p, ci)];
#pragma unroll
      for (int v = V0; v < V1; ++v) {
        // resid = -(dF/dxi + dG/deta) / J   (compresid.f90:41,56), with the
        // viscous divergence added before the scaling (compresid.f90:43-46)
        if (VIS)
          R[v] = (a.visres[(size_t)tile * G::QT + qo<N>(v, p, ci)] - D[s][v]) * ji;
        else
          R[v] = -D[s][v] * ji;
      }
      const double divB = db[s] * ji;
      const double ri = rcp(q[0]);
      const double ux = q[1] * ri, uy = q[2] * ri, uz = q[3] * ri;
      // Powell source and GLM damping, compresid.f90:58-74
      if (V0 <= 1 && 1 < V1) R[1] -= divB * q[5];
      if (V0 <= 2 && 2 < V1) R[2] -= divB * q[6];
      if (V0 <= 3 && 3 < V1) R[3] -= divB * q[7];
      if (V0 <= 4 && 4 < V1) R[4] -= divB * (ux * q[5] + uy * q[6] + uz * q[7]);
      if (V0 <= 5 && 5 < V1) R[5] -= divB * ux;
      if (V0 <= 6 && 6 < V1) R[6] -= divB * uy;
      if (V0 <= 7 && 7 < V1) R[7] -= divB * uz;
      if (V0 <= 8 && 8 < V1) R[8] -= ph.alpha * q[8];
      if (AV && V0 == (AVF == 0 ? 0 : VS)) {
        // the point's Jz^2 replaces its Jz weight in LDS (this thread read
        // that slot in the exchange; the sensor reads it back per cell)
        const double jz = dj[s] * ji;
        rho_s[s] = q[0];
        jz2_s[s] = jz * jz;
        lds[LD::XJ + p * T + ci] = jz2_s[s];
      }
      if (NORMS) {
#pragma unroll
        for (int v = V0; v < V1; ++v) nres = fma(R[v], R[v], nres);
        if (V0 == 0) ndiv = fma(divB, divB, ndiv);
      }
      if (COMPAT) {
        const size_t pf = (size_t)p + (size_t)N2 * a.s2c[cell];
#pragma unroll
        for (int v = V0; v < V1; ++v) a.resid_f[v + 9 * pf] = R[v];
      }
      if (DIVB && V0 == 0) a.divb[(size_t)tile * N2 * T + (size_t)p * T + ci] = divB;
      if (RK) {
        const size_t off = (size_t)tile * G::QT;
        // several registers (FiveStageRk's last stage): their weighted sum,
        // formed before the point's stores (the output may be one of them)
        double xs[9];
        if (NAUX > 1) {
#pragma unroll
          for (int v = V0; v < V1; ++v) xs[v] = 0.0;
#pragma unroll
          for (int i = 0; i < NAUX; ++i) {
            double xv[9];
            load_vars<N, V0, V1>(a.aux[i] + off, p, ci, xv);
#pragma unroll
            for (int v = V0; v < V1; ++v) xs[v] = fma(a.ca[i], xv[v - V0], xs[v]);
            __builtin_amdgcn_sched_barrier(0);   // one register's loads in flight at a time
          }
        }
        double out[9];
#pragma unroll
        for (int v = V0; v < V1; ++v) {
          double acc = a.cin * q[v];
          if (NAUX == 1) acc = fma(a.ca[0], ax[s][v - V0], acc);
          if (NAUX > 1) acc += xs[v];
          out[v - V0] = fma(cr, R[v], acc);
        }
        store_vars<N, V0, V1>(a.qout + off, p, ci, out);
        if (RSTORE) store_vars<N, V0, V1>(a.rstore + off, p, ci, R + V0);
      }
    }
  };
  if (valid) {
    if (fam == 0) points(std::integral_constant<int, 0>{}, std::integral_constant<int, VS>{});
    else points(std::integral_constant<int, VS>{}, std::integral_constant<int, 9>{});
  }
  // (after the points, when the derivative accumulators are free)
  if (COMPAT) {
    if (valid && fam == 0) compat_gradient(false);
    __syncthreads();
    if (valid && fam == 1) compat_gradient(true);
  }

  // ---- AV smoothness sensor (AV.f90:14-83), per cell, by the wave of the
  //      family AVF (the xi-lines, which finish fewer variables in the points
  //      step), before the reduction barrier so that it overlaps the other
  //      wave's points: rho from the staged state, Jz^2 from the points step
  //      in the Jz area, which then takes the line sums (only this wave uses
  //      it after the exchange) --------------------------------------------------------
  if (AV && fam == AVF)
    av_sensor_lines<N, T, AVF == 0, 2>(a, lds + LD::Q0 + qo<N>(0, 0, 0), lds + LD::XJ,
                                       lds + LD::XJ, rho_s, jz2_s, line, ci, cell, valid);

  // ---- block reductions: running maxima (getrusanov.f90:116-123,188-191) and
  //      norms (iterations.f90:37-46) ----------------------------------------------
  {
    constexpr int NW = STAGE_THREADS / 64;
    const int w = tid >> 6, l = tid & 63;
    eig = wave_max(eig);
    um = wave_max(um);
    if (NORMS) {
      nres = wave_sum(nres);
      ndiv = wave_sum(ndiv);
    }
    if (l == 0) {
      lds[LD::RED + w] = eig;
      lds[LD::RED + NW + w] = um;
      lds[LD::RED + 2 * NW + w] = nres;
      lds[LD::RED + 3 * NW + w] = ndiv;
    }
    __syncthreads();
    if (tid == 0) {
      double e = lds[LD::RED], u = lds[LD::RED + NW];
      double s1 = lds[LD::RED + 2 * NW], s2 = lds[LD::RED + 3 * NW];
#pragma unroll
      for (int k = 1; k < NW; ++k) {
        e = lds[LD::RED + k] > e ? lds[LD::RED + k] : e;
        u = lds[LD::RED + NW + k] > u ? lds[LD::RED + NW + k] : u;
        s1 += lds[LD::RED + 2 * NW + k];
        s2 += lds[LD::RED + 3 * NW + k];
      }
      unsigned long long *sl = a.slots + 2 * ((size_t)mem * NSLOT + blockIdx.x % NSLOT);
      atomicMax(sl, (unsigned long long)__double_as_longlong(e));
      atomicMax(sl + 1, (unsigned long long)__double_as_longlong(u));
      if (NORMS) {
        a.normbuf[2 * tile] = s1;
        a.normbuf[2 * tile + 1] = s2;
      }
    }
  }

}

}  // namespace sd2d
