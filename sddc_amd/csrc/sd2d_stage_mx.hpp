// sd2d_stage_mx.hpp - the mixed-precision form of the stage kernel
// (precision SD2D_MIXED).
//
// Same one-thread-per-grid-line structure as stage_kernel, with the split
// of precisions the mixed form is defined by:
//   FP32   interpolation to the flux points, the interior fluxes, the
//          Rusanov face fluxes and boundary ghosts, the line derivatives
//          and the divB / Jz weights (operators and metrics in FP32);
//   FP64   the state in HBM, the pressure recovery inside the fluxes
//          (E minus kinetic minus magnetic energy, formed from the FP32
//          traces and rounded after the cancellation), the Jacobian scaling
//          of the residual, the Powell and GLM sources, the RK stage
//          combination, the norms and the AV sensor.
// The tile's state is staged in LDS as FP32 for the traces and stays there
// through the points step, whose Powell source reads the velocity and the
// field from it (FP64 arithmetic on the FP32-rounded values, the precision
// of the FP32 divB they multiply); from memory the points step reads only
// the FP64 values of the variables the thread updates, for the RK
// combination (one point ahead of their use).  The derivatives are handed
// between the two line families, which split the points' work by
// variable, through an area of their own.  About 1.6 KB of LDS per cell and
// roughly half the registers of the FP64 kernel, so more workgroups share
// a CU.
#pragma once
#include "sd2d_stage.hpp"

namespace sd2d {

// FP32 copies of the operators (rounded from the FP64 Lmat/Mmat).
struct OpsF {
  float L[NMAX + 1][NMAX];
  float M[NMAX + 1][NMAX];
};

template <int N, int MODE> struct StageLdsMx {
  static constexpr int T = Geo<N>::CPW, N2 = N * N, NL = N - 1;
  static constexpr bool AV = (MODE & M_AV) != 0;
  // in floats: the staged state (qo layout); the flux derivatives of up to
  // XV variables (xo layout) handed from one family to the other (eta -> xi:
  // variables 0-3, then xi -> eta: 4-8); the divB and Jz derivatives [p][ci]
  static constexpr int XV = 5;
  static constexpr int QF = 0;
  static constexpr int XD = 9 * N2 * T;
  static constexpr int XB = XD + XV * N2 * T;
  static constexpr int XJ = XB + N2 * T;
  static constexpr int FLOATS = XJ + (AV ? N2 * T : 0);
  // in doubles, after the float area: reduction slots, and the AV sensor's
  // FP64 rho and Jz^2 [p][ci] and row sums [2][line][ci]
  static constexpr int D0 = (FLOATS + 1) / 2;
  static constexpr int RED = D0;
  static constexpr int A_RHO = RED + 4 * (STAGE_THREADS / 64);
  static constexpr int A_JZ = A_RHO + N2 * T;
  static constexpr int A_ROW = A_JZ + N2 * T;
  static constexpr int TOTAL_D = AV ? A_ROW + 2 * N * T : A_RHO;
};

template <int N, int MODE>
constexpr size_t stage_mx_lds_bytes() { return sizeof(double) * StageLdsMx<N, MODE>::TOTAL_D; }

// three waves per SIMD for the common stages; the rarer ones (AV sensor,
// norms, FiveStageRk's residual store and four registers) carry more live
// values and keep two
template <int N, int MODE>
__global__ __launch_bounds__(STAGE_THREADS,
                             (MODE & (M_AV | M_NORMS | M_AUX4 | M_RSTORE)) ? 2 : 3)
void stage_kernel_mx(StageArgs a, OpsF ops, const float *__restrict__ metf) {
  using G = Geo<N>;
  using LD = StageLdsMx<N, MODE>;
  constexpr int T = G::CPW, N2 = G::N2, NF = N + 1;
  constexpr bool AV = (MODE & M_AV) != 0, NORMS = (MODE & M_NORMS) != 0;
  constexpr bool RK = (MODE & M_RK) != 0;
  constexpr int NAUX = (MODE & M_AUX4) ? 4 : ((MODE & M_AUX1) ? 1 : 0);
  constexpr bool RSTORE = (MODE & M_RSTORE) != 0;
  static_assert(RK && (MODE & M_COMPAT) == 0, "the mixed kernel is the stepping kernel");
  static_assert(N * T <= 64 && STAGE_THREADS == 128, "one wavefront per line family");
  extern __shared__ double lds[];
  float *ldf = reinterpret_cast<float *>(lds);
  const int tid = threadIdx.x;
  const int tile = tile_of_block(blockIdx.x, a.ntile, a.rev);   // tiles are numbered in launch order
  // one family per wavefront (wave 0: xi-lines, wave 1: eta-lines), so the
  // branches that depend on the family (in-tile or gathered neighbour
  // traces, the points step) do not diverge inside a wave
  const int fam = tid >> 6;            // 0: xi-line (row js), 1: eta-line (column is)
  const int line = (tid & 63) / T;     // N lines of T cells; the rest of the wave idles
  const int ci = (tid & 63) % T;
  // ---- loads that depend on no other load, issued first (as in
  //      stage_kernel): the face-table words of the line's two end faces,
  //      the tile's info (waited for only after the staging), and the
  //      tile's state -> LDS as FP32, by the wave of the line family with
  //      fewer neighbours in other tiles ----------------------------------------------
  int nb0, nbN, info0, infoN;
  face_words<N>(a, tile, ci, fam, line, 0, nb0, info0);
  face_words<N>(a, tile, ci, fam, line, N, nbN, infoN);
  // the tile's member (0 for one run) and valid cells (its first slots)
  const int ti = a.tinfo[tile];
  if (fam == a.sfam) {
    constexpr int NV2 = 9 * N2 * T / 2;
    constexpr int NSW = (NV2 + 63) / 64;
    const double2 *src = reinterpret_cast<const double2 *>(a.qin + (size_t)tile * G::QT);
    float2 *dst = reinterpret_cast<float2 *>(ldf + LD::QF);
    const int l = tid & 63;
    double2 st[NSW];
#pragma unroll
    for (int k = 0; k < NSW; ++k) {
      const int i = l + k * 64;
      st[k] = src[(k < NSW - 1 || i < NV2) ? i : 0];
    }
#pragma unroll
    for (int k = 0; k < NSW; ++k) {
      const int i = l + k * 64;
      if (k < NSW - 1 || i < NV2) dst[i] = make_float2(float(st[k].x), float(st[k].y));
    }
  }
  const int mem = ti >> 6, nvt = ti & 63;
  const bool valid = line < N && ci < nvt;
  const float *mtf = metf + (size_t)tile * G::NM * T + ci;
  const double *mt = a.met + (size_t)tile * G::NM * T + ci;
  const int p0 = fam ? line : N * line;
  const int ps = fam ? N : 1;
  const int mx = fam ? G::S2X : G::S1X, my = fam ? G::S2Y : G::S1Y;
  const float *qsf = ldf + LD::QF;            // the staged tile (qo layout, FP32)
  // ---- neighbour traces at both end faces (memory or the staged tile) -------
  float qn0[9], qnN[9];
  int lci0 = -1, lciN = -1;
  if (valid) {
    face_gather_pair<N, float>(a, ops, tile, fam, line, nb0, info0, nbN, infoN, qn0, qnN, lci0,
                               lciN);
  }
  // the metric vectors of the line's N+1 flux points, loaded while the
  // gathers are in flight
  float msx[NF], msy[NF];
  if (valid) {
#pragma unroll
    for (int r = 0; r < NF; ++r) {
      msx[r] = mtf[(mx + r + NF * line) * T];
      msy[r] = mtf[(my + r + NF * line) * T];
    }
  }
  // the member's lagged maxima, step and free stream (needed from the face
  // fluxes on)
  const double ge = a.gmax[2 * mem], gu = a.gmax[2 * mem + 1];
  const double cr = a.dtm ? a.cr * a.dtm[mem] : a.cr;   // b * dt of the member
  const Phys &ph = a.phm[mem];                            // the member's free stream
  const float c_h = float(sqrt(ge * (ge - gu)));   // compflux.f90:6
  float eig = 0.f, um = 0.f;
  double nres = 0.0, ndiv = 0.0;
  __syncthreads();
  if (valid && lci0 >= 0) face_trace_lds<N, float>(ldf + LD::QF, ops, info0, fam, line, 0, lci0, qn0);
  if (valid && lciN >= 0) face_trace_lds<N, float>(ldf + LD::QF, ops, infoN, fam, line, N, lciN, qnN);

  float D[N][9], db[N], dj[N];
  if (valid) {
    float F0[9], wb0, wj0;
    {
      float t0[9];
      line_vals<N, float>(qsf, p0, ps, ci, [&](int v, const float *vals) {
        t0[v] = line_trace<N, float>(vals, ops.L[0]);
      });
      const float sx = msx[0], sy = msy[0];
      float Qv[9];
      face_flux<N, float>(info0, fam, 0, ph, c_h, t0, qn0, sx, sy, F0, Qv, eig, um);
      wb0 = Qv[5] * sx + Qv[6] * sy;
      wj0 = Qv[6] * sx - Qv[5] * sy;
    }
    __builtin_amdgcn_sched_barrier(0);
    {
      float tN[9];
      line_vals<N, float>(qsf, p0, ps, ci, [&](int v, const float *vals) {
        tN[v] = line_trace<N, float>(vals, ops.L[N]);
      });
      const float sx = msx[N], sy = msy[N];
      float FN[9], Qv[9];
      face_flux<N, float>(infoN, fam, N, ph, c_h, tN, qnN, sx, sy, FN, Qv, eig, um);
      const float wbN = Qv[5] * sx + Qv[6] * sy;
      const float wjN = Qv[6] * sx - Qv[5] * sy;
#pragma unroll
      for (int s = 0; s < N; ++s) {
#pragma unroll
        for (int v = 0; v < 9; ++v) D[s][v] = fmaf(FN[v], ops.M[N][s], F0[v] * ops.M[0][s]);
        db[s] = fmaf(wbN, ops.M[N][s], wb0 * ops.M[0][s]);
        dj[s] = fmaf(wjN, ops.M[N][s], wj0 * ops.M[0][s]);
      }
    }
    // interior flux points two at a time: one pass over the staged line
    // values forms both traces, then each point is folded in, in order
    auto interior_point = [&](int r, const float *tr) {
      const float sx = msx[r], sy = msy[r];
      float F[9];
      iflux_n<float>(tr, c_h, sx, sy, F, um);
      const float wb = tr[5] * sx + tr[6] * sy;
      const float wj = tr[6] * sx - tr[5] * sy;
#pragma unroll
      for (int s = 0; s < N; ++s) {
#pragma unroll
        for (int v = 0; v < 9; ++v) D[s][v] = fmaf(F[v], ops.M[r][s], D[s][v]);
        db[s] = fmaf(wb, ops.M[r][s], db[s]);
        dj[s] = fmaf(wj, ops.M[r][s], dj[s]);
      }
    };
#pragma unroll
    for (int r = 1; r < N; r += 2) {
      __builtin_amdgcn_sched_barrier(0);
      const bool two = r + 1 < N;
      float tr[9], tr2[9];
      line_vals<N, float>(qsf, p0, ps, ci, [&](int v, const float *vals) {
        tr[v] = line_trace<N, float>(vals, ops.L[r]);
        if (two) tr2[v] = line_trace<N, float>(vals, ops.L[r + 1]);
      });
      interior_point(r, tr);
      if (two) interior_point(r + 1, tr2);
    }
    __builtin_amdgcn_sched_barrier(0);
  }

  // ---- the two families share the solution points' work by variable, as in
  //      stage_kernel: xi-lines finish variables 0..VS-1, eta-lines VS..8.
  //      The FP64 values of a thread's own variables and its RK register at
  //      its first point are loaded now; every later point's one point ahead
  //      of their use, before the stores of the point in hand ----------------------
  constexpr int VS = 4, VM = VS > 9 - VS ? VS : 9 - VS;
  static_assert(VS <= LD::XV && 9 - VS <= LD::XV, "exchange area holds one round");
  const double *qg = a.qin + (size_t)tile * G::QT;
  const double *aux0 = (NAUX == 1 ? a.aux[0] : a.qin) + (size_t)tile * G::QT;
  double qn[9], xn[VM], jis[N];
  if (valid) {
#pragma unroll
    for (int s = 0; s < N; ++s) {   // 1/J at the thread's points
      const int is = fam ? line : s, js = fam ? s : line;
      jis[s] = mt[(G::JI + js + N * is) * T];
    }
    if (fam == 0) {
      load_vars<N, 0, VS>(qg, p0, ci, qn);
      if (NAUX == 1) load_vars<N, 0, VS>(aux0, p0, ci, xn);
    } else {
      load_vars<N, VS, 9>(qg, p0, ci, qn + VS);
      if (NAUX == 1) load_vars<N, VS, 9>(aux0, p0, ci, xn);
    }
  }
  // ---- exchange through the derivative area (one slot per variable):
  //      round A, eta -> xi: variables 0..VS-1 and the divB (Jz) weights;
  //      round B, xi -> eta: variables VS..8 and the summed divB weight.  The
  //      staged state stays as it is for the sources of the points step ----
  if (valid && fam == 1) {
#pragma unroll
    for (int s = 0; s < N; ++s) {
      const int p = p0 + ps * s;
      xput<N, VS>(ldf + LD::XD, p, ci, D[s]);
      ldf[LD::XB + p * T + ci] = db[s];
      if (AV) ldf[LD::XJ + p * T + ci] = dj[s];
    }
  }
  __syncthreads();
  if (valid && fam == 0) {
#pragma unroll
    for (int s = 0; s < N; ++s) {
      const int p = N * line + s;
      xadd<N, VS>(ldf + LD::XD, p, ci, D[s]);
      db[s] += ldf[LD::XB + p * T + ci];
      if (AV) dj[s] += ldf[LD::XJ + p * T + ci];
      // round B into the slots of this point just read (only this thread
      // reads or writes the slots of its points)
      xput<N, 9 - VS>(ldf + LD::XD, p, ci, D[s] + VS);
      ldf[LD::XB + p * T + ci] = db[s];
    }
  }
  __syncthreads();
  if (valid && fam == 1) {
#pragma unroll
    for (int s = 0; s < N; ++s) {
      const int p = p0 + ps * s;
      xadd<N, 9 - VS>(ldf + LD::XD, p, ci, D[s] + VS);
      db[s] = ldf[LD::XB + p * T + ci];
    }
  }

  // ---- solution points, in FP64: point (is, js) = (s, line) of a xi-line,
  //      (line, s) of an eta-line; variables V0..V1-1 of it ----------------------
  double rho_s[N], jz2_s[N];
  auto points = [&](auto v0c, auto v1c) {
    constexpr int V0 = decltype(v0c)::value, V1 = decltype(v1c)::value;
#pragma unroll
    for (int s = 0; s < N; ++s) {
      const int p = p0 + ps * s;
      const double ji = jis[s];
      double R[9], q[9], xc[VM];
#pragma unroll
      for (int v = V0; v < V1; ++v) q[v] = qn[v];
      // the velocity and the field of the sources, from the staged state
      const float2 f01 = *reinterpret_cast<const float2 *>(qsf + qo<N>(0, p, ci));
      const float2 f23 = *reinterpret_cast<const float2 *>(qsf + qo<N>(2, p, ci));
      const float2 f45 = *reinterpret_cast<const float2 *>(qsf + qo<N>(4, p, ci));
      const float2 f67 = *reinterpret_cast<const float2 *>(qsf + qo<N>(6, p, ci));
      const double w0 = f01.x, w1 = f01.y, w2 = f23.x, w3 = f23.y;
      const double w5 = f45.y, w6 = f67.x, w7 = f67.y;
#pragma unroll
      for (int v = 0; v < V1 - V0; ++v) xc[v] = xn[v];
#pragma unroll
      for (int v = V0; v < V1; ++v) R[v] = -double(D[s][v]) * ji;   // compresid.f90:41,56
      const double divB = double(db[s]) * ji;
      if (s + 1 < N) {
        load_vars<N, V0, V1>(qg, p + ps, ci, qn + V0);
        if (NAUX == 1) load_vars<N, V0, V1>(aux0, p + ps, ci, xn);
      }
      const double ri = 1.0 / w0;
      const double ux = w1 * ri, uy = w2 * ri, uz = w3 * ri;
      // Powell source and GLM damping, compresid.f90:58-74
      if (V0 <= 1 && 1 < V1) R[1] -= divB * w5;
      if (V0 <= 2 && 2 < V1) R[2] -= divB * w6;
      if (V0 <= 3 && 3 < V1) R[3] -= divB * w7;
      if (V0 <= 4 && 4 < V1) R[4] -= divB * (ux * w5 + uy * w6 + uz * w7);
      if (V0 <= 5 && 5 < V1) R[5] -= divB * ux;
      if (V0 <= 6 && 6 < V1) R[6] -= divB * uy;
      if (V0 <= 7 && 7 < V1) R[7] -= divB * uz;
      if (V0 <= 8 && 8 < V1) R[8] -= ph.alpha * q[8];
      if (AV && V0 == 0) {
        const double jz = double(dj[s]) * ji;
        rho_s[s] = q[0];
        jz2_s[s] = jz * jz;
        lds[LD::A_RHO + p * T + ci] = rho_s[s];
        lds[LD::A_JZ + p * T + ci] = jz2_s[s];
      }
      if (NORMS) {
#pragma unroll
        for (int v = V0; v < V1; ++v) nres = fma(R[v], R[v], nres);
        if (V0 == 0) ndiv = fma(divB, divB, ndiv);
      }
      const size_t off = (size_t)tile * G::QT;
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
        if (NAUX == 1) acc = fma(a.ca[0], xc[v - V0], acc);
        if (NAUX > 1) acc += xs[v];
        out[v - V0] = fma(cr, R[v], acc);
      }
      store_vars<N, V0, V1>(a.qout + off, p, ci, out);
      if (RSTORE) store_vars<N, V0, V1>(a.rstore + off, p, ci, R + V0);
    }
  };
  if (valid) {
    if (fam == 0) points(std::integral_constant<int, 0>{}, std::integral_constant<int, VS>{});
    else points(std::integral_constant<int, VS>{}, std::integral_constant<int, 9>{});
  }

  // ---- block reductions -----------------------------------------------------------
  {
    constexpr int NW = STAGE_THREADS / 64;
    const int w = tid >> 6, l = tid & 63;
    double e = wave_max(double(eig)), u = wave_max(double(um));
    if (NORMS) {
      nres = wave_sum(nres);
      ndiv = wave_sum(ndiv);
    }
    if (l == 0) {
      lds[LD::RED + w] = e;
      lds[LD::RED + NW + w] = u;
      if (NORMS) {
        lds[LD::RED + 2 * NW + w] = nres;
        lds[LD::RED + 3 * NW + w] = ndiv;
      }
    }
    __syncthreads();
    if (tid == 0) {
      double em = lds[LD::RED], ummax = lds[LD::RED + NW];
      double s1 = NORMS ? lds[LD::RED + 2 * NW] : 0.0, s2 = NORMS ? lds[LD::RED + 3 * NW] : 0.0;
#pragma unroll
      for (int k = 1; k < NW; ++k) {
        em = lds[LD::RED + k] > em ? lds[LD::RED + k] : em;
        ummax = lds[LD::RED + NW + k] > ummax ? lds[LD::RED + NW + k] : ummax;
        if (NORMS) {
          s1 += lds[LD::RED + 2 * NW + k];
          s2 += lds[LD::RED + 3 * NW + k];
        }
      }
      unsigned long long *sl = a.slots + 2 * ((size_t)mem * NSLOT + blockIdx.x % NSLOT);
      atomicMax(sl, (unsigned long long)__double_as_longlong(em));
      atomicMax(sl + 1, (unsigned long long)__double_as_longlong(ummax));
      if (NORMS) {
        a.normbuf[2 * tile] = s1;
        a.normbuf[2 * tile + 1] = s2;
      }
    }
  }

  // ---- AV smoothness sensor (AV.f90:14-83), in FP64, per cell, by the
  //      xi-line wave from the FP64 rho and Jz^2 of the points step -----------------
  if (AV && fam == 0)
    av_sensor_lines<N, T, true, 1>(a, lds + LD::A_RHO, lds + LD::A_JZ, lds + LD::A_ROW, rho_s, jz2_s, line,
                         ci, tile * T + ci, valid);
}

}  // namespace sd2d
