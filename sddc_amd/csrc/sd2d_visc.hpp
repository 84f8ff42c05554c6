// sd2d_visc.hpp - the viscous and resistive fluxes of the residual
// (vismode = 1 with mu or eta0 non-zero), in two kernels that run before
// the stage kernel of every residual evaluation:
//
//   vis_grad_kernel   BR1 gradient of all nine variables at the solution
//                     points from the averaged face states (nablaQsp,
//                     compflux.f90:107-157) and, on a mesh with walls, the
//                     computational derivatives of the internal energy
//                     (comp_grad_internal_energy, compflux.f90:260-299);
//   vis_flux_kernel   the gradient at every flux point (interpolated inside
//                     the cell, nablaQ_interior_FP, compflux.f90:159-196;
//                     averaged with the neighbour's trace on interior and
//                     periodic faces, interfaceflux_all.f90:114-195 and
//                     cyclocflux_all.f90:114-204; boundary gradient ghosts
//                     with the adiabatic-wall energy gradient,
//                     bcflux.f90:380-666), the viscous/resistive flux normal
//                     to the point's metric vector (getVfluxvectors,
//                     getrusanov.f90:197-294; comp_viscous_flux,
//                     compflux.f90:198-258) and its divergence dFv + dGv at
//                     the solution points, which the stage kernel adds to
//                     the residual before the Jacobian scaling
//                     (compresid.f90:43-46).
//
// Both use the stage kernel's thread layout: one 128-thread workgroup per
// tile of cells, one thread per grid line.  The gradient (18 values per
// point) lives in HBM between the two kernels in the tile layout
// [tile][(v*2 + d)*N^2 + p][cell], twice the size of a state.
//
// The artificial-viscosity part of comp_viscous_flux adds
// getVfluxvectors_AV with mu_art = eta_art = 0 (getrusanov.f90:325-328),
// which is zero for finite gradients, and is not formed.
#pragma once
#include "sd2d_stage.hpp"

namespace sd2d {

constexpr double LAMBDA_V = -2.0 / 3.0;   // setup2d.f90:69
constexpr double PRANDTL = 0.72;          // setup2d.f90:70

struct VisArgs {
  const double *qin;       // state (tile layout)
  const int *ftab;         // per-cell face table (tile layout)
  const double *met;       // metrics (tile layout)
  const double *metw;      // other-family metrics at the line ends, for walls
  double *gq;              // gradient [tile][(v*2+d)*N2 + p][ci]
  double *dei;             // d(e)/dxi, d(e)/deta [tile][k*N2 + p][ci] (walls)
  double *visres;          // dFv + dGv, state tile layout (qo)
  const int *s2c;          // slot -> mesh cell, -1 for a padding slot
  double mu, eta0;
  int ncell, ntile, walls;
};

// Slots of the other-family metric table metw, per cell: for each family,
// line end (0: rfp = 0, 1: rfp = N), component and line, S1(2,c,rfp,js)
// on xi-lines and S2(1,c,is,rfp) on eta-lines (bcflux.f90:547-571).
template <int N> struct VisGeo {
  static constexpr int NW = 8 * N;
  static __host__ __device__ constexpr int slot(int fam, int end, int comp, int line) {
    return fam * 4 * N + end * 2 * N + comp * N + line;
  }
};

// Flux-point values of a tile in LDS: [fam][line][rfp][k][ci] for the
// averaged states (k = variable) or the viscous fluxes.
template <int N> struct VisLds {
  static constexpr int T = Geo<N>::CPW, NF = N + 1;
  static __device__ __forceinline__ int at(int fam, int line, int r, int k, int ci) {
    return (((fam * N + line) * NF + r) * 9 + k) * T + ci;
  }
  static constexpr int FP = 2 * N * NF * 9 * T;     // flux-point values
  static constexpr int EF = FP;                     // internal energy [fam][line][rfp][ci]
  static constexpr int TOTAL = FP + 2 * N * NF * T;
};

// The averaged state at end rfp of a line (BR1: interfaceflux_all.f90:79-91,
// bcflux.f90 with vismode = 1): own trace t, neighbour trace qn.
__device__ __forceinline__ void vis_end_state(int info, const Phys &ph, const double *t,
                                              const double *qn, double sx, double sy,
                                              double *Qv) {
  const int bc = (info >> 3) & 7;
  if (bc == FT_INTERIOR) {
    const int left = (info >> 2) & 1;
#pragma unroll
    for (int v = 0; v < 9; ++v) {
      const double QL = left ? t[v] : qn[v], QR = left ? qn[v] : t[v];
      Qv[v] = 0.5 * QL + 0.5 * QR;
    }
  } else {
    double QR[9];
    bc_ghost<double>(bc, t, sx, sy, ph, QR);
#pragma unroll
    for (int v = 0; v < 9; ++v) Qv[v] = 0.5 * t[v] + 0.5 * QR[v];
  }
}

// Internal energy of a flux-point state as comp_grad_internal_energy forms
// it (compflux.f90:275-280).
__device__ __forceinline__ double enei_of(const double *q) {
  const double u = q[1] / q[0], v = q[2] / q[0], w = q[3] / q[0];
  return (((q[4] / q[0] - 0.5 * (u * u)) - 0.5 * (v * v)) - 0.5 * (w * w)) -
         0.5 * (((q[5] * q[5] + q[6] * q[6]) + q[7] * q[7]) + q[8] * q[8]) / q[0];
}

// Trace at flux point r of the line through point p0 (stride ps) of the
// nc-component field f (tile layout with component stride N^2*T) of the
// cell at f, components c0 .. c0+NC-1 -> out.
template <int N, int NC>
__device__ __forceinline__ void vis_trace(const double *f, int c0, int p0, int ps,
                                          const double *Lr, double *out) {
  constexpr int T = Geo<N>::CPW, N2 = Geo<N>::N2;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    double vals[N];
#pragma unroll
    for (int s = 0; s < N; ++s) vals[s] = f[(size_t)((c0 + c) * N2 + p0 + ps * s) * T];
    out[c] = line_trace<N, double>(vals, Lr);
    __builtin_amdgcn_sched_barrier(0);   // one component's loads in flight at a time
  }
}

// Trace of the 9 state variables at flux point row Lr of the line through
// point p0 (stride ps) of slot ci of the state tile q (qo layout).
template <int N>
__device__ __forceinline__ void vis_trace_q(const double *q, int ci, int p0, int ps,
                                            const double *Lr, double *out) {
#pragma unroll
  for (int v = 0; v < 9; ++v) {
    double vals[N];
#pragma unroll
    for (int s = 0; s < N; ++s) vals[s] = q[qo<N>(v, p0 + ps * s, ci)];
    out[v] = line_trace<N, double>(vals, Lr);
    __builtin_amdgcn_sched_barrier(0);   // one variable's loads in flight at a time
  }
}

// Averaged state at flux point r of a line (compflux.f90:71-76,89-94 in the
// cell, the face/boundary average at the two ends), all read from memory.
template <int N>
__device__ __forceinline__ void vis_point_state(const VisArgs &va, const Ops &ops,
                                                const Phys &ph, int tile, int ci, int fam,
                                                int line, int r, double *Qv) {
  constexpr int T = Geo<N>::CPW, QT = Geo<N>::QT, NF = N + 1;
  const int p0 = fam ? line : N * line, ps = fam ? N : 1;
  double Lr[N];
#pragma unroll
  for (int s = 0; s < N; ++s) Lr[s] = ops.L[r][s];
  double t[9];
  vis_trace_q<N>(va.qin + (size_t)tile * QT, ci, p0, ps, Lr, t);
  if (r != 0 && r != N) {
#pragma unroll
    for (int v = 0; v < 9; ++v) Qv[v] = t[v];
    return;
  }
  int k, nfp;
  end_face<N>(fam, line, r, k, nfp);
  const int *ft = va.ftab + ((size_t)tile * 8 + k * 2) * T + ci;
  const int nbase = ft[0], info = ft[T];
  double qn[9];
  if (((info >> 3) & 7) == FT_INTERIOR) {
    // the neighbour's line through the shared point (interfaceflux_all.f90:59-76)
    int np0, nst;
    bool endN;
    neighbour_line<N>(info & 3, nfp, np0, nst, endN);
    double Ln[N];
#pragma unroll
    for (int s = 0; s < N; ++s) Ln[s] = endN ? ops.L[N][s] : ops.L[0][s];
    const int nt = nbase / Geo<N>::QT;
    vis_trace_q<N>(va.qin + (size_t)nt * Geo<N>::QT, nbase - nt * Geo<N>::QT, np0, nst, Ln, qn);
  }
  const int mx = fam ? Geo<N>::S2X : Geo<N>::S1X, my = fam ? Geo<N>::S2Y : Geo<N>::S1Y;
  const double *mt = va.met + (size_t)tile * Geo<N>::NM * T + ci;
  vis_end_state(info, ph, t, qn, mt[(mx + r + NF * line) * T], mt[(my + r + NF * line) * T], Qv);
}

// nablaQsp (compflux.f90:107-157) and comp_grad_internal_energy
// (compflux.f90:260-299).  Every line thread forms the averaged states at
// its N+1 flux points into LDS; the xi-line threads then form, at their
// points, nablaQs = (dps + det)/J from both families' states in the
// reference's order, and the internal-energy derivatives.
template <int N>
__global__ __launch_bounds__(STAGE_THREADS) void vis_grad_kernel(VisArgs va, Ops ops, Phys ph) {
  using G = Geo<N>;
  using LD = VisLds<N>;
  constexpr int T = G::CPW, N2 = G::N2, NF = N + 1, NT = 2 * N * T;
  extern __shared__ double lds[];
  const int tid = threadIdx.x, tile = blockIdx.x;
  const int ci = tid % T, fl = tid / T;
  const int fam = fl >= N ? 1 : 0, line = fl - fam * N;
  const int cell = tile * T + ci;   // the slot
  const bool valid = tid < NT && va.s2c[cell] >= 0;
  if (valid) {
#pragma unroll 1
    for (int r = 0; r < NF; ++r) {
      double Qv[9];
      switch (r) {   // compile-time operator rows
#define SD2D_R(i) case i: if (i <= N) vis_point_state<N>(va, ops, ph, tile, ci, fam, line, i, Qv); break;
        SD2D_R(0) SD2D_R(1) SD2D_R(2) SD2D_R(3) SD2D_R(4) SD2D_R(5)
#undef SD2D_R
      }
#pragma unroll
      for (int v = 0; v < 9; ++v) lds[LD::at(fam, line, r, v, ci)] = Qv[v];
      if (va.walls) lds[LD::EF + ((fam * N + line) * NF + r) * T + ci] = enei_of(Qv);
    }
  }
  __syncthreads();
  if (!valid || fam != 0) return;
  const double *mt = va.met + (size_t)tile * G::NM * T + ci;
  double *gq = va.gq + (size_t)tile * 2 * G::QT + ci;
  const int js = line;
#pragma unroll
  for (int is = 0; is < N; ++is) {
    const int p = is + N * js;
    const double ji = mt[(G::JI + js + N * is) * T];
#pragma unroll
    for (int v = 0; v < 9; ++v) {
#pragma unroll
      for (int d = 0; d < 2; ++d) {
        // dps = dps + Qvfi*Mmat*S1(1,d), det = det + Qvfj*Mmat*S2(2,d)
        // (compflux.f90:128-150)
        double dps = 0.0, det = 0.0;
        const int s1 = d ? G::S1Y : G::S1X, s2 = d ? G::S2Y : G::S2X;
#pragma unroll
        for (int r = 0; r < NF; ++r)
          dps = dps + (lds[LD::at(0, js, r, v, ci)] * ops.M[r][is]) * mt[(s1 + r + NF * js) * T];
#pragma unroll
        for (int r = 0; r < NF; ++r)
          det = det + (lds[LD::at(1, is, r, v, ci)] * ops.M[r][js]) * mt[(s2 + r + NF * is) * T];
        gq[(size_t)((2 * v + d) * N2 + p) * T] = (dps + det) * ji;
      }
    }
    if (va.walls) {
      double dx = 0.0, de = 0.0;
#pragma unroll
      for (int r = 0; r < NF; ++r) dx = dx + lds[LD::EF + ((0 * N + js) * NF + r) * T + ci] * ops.M[r][is];
#pragma unroll
      for (int r = 0; r < NF; ++r) de = de + lds[LD::EF + ((1 * N + is) * NF + r) * T + ci] * ops.M[r][js];
      double *o = va.dei + (size_t)tile * 2 * N2 * T + ci;
      o[(size_t)p * T] = dx;
      o[(size_t)(N2 + p) * T] = de;
    }
  }
}

// getVfluxvectors (getrusanov.f90:197-294) contracted with the metric
// vector (sx, sy): Fn = Fv*sx + Gv*sy (compflux.f90:215-216,243-244).
// g[v] is d(q_v)/dx, g[9+v] is d(q_v)/dy.
__device__ __forceinline__ void vflux_n(const double *Qv, const double *g, double mu,
                                        double eta, double sx, double sy, double *Fn) {
  const double rho = Qv[0], rhoE = Qv[4];
  const double Bx = Qv[5], By = Qv[6], Bz = Qv[7], psi = Qv[8];
  const double u = Qv[1] / rho, v = Qv[2] / rho, w = Qv[3] / rho;
  const double nu = (u * u + v * v) + w * w;
  const double nb = (Bx * Bx + By * By) + Bz * Bz;
  const double inte = (((rhoE - 0.5 * nu * rho) - 0.5 * nb) - 0.5 * (psi * psi)) / rho;
  double du[2], dv[2], dw[2], de[2];
#pragma unroll
  for (int d = 0; d < 2; ++d) {
    const double *gd = g + 9 * d;
    const double dr = gd[0];
    du[d] = (gd[1] - dr * u) / rho;
    dv[d] = (gd[2] - dr * v) / rho;
    dw[d] = (gd[3] - dr * w) / rho;
    const double dke = 0.5 * nu * dr + rho * ((u * du[d] + v * dv[d]) + w * dw[d]);
    const double dpb = (Bx * gd[5] + By * gd[6]) + Bz * gd[7];
    const double dps = psi * gd[8];
    de[d] = ((((gd[4] - dke) - dpb) - dr * inte) - dps) / rho;
  }
  const double ux = du[0], uy = du[1], vx = dv[0], vy = dv[1], wx = dw[0], wy = dw[1];
  const double Bxy = g[9 + 5], Byx = g[6], Bzx = g[7], Bzy = g[9 + 7];
  double Fv[9], Gv[9];
  Fv[0] = 0.0;
  Fv[1] = mu * (2.0 * ux + LAMBDA_V * (ux + vy));
  Fv[2] = mu * (vx + uy);
  Fv[3] = mu * wx;
  Fv[4] = ((u * Fv[1] + v * Fv[2]) + w * Fv[3]) + mu * GAM * de[0] / PRANDTL;
  Gv[0] = 0.0;
  Gv[1] = mu * (vx + uy);
  Gv[2] = mu * (2.0 * vy + LAMBDA_V * (ux + vy));
  Gv[3] = mu * wy;
  Gv[4] = ((Gv[1] * u + Gv[2] * v) + Gv[3] * w) + mu * GAM * de[1] / PRANDTL;
  // resistive part: J = curl B in 2-D (getrusanov.f90:271-292)
  const double Jx = Bzy, Jy = -Bzx, Jz = Byx - Bxy;
  Fv[5] = 0.0;
  Fv[6] = eta * Jz;
  Fv[7] = -eta * Jy;
  Fv[4] = (Fv[4] + Fv[6] * By) + Fv[7] * Bz;
  Gv[5] = -eta * Jz;
  Gv[6] = 0.0;
  Gv[7] = eta * Jx;
  Gv[4] = (Gv[4] + Gv[5] * Bx) + Gv[7] * Bz;
  Fv[8] = 0.0;
  Gv[8] = 0.0;
#pragma unroll
  for (int k = 0; k < 9; ++k) Fn[k] = Fv[k] * sx + Gv[k] * sy;
}

// Gradient ghost of the adiabatic wall (bcflux.f90:484-571) for the wall
// trace Qo of the state, the own gradient traces gl (x: gl[0..8], y:
// gl[9..17]), the internal-energy derivative across the wall family dface
// (dedeta on xi-lines, dedxi on eta-lines) and the other-family metric
// (ox, oy).  Only the energy row differs from the own traces; the second
// direction reuses the first direction's magnetic term, as the reference
// does (bcflux.f90:563).
__device__ __forceinline__ void wall_grad_ghost(const double *Qo, const double *gl, double dface,
                                                double ox, double oy, double *gr) {
  const double ur = Qo[1] / Qo[0], vr = Qo[2] / Qo[0], wr = Qo[3] / Qo[0];
  const double inte = (Qo[4] / Qo[0] - 0.5 * ((ur * ur + vr * vr) + wr * wr)) -
                      0.5 * (((Qo[5] * Qo[5] + Qo[6] * Qo[6]) + Qo[7] * Qo[7]) + Qo[8] * Qo[8]) /
                          Qo[0];
#pragma unroll
  for (int k = 0; k < 18; ++k) gr[k] = gl[k];
  const double dBp = ((Qo[5] * gl[5] + Qo[6] * gl[6]) + Qo[7] * gl[7]) + Qo[8] * gl[8];
#pragma unroll
  for (int d = 0; d < 2; ++d) {
    const double *g = gl + 9 * d;
    const double dur = (g[1] - ur * g[0]) / Qo[0];
    const double dvr = (g[2] - vr * g[0]) / Qo[0];
    const double dwr = (g[3] - wr * g[0]) / Qo[0];
    const double dKE = 0.5 * ((ur * ur + vr * vr) + wr * wr) * g[0] +
                       Qo[0] * ((ur * dur + vr * dvr) + wr * dwr);
    gr[9 * d + 4] = (((-g[4] + 2.0 * inte * g[0]) + 2.0 * dKE) + dBp) +
                    2.0 * dface * (d == 0 ? ox : oy) * Qo[0];
  }
}

// The flux-point gradient at end rfp of a line: own trace gl; BR1 average
// with the neighbour's trace on interior and periodic faces
// (interfaceflux_all.f90:140-190), the ghost on boundaries (a copy, or the
// adiabatic wall: bcflux.f90:380-666).
template <int N>
__device__ __forceinline__ void vis_end_grad(const VisArgs &va, const Ops &ops, int tile,
                                             int ci, int fam, int line, int rfp,
                                             const double *gl, double *g) {
  constexpr int T = Geo<N>::CPW, N2 = Geo<N>::N2, QT = Geo<N>::QT;
  int k, nfp;
  end_face<N>(fam, line, rfp, k, nfp);
  const int *ft = va.ftab + ((size_t)tile * 8 + k * 2) * T + ci;
  const int nbase = ft[0], info = ft[T];
  const int bc = (info >> 3) & 7;
  if (bc == FT_INTERIOR) {
    const int nt = nbase / QT, nci = nbase - nt * QT;
    int p0, stride;
    bool endN;
    neighbour_line<N>(info & 3, nfp, p0, stride, endN);
    const double *gn = va.gq + (size_t)nt * 2 * QT + nci;
    const int left = (info >> 2) & 1;
    double Lr[N];
#pragma unroll
    for (int s = 0; s < N; ++s) Lr[s] = endN ? ops.L[N][s] : ops.L[0][s];
#pragma unroll
    for (int c = 0; c < 18; ++c) {
      const int v = c % 9, d = c / 9;
      double gr;
      vis_trace<N, 1>(gn, 2 * v + d, p0, stride, Lr, &gr);
      const double a = left ? gl[c] : gr, b = left ? gr : gl[c];
      g[c] = 0.5 * a + 0.5 * b;
    }
  } else if (bc == FT_WALL) {
    // own state trace and internal-energy derivative at the wall point
    const int p0 = fam ? line : N * line, ps = fam ? N : 1;
    double Lr[N];
#pragma unroll
    for (int s = 0; s < N; ++s) Lr[s] = ops.L[rfp][s];
    double Qo[9], dface;
    vis_trace_q<N>(va.qin + (size_t)tile * QT, ci, p0, ps, Lr, Qo);
    // d(e)/deta along a xi-line, d(e)/dxi along an eta-line
    vis_trace<N, 1>(va.dei + (size_t)tile * 2 * N2 * T + ci, 1 - fam, p0, ps, Lr, &dface);
    const double *mw = va.metw + (size_t)tile * VisGeo<N>::NW * T + ci;
    const int end = rfp == 0 ? 0 : 1;
    const double ox = mw[VisGeo<N>::slot(fam, end, 0, line) * T];
    const double oy = mw[VisGeo<N>::slot(fam, end, 1, line) * T];
    double gr[18];
    wall_grad_ghost(Qo, gl, dface, ox, oy, gr);
#pragma unroll
    for (int c = 0; c < 18; ++c) g[c] = 0.5 * gl[c] + 0.5 * gr[c];
  } else {
#pragma unroll
    for (int c = 0; c < 18; ++c) g[c] = 0.5 * gl[c] + 0.5 * gl[c];
  }
}

// Viscous flux at flux point r of a line (comp_viscous_flux,
// compflux.f90:198-258), normal to the point's own metric vector.
template <int N>
__device__ __forceinline__ void vis_point_flux(const VisArgs &va, const Ops &ops, const Phys &ph,
                                               int tile, int ci, int fam, int line, int r,
                                               double *Fn) {
  constexpr int T = Geo<N>::CPW, QT = Geo<N>::QT, NF = N + 1;
  const int p0 = fam ? line : N * line, ps = fam ? N : 1;
  double Qv[9];
  vis_point_state<N>(va, ops, ph, tile, ci, fam, line, r, Qv);
  // own gradient trace (nablaQ_interior_FP, compflux.f90:159-196)
  double Lr[N];
#pragma unroll
  for (int s = 0; s < N; ++s) Lr[s] = ops.L[r][s];
  const double *go = va.gq + (size_t)tile * 2 * QT + ci;
  double gl[18], g[18];
#pragma unroll
  for (int c = 0; c < 18; ++c) vis_trace<N, 1>(go, 2 * (c % 9) + c / 9, p0, ps, Lr, gl + c);
  if (r == 0 || r == N)
    vis_end_grad<N>(va, ops, tile, ci, fam, line, r, gl, g);
  else
#pragma unroll
    for (int c = 0; c < 18; ++c) g[c] = gl[c];
  const int mx = fam ? Geo<N>::S2X : Geo<N>::S1X, my = fam ? Geo<N>::S2Y : Geo<N>::S1Y;
  const double *mt = va.met + (size_t)tile * Geo<N>::NM * T + ci;
  vflux_n(Qv, g, va.mu, va.eta0, mt[(mx + r + NF * line) * T], mt[(my + r + NF * line) * T], Fn);
}

// Viscous flux divergence dFv + dGv at the solution points: every line
// thread forms the fluxes at its N+1 flux points into LDS; the xi-line
// threads then differentiate both families at their points
// (compresid.f90:24-46).
template <int N>
__global__ __launch_bounds__(STAGE_THREADS) void vis_flux_kernel(VisArgs va, Ops ops, Phys ph) {
  using G = Geo<N>;
  using LD = VisLds<N>;
  constexpr int T = G::CPW, N2 = G::N2, NF = N + 1, NT = 2 * N * T;
  extern __shared__ double lds[];
  const int tid = threadIdx.x, tile = blockIdx.x;
  const int ci = tid % T, fl = tid / T;
  const int fam = fl >= N ? 1 : 0, line = fl - fam * N;
  const int cell = tile * T + ci;   // the slot
  const bool valid = tid < NT && va.s2c[cell] >= 0;
  if (valid) {
#pragma unroll 1
    for (int r = 0; r < NF; ++r) {
      double Fn[9];
      switch (r) {   // compile-time operator rows
#define SD2D_R(i) case i: if (i <= N) vis_point_flux<N>(va, ops, ph, tile, ci, fam, line, i, Fn); break;
        SD2D_R(0) SD2D_R(1) SD2D_R(2) SD2D_R(3) SD2D_R(4) SD2D_R(5)
#undef SD2D_R
      }
#pragma unroll
      for (int v = 0; v < 9; ++v) lds[LD::at(fam, line, r, v, ci)] = Fn[v];
    }
  }
  __syncthreads();
  if (!valid || fam != 0) return;
  double *o = va.visres + (size_t)tile * G::QT;
  const int js = line;
#pragma unroll
  for (int is = 0; is < N; ++is)
#pragma unroll
    for (int v = 0; v < 9; ++v) {
      double dF = 0.0, dG = 0.0;
#pragma unroll
      for (int r = 0; r < NF; ++r) dF = dF + lds[LD::at(0, js, r, v, ci)] * ops.M[r][is];
#pragma unroll
      for (int r = 0; r < NF; ++r) dG = dG + lds[LD::at(1, is, r, v, ci)] * ops.M[r][js];
      o[qo<N>(v, is + N * js, ci)] = dF + dG;
    }
}

template <int N>
constexpr size_t vis_lds_bytes() { return sizeof(double) * VisLds<N>::TOTAL; }

}  // namespace sd2d
