// sd2d_adj.hpp - device kernels of the adjoint (vector-Jacobian product)
// of the residual evaluation and of the wall-drag objective.
//
// For a residual evaluation R(Q; c_h, free stream) on the inviscid path
// (compresid.f90:1-80 with the fluxes of compflux.f90:35-104,
// interfaceflux_all.f90:1-112, cyclocflux_all.f90:1-112, bcflux.f90:1-377)
// and a cotangent Rbar, the kernels accumulate
//   Qbar += (dR/dQ)^T Rbar,  and the partial sums of
//   c_hbar = (dR/dc_h)^T Rbar and pbar = (dR/d free stream)^T Rbar.
// Reverse sweep, written by hand over the structure of the forward kernel:
//   adj_point_kernel  per solution point: resid = -(dF/dxi + dG/deta)/J -
//                     Powell(divB) - alpha psi, reversed: the cotangent of
//                     the flux derivatives Dbar = -Rbar/J and of the divB
//                     weight wbar = (dR/d divB)^T Rbar / J, and the direct
//                     (pointwise) part of Qbar;
//   adj_line_kernel   per grid line and flux point: the cotangent of the
//                     flux there (Mmat^T Dbar along the line) goes through
//                     the transposed Jacobian of the interior flux, of the
//                     Rusanov flux at faces (with the neighbour's own
//                     cotangent of the same face flux, gathered from its
//                     line, so each cell updates only its own Qbar) or of
//                     the boundary flux with its ghost state; the divB
//                     weights of the averaged face states likewise; the
//                     trace cotangents go back to the solution points
//                     through Lmat^T.
// The local Jacobians come from the same physics templates instantiated
// with dual numbers (sd2d_dual.hpp), one input direction per evaluation.
// The AV sensor does not act on the state and has no adjoint.
#pragma once
#include "sd2d_dual.hpp"
#include "sd2d_stage.hpp"

namespace sd2d {

struct AdjArgs {
  const double *q;         // state the residual is taken of (tile layout)
  const double *rbar;      // cotangent of the residual (tile layout)
  double rscale;           // the cotangent used is rscale * rbar
  const double *divb;      // divB of q per solution point [tile][p][ci]
  const double *met;
  const int *ftab;
  const int *s2c;
  const Phys *phm;         // free stream
  const double *gm;        // eigvmax_g, umax_g that the evaluation's c_h came from
  double *g;               // [tile][k*N2 + p][ci]: k < 9 Dbar, k = 9 wbar
  double *qbar;            // accumulated (tile layout)
  double *part;            // per tile: c_hbar, pbar[6] (rinf pinf uinf vinf bx by)
  double alpha;
  int ncell, ntile;
};

constexpr int ADJ_NPART = 7;

template <int N> struct AdjLds {
  static constexpr int T = Geo<N>::CPW, NF = N + 1;
  // trace cotangents [fam][line][rfp][v][ci], then the reduction slots
  static __device__ __forceinline__ int at(int fam, int line, int r, int v, int ci) {
    return (((fam * N + line) * NF + r) * 9 + v) * T + ci;
  }
  static constexpr int RED = 2 * N * NF * 9 * T;
  static constexpr int OPS = RED + ADJ_NPART * (STAGE_THREADS / 64);
  static constexpr int TOTAL = OPS + (int)(sizeof(Ops) / sizeof(double));
};

// The operators copied to LDS, where the reverse sweeps index their rows
// with run-time flux-point numbers.
template <int OFF>
__device__ __forceinline__ const Ops &ops_to_lds(double *lds, const Ops &ops) {
  Ops *o = reinterpret_cast<Ops *>(lds + OFF);
  const double *src = reinterpret_cast<const double *>(&ops);
  for (int i = threadIdx.x; i < (int)(sizeof(Ops) / sizeof(double)); i += blockDim.x)
    reinterpret_cast<double *>(o)[i] = src[i];
  __syncthreads();
  return *o;
}

template <int N>
constexpr size_t adj_lds_bytes() { return sizeof(double) * AdjLds<N>::TOTAL; }

// ---- per solution point ------------------------------------------------------
// Forward (stage kernel, points step): R = -D/J, then
//   R1..3 -= divB*B, R4 -= divB*(u.B), R5..7 -= divB*u, R8 -= alpha*psi,
// divB = db/J with db the summed divB-weight derivatives.
template <int N>
__global__ __launch_bounds__(STAGE_THREADS) void adj_point_kernel(AdjArgs a) {
  using G = Geo<N>;
  constexpr int T = G::CPW, N2 = G::N2;
  const int tile = blockIdx.x;
  const double *mt = a.met + (size_t)tile * G::NM * T;
  for (int i = threadIdx.x; i < N2 * T; i += STAGE_THREADS) {
    const int p = i / T, ci = i % T;
    const int slot = tile * T + ci;
    const int cell = a.s2c[slot];
    if (cell < 0) continue;   // padding slot
    const int is = p % N, js = p / N;
    const double ji = mt[(G::JI + js + N * is) * T + ci];
    const size_t off = (size_t)tile * G::QT;
    double q[9], rb[9];
#pragma unroll
    for (int v = 0; v < 9; ++v) {
      q[v] = a.q[off + qo<N>(v, p, ci)];
      rb[v] = a.rscale * a.rbar[off + qo<N>(v, p, ci)];
    }
    const double divB = a.divb[(size_t)tile * N2 * T + (size_t)p * T + ci];
    const double ri = 1.0 / q[0];
    const double ux = q[1] * ri, uy = q[2] * ri, uz = q[3] * ri;
    // cotangent of divB
    const double dd = -(rb[1] * q[5] + rb[2] * q[6] + rb[3] * q[7] +
                        rb[4] * (ux * q[5] + uy * q[6] + uz * q[7]) + rb[5] * ux + rb[6] * uy +
                        rb[7] * uz);
    double *gp = a.g + (size_t)tile * 10 * N2 * T + ci;
#pragma unroll
    for (int v = 0; v < 9; ++v) gp[(size_t)(v * N2 + p) * T] = -ji * rb[v];
    gp[(size_t)(9 * N2 + p) * T] = ji * dd;
    // direct part: Powell source and GLM damping at the point
    double qb[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    qb[5] = -divB * (rb[1] + rb[4] * ux);
    qb[6] = -divB * (rb[2] + rb[4] * uy);
    qb[7] = -divB * (rb[3] + rb[4] * uz);
    const double ubx = -divB * (rb[4] * q[5] + rb[5]);
    const double uby = -divB * (rb[4] * q[6] + rb[6]);
    const double ubz = -divB * (rb[4] * q[7] + rb[7]);
    qb[1] = ubx * ri;
    qb[2] = uby * ri;
    qb[3] = ubz * ri;
    qb[0] = -(ubx * ux + uby * uy + ubz * uz) * ri;
    qb[8] = -a.alpha * rb[8];
#pragma unroll
    for (int v = 0; v < 9; ++v) a.qbar[off + qo<N>(v, p, ci)] += qb[v];
  }
}

// Cotangent of the flux and of the divB weight at flux point r of a line:
// Mmat^T applied to the line's Dbar / wbar (g at the cell `gb`).
template <int N>
__device__ __forceinline__ void adj_flux_cot(const double *gb, int p0, int ps, int r,
                                             const Ops &ops, double *Fb, double &Wb) {
  constexpr int T = Geo<N>::CPW, N2 = Geo<N>::N2;
#pragma unroll
  for (int v = 0; v < 10; ++v) {
    double acc = 0.0;
#pragma unroll
    for (int s = 0; s < N; ++s) acc += ops.M[r][s] * gb[(size_t)(v * N2 + p0 + ps * s) * T];
    if (v < 9) Fb[v] = acc; else Wb = acc;
  }
}

// Trace of the 9 variables along the line through point p0 (stride ps) of
// the slot `base` = tile * QT + ci (the face-table encoding) of the
// state-layout array q.
template <int N>
__device__ __forceinline__ void adj_trace(const double *q, int base, int p0, int ps,
                                          const double *Lr, double *t) {
  constexpr int QT = Geo<N>::QT;
  const int nt = base / QT, ci = base - nt * QT;
  const double *qt = q + (size_t)nt * QT;
#pragma unroll
  for (int v = 0; v < 9; ++v) {
    double vals[N];
#pragma unroll
    for (int s = 0; s < N; ++s) vals[s] = qt[qo<N>(v, p0 + ps * s, ci)];
    t[v] = line_trace<N, double>(vals, Lr);
  }
}

// ---- reverse sweeps of the pointwise fluxes (FP64) -------------------------------
// With Fb the cotangent of a flux, these accumulate (dF/dq)^T Fb and
// (dF/dc_h)^T Fb, the forward values being formed as iflux_n / rusanov
// form them.  Derivatives of |x| take the sign of x (+1 at 0), of sqrt(x)
// 1/(2 sqrt(x)), as the dual numbers of sd2d_dual.hpp do; the running
// maxima outputs are not differentiated here (adj_argmax_kernel).

// iflux_n (getrusanov.f90:127-193 contracted with (sx, sy))
__device__ __forceinline__ void iflux_n_vjp(const double *q, double c_h, double sx, double sy,
                                            const double *Fb, double *qb, double &chb) {
  constexpr double g1 = GAM - 1.0;
  const double rho = q[0], rhou = q[1], rhov = q[2], rhow = q[3], rhoE = q[4];
  const double Bx = q[5], By = q[6], Bz = q[7], psi = q[8];
  const double ri = rcp(rho);
  const double u = rhou * ri, v = rhov * ri, w = rhow * ri;
  const double B2 = Bx * Bx + By * By + Bz * Bz;
  const double pr = g1 * (rhoE - 0.5 * (rhou * u + rhov * v + rhow * w) - 0.5 * B2 - 0.5 * psi * psi);
  const double pt = pr + 0.5 * B2;
  const double uB = u * Bx + v * By + w * Bz;
  const double ef = rhoE - 0.5 * psi * psi + pt;
  const double un = u * sx + v * sy;
  const double bn = Bx * sx + By * sy;
  const double cp = c_h * psi;
  // the nine outputs (iflux_n), backwards
  double rhob = 0.0, rhoub = Fb[0] * sx + Fb[1] * un, rhovb = Fb[0] * sy + Fb[2] * un;
  double rhowb = Fb[3] * un, rhoEb = 0.0, psib = 0.0;
  double Bxb = -Fb[1] * bn, Byb = -Fb[2] * bn, Bzb = -Fb[3] * bn + Fb[7] * un;
  double ub = 0.0, vb = 0.0, wb = -Fb[7] * bn;
  double unb = Fb[1] * rhou + Fb[2] * rhov + Fb[3] * rhow + Fb[4] * ef + Fb[7] * Bz;
  double ptb = Fb[1] * sx + Fb[2] * sy;
  double bnb = -Fb[1] * Bx - Fb[2] * By - Fb[3] * Bz + Fb[4] * (cp - uB) - Fb[7] * w + Fb[8] * c_h;
  const double efb = Fb[4] * un;
  const double uBb = -Fb[4] * bn;
  const double cpb = Fb[4] * bn + Fb[5] * sx + Fb[6] * sy;
  const double xb = Fb[5] * sy - Fb[6] * sx;        // x = v*Bx - u*By
  double ch = Fb[8] * bn;
  // intermediates, backwards
  vb += xb * Bx; Bxb += xb * v; ub -= xb * By; Byb -= xb * u;
  ch += cpb * psi; psib += cpb * c_h;
  Bxb += bnb * sx; Byb += bnb * sy;
  ub += unb * sx; vb += unb * sy;
  rhoEb += efb; psib -= efb * psi; ptb += efb;
  ub += uBb * Bx; Bxb += uBb * u; vb += uBb * By; Byb += uBb * v; wb += uBb * Bz; Bzb += uBb * w;
  const double prb = ptb;
  double B2b = 0.5 * ptb;
  rhoEb += g1 * prb;
  rhoub -= 0.5 * g1 * prb * u; ub -= 0.5 * g1 * prb * rhou;
  rhovb -= 0.5 * g1 * prb * v; vb -= 0.5 * g1 * prb * rhov;
  rhowb -= 0.5 * g1 * prb * w; wb -= 0.5 * g1 * prb * rhow;
  B2b -= 0.5 * g1 * prb; psib -= g1 * prb * psi;
  Bxb += 2.0 * Bx * B2b; Byb += 2.0 * By * B2b; Bzb += 2.0 * Bz * B2b;
  rhoub += ub * ri; rhovb += vb * ri; rhowb += wb * ri;
  rhob -= (ub * rhou + vb * rhov + wb * rhow) * ri * ri;
  qb[0] = rhob; qb[1] = rhoub; qb[2] = rhovb; qb[3] = rhowb; qb[4] = rhoEb;
  qb[5] = Bxb; qb[6] = Byb; qb[7] = Bzb; qb[8] = psib;
  chb += ch;
}

// one side's physical flux of rusanov (Fl or Fr) backwards, with Fb its
// cotangent and pgb / Vnb the cotangents its gas pressure and normal
// velocity receive from the wave speed; accumulates into Qb and chb
__device__ __forceinline__ void rusanov_side_vjp(const double *Q, double ux, double uy, double c_h,
                                                 const double *Fb, double pgb, double Vnb,
                                                 double *Qb, double &chb) {
  constexpr double g1 = GAM - 1.0;
  const double ri = rcp(Q[0]);
  const double u = ri * Q[1], v = ri * Q[2], w = ri * Q[3];
  const double B2 = Q[5] * Q[5] + Q[6] * Q[6] + Q[7] * Q[7];
  const double pg = g1 * (Q[4] - 0.5 * Q[0] * (u * u + v * v + w * w) - B2 * 0.5 - 0.5 * Q[8] * Q[8]);
  const double pt = pg + 0.5 * B2;
  const double Vn = u * ux + v * uy;
  const double Bn = Q[5] * ux + Q[6] * uy;
  const double e = Q[4] - 0.5 * Q[8] * Q[8] + pt;
  const double uB = u * Q[5] + v * Q[6] + w * Q[7];
  Vnb += Fb[0] * Q[0] + Fb[1] * Q[1] + Fb[2] * Q[2] + Fb[3] * Q[3] + Fb[4] * e + Fb[7] * Q[7];
  Qb[0] += Fb[0] * Vn; Qb[1] += Fb[1] * Vn; Qb[2] += Fb[2] * Vn; Qb[3] += Fb[3] * Vn;
  Qb[7] += Fb[7] * Vn;
  double ptb = Fb[1] * ux + Fb[2] * uy;
  const double eb = Fb[4] * Vn;
  Qb[4] += eb; Qb[8] -= eb * Q[8]; ptb += eb;
  const double Bnb = -Fb[1] * Q[5] - Fb[2] * Q[6] - Fb[3] * Q[7] - Fb[4] * uB + Fb[4] * c_h * Q[8] -
                     Fb[7] * w + Fb[8] * c_h;
  Qb[5] -= Fb[1] * Bn; Qb[6] -= Fb[2] * Bn; Qb[7] -= Fb[3] * Bn;
  const double uBb = -Fb[4] * Bn;
  chb += Fb[4] * Q[8] * Bn + Fb[5] * Q[8] * ux + Fb[6] * Q[8] * uy + Fb[8] * Bn;
  Qb[8] += Fb[4] * c_h * Bn + Fb[5] * c_h * ux + Fb[6] * c_h * uy;
  const double xb = Fb[5] * uy - Fb[6] * ux;        // x = v*Q5 - u*Q6
  double ub = 0.0, vb = 0.0, wb = -Fb[7] * Bn;
  vb += xb * Q[5]; Qb[5] += xb * v; ub -= xb * Q[6]; Qb[6] -= xb * u;
  ub += uBb * Q[5]; Qb[5] += uBb * u; vb += uBb * Q[6]; Qb[6] += uBb * v;
  wb += uBb * Q[7]; Qb[7] += uBb * w;
  Qb[5] += Bnb * ux; Qb[6] += Bnb * uy;
  ub += Vnb * ux; vb += Vnb * uy;
  pgb += ptb;
  double B2b = 0.5 * ptb;
  Qb[4] += g1 * pgb;
  Qb[0] -= 0.5 * g1 * pgb * (u * u + v * v + w * w);
  ub -= g1 * pgb * Q[0] * u; vb -= g1 * pgb * Q[0] * v; wb -= g1 * pgb * Q[0] * w;
  B2b -= 0.5 * g1 * pgb; Qb[8] -= g1 * pgb * Q[8];
  Qb[5] += 2.0 * Q[5] * B2b; Qb[6] += 2.0 * Q[6] * B2b; Qb[7] += 2.0 * Q[7] * B2b;
  Qb[1] += ub * ri; Qb[2] += vb * ri; Qb[3] += wb * ri;
  Qb[0] -= (ub * Q[1] + vb * Q[2] + wb * Q[3]) * ri * ri;
}

// rusanov (getrusanov.f90:1-125): Fn = sc (Fl + Fr - ev (QR - QL)) with the
// fast magneto-acoustic speed ev of the averaged state; QLb, QRb are set
__device__ __forceinline__ void rusanov_vjp(const double *QL, const double *QR, double nx,
                                            double ny, int sign_l, double c_h, const double *Fnb,
                                            double *QLb, double *QRb, double &chb) {
  constexpr double g1 = GAM - 1.0;
  const double mag = sqrt(nx * nx + ny * ny);
  const double im = double(sign_l) * rcp(mag);
  const double ux = nx * im, uy = ny * im;
  const double ril = rcp(QL[0]), rir = rcp(QR[0]);
  const double ul = ril * QL[1], vl = ril * QL[2], wl = ril * QL[3];
  const double ur = rir * QR[1], vr = rir * QR[2], wr = rir * QR[3];
  const double B2l = QL[5] * QL[5] + QL[6] * QL[6] + QL[7] * QL[7];
  const double B2r = QR[5] * QR[5] + QR[6] * QR[6] + QR[7] * QR[7];
  const double pgl = g1 * (QL[4] - 0.5 * QL[0] * (ul * ul + vl * vl + wl * wl) - B2l * 0.5 -
                           0.5 * QL[8] * QL[8]);
  const double pgr = g1 * (QR[4] - 0.5 * QR[0] * (ur * ur + vr * vr + wr * wr) - B2r * 0.5 -
                           0.5 * QR[8] * QR[8]);
  const double Vnl = ul * ux + vl * uy, Vnr = ur * ux + vr * uy;
  const double bx = 0.5 * (QL[5] + QR[5]), by = 0.5 * (QL[6] + QR[6]), bz = 0.5 * (QL[7] + QR[7]);
  const double irs = rcp(QL[0] + QR[0]);
  const double ca = GAM * (pgl + pgr) * irs;
  const double bb = bx * bx + by * by + bz * bz;
  const double cd = 2.0 * bb * irs;
  const double bn = bx * ux + by * uy;
  const double cn = 2.0 * bn * bn * irs;
  const double cad = ca + cd;
  const double ct = sqrt(cad * cad - 4.0 * ca * cn);
  const double cf = sqrt(0.5 * (cad + ct));
  const double sV = Vnl + Vnr;
  const double ev = 0.5 * fabs(sV) + cf;
  const double sc = 0.5 * mag * double(sign_l);
  double G[9], evb = 0.0;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    G[k] = sc * Fnb[k];
    evb -= G[k] * (QR[k] - QL[k]);
    QLb[k] = G[k] * ev;
    QRb[k] = -G[k] * ev;
  }
  // ev = |Vnl + Vnr|/2 + cf, cf = sqrt((cad + ct)/2), ct = sqrt(cad^2 - 4 ca cn)
  const double sb = 0.5 * (sV < 0.0 ? -evb : evb);
  const double yb = evb / (2.0 * cf);
  double cadb = 0.5 * yb;
  const double ctb = 0.5 * yb;
  const double zb = ctb / (2.0 * ct);
  cadb += 2.0 * cad * zb;
  const double cab = -4.0 * cn * zb + cadb;
  const double cnb = -4.0 * ca * zb;
  const double cdb = cadb;
  const double bnb = 4.0 * bn * irs * cnb;
  double irsb = 2.0 * bn * bn * cnb + 2.0 * bb * cdb + GAM * (pgl + pgr) * cab;
  const double bxb = bnb * ux + 4.0 * bx * irs * cdb;
  const double byb = bnb * uy + 4.0 * by * irs * cdb;
  const double bzb = 4.0 * bz * irs * cdb;
  const double pgb = GAM * irs * cab;
  const double d0 = -irs * irs * irsb;
  QLb[0] += d0; QRb[0] += d0;
  QLb[5] += 0.5 * bxb; QRb[5] += 0.5 * bxb;
  QLb[6] += 0.5 * byb; QRb[6] += 0.5 * byb;
  QLb[7] += 0.5 * bzb; QRb[7] += 0.5 * bzb;
  rusanov_side_vjp(QL, ux, uy, c_h, G, pgb, sb, QLb, chb);
  rusanov_side_vjp(QR, ux, uy, c_h, G, pgb, sb, QRb, chb);
}

// The trace cotangent of flux point r of a line (own side only), and the
// c_h / free-stream partial sums of the points this cell owns.
template <int N>
__device__ void adj_point_flux(const AdjArgs &a, const Ops &ops, double c_h, int tile, int ci,
                               int fam, int line, int r, double *Qt, double &chb, double *pb) {
  using G = Geo<N>;
  constexpr int T = G::CPW, N2 = G::N2, NF = N + 1, QT = G::QT;
  const int p0 = fam ? line : N * line, ps = fam ? N : 1;
  const int mx = fam ? G::S2X : G::S1X, my = fam ? G::S2Y : G::S1Y;
  const double *mt = a.met + (size_t)tile * G::NM * T + ci;
  const double sx = mt[(mx + r + NF * line) * T], sy = mt[(my + r + NF * line) * T];
  double Lr[N];
#pragma unroll
  for (int s = 0; s < N; ++s) Lr[s] = ops.L[r][s];
  double t[9], Fb[9], Wb;
  adj_trace<N>(a.q, tile * QT + ci, p0, ps, Lr, t);
  adj_flux_cot<N>(a.g + (size_t)tile * 10 * N2 * T + ci, p0, ps, r, ops, Fb, Wb);
  if (r != 0 && r != N) {
    // interior flux point (compflux.f90:53-104) and the divB weight of the trace
    iflux_n_vjp(t, c_h, sx, sy, Fb, Qt, chb);
    Qt[5] += Wb * sx;
    Qt[6] += Wb * sy;
    return;
  }
  int k, nfp;
  end_face<N>(fam, line, r, k, nfp);
  const int *ft = a.ftab + ((size_t)tile * 8 + k * 2) * T + ci;
  const int nbase = ft[0], info = ft[T];
  const int bc = (info >> 3) & 7;
  const int own_sl = (k == 0 || k == 3) ? -1 : 1;
  if (bc == FT_INTERIOR) {
    // the face as interfaceflux / CYCLOCFLUX evaluate it, and the
    // neighbour's cotangent of its own copy of the face flux
    const int kn = info & 3, left = (info >> 2) & 1, sneg = (info >> 6) & 1;
    const int nt = nbase / QT, nci = nbase - nt * QT;
    int np0, nst;
    bool endN;
    neighbour_line<N>(kn, nfp, np0, nst, endN);
    double Ln[N];
#pragma unroll
    for (int s = 0; s < N; ++s) Ln[s] = endN ? ops.L[N][s] : ops.L[0][s];
    double qn[9], Fnb[9], Wnb;
    adj_trace<N>(a.q, nbase, np0, nst, Ln, qn);
    adj_flux_cot<N>(a.g + (size_t)nt * 10 * N2 * T + nci, np0, nst, endN ? N : 0, ops, Fnb, Wnb);
    const int sl = left ? own_sl : ((kn == 0 || kn == 3) ? -1 : 1);
    const int sr = left ? ((kn == 1 || kn == 2) ? -1 : 1) : ((k == 1 || k == 2) ? -1 : 1);
    const double nsg = (left || !sneg) ? 1.0 : -1.0;
    const double fo = left ? 1.0 : double(sl * sr), fn = left ? double(sl * sr) : 1.0;
    // the neighbour's own metric at the point is sig times this cell's
    int sgn_nb = sneg;
    if (left) sgn_nb = (a.ftab[((size_t)nt * 8 + kn * 2 + 1) * T + nci] >> 6) & 1;
    const double sig = sgn_nb ? -1.0 : 1.0;
    double Gb[9];
#pragma unroll
    for (int v = 0; v < 9; ++v) Gb[v] = fo * Fb[v] + fn * Fnb[v];
    {
      // the face flux backwards; this cell keeps its own side's cotangent,
      // and the left cell also takes the c_h part
      double QLb[9], QRb[9], chf = 0.0;
      rusanov_vjp(left ? t : qn, left ? qn : t, nsg * sx, nsg * sy, sl, c_h, Gb, QLb, QRb, chf);
#pragma unroll
      for (int v = 0; v < 9; ++v) Qt[v] = left ? QLb[v] : QRb[v];
      if (left) chb += chf;
    }
    // divB weights of the averaged state Qv = (QL + QR)/2 seen by both cells
    const double wq = 0.5 * (Wb + sig * Wnb);
    Qt[5] += wq * sx;
    Qt[6] += wq * sy;
    return;
  }
  // boundary face: Rusanov with the ghost state of the own trace (bcflux.f90)
  const Phys &ph = a.phm[0];
  const double par[6] = {ph.rinf, ph.pinf, ph.uinf, ph.vinf, ph.bxinf, ph.byinf};
#pragma unroll 1
  for (int kk = 0; kk < 16; ++kk) {   // 9 state, c_h, 6 free-stream directions
    Dual Qo[9], QR[9], F[9], eig(0.0), um(0.0), pd[6];
#pragma unroll
    for (int v = 0; v < 9; ++v) Qo[v] = Dual(t[v], v == kk ? 1.0 : 0.0);
#pragma unroll
    for (int j = 0; j < 6; ++j) pd[j] = Dual(par[j], kk == 10 + j ? 1.0 : 0.0);
    bc_ghost_p<Dual>(bc, Qo, Dual(sx), Dual(sy), pd[0], pd[1], pd[2], pd[3], pd[4], pd[5],
                     ph.vismode, QR);
    rusanov<Dual>(Qo, QR, Dual(sx), Dual(sy), own_sl, Dual(c_h, kk == 9 ? 1.0 : 0.0), F, eig,
                  um);
    double acc = 0.0;
#pragma unroll
    for (int v = 0; v < 9; ++v) acc += Fb[v] * F[v].d;
    // divB weight of Qv = (Qo + ghost)/2
    acc += Wb * 0.5 * ((Qo[5].d + QR[5].d) * sx + (Qo[6].d + QR[6].d) * sy);
    if (kk < 9) Qt[kk] = acc;
    else if (kk == 9) chb += acc;
    else pb[kk - 10] += acc;
  }
}

// ---- per grid line -------------------------------------------------------------
// Every line thread forms the trace cotangents of its N+1 flux points into
// LDS; the xi-line threads then send both families' trace cotangents back
// to their points through Lmat^T and add them to Qbar.
template <int N>
__global__ __launch_bounds__(STAGE_THREADS) void adj_line_kernel(AdjArgs a, Ops ops_arg) {
  using G = Geo<N>;
  using LD = AdjLds<N>;
  constexpr int T = G::CPW, N2 = G::N2, NF = N + 1, NT = 2 * N * T;
  extern __shared__ double lds[];
  const int tid = threadIdx.x, tile = blockIdx.x;
  const int ci = tid % T, fl = tid / T;
  const int fam = fl >= N ? 1 : 0, line = fl - fam * N;
  const bool valid = tid < NT && a.s2c[tile * T + ci] >= 0;
  const Ops &ops = ops_to_lds<LD::OPS>(lds, ops_arg);
  const double ge = a.gm[0], gu = a.gm[1];
  const double c_h = sqrt(ge * (ge - gu));   // as the forward forms it (compflux.f90:6)
  double chb = 0.0, pb[6] = {0, 0, 0, 0, 0, 0};
  if (valid) {
#pragma unroll 1
    for (int r = 0; r < NF; ++r) {
      double Qt[9];
      adj_point_flux<N>(a, ops, c_h, tile, ci, fam, line, r, Qt, chb, pb);
#pragma unroll
      for (int v = 0; v < 9; ++v) lds[LD::at(fam, line, r, v, ci)] = Qt[v];
    }
  }
  // partial sums of the block
  {
    constexpr int NW = STAGE_THREADS / 64;
    const int w = tid >> 6, l = tid & 63;
    double x[ADJ_NPART] = {chb, pb[0], pb[1], pb[2], pb[3], pb[4], pb[5]};
#pragma unroll
    for (int j = 0; j < ADJ_NPART; ++j) {
      x[j] = wave_sum(x[j]);
      if (l == 0) lds[LD::RED + j * NW + w] = x[j];
    }
  }
  __syncthreads();
  if (tid < ADJ_NPART) {
    constexpr int NW = STAGE_THREADS / 64;
    double s = 0.0;
    for (int w = 0; w < NW; ++w) s += lds[LD::RED + tid * NW + w];
    a.part[(size_t)tile * ADJ_NPART + tid] = s;
  }
  if (!valid || fam != 0) return;
  const int js = line;
  double *qb = a.qbar + (size_t)tile * G::QT;
#pragma unroll
  for (int is = 0; is < N; ++is)
#pragma unroll
    for (int v = 0; v < 9; ++v) {
      double acc = 0.0;
#pragma unroll
      for (int r = 0; r < NF; ++r) acc += ops.L[r][is] * lds[LD::at(0, js, r, v, ci)];
#pragma unroll
      for (int r = 0; r < NF; ++r) acc += ops.L[r][js] * lds[LD::at(1, is, r, v, ci)];
      qb[qo<N>(v, is + N * js, ci)] += acc;
    }
}

// ---- wall drag ------------------------------------------------------------------
// D = sum over WALL faces and their N points of w_i * p * (outward S)_x,
// p = (gamma-1)(E - |m|^2/(2 rho) - |B|^2/2 - psi^2/2) of the cell's own
// trace, S the cell's metric vector at the point (S1(1,:) or S2(2,:),
// whose length is the face's arc length per unit reference coordinate),
// signed outwards, and w_i the weights of the interpolatory quadrature on
// the solution-point positions Xs of the face.  One thread per cell: the
// value per cell, and with qbar set, Qbar += scale * dD/dQ.
template <int N>
__global__ void drag_kernel(const double *q, const int *ftab, const double *met, const double *wq,
                            Ops ops_arg, int ncell, int ntile, double *dcell, double *qbar,
                            double scale) {
  using G = Geo<N>;
  constexpr int T = G::CPW, N2 = G::N2, NF = N + 1;
  __shared__ double so[sizeof(Ops) / sizeof(double)];
  const Ops &ops = ops_to_lds<0>(so, ops_arg);
  const int slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= ntile * T) return;
  const int tile = slot / T, ci = slot % T;
  double d = 0.0;
  {   // (a padding slot has no wall face: d stays 0)
    const double *qc = q + (size_t)tile * G::QT;
    const double *mt = met + (size_t)tile * G::NM * T + ci;
    for (int k = 0; k < 4; ++k) {
      const int info = ftab[((size_t)tile * 8 + k * 2 + 1) * T + ci];
      if (((info >> 3) & 7) != FT_WALL) continue;
      // the face's points: own lines of the family crossing it
      const int fam = (k == 1 || k == 3) ? 0 : 1;
      const int rfp = (k == 0 || k == 3) ? 0 : N;
      const int mx = fam ? G::S2X : G::S1X;
      const double sgn = (k == 0 || k == 3) ? -1.0 : 1.0;
      for (int line = 0; line < N; ++line) {
        const int p0 = fam ? line : N * line, ps = fam ? N : 1;
        double t[9];
        for (int v = 0; v < 9; ++v) {
          double acc = 0.0;
          for (int s = 0; s < N; ++s) acc = fma(qc[qo<N>(v, p0 + ps * s, ci)], ops.L[rfp][s], acc);
          t[v] = acc;
        }
        const double g1 = GAM - 1.0;
        const double ke = 0.5 * (t[1] * t[1] + t[2] * t[2] + t[3] * t[3]) / t[0];
        const double pr = g1 * (t[4] - ke - 0.5 * (t[5] * t[5] + t[6] * t[6] + t[7] * t[7]) -
                                0.5 * t[8] * t[8]);
        const double nx = sgn * mt[(mx + rfp + NF * line) * T];
        d += wq[line] * pr * nx;
        if (qbar) {
          const double c = scale * wq[line] * nx * g1;
          double pt[9];
          pt[0] = c * ke / t[0];
          pt[1] = -c * t[1] / t[0];
          pt[2] = -c * t[2] / t[0];
          pt[3] = -c * t[3] / t[0];
          pt[4] = c;
          pt[5] = -c * t[5];
          pt[6] = -c * t[6];
          pt[7] = -c * t[7];
          pt[8] = -c * t[8];
          double *qb = qbar + (size_t)tile * G::QT;
          for (int v = 0; v < 9; ++v)
            for (int s = 0; s < N; ++s) qb[qo<N>(v, p0 + ps * s, ci)] += pt[v] * ops.L[rfp][s];
        }
      }
    }
  }
  dcell[slot] = d;
}

// ---- the running maxima behind c_h -------------------------------------------------
// c_h of an evaluation is formed from the running maxima of eigvmax and
// umax over all earlier evaluations (compflux.f90:6; getrusanov.f90:116-123,
// 184-191).  Where an evaluation raised a maximum, the cotangent of the
// maximum goes to the wave speed of the one flux point that attained it.
// Location code of a flux point: ((slot*2 + fam)*N + line)*(N+1) + rfp.

// wave speed eig (face points) and velocity bound um of flux point r of a
// line, with the forward's arithmetic; T is double or Dual, with the own
// trace t and the neighbour trace qn (interior faces) as inputs
template <int N, typename R>
__device__ __forceinline__ void adj_point_speeds(const AdjArgs &a, int info, int k, int r,
                                                 bool face, double sx, double sy,
                                                 const R *t, const R *qn, R &eig, R &um) {
  eig = R(0.0);
  um = R(0.0);
  if (!face) {
    R F[9];
    iflux_n<R>(t, R(0.0), R(sx), R(sy), F, um);
    return;
  }
  const int bc = (info >> 3) & 7;
  const int own_sl = (k == 0 || k == 3) ? -1 : 1;
  R F[9];
  if (bc == FT_INTERIOR) {
    const int kn = info & 3, left = (info >> 2) & 1, sneg = (info >> 6) & 1;
    R QL[9], QR[9];
#pragma unroll
    for (int v = 0; v < 9; ++v) { QL[v] = left ? t[v] : qn[v]; QR[v] = left ? qn[v] : t[v]; }
    const int sl = left ? own_sl : ((kn == 0 || kn == 3) ? -1 : 1);
    const double nsg = (left || !sneg) ? 1.0 : -1.0;
    rusanov<R>(QL, QR, R(nsg * sx), R(nsg * sy), sl, R(0.0), F, eig, um);
  } else {
    const Phys &ph = a.phm[0];
    R QR[9];
    bc_ghost_p<R>(bc, t, R(sx), R(sy), R(ph.rinf), R(ph.pinf), R(ph.uinf), R(ph.vinf),
                  R(ph.bxinf), R(ph.byinf), ph.vismode, QR);
    rusanov<R>(t, QR, R(sx), R(sy), own_sl, R(0.0), F, eig, um);
  }
}

// the geometry of flux point r of a line: metric, face word and the
// neighbour's line (for a face point of an interior face)
template <int N>
__device__ __forceinline__ void adj_point_geom(const AdjArgs &a, const Ops &ops, int tile, int ci,
                                               int fam, int line, int r, double &sx, double &sy,
                                               int &info, int &k, int &nbase, int &np0, int &nst,
                                               bool &endN) {
  using G = Geo<N>;
  constexpr int T = G::CPW, NF = N + 1;
  const int mx = fam ? G::S2X : G::S1X, my = fam ? G::S2Y : G::S1Y;
  const double *mt = a.met + (size_t)tile * G::NM * T + ci;
  sx = mt[(mx + r + NF * line) * T];
  sy = mt[(my + r + NF * line) * T];
  info = 0; k = 0; nbase = 0; np0 = 0; nst = 1; endN = false;
  if (r != 0 && r != N) return;
  int nfp;
  end_face<N>(fam, line, r, k, nfp);
  const int *ft = a.ftab + ((size_t)tile * 8 + k * 2) * T + ci;
  nbase = ft[0];
  info = ft[T];
  if (((info >> 3) & 7) == FT_INTERIOR) neighbour_line<N>(info & 3, nfp, np0, nst, endN);
}

// pass 0: the maxima of eig and um over all flux points (bit patterns of
// non-negative doubles, atomicMax); pass 1: the smallest location code
// attaining each maximum (atomicMin)
template <int N>
__global__ __launch_bounds__(STAGE_THREADS) void adj_argmax_kernel(AdjArgs a, Ops ops_arg,
                                                                   unsigned long long *amax,
                                                                   int *aloc, int pass) {
  using G = Geo<N>;
  constexpr int T = G::CPW, NF = N + 1, NT = 2 * N * T, QT = G::QT;
  __shared__ double so[sizeof(Ops) / sizeof(double)];
  const Ops &ops = ops_to_lds<0>(so, ops_arg);
  const int tid = threadIdx.x, tile = blockIdx.x;
  const int ci = tid % T, fl = tid / T;
  const int fam = fl >= N ? 1 : 0, line = fl - fam * N;
  if (!(tid < NT && a.s2c[tile * T + ci] >= 0)) return;
  const int p0 = fam ? line : N * line, ps = fam ? N : 1;
  double be = 0.0, bu = 0.0;
  for (int r = 0; r < NF; ++r) {
    double sx, sy;
    int info, k, nbase, np0, nst;
    bool endN;
    adj_point_geom<N>(a, ops, tile, ci, fam, line, r, sx, sy, info, k, nbase, np0, nst, endN);
    const bool face = r == 0 || r == N;
    double Lr[N], t[9], qn[9];
    for (int s = 0; s < N; ++s) Lr[s] = ops.L[r][s];
    adj_trace<N>(a.q, tile * QT + ci, p0, ps, Lr, t);
    if (face && ((info >> 3) & 7) == FT_INTERIOR) {
      double Ln[N];
      for (int s = 0; s < N; ++s) Ln[s] = endN ? ops.L[N][s] : ops.L[0][s];
      adj_trace<N>(a.q, nbase, np0, nst, Ln, qn);
    }
    double e, u;
    adj_point_speeds<N, double>(a, info, k, r, face, sx, sy, t, qn, e, u);
    const int loc = (((tile * T + ci) * 2 + fam) * N + line) * NF + r;
    if (pass == 0) {
      be = e > be ? e : be;
      bu = u > bu ? u : bu;
    } else {
      if (face && __double_as_longlong(e) == (long long)amax[0]) atomicMin(aloc, loc);
      if (__double_as_longlong(u) == (long long)amax[1]) atomicMin(aloc + 1, loc);
    }
  }
  if (pass == 0) {
    atomicMax(amax, (unsigned long long)__double_as_longlong(be));
    atomicMax(amax + 1, (unsigned long long)__double_as_longlong(bu));
  }
}

// Qbar += w * d(speed)/dQ at one flux point (kind 0: eig, 1: um): one
// thread, both cells of a face.
template <int N>
__global__ void adj_argmax_apply(AdjArgs a, Ops ops_arg, int loc, int kind, double w) {
  using G = Geo<N>;
  constexpr int T = G::CPW, N2 = G::N2, NF = N + 1, QT = G::QT;
  __shared__ double so[sizeof(Ops) / sizeof(double)];
  const Ops &ops = ops_to_lds<0>(so, ops_arg);
  if (threadIdx.x != 0) return;
  const int r = loc % NF, line = (loc / NF) % N, fam = (loc / (NF * N)) % 2;
  const int slot = loc / (NF * N * 2), tile = slot / T, ci = slot % T;
  const int p0 = fam ? line : N * line, ps = fam ? N : 1;
  double sx, sy;
  int info, k, nbase, np0, nst;
  bool endN;
  adj_point_geom<N>(a, ops, tile, ci, fam, line, r, sx, sy, info, k, nbase, np0, nst, endN);
  const bool face = r == 0 || r == N;
  const bool nbr = face && ((info >> 3) & 7) == FT_INTERIOR;
  double Lr[N], Ln[N], t[9], qn[9];
  for (int s = 0; s < N; ++s) {
    Lr[s] = ops.L[r][s];
    Ln[s] = endN ? ops.L[N][s] : ops.L[0][s];
  }
  adj_trace<N>(a.q, tile * QT + ci, p0, ps, Lr, t);
  if (nbr) adj_trace<N>(a.q, nbase, np0, nst, Ln, qn);
  double *qob = a.qbar + (size_t)tile * QT;
  const int nbt = nbase / QT, nbc = nbase - nbt * QT;
  double *qnb = a.qbar + (size_t)nbt * QT;
  for (int kk = 0; kk < (nbr ? 18 : 9); ++kk) {
    Dual td[9], nd[9], e, u;
    for (int v = 0; v < 9; ++v) {
      td[v] = Dual(t[v], kk == v ? 1.0 : 0.0);
      nd[v] = Dual(nbr ? qn[v] : 0.0, kk == 9 + v ? 1.0 : 0.0);
    }
    adj_point_speeds<N, Dual>(a, info, k, r, face, sx, sy, td, nd, e, u);
    const double d = w * (kind == 0 ? e.d : u.d);
    if (kk < 9) {
      for (int s = 0; s < N; ++s) qob[qo<N>(kk, p0 + ps * s, ci)] += d * Lr[s];
    } else {
      for (int s = 0; s < N; ++s) qnb[qo<N>(kk - 9, np0 + nst * s, nbc)] += d * Ln[s];
    }
  }
}

}  // namespace sd2d
