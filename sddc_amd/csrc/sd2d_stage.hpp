// sd2d_stage.hpp - the fused residual + Runge-Kutta stage kernel.
//
// One 256-thread workgroup owns a tile of T cells (T = 8/12/20 for
// N = 5/4/3) and takes it through short phases that meet in LDS:
//   P0  the tile's state -> LDS (one contiguous, coalesced block)
//   P1  traces: every grid line of every cell is interpolated to its N+1
//       flux points (item = cell, family, line, variable)
//   P2  fluxes: at every flux point the ideal-MHD + GLM flux (interior) or
//       the Rusanov flux with the neighbour's trace or a boundary ghost
//       (face points); item = one flux point
//   P3  d/dxi and d/deta of the fluxes along each line, and the gradient
//       partials of Bx, By (all 9 variables for the compresid output)
//   P4  per solution point: residual, Powell + GLM sources, RK update,
//       norms, and the AV inputs
//   P5  the AV smoothness sensor per cell
// A face point is evaluated by both adjacent cells from the same two
// traces in the same left/right roles (the neighbour's line is read
// through L2), so both cells store bitwise the same flux and only the
// state, the metrics and the RK registers move through HBM.
#pragma once
#include "sd2d_device.hpp"

namespace sd2d {

constexpr int STAGE_THREADS = 256;

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
  const double *avtab;      // hlow[isl][is] then hhigh[is][isl] (AV.f90:18-63)
  int ncell, ntile;
};

// XCD-aware tile order: consecutive tiles go to the same XCD (and its L2)
// so neighbour lines read by adjacent tiles hit in cache.  Bijective for
// any tile count.
__device__ __forceinline__ int tile_of_block(int b, int nt) {
  const int q = nt >> 3, r = nt & 7;
  const int x = b & 7, i = b >> 3;
  return (x < r) ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

// LDS layout of one tile (all in doubles).
// TL cells of a tile are handled per workgroup: the whole tile for the
// stepping kernel, half a tile for the all-variable compresid variant.
template <int N, int GV> struct Lds {
  static constexpr int NSUB = GV == 9 ? 2 : 1;
  static constexpr int T = Geo<N>::CPW / NSUB;
  static constexpr int N2 = N * N;
  static constexpr int NF = N + 1;
  // state: [v][p][ci]
  static constexpr int Q0 = 0;
  static constexpr int QN = 9 * N2 * T;
  // traces -> fluxes -> line derivatives: [fam][line][rfp][v][ci]
  static constexpr int TR0 = Q0 + QN;
  static constexpr int TRN = 2 * N * NF * 9 * T;
  // averaged states of the gradient variables -> gradient partials:
  // [fam][line][slot][ci], slot < 2*GV*N fits in NF*GV
  static constexpr int GS = (2 * GV * N > NF * GV) ? 2 * GV * N : NF * GV;
  static constexpr int QV0 = TR0 + TRN;
  static constexpr int QVN = 2 * N * GS * T;
  // AV input |Jz|^2 per point: [p][ci]
  static constexpr int JZ0 = QV0 + QVN;
  static constexpr int JZN = N2 * T;
  // block reductions
  static constexpr int RED0 = JZ0 + JZN;
  static constexpr int REDN = 16;
  static constexpr int TOTAL = RED0 + REDN;
  static __host__ __device__ constexpr int tr(int fam, int line, int rfp, int v, int ci) {
    return TR0 + (((fam * N + line) * NF + rfp) * 9 + v) * T + ci;
  }
  static __host__ __device__ constexpr int qv(int fam, int line, int slot, int ci) {
    return QV0 + ((fam * N + line) * GS + slot) * T + ci;
  }
  static __host__ __device__ constexpr int q(int v, int p, int ci) {
    return Q0 + (v * N2 + p) * T + ci;
  }
};

template <int N, int GV>
constexpr size_t stage_lds_bytes() { return sizeof(double) * Lds<N, GV>::TOTAL; }

// Point index p = is + N*js of point s on line `line` of family fam
// (fam 0: row js = line, s = is; fam 1: column is = line, s = js).
template <int N>
__device__ __forceinline__ int line_point(int fam, int line, int s) {
  return fam == 0 ? s + N * line : line + N * s;
}

// Trace of one line of N values with interpolation row Lr:
// acc = sum_s q[s]*Lr[s], summed in order s = 0..N-1 with FMA.  The own
// trace (from LDS) and the neighbour trace (from HBM) use this same
// formula so that both cells form bitwise the same face state.
template <int N>
__device__ __forceinline__ double line_trace(const double *vals, const double *Lr) {
  double acc = 0.0;
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

// MODE bits
enum { M_RK = 1, M_AV = 2, M_NORMS = 4, M_COMPAT = 8 };

// P2 for one face point: own trace Qo (left or right cell of the face),
// neighbour trace from HBM, Rusanov in the reference's roles
// (interfaceflux_all.f90:31-106, cyclocflux_all.f90:226-301,
// bcflux.f90).  Returns the flux stored for this cell and the averaged
// state Qv.
template <int N>
__device__ __forceinline__ void face_point(const StageArgs &a, const Ops &ops,
                                           const Phys &ph, double c_h, int tile,
                                           int cig, int fam, int line, int rfp,
                                           const double *Qo, double sx, double sy,
                                           double *Fh, double *Qv, double &eig,
                                           double &um) {
  constexpr int T = Geo<N>::CPW, N2 = Geo<N>::N2;
  // own local face (0-based) and own point index nfp (1-based), from
  // face2fpconnec (initsetup.f90:356-380)
  const int k = fam == 0 ? (rfp == 0 ? 3 : 1) : (rfp == 0 ? 0 : 2);
  const int nfp = fam == 0 ? (rfp == 0 ? N - line : line + 1)
                           : (rfp == 0 ? line + 1 : N - line);
  const int *ft = a.ftab + ((size_t)tile * 8 + k * 2) * T + cig;
  const int nbase = ft[0];
  const int info = ft[T];
  const int bc = (info >> 3) & 7;
  const int own_sl = (k == 0 || k == 3) ? -1 : 1;  // sign_l if this cell is left
  if (bc == FT_INTERIOR) {
    const int kn = info & 3, left = (info >> 2) & 1, sneg = (info >> 6) & 1;
    const int m = N - nfp + 1;  // neighbour point index on its face (1-based)
    int p0, stride;
    if (kn == 0) { p0 = m - 1; stride = N; }
    else if (kn == 1) { p0 = N * (m - 1); stride = 1; }
    else if (kn == 2) { p0 = N - m; stride = N; }
    else { p0 = N * (N - m); stride = 1; }
    const bool endN = (kn == 1 || kn == 2);
    double Lr[N];
#pragma unroll
    for (int s = 0; s < N; ++s) Lr[s] = endN ? ops.L[N][s] : ops.L[0][s];
    double Qn[9];
#pragma unroll
    for (int v = 0; v < 9; ++v) {
      double vals[N];
#pragma unroll
      for (int s = 0; s < N; ++s) vals[s] = a.qin[nbase + (v * N2 + p0 + stride * s) * T];
      Qn[v] = line_trace<N>(vals, Lr);
    }
    double QL[9], QR[9];
#pragma unroll
    for (int v = 0; v < 9; ++v) {
      QL[v] = left ? Qo[v] : Qn[v];
      QR[v] = left ? Qn[v] : Qo[v];
    }
    const int sl = left ? own_sl : ((kn == 0 || kn == 3) ? -1 : 1);
    const int sr = left ? ((kn == 1 || kn == 2) ? -1 : 1) : ((k == 1 || k == 2) ? -1 : 1);
    const double nsg = (left || !sneg) ? 1.0 : -1.0;
    double Fl[9];
    rusanov<double>(QL, QR, nsg * sx, nsg * sy, sl, c_h, Fl, eig, um);
    const double fac = left ? 1.0 : double(sl * sr);
#pragma unroll
    for (int v = 0; v < 9; ++v) {
      Fh[v] = fac * Fl[v];
      Qv[v] = 0.5 * QL[v] + 0.5 * QR[v];
    }
  } else {
    double QR[9];
    bc_ghost<double>(bc, Qo, sx, sy, ph, QR);
    rusanov<double>(Qo, QR, sx, sy, own_sl, c_h, Fh, eig, um);
#pragma unroll
    for (int v = 0; v < 9; ++v) Qv[v] = 0.5 * Qo[v] + 0.5 * QR[v];
  }
}

template <int N, int GV, int MODE>
__global__ __launch_bounds__(STAGE_THREADS) void stage_kernel(StageArgs a, Ops ops, Phys ph) {
  using G = Geo<N>;
  using LD = Lds<N, GV>;
  constexpr int T = G::CPW;          // cells per tile in HBM
  constexpr int TL = LD::T;          // cells of this workgroup (T / NSUB)
  constexpr int N2 = G::N2, NF = N + 1;
  extern __shared__ double lds[];
  const int tid = threadIdx.x;
  const int tile = LD::NSUB == 1 ? tile_of_block(blockIdx.x, a.ntile) : blockIdx.x / LD::NSUB;
  const int c0 = (LD::NSUB == 1 ? 0 : blockIdx.x % LD::NSUB) * TL;  // first cell of the tile
  const int cell0 = tile * T + c0;
  const int ncv = (a.ncell - cell0) < TL ? (a.ncell - cell0) : TL;  // valid cells (may be <= 0)
  const double ge = a.gmax[0], gu = a.gmax[1];
  const double c_h = sqrt(ge * (ge - gu));   // lagged GLM speed, compflux.f90:6
  double eig = 0.0, um = 0.0;
  const double *qt = a.qin + (size_t)tile * G::QT + c0;
  const double *met = a.met + (size_t)tile * G::NM * T + c0;

  // ---- P0: state of the cells -> LDS ------------------------------------------
  for (int i = tid; i < 9 * N2 * TL; i += STAGE_THREADS) {
    const int ci = i % TL, vp = i / TL;
    lds[LD::Q0 + i] = qt[vp * T + ci];
  }
  __syncthreads();

  // ---- P1: traces, item = (fam, line, v, ci) ----------------------------------
  for (int it = tid; it < 2 * N * 9 * TL; it += STAGE_THREADS) {
    const int ci = it % TL;
    const int r = it / TL;
    const int v = r % 9;
    const int fl = r / 9;
    const int fam = fl / N, line = fl % N;
    if (ci >= ncv) continue;
    double vals[N];
#pragma unroll
    for (int s = 0; s < N; ++s) vals[s] = lds[LD::q(v, line_point<N>(fam, line, s), ci)];
#pragma unroll
    for (int rfp = 0; rfp < NF; ++rfp)
      lds[LD::tr(fam, line, rfp, v, ci)] = line_trace<N>(vals, ops.L[rfp]);
  }
  __syncthreads();

  // ---- P2: fluxes at flux points; face points first, then interior -----------
  {
    constexpr int NFACEIT = 2 * N * 2 * TL;           // (fam, line, end, ci)
    constexpr int NALL = 2 * N * NF * TL;
    for (int it = tid; it < NALL; it += STAGE_THREADS) {
      int ci, fam, line, rfp;
      if (it < NFACEIT) {
        ci = it % TL;
        const int r = it / TL;
        rfp = (r % 2) * N;
        const int fl = r / 2;
        fam = fl / N; line = fl % N;
      } else {
        const int j = it - NFACEIT;
        ci = j % TL;
        const int r = j / TL;
        rfp = 1 + r % (N - 1);
        const int fl = r / (N - 1);
        fam = fl / N; line = fl % N;
      }
      if (ci >= ncv) continue;
      const int mx = fam == 0 ? G::S1X : G::S2X, my = fam == 0 ? G::S1Y : G::S2Y;
      const double sx = met[(mx + rfp + NF * line) * T + ci];
      const double sy = met[(my + rfp + NF * line) * T + ci];
      double Qo[9], Fh[9], Qv[9];
#pragma unroll
      for (int v = 0; v < 9; ++v) Qo[v] = lds[LD::tr(fam, line, rfp, v, ci)];
      if (it < NFACEIT) {
        face_point<N>(a, ops, ph, c_h, tile, c0 + ci, fam, line, rfp, Qo, sx, sy, Fh, Qv,
                      eig, um);
      } else {
        double F[9], Gf[9];
        iflux<double>(Qo, c_h, F, Gf, um);   // inviscid_flux_interior, compflux.f90:53-104
#pragma unroll
        for (int v = 0; v < 9; ++v) {
          Fh[v] = F[v] * sx + Gf[v] * sy;
          Qv[v] = Qo[v];
        }
      }
#pragma unroll
      for (int v = 0; v < 9; ++v) lds[LD::tr(fam, line, rfp, v, ci)] = Fh[v];
#pragma unroll
      for (int g = 0; g < GV; ++g)
        lds[LD::qv(fam, line, rfp * GV + g, ci)] = Qv[GV == 9 ? g : 5 + g];
    }
  }
  __syncthreads();

  // ---- P3: line derivatives (compresid.f90:24-39) and gradient partials
  //          (nablaQsp, compflux.f90:107-157) ---------------------------------
  {
    constexpr int ND = 2 * N * 9 * TL;   // derivative items (fam, line, v, ci)
    constexpr int NG = 2 * N * TL;       // gradient items (fam, line, ci)
    for (int it = tid; it < ND + NG; it += STAGE_THREADS) {
      if (it < ND) {
        const int ci = it % TL;
        const int r = it / TL;
        const int v = r % 9;
        const int fl = r / 9;
        const int fam = fl / N, line = fl % N;
        if (ci >= ncv) continue;
        double f[NF];
#pragma unroll
        for (int rfp = 0; rfp < NF; ++rfp) f[rfp] = lds[LD::tr(fam, line, rfp, v, ci)];
#pragma unroll
        for (int s = 0; s < N; ++s) {
          double acc = 0.0;
#pragma unroll
          for (int rfp = 0; rfp < NF; ++rfp) acc = fma(f[rfp], ops.M[rfp][s], acc);
          lds[LD::tr(fam, line, s, v, ci)] = acc;
        }
      } else {
        const int j = it - ND;
        const int ci = j % TL;
        const int fl = j / TL;
        const int fam = fl / N, line = fl % N;
        if (ci >= ncv) continue;
        const int mx = fam == 0 ? G::S1X : G::S2X, my = fam == 0 ? G::S1Y : G::S2Y;
        double qv[NF][GV], sxv[NF], syv[NF];
#pragma unroll
        for (int rfp = 0; rfp < NF; ++rfp) {
          sxv[rfp] = met[(mx + rfp + NF * line) * T + ci];
          syv[rfp] = met[(my + rfp + NF * line) * T + ci];
#pragma unroll
          for (int g = 0; g < GV; ++g) qv[rfp][g] = lds[LD::qv(fam, line, rfp * GV + g, ci)];
        }
#pragma unroll
        for (int s = 0; s < N; ++s)
#pragma unroll
          for (int g = 0; g < GV; ++g) {
            double ax = 0.0, ay = 0.0;
#pragma unroll
            for (int rfp = 0; rfp < NF; ++rfp) {
              const double t = qv[rfp][g] * ops.M[rfp][s];
              ax = fma(t, sxv[rfp], ax);
              ay = fma(t, syv[rfp], ay);
            }
            lds[LD::qv(fam, line, (s * GV + g) * 2 + 0, ci)] = ax;
            lds[LD::qv(fam, line, (s * GV + g) * 2 + 1, ci)] = ay;
          }
      }
    }
  }
  __syncthreads();

  // ---- P4: solution points, item = (p, ci) ------------------------------------
  double nres = 0.0, ndiv = 0.0;
  for (int it = tid; it < N2 * TL; it += STAGE_THREADS) {
    const int ci = it % TL;
    const int p = it / TL;
    const int is = p % N, js = p / N;
    if (ci >= ncv) continue;
    const int cell = cell0 + ci;
    const double ji = met[(G::JI + js + N * is) * T + ci];
    double R[9], q[9];
#pragma unroll
    for (int v = 0; v < 9; ++v) {
      q[v] = lds[LD::q(v, p, ci)];
      // resid = -(dF/dxi + dG/deta) / J   (compresid.f90:41,56)
      R[v] = -(lds[LD::tr(0, js, is, v, ci)] + lds[LD::tr(1, is, js, v, ci)]) * ji;
    }
    double nab[2 * GV];
#pragma unroll
    for (int g = 0; g < GV; ++g)
#pragma unroll
      for (int d = 0; d < 2; ++d)
        nab[2 * g + d] = (lds[LD::qv(0, js, (is * GV + g) * 2 + d, ci)] +
                          lds[LD::qv(1, is, (js * GV + g) * 2 + d, ci)]) * ji;
    const int ibx = GV == 9 ? 5 : 0, iby = GV == 9 ? 6 : 1;
    const double divB = nab[2 * ibx] + nab[2 * iby + 1];
    const double ri = 1.0 / q[0];
    const double ux = q[1] * ri, uy = q[2] * ri, uz = q[3] * ri;
    // Powell source and GLM damping, compresid.f90:58-74
    R[1] -= divB * q[5];
    R[2] -= divB * q[6];
    R[3] -= divB * q[7];
    R[4] -= divB * (ux * q[5] + uy * q[6] + uz * q[7]);
    R[5] -= divB * ux;
    R[6] -= divB * uy;
    R[7] -= divB * uz;
    R[8] -= ph.alpha * q[8];
    if (MODE & M_AV) {
      const double jz = nab[2 * iby] - nab[2 * ibx + 1];
      lds[LD::JZ0 + p * TL + ci] = jz * jz;
    }
    if (MODE & M_NORMS) {
#pragma unroll
      for (int v = 0; v < 9; ++v) nres = fma(R[v], R[v], nres);
      ndiv = fma(divB, divB, ndiv);
    }
    if (MODE & M_COMPAT) {
      const size_t pf = (size_t)p + (size_t)N2 * cell;
#pragma unroll
      for (int v = 0; v < 9; ++v) a.resid_f[v + 9 * pf] = R[v];
      if (GV == 9) {
#pragma unroll
        for (int v = 0; v < 9; ++v) {
          a.nabla_f[v + 9 * (0 + 2 * pf)] = nab[2 * v];
          a.nabla_f[v + 9 * (1 + 2 * pf)] = nab[2 * v + 1];
        }
      }
    }
    if (MODE & M_RK) {
      const size_t off = (size_t)tile * G::QT + c0 + ci;
#pragma unroll
      for (int v = 0; v < 9; ++v) {
        const size_t o = off + (size_t)(v * N2 + p) * T;
        double acc = a.cin * q[v];
#pragma unroll
        for (int i = 0; i < 4; ++i)
          if (i < a.naux) acc = fma(a.ca[i], a.aux[i][o], acc);
        acc = fma(a.cr, R[v], acc);
        a.qout[o] = acc;
        if (a.rstore) a.rstore[o] = R[v];
      }
    }
  }

  // ---- block reductions: running maxima (getrusanov.f90:116-123,188-191) and
  //      norms (iterations.f90:37-46) ----------------------------------------------
  {
    const int w = tid >> 6, l = tid & 63;
    eig = wave_max(eig);
    um = wave_max(um);
    if (MODE & M_NORMS) {
      nres = wave_sum(nres);
      ndiv = wave_sum(ndiv);
    }
    if (l == 0) {
      lds[LD::RED0 + w] = eig;
      lds[LD::RED0 + 4 + w] = um;
      lds[LD::RED0 + 8 + w] = nres;
      lds[LD::RED0 + 12 + w] = ndiv;
    }
    __syncthreads();
    if (tid == 0) {
      double e = lds[LD::RED0], u = lds[LD::RED0 + 4];
      double s1 = lds[LD::RED0 + 8], s2 = lds[LD::RED0 + 12];
#pragma unroll
      for (int k = 1; k < STAGE_THREADS / 64; ++k) {
        e = lds[LD::RED0 + k] > e ? lds[LD::RED0 + k] : e;
        u = lds[LD::RED0 + 4 + k] > u ? lds[LD::RED0 + 4 + k] : u;
        s1 += lds[LD::RED0 + 8 + k];
        s2 += lds[LD::RED0 + 12 + k];
      }
      unsigned long long *sl = a.slots + 2 * (blockIdx.x % NSLOT);
      atomicMax(sl, (unsigned long long)__double_as_longlong(e));
      atomicMax(sl + 1, (unsigned long long)__double_as_longlong(u));
      if (MODE & M_NORMS) {
        a.normbuf[2 * tile] = s1;
        a.normbuf[2 * tile + 1] = s2;
      }
    }
  }

  // ---- P5: AV smoothness sensor (AV.f90:14-83) --------------------------------
  if (MODE & M_AV) {
    constexpr int NL = N - 1, NL2 = NL * NL;
    const double *hlow = a.avtab;            // [isl][is]
    const double *hhigh = a.avtab + NL * N;  // [is][isl]
    // scratch in the trace area, free after P4
    double *lowb = lds + LD::TR0;                 // [2][NL2][ci]
    double *seb = lds + LD::TR0 + 2 * NL2 * TL;   // [2][N2][ci]
    for (int it = tid; it < NL2 * TL; it += STAGE_THREADS) {
      const int ci = it % TL, t = it / TL;
      const int il = t % NL, jl = t / NL;
      if (ci >= ncv) continue;
      double s1 = 0.0, s2 = 0.0;
      for (int js = 0; js < N; ++js)
        for (int is = 0; is < N; ++is) {
          const double h = hlow[il * N + is] * hlow[jl * N + js];
          s1 = fma(lds[LD::q(0, is + N * js, ci)], h, s1);
          s2 = fma(lds[LD::JZ0 + (is + N * js) * TL + ci], h, s2);
        }
      lowb[(0 * NL2 + t) * TL + ci] = s1;
      lowb[(1 * NL2 + t) * TL + ci] = s2;
    }
    __syncthreads();
    for (int it = tid; it < N2 * TL; it += STAGE_THREADS) {
      const int ci = it % TL, p = it / TL;
      const int is = p % N, js = p / N;
      if (ci >= ncv) continue;
      double h1 = 0.0, h2 = 0.0;
      for (int jl = 0; jl < NL; ++jl)
        for (int il = 0; il < NL; ++il) {
          const double h = hhigh[is * NL + il] * hhigh[js * NL + jl];
          h1 = fma(lowb[(0 * NL2 + il + NL * jl) * TL + ci], h, h1);
          h2 = fma(lowb[(1 * NL2 + il + NL * jl) * TL + ci], h, h2);
        }
      const double rho = lds[LD::q(0, p, ci)], jz2 = lds[LD::JZ0 + p * TL + ci];
      const double d1 = h1 - rho, d2 = h2 - jz2;
      seb[(0 * N2 + p) * TL + ci] = (d1 * d1) / (rho * rho);
      seb[(1 * N2 + p) * TL + ci] = (d2 * d2) / (jz2 * jz2);
    }
    __syncthreads();
    for (int ci = tid; ci < ncv; ci += STAGE_THREADS) {
      double s1 = 0.0, s2 = 0.0;
      for (int p = 0; p < N2; ++p) {
        s1 += seb[(0 * N2 + p) * TL + ci];
        s2 += seb[(1 * N2 + p) * TL + ci];
      }
      a.sm_ind[2 * (size_t)(cell0 + ci)] = log10(s1);
      a.sm_ind[2 * (size_t)(cell0 + ci) + 1] = log10(s2);
    }
  }
}

}  // namespace sd2d
