// sd2d_device.hpp - device-side building blocks of the SD2D-MHD residual:
// pointwise ideal-MHD + GLM physics, the Rusanov interface flux, boundary
// ghost states, and the tile geometry of the fused stage kernel.
//
// Arithmetic follows the reference routines cited per function; the GPU
// versions replace repeated divisions by one reciprocal and let the
// compiler form fused multiply-adds, so results agree with the reference
// to rounding (about 1e-15 relative), not bit for bit.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sd2d {

constexpr int NMAX = 5;
constexpr int NSLOT = 256;          // running-maximum slots (spread atomics)

// Cells per tile for each order.  One 128-thread workgroup of the stage
// kernel owns one tile and runs one thread per grid line of each cell (2N
// lines per cell): 120, 128 and 120 busy threads for N = 5, 4, 3.  Every
// tile of state is a multiple of 16 bytes.
__host__ __device__ constexpr int cpw_of(int n) {
  return n == 5 ? 12 : (n == 4 ? 16 : 20);
}

template <int N> struct Geo {
  static constexpr int CPW = cpw_of(N);
  static constexpr int N2 = N * N;
  // per-cell metric slots: S1(1,1:2,rfp,js), S2(2,1:2,is,rfp), 1/Jac
  static constexpr int S1X = 0;
  static constexpr int S1Y = N * (N + 1);
  static constexpr int S2X = 2 * N * (N + 1);
  static constexpr int S2Y = 3 * N * (N + 1);
  static constexpr int JI = 4 * N * (N + 1);
  static constexpr int NM = 4 * N * (N + 1) + N * N;
  static constexpr int QT = 9 * N2 * CPW;   // doubles per state tile
};

// Offset of variable v at solution point p (= is + N js) of slot ci inside a
// state tile of T = CPW cells (9 N^2 T doubles, in HBM and staged in LDS
// alike): variables 0-7 in pairs, [v/2][p][ci][v%2], so that one 16-byte
// access moves two variables of a point of a cell, then variable 8 alone,
// [p][ci].  Every kernel that reads or writes a state-shaped tile array
// (state, RK registers, viscous divergence, adjoint cotangents) indexes it
// through qo, and qo_inv is its inverse.
template <int N>
__host__ __device__ constexpr int qo(int v, int p, int ci) {
  constexpr int T = cpw_of(N), N2 = N * N;
  return v < 8 ? (((v >> 1) * N2 + p) * T + ci) * 2 + (v & 1) : 8 * N2 * T + p * T + ci;
}
template <int N>
__host__ __device__ inline void qo_inv(int off, int &v, int &p, int &ci) {
  constexpr int T = cpw_of(N), N2 = N * N;
  if (off < 8 * N2 * T) {
    const int k = off / (2 * N2 * T), r = off - k * 2 * N2 * T;
    p = r / (2 * T);
    const int r2 = r - p * 2 * T;
    ci = r2 >> 1;
    v = 2 * k + (r2 & 1);
  } else {
    const int r = off - 8 * N2 * T;
    p = r / T;
    ci = r - p * T;
    v = 8;
  }
}
// the same offset for an order known only at run time (host code)
inline int qo_rt(int N, int v, int p, int ci) {
  const int T = cpw_of(N), N2 = N * N;
  return v < 8 ? (((v >> 1) * N2 + p) * T + ci) * 2 + (v & 1) : 8 * N2 * T + p * T + ci;
}

// Values handed between lanes of one wavefront through LDS: the LDS unit
// performs a wavefront's operations in order, so no wait is needed between
// one lane's store and another lane's load; the fences keep the compiler
// from moving LDS accesses across this point.  Every lane of the wavefront
// must reach it (the CPU emulation of the kernels syncs the wave here).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Interpolation / differentiation operators (initsetup.f90:262-319),
// passed by value in the kernarg.
struct Ops {
  double L[NMAX + 1][NMAX];      // L[ifp][is] = Lmat(ifp,is)
  double M[NMAX + 1][NMAX];      // M[ifp][is] = Mmat(ifp,is)
};

struct Phys {
  double alpha, rinf, pinf, uinf, vinf, bxinf, byinf;
  int vismode;
};

// Boundary types in the per-cell face table.
enum { FT_INTERIOR = 0, FT_INLE = 1, FT_OUTL = 2, FT_WALL = 3, FT_SYMP = 4,
       FT_FREE = 5 };

// face-info word: bits 0-1 neighbour local face (0-based), bit 2 this
// cell is the left cell, bits 3-5 boundary type, bit 6 the left cell's
// normal is the negative of this cell's own metric vector.
__host__ __device__ inline int finfo_pack(int kn, int left, int bc, int sneg) {
  return (kn & 3) | ((left & 1) << 2) | ((bc & 7) << 3) | ((sneg & 1) << 6);
}

constexpr double GAM = 5.0 / 3;   // setup2d.f90:67

// 1/x.  FP64: the hardware estimate refined by two Newton-Raphson steps
// r <- r + r (1 - x r), within an ulp of the quotient (the division
// sequence costs twice the instructions); FP32 (the mixed-precision path):
// the hardware reciprocal, accurate to 1 ulp.
__device__ __forceinline__ double rcp(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}
__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// Gas pressure (gamma-1)*(E - |m|^2/(2 rho) - |B|^2/2 - psi^2/2), the
// recovery that cancels at low plasma beta.  The FP64 path keeps the
// operation order of each caller (iflux_n, rusanov); the mixed path
// (T = float) forms it in FP64 from the FP32 traces and rounds once, after
// the cancellation.
__device__ __forceinline__ float pressure_mx(float rho, float m1, float m2, float m3, float E,
                                             float b1, float b2, float b3, float psi) {
  const double r = rho, x = m1, y = m2, z = m3, bx = b1, by = b2, bz = b3, ps = psi;
  const double ke = 0.5 * (x * x + y * y + z * z) / r;
  const double me = 0.5 * (bx * bx + by * by + bz * bz);
  return float((GAM - 1.0) * (((double(E) - ke) - me) - 0.5 * ps * ps));
}

// ---------------------------------------------------------------------------
// getIfluxvectors (getrusanov.f90:127-193) contracted with the metric
// vector (sx, sy) of the flux point, as inviscid_flux_interior stores it
// (F1 = F*S1(1,1) + G*S1(1,2), compflux.f90:79,97): the ideal-MHD + GLM
// flux normal to (sx, sy), formed directly from the normal velocity and
// normal field.  The running maximum of |u|,|v|,|w| is returned through um.
template <typename T>
__device__ __forceinline__ void iflux_n(const T *q, T c_h, T sx, T sy, T *Fn, T &um) {
  const T rho = q[0], rhou = q[1], rhov = q[2], rhow = q[3], rhoE = q[4];
  const T Bx = q[5], By = q[6], Bz = q[7], psi = q[8];
  const T ri = rcp(rho);
  const T u = rhou * ri, v = rhov * ri, w = rhow * ri;
  const T B2 = Bx * Bx + By * By + Bz * Bz;
  T pr;
  if constexpr (sizeof(T) == 4)
    pr = pressure_mx(rho, rhou, rhov, rhow, rhoE, Bx, By, Bz, psi);
  else
    pr = T(GAM - 1.0) *
         (rhoE - T(0.5) * (rhou * u + rhov * v + rhow * w) - T(0.5) * B2 - T(0.5) * psi * psi);
  const T pt = pr + T(0.5) * B2;
  const T ub = u * Bx + v * By + w * Bz;
  const T ef = rhoE - T(0.5) * psi * psi + pt;
  const T un = u * sx + v * sy;
  const T bn = Bx * sx + By * sy;
  const T cp = c_h * psi;
  Fn[0] = rhou * sx + rhov * sy;
  Fn[1] = rhou * un + pt * sx - Bx * bn;
  Fn[2] = rhov * un + pt * sy - By * bn;
  Fn[3] = rhow * un - Bz * bn;
  Fn[4] = un * ef - ub * bn + cp * bn;
  Fn[5] = cp * sx + (v * Bx - u * By) * sy;
  Fn[6] = (u * By - v * Bx) * sx + cp * sy;
  Fn[7] = Bz * un - w * bn;
  Fn[8] = c_h * bn;
  T ut = fmax(fmax(fabs(u), fabs(v)), fabs(w));
  if (ut > um) um = ut;
}

// getrusanovflux (getrusanov.f90:1-125).  Returns the flux stored for the
// left cell (Fnl) and the factor sign_l*sign_r that gives the right one.
template <typename T>
__device__ __forceinline__ void rusanov(const T *QL, const T *QR, T nx, T ny,
                                        int sign_l, T c_h, T *Fn, T &eig,
                                        T &um) {
  const T mag = sqrt(nx * nx + ny * ny);
  const T im = T(sign_l) * rcp(mag);
  const T ux = nx * im, uy = ny * im;
  const T ril = rcp(QL[0]), rir = rcp(QR[0]);
  const T ul = ril * QL[1], vl = ril * QL[2], wl = ril * QL[3];
  const T ur = rir * QR[1], vr = rir * QR[2], wr = rir * QR[3];
  const T B2l = QL[5] * QL[5] + QL[6] * QL[6] + QL[7] * QL[7];
  const T B2r = QR[5] * QR[5] + QR[6] * QR[6] + QR[7] * QR[7];
  T pgl, pgr;
  if constexpr (sizeof(T) == 4) {
    pgl = pressure_mx(QL[0], QL[1], QL[2], QL[3], QL[4], QL[5], QL[6], QL[7], QL[8]);
    pgr = pressure_mx(QR[0], QR[1], QR[2], QR[3], QR[4], QR[5], QR[6], QR[7], QR[8]);
  } else {
    pgl = T(GAM - 1.0) * (QL[4] - T(0.5) * QL[0] * (ul * ul + vl * vl + wl * wl) -
                          B2l * T(0.5) - T(0.5) * QL[8] * QL[8]);
    pgr = T(GAM - 1.0) * (QR[4] - T(0.5) * QR[0] * (ur * ur + vr * vr + wr * wr) -
                          B2r * T(0.5) - T(0.5) * QR[8] * QR[8]);
  }
  const T ptl = pgl + T(0.5) * B2l, ptr = pgr + T(0.5) * B2r;
  const T Vnl = ul * ux + vl * uy, Vnr = ur * ux + vr * uy;
  const T Bnl = QL[5] * ux + QL[6] * uy, Bnr = QR[5] * ux + QR[6] * uy;
  T Fl[9], Fr[9];
  Fl[0] = QL[0] * Vnl;
  Fl[1] = QL[1] * Vnl + ptl * ux - QL[5] * Bnl;
  Fl[2] = QL[2] * Vnl + ptl * uy - QL[6] * Bnl;
  Fl[3] = QL[3] * Vnl - QL[7] * Bnl;
  Fl[4] = (QL[4] - T(0.5) * QL[8] * QL[8] + ptl) * Vnl -
          Bnl * (ul * QL[5] + vl * QL[6] + wl * QL[7]) + c_h * QL[8] * Bnl;
  Fl[5] = (vl * QL[5] - ul * QL[6]) * uy + c_h * QL[8] * ux;
  Fl[6] = (ul * QL[6] - vl * QL[5]) * ux + c_h * QL[8] * uy;
  Fl[7] = QL[7] * Vnl - wl * Bnl;
  Fl[8] = c_h * Bnl;
  Fr[0] = QR[0] * Vnr;
  Fr[1] = QR[1] * Vnr + ptr * ux - QR[5] * Bnr;
  Fr[2] = QR[2] * Vnr + ptr * uy - QR[6] * Bnr;
  Fr[3] = QR[3] * Vnr - QR[7] * Bnr;
  Fr[4] = (QR[4] - T(0.5) * QR[8] * QR[8] + ptr) * Vnr -
          Bnr * (ur * QR[5] + vr * QR[6] + wr * QR[7]) + c_h * QR[8] * Bnr;
  Fr[5] = (vr * QR[5] - ur * QR[6]) * uy + c_h * QR[8] * ux;
  Fr[6] = (ur * QR[6] - vr * QR[5]) * ux + c_h * QR[8] * uy;
  Fr[7] = QR[7] * Vnr - wr * Bnr;
  Fr[8] = c_h * Bnr;
  // fast magneto-acoustic speed from averaged B and rho (:92-105)
  const T bx = T(0.5) * (QL[5] + QR[5]), by = T(0.5) * (QL[6] + QR[6]),
          bz = T(0.5) * (QL[7] + QR[7]);
  const T irs = rcp(QL[0] + QR[0]);
  const T ca = T(GAM) * (pgl + pgr) * irs;
  const T cd = T(2) * (bx * bx + by * by + bz * bz) * irs;
  const T bn = bx * ux + by * uy;
  const T cn = T(2) * bn * bn * irs;
  const T cad = ca + cd;
  const T ct = sqrt(cad * cad - T(4) * ca * cn);
  const T cf = sqrt(T(0.5) * (cad + ct));
  const T ev = T(0.5) * fabs(Vnl + Vnr) + cf;
  const T sc = T(0.5) * mag * T(sign_l);
#pragma unroll
  for (int k = 0; k < 9; ++k) Fn[k] = sc * (Fl[k] + Fr[k] - ev * (QR[k] - QL[k]));
  if (ev > eig) eig = ev;
  T ut = fmax(fmax(T(0.5) * fabs(ul + ur), T(0.5) * fabs(vl + vr)),
              T(0.5) * fabs(wl + wr));
  if (ut > um) um = ut;
}

// Ghost state of BCFLUX (bcflux.f90:17-375) for the left trace Ql with the
// cell's own metric normal (nx,ny).  Every boundary type's value is formed
// and the one of type `bc` is selected element by element, so the ghost
// state stays in registers (boundary points are a vanishing fraction of
// all face points, so the extra arithmetic costs nothing measurable).
// The free stream enters as values of T (rinf, pinf, uinf, vinf, bxinf,
// byinf) so that the adjoint can differentiate the ghost with respect to it.
template <typename T>
__device__ __forceinline__ void bc_ghost_p(int bc, const T *Ql, T nx, T ny, T rinf, T pinf,
                                           T ui, T vi, T bxi, T byi, int vismode, T *Qr) {
  const T g1 = T(GAM - 1.0);
  const T ri = rcp(Ql[0]);
  const T uw = Ql[1] * ri, vw = Ql[2] * ri, ww = Ql[3] * ri;
  const T im = rcp(sqrt(nx * nx + ny * ny));
  const T vn = (uw * nx + vw * ny) * im;
  const T ur = uw - T(2) * vn * nx * im, vr = vw - T(2) * vn * ny * im;  // reflected
  const T p_e = pinf / g1;
  const T half_psi2 = T(0.5) * Ql[8] * Ql[8];
  // INLE (bcflux.f90:51-72)
  const T e_in = p_e + T(0.5) * rinf * (ui * ui + vi * vi) +
                 T(0.5) * (bxi * bxi + byi * byi) + half_psi2;
  // OUTL (bcflux.f90:122-143)
  const T e_out = p_e + T(0.5) * Ql[0] * (uw * uw + vw * vw) +
                  T(0.5) * (bxi * bxi + byi * byi) + half_psi2;
  // SYMP (bcflux.f90:272-299)
  const T rs = rinf * T(2) - Ql[0];
  const T ms1 = rs * ur, ms2 = rs * vr, ms3 = -rs * ww;
  const T e_sy = p_e + T(0.5) * rcp(rs) * (ms1 * ms1 + ms2 * ms2 + ms3 * ms3) +
                 T(0.5) * (bxi * bxi + byi * byi + Ql[7] * Ql[7] + Ql[8] * Ql[8]);
  // WALL (bcflux.f90:196-227): slip reflection for vismode 0, mirror else
  const bool slip = vismode == 0;
  const T mw1 = slip ? Ql[0] * ur : -Ql[1];
  const T mw2 = slip ? Ql[0] * vr : -Ql[2];
  const T mw3 = -Ql[3];
  const bool in = bc == FT_INLE, out = bc == FT_OUTL, wall = bc == FT_WALL,
             sy = bc == FT_SYMP;
  Qr[0] = in ? rinf : (sy ? rs : Ql[0]);
  Qr[1] = in ? rinf * ui : (out ? Ql[0] * uw : (wall ? mw1 : (sy ? ms1 : Ql[1])));
  Qr[2] = in ? rinf * vi : (out ? Ql[0] * vw : (wall ? mw2 : (sy ? ms2 : Ql[2])));
  Qr[3] = (in || out) ? T(0) : (wall ? mw3 : (sy ? ms3 : Ql[3]));
  Qr[4] = in ? e_in : (out ? e_out : (sy ? e_sy : Ql[4]));
  Qr[5] = (in || out || sy) ? bxi : (wall ? -Ql[5] : Ql[5]);
  Qr[6] = (in || out || sy) ? byi : (wall ? -Ql[6] : Ql[6]);
  Qr[7] = (in || out) ? T(0) : ((wall || sy) ? -Ql[7] : Ql[7]);
  Qr[8] = out ? T(0) : Ql[8];
}

template <typename T>
__device__ __forceinline__ void bc_ghost(int bc, const T *Ql, T nx, T ny,
                                         const Phys &ph, T *Qr) {
  bc_ghost_p<T>(bc, Ql, nx, ny, T(ph.rinf), T(ph.pinf), T(ph.uinf), T(ph.vinf), T(ph.bxinf),
                T(ph.byinf), ph.vismode, Qr);
}

}  // namespace sd2d
