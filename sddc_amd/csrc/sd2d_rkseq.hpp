// sd2d_rkseq.hpp - the stage sequences of the two Runge-Kutta schemes
// (host side; `launch(args, mode)` runs one stage kernel and the
// running-maximum update).  The state rotates through three buffers so no
// copy is made; FiveStageRk keeps its registers qj2, qj3, resid0 in three
// more.
#pragma once
#include "sd2d_stage.hpp"

namespace sd2d {

struct RkBuffers {
  double *q[3];      // state buffers; q[cur] holds the state
  double *fsr[3];    // FiveStageRk: qj2, qj3, resid0
  int cur;
};

// RK_S5P3 (iterations.f90:158-213), Spiteri-Ruuth SSPRK(5,3):
//   s1 A->B, s2 B->C, s3 C(+A)->B, s4 B(+A)->A in place, s5 A(+C)->C in place
template <class Launch>
int s5p3_sequence(RkBuffers &rb, double dt, bool norms, int av, const StageArgs &base,
                  Launch &&launch) {
  const double a30 = 0.355909775063327, a32 = 0.644090224936674;
  const double a40 = 0.367933791638137, a43 = 0.632066208361863;
  const double a52 = 0.237593836598569, a54 = 0.762406163401431;
  const double b10 = 0.377268915331368, b21 = 0.377268915331368;
  const double b32 = 0.242995220537396, b43 = 0.238458932846290;
  const double b54 = 0.287632146308408;
  double *A = rb.q[rb.cur], *B = rb.q[(rb.cur + 1) % 3], *C = rb.q[(rb.cur + 2) % 3];
  struct S { double *in, *aux, *out; double ca, cin, cr; } st[5] = {
      {A, nullptr, B, 0.0, 1.0, b10 * dt},
      {B, nullptr, C, 0.0, 1.0, b21 * dt},
      {C, A, B, a30, a32, b32 * dt},
      {B, A, A, a40, a43, b43 * dt},
      {A, C, C, a52, a54, b54 * dt}};
  for (int s = 0; s < 5; ++s) {
    StageArgs a = base;
    a.qin = st[s].in;
    a.qout = st[s].out;
    a.cin = st[s].cin;
    a.cr = st[s].cr;
    a.naux = st[s].aux ? 1 : 0;
    a.aux[0] = st[s].aux;
    a.ca[0] = st[s].ca;
    // the AV sensor is a diagnostic (it does not feed back), so it is
    // formed on the last stage only: after every step SM_IND is the last
    // residual evaluation's, as in the reference
    const int mode = M_RK | (a.naux ? M_AUX1 : 0) | (s == 4 ? av : 0) |
                     ((norms && s == 4) ? M_NORMS : 0);
    const int r = launch(a, mode);
    if (r) return r;
  }
  rb.cur = (rb.cur + 2) % 3;
  return 0;
}

// FiveStageRk (iterations.f90:88-156), SSPRK(5,4)
template <class Launch>
int fsrk_sequence(RkBuffers &rb, double dt, bool norms, int av, const StageArgs &base,
                  Launch &&launch) {
  const double a11 = 0.39175222700392;
  const double a21 = 0.44437049406734, a22 = 0.55562950593266, a23 = 0.36841059262959;
  const double a31 = 0.6201018513854, a32 = 0.3798981486146, a33 = 0.25189177424738;
  const double a41 = 0.17807995410773, a42 = 0.82192004589227, a43 = 0.54497475021237;
  const double a51 = 0.00683325884039, a52 = 0.51723167208978, a53 = 0.12759831133288;
  const double a54 = 0.34833675773694, a55 = 0.08460416338212, a56 = 0.22600748319395;
  double *Q0 = rb.q[rb.cur], *X = rb.q[(rb.cur + 1) % 3], *Y = rb.q[(rb.cur + 2) % 3];
  double *qj2 = rb.fsr[0], *qj3 = rb.fsr[1], *res0 = rb.fsr[2];
  for (int s = 0; s < 5; ++s) {
    StageArgs a = base;
    int mode = M_RK | (s == 4 ? av : 0);   // AV sensor on the last stage only
    switch (s) {
      case 0: a.qin = Q0; a.qout = X; a.cin = 1.0; a.cr = dt * a11; break;
      case 1: a.qin = X; a.qout = qj2; a.cin = a22; a.cr = a23 * dt;
              a.naux = 1; a.aux[0] = Q0; a.ca[0] = a21; break;
      case 2: a.qin = qj2; a.qout = qj3; a.cin = a32; a.cr = a33 * dt;
              a.naux = 1; a.aux[0] = Q0; a.ca[0] = a31; break;
      case 3: a.qin = qj3; a.qout = Y; a.cin = a42; a.cr = a43 * dt;
              a.naux = 1; a.aux[0] = Q0; a.ca[0] = a41; a.rstore = res0; break;
      default:
        a.qin = Y; a.qout = X; a.cin = a54; a.cr = a56 * dt;
        a.naux = 4;
        a.aux[0] = Q0; a.ca[0] = a51;
        a.aux[1] = qj2; a.ca[1] = a52;
        a.aux[2] = qj3; a.ca[2] = a53;
        a.aux[3] = res0; a.ca[3] = a55 * dt;
        if (norms) mode |= M_NORMS;
        break;
    }
    mode |= a.naux == 4 ? M_AUX4 : (a.naux == 1 ? M_AUX1 : 0);
    if (a.rstore) mode |= M_RSTORE;
    const int r = launch(a, mode);
    if (r) return r;
  }
  rb.cur = (rb.cur + 1) % 3;
  return 0;
}

}  // namespace sd2d
