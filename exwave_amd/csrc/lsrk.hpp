// lsrk.hpp -- low-storage RK stage schedule with the RKVectorUpdater fused
// into each operator sweep, shared by the single-solve path and the
// multi-instance serving pools.
#pragma once

namespace exw {

// LowStorageRK45Reg2 coefficients (source/time_integrators.cc:243-252)
struct Lsrk45 {
  static constexpr int S = 5;
  static constexpr double a[5] = {970286171893. / 4311952581923., 6584761158862. / 12103376702013.,
                                  2251764453980. / 15575788980749.,
                                  26877169314380. / 34165994151039., 0.0};
  static constexpr double b[5] = {1153189308089. / 22510343858157., 1772645290293. / 4653164025191.,
                                  -1672844663538. / 4480602732383., 2114624349019. / 3568978502595.,
                                  5198255086312. / 14908931495163.};
};

// Registers: R_n (vec_n), R_np (vec_np), R_t (work).  Stage 1 reads the state
// from R_n and writes the next stage input to X1 and the accumulator to X2
// ({X1,X2} = {R_t,R_np}); every later non-last stage reads (input, acc),
// writes the new input over acc (pointwise, in place) and the new acc into
// the free register.  The last stage writes acc + f1*k into whichever of
// {acc, free} is R_np.  X1/X2 are chosen so that the last stage's input is not
// R_np.  stage(src, acc, y_out, b_out, f1, f2, is_last) launches one sweep.
template <class Stage>
int lsrk_fused_schedule(int S, const double *a, const double *b, double dt, double *vec_n,
                        double *vec_np, double *work, Stage stage) {
  enum { RN = 0, X1 = 1, X2 = 2 };
  int src = RN, acc = RN, fre = -1;
  for (int s = 0; s < S - 1; ++s) {
    if (s == 0) {
      src = X1;
      acc = X2;
      fre = RN;
    } else {
      const int ns = acc, na = fre, nf = src;
      src = ns;
      acc = na;
      fre = nf;
    }
  }
  double *reg[3];
  reg[RN] = vec_n;
  if (S == 1 || src == X1 || src == RN) {
    reg[X1] = work;
    reg[X2] = vec_np;
  } else {
    reg[X1] = vec_np;
    reg[X2] = work;
  }
  int cs = RN, ca = RN, cf = -1;
  for (int s = 0; s < S; ++s) {
    const bool last = s == S - 1;
    int rc;
    if (!last) {
      const int y_out = s == 0 ? X1 : ca, b_out = s == 0 ? X2 : cf;
      rc = stage(reg[cs], reg[ca], reg[y_out], reg[b_out], a[s] * dt, (b[s] - a[s]) * dt, false);
      if (rc) return rc;
      if (s == 0) {
        cs = X1;
        ca = X2;
        cf = RN;
      } else {
        const int ns = ca, na = cf, nf = cs;
        cs = ns;
        ca = na;
        cf = nf;
      }
    } else {
      double *out = (S == 1 || reg[ca] == vec_np) ? vec_np : reg[cf];
      if (out != vec_np) return -1;
      rc = stage(reg[cs], reg[ca], out, nullptr, b[s] * dt, 0.0, true);
      if (rc) return rc;
    }
  }
  return 0;
}

}  // namespace exw
