// Host-side state of one single-root Davidson solve (reference:
// chemistry/tpb/davidson.h:707-970), advanced one device pass at a time.
//
// The solver is written as a state machine so that one implementation serves
// two drivers: the synchronous single-solve driver (sbd_davidson / sbd_diag)
// and the batching engine, which advances many solves together and runs each
// kind of pass for all of them in one launch.  A driver asks next() which pass
// to run, runs it on the device with the parameters the job exposes, and
// hands the reduced results back through the matching *_done() call.
//
// Step structure (subspace size k = ib + 1):
//   kSigma  : HC[ib] = H C[ib] and <C[ib], HC[ib]>        -> sigma_done
//   (host)    projected column from the carried <HC_j, C[ib]>, Ritz problem
//   kRitz   : t = (E0 - d)^-1 (sum u HC - E0 sum u C) and its sums -> ritz_done
//   kGs     : classical Gram-Schmidt of t against C[0..ib] -> gs_done
//   kDots   : (only when 99.9% of |t| cancels) <C_j, t>    -> dots_done
//   kLincomb: restart vector sum u C into the spare buffer  -> lincomb_done
//             (not reached from a Ritz pass run with ritz_restart(): at the
//             last step of a cycle that pass itself stores sum u C in the
//             spare buffer, t's place, with its |.|^2 in ritz sum 2)
//   kScale  : renormalise the restart vector                 -> scale_done
//   kDone
#pragma once

#include <cmath>
#include <vector>

namespace sbdx {

void syev_jacobi(int n, double* A, int lda, double* w);

template <typename VT>
class DavidsonJob {
 public:
  enum Op { kSigma, kRitz, kGs, kDots, kLincomb, kScale, kDone };

  // device buffers, owned by the driver: C[nb], HC[nb], spare; dii = diagonal
  std::vector<VT*> C, HC;
  VT* spare = nullptr;
  const VT* dii = nullptr;
  uint64_t N = 0;
  int nb = 0, max_it = 0;
  double eps = 0.0;

  // progress
  int it = 0, ib = 0, steps = 0, last_k = 1;
  bool converged = false;
  double E0 = 0.0, normR = 0.0;
  std::vector<double> Evals;  // Ritz values of the last step, ascending

  void init(const std::vector<VT*>& C_, const std::vector<VT*>& HC_, VT* spare_, const VT* dii_,
            uint64_t N_, int max_it_, double eps_) {
    C = C_;
    HC = HC_;
    spare = spare_;
    dii = dii_;
    N = N_;
    nb = (int)C.size();
    max_it = max_it_;
    eps = eps_;
    H.assign((size_t)nb * nb, 0.0);
    U.assign((size_t)nb * nb, 0.0);
    Evals.assign(nb, 0.0);
    u.assign(nb, 0.0);
    sums.assign(2 * nb + 3, 0.0);
    hnext.assign(nb, 0.0);
    hcur.assign(nb, 0.0);
    o2.assign(nb, 0.0);
    op = kSigma;
  }

  Op next() const { return op; }
  void stop() { op = kDone; }

  // ---- parameters of the pending pass
  int k() const { return ib + 1; }
  VT* sigma_in() const { return C[ib]; }
  VT* sigma_out() const { return HC[ib]; }
  const double* ritz_u() const { return u.data(); }
  VT* t() const { return ib < nb - 1 ? C[ib + 1] : spare; }
  // the Ritz pass of this step writes the restart vector into t() (the spare)
  bool ritz_restart() const { return ib == nb - 1; }
  double* ritz_sums() { return sums.data(); }  // 3 + 2k values
  const double* gs_coef() const { return gs_second ? o2.data() : sums.data() + 3; }
  double gs_scale() const { return gs_sc; }
  double* dots_out() { return o2.data(); }
  int lincomb_k() const { return last_k; }
  VT* lincomb_out() const { return spare; }
  double scale_factor() const { return sc_restart; }

  // ---- results of the pass
  void sigma_done(double tdiag) {
    for (int jb = 0; jb < ib; jb++) {
      H[jb + (size_t)nb * ib] = hnext[jb];
      H[ib + (size_t)nb * jb] = hnext[jb];
    }
    H[ib + (size_t)nb * ib] = tdiag;
    for (int jb = 0; jb <= ib; jb++)
      for (int kb = 0; kb <= ib; kb++) U[jb + (size_t)nb * kb] = H[jb + (size_t)nb * kb];
    syev_jacobi(ib + 1, U.data(), nb, Evals.data());
    E0 = Evals[0];
    for (int q = 0; q <= ib; q++) u[q] = U[q];
    last_k = ib + 1;
    op = kRitz;
  }

  // sums as vec_ritz_precond returns them; true when this step converged
  bool ritz_done() {
    normR = std::sqrt(sums[0]);
    steps++;
    if (normR < eps) {
      converged = true;
      op = kDone;
      return true;
    }
    if (ib < nb - 1) {
      const int k1 = ib + 1;
      const double tt = sums[2];
      double proj = tt;
      for (int q = 0; q < k1; q++) proj -= sums[3 + q] * sums[3 + q];
      if (tt > 0 && proj > 1e-3 * tt) {
        gs_second = false;
        gs_sc = 1.0 / std::sqrt(proj);
        gs_pending_reorth = false;
      } else {
        // the direction is nearly inside the subspace: orthogonalise twice
        gs_second = false;
        gs_sc = 1.0;
        gs_pending_reorth = true;
      }
      op = kGs;
    } else {
      end_of_block();
      // the pass stored the restart vector sum u C in the spare buffer
      if (op == kLincomb) lincomb_done(sums[2]);
    }
    return false;
  }

  void gs_done(double n2) {
    const int k1 = ib + 1;
    const double* oc = sums.data() + 3;
    const double* oh = sums.data() + 3 + k1;
    if (!gs_second) {
      // <HC_j, t - sum_i c_i C_i> = oh_j - sum_i c_i H[i][j]
      project(oc, oh, hcur.data());
      if (gs_pending_reorth) {
        n2_first = n2;
        op = kDots;
        return;
      }
    } else {
      project(o2.data(), hcur.data(), hcur.data());
    }
    for (int j = 0; j < k1; j++) hnext[j] = hcur[j] * gs_sc;
    ib++;
    op = kSigma;
  }

  void dots_done() {
    const int k1 = ib + 1;
    double p2 = n2_first;
    for (int q = 0; q < k1; q++) p2 -= o2[q] * o2[q];
    if (!(p2 > 0.0)) {
      // the new direction lies in the subspace: nothing left to add
      op = kDone;
      return;
    }
    gs_second = true;
    gs_pending_reorth = false;
    gs_sc = 1.0 / std::sqrt(p2);
    op = kGs;
  }

  // restart vector sum u C computed into spare with |.|^2 = n2
  void lincomb_done(double n2) {
    if (std::fabs(n2 - 1.0) > 1e-12) {
      sc_restart = 1.0 / std::sqrt(n2);
      op = kScale;
    } else {
      restart();
    }
  }
  void scale_done() { restart(); }

 private:
  Op op = kSigma;
  std::vector<double> H, U, u, sums, hnext, hcur, o2;
  bool gs_second = false, gs_pending_reorth = false;
  double gs_sc = 1.0, n2_first = 0.0, sc_restart = 1.0;

  void project(const double* c, const double* h, double* out) {
    const int k1 = ib + 1;
    std::vector<double> tmp(k1);
    for (int j = 0; j < k1; j++) {
      double x = h[j];
      for (int i = 0; i < k1; i++) x -= c[i] * H[i + (size_t)nb * j];
      tmp[j] = x;
    }
    for (int j = 0; j < k1; j++) out[j] = tmp[j];
  }

  void end_of_block() {
    it++;
    if (it >= max_it) {
      op = kDone;
      return;
    }
    // restart with C0 = W (davidson.h:958-962)
    op = kLincomb;
  }

  void restart() {
    // the restart vector becomes C[0]; the old C[0] buffer becomes the spare
    VT* old = C[0];
    C[0] = spare;
    spare = old;
    ib = 0;
    op = kSigma;
  }
};

}  // namespace sbdx
