// Restarted Lanczos on device vectors (method 2 of sbd::tpb::diag).
//
// Reference: chemistry/tpb/lanczos.h:162-327 (the operator form that
// sbd::tpb::diag calls for method 2, chemistry/tpb/sbdiag.h:461-468).  Each
// iteration restarts the three-term recurrence from C[0] = W, with no
// reorthogonalisation:
//   HC = 0; for ib: HC += H C[ib]; alpha = <C[ib], HC>; E = eig(A[0:n, 0:n]);
//   stop if |E[0] - E_old| < eps; HC -= alpha C[ib]; beta = |HC|;
//   C[ib+1] = HC / beta; break if beta < eps; HC = -beta C[ib];
// then W = sum_ib U[ib, 0] C[ib] (not renormalised, as the reference).
//
// Device work per step: one sigma accumulating into HC, whose fused <T, W>
// sum is alpha; one Gram-Schmidt pass HC -= alpha C[ib] returning |HC|^2 (the
// HC buffer then becomes C[ib+1]); a scale of C[ib+1] by 1/beta; and the next
// HC = -beta C[ib] written into a free buffer.  The (n x n) tridiagonal
// eigenproblem is solved on the host.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "sbd_internal.hpp"

namespace sbdx {

void syev_jacobi(int n, double* A, int lda, double* w);

namespace {
struct Buf {
  double* p = nullptr;
  ~Buf() {
    if (p) (void)hipFree(p);
  }
};
}  // namespace

int lanczos_core(sbd_ctx* ctx, const double* hii, double* W, int max_it, int nb, double eps,
                 int verbose, double* energy, std::vector<double>* hist, sbd_report* rep,
                 double t_start) {
  const uint64_t N = ctx->N;
  hipStream_t st = ctx->stream;
  if (nb < 1 || nb > kMaxVecs) return set_err(ctx, SBD_ERR_ARG, "max_nb must be in [1, 40]");
  if (max_it < 1) return set_err(ctx, SBD_ERR_ARG, "max_it must be >= 1");
  // nb Lanczos vectors and HC; the HC buffer of step ib becomes C[ib+1] and
  // the buffer kept for C[ib+1] takes the next HC
  std::vector<Buf> pool(nb + 1);
  for (auto& b : pool)
    if (hipMalloc(&b.p, 8 * N) != hipSuccess)
      return set_err(ctx, SBD_ERR_MEMORY, "device allocation of Lanczos vectors failed");
  std::vector<double*> C(nb, nullptr);
  std::vector<double> A((size_t)nb * nb, 0.0), U((size_t)nb * nb, 0.0), E(nb, 0.0);
  int steps = 0, n = 0;
  bool stop = false;
  double E0 = 0.0, dE = 0.0, t_sigma = 0.0;
  hipEvent_t ev0, ev1;
  SBD_HIP_TRY(ctx, hipEventCreate(&ev0));
  SBD_HIP_TRY(ctx, hipEventCreate(&ev1));
  int rc = 0;
  for (int it = 0; it < max_it && !rc; it++) {
    for (int k = 0; k < nb; k++) C[k] = pool[k].p;
    double* HC = pool[nb].p;
    SBD_HIP_TRY(ctx, hipMemcpyAsync(C[0], W, 8 * N, hipMemcpyDeviceToDevice, st));
    // HC = 0 then HC += H C[0]: the first sigma of a pass overwrites HC
    // (accumulate = 0 is bit-identical to accumulating onto zeros)
    n = 0;
    double E_old = 1.0e+8;
    stop = false;
    for (int ib = 0; ib < nb; ib++) {
      n++;
      double alpha = 0.0;
      SBD_HIP_TRY(ctx, hipEventRecord(ev0, st));
      rc = launch_sigma<double>(ctx, hii, C[ib], HC, ib == 0 ? 0 : 1, st, &alpha);
      if (rc) break;
      SBD_HIP_TRY(ctx, hipEventRecord(ev1, st));
      SBD_HIP_TRY(ctx, hipEventSynchronize(ev1));
      float ms = 0.f;
      SBD_HIP_TRY(ctx, hipEventElapsedTime(&ms, ev0, ev1));
      t_sigma += 1e-3 * ms;
      A[ib + (size_t)nb * ib] = alpha;
      for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) U[i + (size_t)nb * j] = A[i + (size_t)nb * j];
      syev_jacobi(n, U.data(), nb, E.data());
      steps++;
      dE = std::fabs(E[0] - E_old);
      E0 = E[0];
      double beta = 0.0;
      const bool converged = dE < eps;
      const bool last = ib + 1 == nb;
      if (!converged && !last) {
        // HC -= alpha C[ib]; beta = |HC|; C[ib+1] = HC / beta
        const double* Cc[1] = {C[ib]};
        const double coef[1] = {alpha};
        double n2 = 0.0;
        rc = vec_gs_update<double>(ctx, Cc, 1, coef, HC, 1.0, N, &n2, st);
        if (rc) break;
        beta = std::sqrt(n2);
        A[ib + (size_t)nb * (ib + 1)] = beta;
        A[ib + 1 + (size_t)nb * ib] = beta;
        std::swap(C[ib + 1], HC);
        rc = vec_scale<double>(ctx, C[ib + 1], 1.0 / beta, N, st);
        if (rc) break;
      }
      if (hist) {
        hist->push_back(E0);
        hist->push_back(beta);
        hist->push_back(now_s() - t_start);
        hist->push_back(64.0);
      }
      if (verbose) {
        printf(" Lanczos iteration %d, step %d: (A,B)=(%.16g,%.16g):", it, ib, alpha, beta);
        for (int p = 0; p < std::min(n, 4); p++) printf(" %.16g", E[p]);
        printf("\n");
        fflush(stdout);
      }
      if (converged) {
        stop = true;
        break;
      }
      E_old = E[0];
      if (last || std::fabs(beta) < eps) break;
      // HC = -beta C[ib]
      const double* Cc[1] = {C[ib]};
      const double coef[1] = {-beta};
      double n2 = 0.0;
      rc = vec_lincomb<double>(ctx, Cc, 1, coef, 1.0, HC, N, &n2, st);
      if (rc) break;
    }
    if (rc) break;
    // W = sum_ib U[ib, 0] C[ib] over the n steps of this pass
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) U[i + (size_t)nb * j] = A[i + (size_t)nb * j];
    syev_jacobi(n, U.data(), nb, E.data());
    std::vector<const double*> Cc(C.begin(), C.begin() + n);
    double n2 = 0.0;
    rc = vec_lincomb<double>(ctx, Cc.data(), n, U.data(), 1.0, W, N, &n2, st);
    if (rc) break;
    // (C and HC only permute the pool's buffers, which the pool keeps owning;
    // every pass takes them afresh)
    if (stop) break;
  }
  (void)hipEventDestroy(ev0);
  (void)hipEventDestroy(ev1);
  if (!rc && hipStreamSynchronize(st) != hipSuccess) rc = SBD_ERR_DEVICE;
  if (energy) *energy = E0;
  if (rep) {
    rep->steps += steps;
    rep->converged = stop ? 1 : 0;
    rep->residual = dE;
    rep->energy = E0;
    rep->t_sigma += t_sigma;
  }
  return rc;
}

}  // namespace sbdx
