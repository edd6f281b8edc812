// C ABI of libsbd_tpb: context, input handling, Davidson driver, diag.
//
// Host-side restatements (reference paths under include/sbd/):
//   LoadFCIDump     framework/fcidump.h:29-95
//   SetupIntegrals  chemistry/basic/makeintegrals.h:124-174 (RHF)
//   from_string     framework/bit_manipulation.h:219-235 (via sbd_pack_strings)
//   Davidson        chemistry/tpb/davidson.h:707-970 (single root)
//   diag            chemistry/tpb/sbdiag.h:151-526 (method 0, one rank)
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "sbd_internal.hpp"
#include "davidson_job.hpp"

namespace sbdx {
int sigma_set_lds_limit(sbd_ctx* ctx);

// Symmetric eigenproblem of order n (column-major, leading dimension lda):
// cyclic Jacobi rotations, then eigenvalues sorted ascending with their
// eigenvectors in the columns of A.  Replaces LAPACK dsyev
// (framework/hp_numeric.h:503-519) for the Davidson subspace problem.
void syev_jacobi(int n, double* A, int lda, double* w) {
  std::vector<double> V((size_t)n * n, 0.0);
  for (int i = 0; i < n; i++) V[i + (size_t)i * n] = 1.0;
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0.0, tot = 0.0;
    for (int j = 0; j < n; j++)
      for (int i = 0; i < n; i++) {
        const double x = A[i + (size_t)j * lda] * A[i + (size_t)j * lda];
        tot += x;
        if (i != j) off += x;
      }
    if (off <= 1e-30 * (tot > 0 ? tot : 1.0)) break;
    for (int p = 0; p < n - 1; p++)
      for (int q = p + 1; q < n; q++) {
        const double apq = A[p + (size_t)q * lda];
        if (std::fabs(apq) < 1e-300) continue;
        const double app = A[p + (size_t)p * lda], aqq = A[q + (size_t)q * lda];
        const double theta = (aqq - app) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; k++) {
          const double akp = A[k + (size_t)p * lda], akq = A[k + (size_t)q * lda];
          A[k + (size_t)p * lda] = c * akp - s * akq;
          A[k + (size_t)q * lda] = s * akp + c * akq;
        }
        for (int k = 0; k < n; k++) {
          const double apk = A[p + (size_t)k * lda], aqk = A[q + (size_t)k * lda];
          A[p + (size_t)k * lda] = c * apk - s * aqk;
          A[q + (size_t)k * lda] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; k++) {
          const double vkp = V[k + (size_t)p * n], vkq = V[k + (size_t)q * n];
          V[k + (size_t)p * n] = c * vkp - s * vkq;
          V[k + (size_t)q * n] = s * vkp + c * vkq;
        }
      }
  }
  std::vector<int> idx(n);
  for (int i = 0; i < n; i++) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) {
    return A[x + (size_t)x * lda] < A[y + (size_t)y * lda];
  });
  for (int i = 0; i < n; i++) w[i] = A[idx[i] + (size_t)idx[i] * lda];
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) A[i + (size_t)j * lda] = V[i + (size_t)idx[j] * n];
}

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int set_err(sbd_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->last_error = msg;
  return code;
}

// Randomize, framework/dm_vector.h:101-191 at one rank and one thread (the
// reference's per-thread compensated sums added in thread order reduce to one
// compensated sum in index order).
void randomize(double* W, uint64_t N, uint64_t seed) {
  if (seed == 0) {
    const uint64_t ns = (uint64_t)std::chrono::high_resolution_clock::now()
                            .time_since_epoch()
                            .count();
    seed = elem_hash(ns, 0);  // sbd_time_seed at rank 0
  }
  double sum = 0.0, eps = 0.0;
  for (uint64_t i = 0; i < N; i++) {
    W[i] = uniform_pm1(elem_hash(seed, i));
    const double val = W[i] * W[i] - eps;
    const double tmp = sum + val;
    eps = (tmp - sum) - val;
    sum = tmp;
  }
  const double factor = 1.0 / std::sqrt(sum);
  for (uint64_t i = 0; i < N; i++) W[i] *= factor;
}

}  // namespace sbdx

using namespace sbdx;

namespace {

// SetupIntegrals (makeintegrals.h:124-174): entries in order, later overwrite.
void setup_integrals(int L, size_t n, const double* val, const int* idx, double& I0,
                     std::vector<double>& I1, std::vector<double>& eri) {
  const int L2 = 2 * L;
  const int64_t B = (int64_t)L * (L + 1) / 2;
  I1.assign((size_t)L2 * L2, 0.0);
  eri.assign((size_t)(B * (B + 1) / 2), 0.0);
  I0 = 0.0;
  for (size_t e = 0; e < n; e++) {
    const int i = idx[4 * e], j = idx[4 * e + 1], k = idx[4 * e + 2], l = idx[4 * e + 3];
    const double v = val[e];
    if (i == 0 && k == 0 && j == 0 && l == 0) {
      I0 = v;
    } else if (k == l && k == 0) {
      I1[(size_t)(2 * (i - 1)) * L2 + 2 * (j - 1)] = v;
      I1[(size_t)(2 * (i - 1) + 1) * L2 + 2 * (j - 1) + 1] = v;
      I1[(size_t)(2 * (j - 1)) * L2 + 2 * (i - 1)] = v;
      I1[(size_t)(2 * (j - 1) + 1) * L2 + 2 * (i - 1) + 1] = v;
    } else {
      eri[eri_index(i - 1, j - 1, k - 1, l - 1)] = v;
    }
  }
}

bool check_entries(int L, size_t n, const int* idx, std::string& msg) {
  for (size_t e = 0; e < n; e++) {
    for (int c = 0; c < 4; c++) {
      const int x = idx[4 * e + c];
      if (x < 0 || x > L) {
        msg = "FCIDUMP orbital index out of range at entry " + std::to_string(e);
        return false;
      }
    }
    const int i = idx[4 * e], j = idx[4 * e + 1], k = idx[4 * e + 2], l = idx[4 * e + 3];
    const bool core = (i == 0 && j == 0 && k == 0 && l == 0);
    const bool one = (k == 0 && l == 0 && i > 0 && j > 0);
    const bool two = (i > 0 && j > 0 && k > 0 && l > 0);
    if (!core && !one && !two) {
      msg = "unsupported FCIDUMP entry (orbital energies / partial zeros) at entry " +
            std::to_string(e);
      return false;
    }
  }
  return true;
}

sbd_ctx* create_common(int norb, double I0, std::vector<double>&& I1, std::vector<double>&& eri,
                       const uint64_t* a, size_t na, const uint64_t* b, size_t nb, int* status) {
  auto fail = [&](int code) -> sbd_ctx* {
    if (status) *status = code;
    return nullptr;
  };
  if (norb < 1 || norb > 64) return fail(SBD_ERR_ARG);
  if (!a || !b || na == 0 || nb == 0) return fail(SBD_ERR_ARG);
  if (na > 0xFFFFFFF0ULL || nb > 0xFFFFFFF0ULL) return fail(SBD_ERR_ARG);
  const uint64_t lim = norb == 64 ? ~0ULL : ((1ULL << norb) - 1ULL);
  for (int s = 0; s < 2; s++) {
    const uint64_t* v = s ? b : a;
    const size_t n = s ? nb : na;
    for (size_t i = 0; i < n; i++) {
      if (v[i] & ~lim) return fail(SBD_ERR_ARG);
      if (i > 0 && v[i] <= v[i - 1]) return fail(SBD_ERR_ARG);  // sorted, unique
    }
  }
  sbd_ctx* ctx = new sbd_ctx();
  ctx->norb = norb;
  ctx->na = na;
  ctx->nb = nb;
  ctx->N = (uint64_t)na * nb;
  ctx->I0 = I0;
  ctx->a_host.assign(a, a + na);
  ctx->b_host.assign(b, b + nb);
  ctx->eri = std::move(eri);
  const int L = norb, L2 = 2 * norb;
  ctx->h1.assign((size_t)L * L, 0.0);
  for (int p = 0; p < L; p++)
    for (int q = 0; q < L; q++) ctx->h1[(size_t)p * L + q] = I1[(size_t)(2 * p) * L2 + 2 * q];
  std::vector<double> J((size_t)L * L), K((size_t)L * L);
  for (int p = 0; p < L; p++)
    for (int q = 0; q < L; q++) {
      J[(size_t)p * L + q] = ctx->eri[eri_index(p, p, q, q)];
      K[(size_t)p * L + q] = ctx->eri[eri_index(p, q, q, p)];
    }
  auto bad = [&](hipError_t e) {
    if (e != hipSuccess) {
      if (status) *status = e == hipErrorOutOfMemory ? SBD_ERR_MEMORY : SBD_ERR_DEVICE;
      sbd_ctx_destroy(ctx);
      return true;
    }
    return false;
  };
  if (bad(hipGetDevice(&ctx->device))) return nullptr;
  hipDeviceProp_t prop;
  if (bad(hipGetDeviceProperties(&prop, ctx->device))) return nullptr;
  ctx->cu_count = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
  // split-sigma slab width: 1024 columns for wide bases (FP32 at 20,000 beta
  // strings: 1.94 against 2.25 ms for the streaming part), 512 below
  ctx->stream_cols = nb >= 8192 ? 1024 : 512;
  if (bad(hipMalloc(&ctx->d_a, 8 * na))) return nullptr;
  if (bad(hipMalloc(&ctx->d_b, 8 * nb))) return nullptr;
  if (bad(hipMalloc(&ctx->d_I1, 8 * I1.size()))) return nullptr;
  if (bad(hipMalloc(&ctx->d_eri, 8 * (ctx->eri.size() + 1)))) return nullptr;
  if (bad(hipMalloc(&ctx->d_J, 8 * J.size()))) return nullptr;
  if (bad(hipMalloc(&ctx->d_K, 8 * K.size()))) return nullptr;
  if (bad(hipMemcpy(ctx->d_a, a, 8 * na, hipMemcpyHostToDevice))) return nullptr;
  if (bad(hipMemcpy(ctx->d_b, b, 8 * nb, hipMemcpyHostToDevice))) return nullptr;
  if (bad(hipMemcpy(ctx->d_I1, I1.data(), 8 * I1.size(), hipMemcpyHostToDevice))) return nullptr;
  if (bad(hipMemcpy(ctx->d_eri, ctx->eri.data(), 8 * ctx->eri.size(), hipMemcpyHostToDevice)))
    return nullptr;
  if (bad(hipMemcpy(ctx->d_J, J.data(), 8 * J.size(), hipMemcpyHostToDevice))) return nullptr;
  if (bad(hipMemcpy(ctx->d_K, K.data(), 8 * K.size(), hipMemcpyHostToDevice))) return nullptr;
  int rc = build_gtable(ctx, nullptr);
  if (!rc) rc = sigma_set_lds_limit(ctx);
  if (!rc && hipStreamSynchronize(nullptr) != hipSuccess) rc = SBD_ERR_DEVICE;
  if (rc) {
    if (status) *status = rc;
    sbd_ctx_destroy(ctx);
    return nullptr;
  }
  if (status) *status = SBD_OK;
  return ctx;
}

template <typename VT>
struct DevBuf {
  VT* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

// Largest subspace size up to nb whose Davidson vectors fit the free device
// memory: 2 nb vectors besides W and the diagonal (C[1..nb-1], HC[0..nb-1] and
// the spare; C[0] is W itself), with 2% of the memory left for the rest.
static int fit_subspace(int nb, uint64_t vec_bytes) {
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) != hipSuccess || vec_bytes == 0) return nb;
  const double usable = (double)fr - 0.02 * (double)tot;
  const int fit = (int)std::floor(usable / (double)vec_bytes / 2.0);
  return std::max(1, std::min(nb, fit));
}

// Single-root Davidson on device vectors (davidson.h:707-970), driven pass
// by pass through DavidsonJob (davidson_job.hpp) with the fused vector passes
// of vecops.hip.  W: device, in/out (normalised Ritz vector on exit); it is
// the first basis vector itself, so a solve holds 2 nb + 2 vectors with the
// diagonal.  The subspace is cut below nb when those would not fit the free
// device memory (rep->nb_used tells the size taken).
template <typename VT>
int davidson_core(sbd_ctx* ctx, const VT* dii, VT* W, int max_it, int nb, double eps,
                  double max_time, int verbose, double* energy, std::vector<double>* hist,
                  sbd_report* rep, double t_start) {
  const uint64_t N = ctx->N;
  hipStream_t st = ctx->stream;
  if (nb < 1 || nb > kMaxVecs) return set_err(ctx, SBD_ERR_ARG, "max_nb must be in [1, 40]");
  if (max_it < 1) return set_err(ctx, SBD_ERR_ARG, "max_it must be >= 1");
  nb = fit_subspace(nb, sizeof(VT) * N);
  // C[0] is W; the other vectors are allocated when a pass first needs them
  // (a solve that converges in a few steps holds a few vectors) and freed on
  // every return.  Restarts only swap the roles of C[0] and the spare, so the
  // allocated buffers (all but W's) stay owned by Cb, HCb and spare.
  std::vector<DevBuf<VT>> Cb(nb), HCb(nb);
  DevBuf<VT> spare;
  std::vector<VT*> C(nb, nullptr), HC(nb, nullptr);
  C[0] = W;
  if (rep) {
    if (sizeof(VT) == 4) rep->nb_used_fp32 = nb;
    else rep->nb_used = nb;
  }
  DavidsonJob<VT> job;
  job.init(C, HC, nullptr, dii, N, max_it, eps);
  // the buffer behind a pointer the next pass writes, allocated on first use
  auto need = [&](VT*& slot, DevBuf<VT>& owner) -> int {
    if (slot) return SBD_OK;
    if (hipMalloc(&owner.p, sizeof(VT) * N) != hipSuccess)
      return set_err(ctx, SBD_ERR_MEMORY, "device allocation of Davidson vectors failed");
    slot = owner.p;
    return SBD_OK;
  };
  auto need_t = [&]() -> int {  // the Ritz pass writes C[ib + 1] or, last, the spare
    const int i = job.ib + 1;
    return i < nb ? need(job.C[i], Cb[i]) : need(job.spare, spare);
  };
  int rc = 0;
  double t_sigma = 0.0, step_t0 = now_s();
  std::vector<double> step_times;
  hipEvent_t ev0, ev1;
  SBD_HIP_TRY(ctx, hipEventCreate(&ev0));
  SBD_HIP_TRY(ctx, hipEventCreate(&ev1));
  auto ptrs = [](const std::vector<VT*>& v, int k) {
    return std::vector<const VT*>(v.begin(), v.begin() + k);
  };
  while (!rc && job.next() != DavidsonJob<VT>::kDone) {
    switch (job.next()) {
      case DavidsonJob<VT>::kSigma: {
        step_t0 = now_s();
        double td = 0.0;
        rc = need(job.HC[job.ib], HCb[job.ib]);
        if (rc) break;
        SBD_HIP_TRY(ctx, hipEventRecord(ev0, st));
        rc = launch_sigma<VT>(ctx, dii, job.sigma_in(), job.sigma_out(), 0, st, &td);
        if (rc) break;
        SBD_HIP_TRY(ctx, hipEventRecord(ev1, st));
        SBD_HIP_TRY(ctx, hipEventSynchronize(ev1));
        float ms = 0.f;
        SBD_HIP_TRY(ctx, hipEventElapsedTime(&ms, ev0, ev1));
        t_sigma += 1e-3 * ms;
        job.sigma_done(td);
        break;
      }
      case DavidsonJob<VT>::kRitz: {
        const int k = job.k();
        rc = need_t();
        if (rc) break;
        const auto Cc = ptrs(job.C, k), HCc = ptrs(job.HC, k);
        rc = vec_ritz_precond<VT>(ctx, Cc.data(), HCc.data(), k, job.ritz_u(), job.E0, dii, job.t(),
                                  N, job.ritz_sums(), st, job.ritz_restart());
        if (rc) break;
        const int ib = job.ib, it = job.it;
        job.ritz_done();
        if (hist) {
          hist->push_back(job.E0);
          hist->push_back(job.normR);
          hist->push_back(now_s() - t_start);
          hist->push_back(sizeof(VT) == 4 ? 32.0 : 64.0);
        }
        if (verbose) {
          printf(" Davidson iteration %d.%d (tol=%g):", it, ib, job.normR);
          for (int pp = 0; pp < std::min(ib + 1, 4); pp++) printf(" %.16g", job.Evals[pp]);
          printf("\n");
          fflush(stdout);
        }
        break;
      }
      case DavidsonJob<VT>::kGs: {
        const int k = job.k();
        const auto Cc = ptrs(job.C, k);
        double n2 = 0.0;
        rc = vec_gs_update<VT>(ctx, Cc.data(), k, job.gs_coef(), job.t(), job.gs_scale(), N, &n2,
                               st);
        if (!rc) job.gs_done(n2);
        break;
      }
      case DavidsonJob<VT>::kDots: {
        const int k = job.k();
        const auto Cc = ptrs(job.C, k);
        rc = vec_multidot<VT>(ctx, Cc.data(), k, job.t(), N, job.dots_out(), st);
        if (!rc) job.dots_done();
        break;
      }
      case DavidsonJob<VT>::kLincomb: {
        const int k = job.lincomb_k();
        rc = need(job.spare, spare);
        if (rc) break;
        const auto Cc = ptrs(job.C, k);
        double n2 = 0.0;
        rc = vec_lincomb<VT>(ctx, Cc.data(), k, job.ritz_u(), 1.0, job.lincomb_out(), N, &n2, st);
        if (!rc) job.lincomb_done(n2);
        break;
      }
      case DavidsonJob<VT>::kScale:
        rc = vec_scale<VT>(ctx, job.lincomb_out(), job.scale_factor(), N, st);
        if (!rc) job.scale_done();
        break;
      default:
        break;
    }
    if (rc) break;
    // max_time guard after each step (davidson.h:926-949): stop if the next
    // step would end late
    const auto nx = job.next();
    if (nx == DavidsonJob<VT>::kSigma || nx == DavidsonJob<VT>::kLincomb) {
      step_times.push_back(now_s() - step_t0);
      step_t0 = now_s();
      double avg = 0.0;
      for (double x : step_times) avg += x;
      avg /= step_times.size();
      if ((now_s() - t_start) + avg > max_time) {
        if (verbose) fprintf(stderr, "halting iterations due to max_time limit\n");
        job.stop();
      }
    }
  }
  if (!rc) {
    // W = normalised Ritz vector of the last step, formed in the spare (W's
    // buffer may be one of the basis vectors it is formed from) and copied
    // into W's buffer when they differ
    double n2 = 0.0;
    const auto Cc = ptrs(job.C, job.last_k);
    rc = need(job.spare, spare);
    VT* out = job.spare;
    if (!rc) rc = vec_lincomb<VT>(ctx, Cc.data(), job.last_k, job.ritz_u(), 1.0, out, N, &n2, st);
    if (!rc) rc = vec_scale<VT>(ctx, out, 1.0 / std::sqrt(n2), N, st);
    if (!rc && out != W && hipMemcpyAsync(W, out, sizeof(VT) * N, hipMemcpyDeviceToDevice, st) !=
                               hipSuccess)
      rc = SBD_ERR_DEVICE;
    if (!rc && hipStreamSynchronize(st) != hipSuccess) rc = SBD_ERR_DEVICE;
  }
  (void)hipEventDestroy(ev0);
  (void)hipEventDestroy(ev1);
  if (energy) *energy = job.E0;
  if (rep) {
    rep->steps += job.steps;
    rep->converged = job.converged ? 1 : 0;
    rep->residual = job.normR;
    rep->energy = job.E0;
    rep->t_sigma += t_sigma;
  }
  return rc;
}

}  // namespace

namespace sbdx {
// W[0, N) = the start vector (BasisInitVector, davidson.h:144-184) in the
// element type of W, built on the device: the caller's (host, FP64; a loaded
// wavefunction, sbdiag.h:339-347), else e_0 (init 0) or the seeded Randomize
// vector (init 1; seed 0 draws a seed from the clock, as the reference does).
template <typename VT>
int start_vector(sbd_ctx* ctx, const sbd_options* opt, const double* W_init, VT* W,
                 int64_t unit) {
  const uint64_t N = ctx->N;
  hipStream_t st = ctx->stream;
  int rc = SBD_OK;
  if (W_init) {
    if (sizeof(VT) == 8) {
      SBD_HIP_TRY(ctx, hipMemcpyAsync(W, W_init, 8 * N, hipMemcpyHostToDevice, st));
    } else {
      // through an FP64 staging chunk, converted in place
      const uint64_t chunk = std::min<uint64_t>(N, 1ULL << 25);
      DevBuf<double> stage;
      if (hipMalloc(&stage.p, 8 * chunk) != hipSuccess)
        return set_err(ctx, SBD_ERR_MEMORY, "device allocation failed");
      for (uint64_t i0 = 0; i0 < N && !rc; i0 += chunk) {
        const uint64_t n = std::min(chunk, N - i0);
        SBD_HIP_TRY(ctx, hipMemcpyAsync(stage.p, W_init + i0, 8 * n, hipMemcpyHostToDevice, st));
        rc = vec_convert<double, VT>(ctx, stage.p, W + i0, n, st);
        if (!rc && hipStreamSynchronize(st) != hipSuccess) rc = SBD_ERR_DEVICE;
      }
    }
  } else if (unit >= 0 || opt->init == 0) {
    // e_unit (the explicit initial determinant, davidson.h:89-127) or e_0
    // (init 0, davidson.h:156-166)
    rc = vec_fill<VT>(ctx, W, N, 0.0, st);
    const VT one = (VT)1;
    const uint64_t k = unit >= 0 ? (uint64_t)unit : 0;
    if (!rc) SBD_HIP_TRY(ctx, hipMemcpyAsync(W + k, &one, sizeof(VT), hipMemcpyHostToDevice, st));
  } else {
    uint64_t seed = opt->seed;
    if (seed == 0) {
      const uint64_t ns = (uint64_t)std::chrono::high_resolution_clock::now()
                              .time_since_epoch()
                              .count();
      seed = elem_hash(ns, 0);  // sbd_time_seed at rank 0
    }
    rc = vec_random<VT>(ctx, W, N, seed, st);
  }
  if (!rc && hipStreamSynchronize(st) != hipSuccess) rc = SBD_ERR_DEVICE;
  return rc;
}
template int start_vector<double>(sbd_ctx*, const sbd_options*, const double*, double*, int64_t);
template int start_vector<float>(sbd_ctx*, const sbd_options*, const double*, float*, int64_t);
}  // namespace sbdx

extern "C" {

const char* sbd_status_string(int s) {
  switch (s) {
    case SBD_OK: return "ok";
    case SBD_ERR_ARG: return "invalid argument";
    case SBD_ERR_IO: return "i/o error";
    case SBD_ERR_DEVICE: return "device error";
    case SBD_ERR_MEMORY: return "out of device memory";
    case SBD_ERR_STATE: return "call out of order";
    case SBD_ERR_NOT_CONVERGED: return "not converged";
    default: return "unknown status";
  }
}

const char* sbd_last_error(const sbd_ctx* ctx) { return ctx ? ctx->last_error.c_str() : ""; }

void sbd_options_default(sbd_options* o) {
  if (!o) return;
  o->method = 0;
  o->max_it = 1;
  o->max_nb = 10;
  o->eps = 1.0e-4;
  o->max_time = 86400.0;
  o->init = 0;
  o->seed = 1729;
  o->precision = 0;
  o->fp32_eps = 1.0e-2;
  o->verbose = 0;
}

int sbd_pack_strings(const size_t* words, size_t n, size_t bit_length, int norb, uint64_t* out) {
  if (!words || !out || bit_length == 0 || bit_length > 64 || norb < 1 || norb > 64)
    return SBD_ERR_ARG;
  const size_t wps = (norb + bit_length - 1) / bit_length;
  for (size_t s = 0; s < n; s++) {
    uint64_t v = 0;
    for (int p = 0; p < norb; p++) {
      const size_t w = words[s * wps + p / bit_length];
      if ((w >> (p % bit_length)) & 1ULL) v |= 1ULL << p;
    }
    out[s] = v;
  }
  return SBD_OK;
}

namespace {
int fcidump_parse(const char* path, int* norb, int* nelec, std::vector<double>* vals,
                  std::vector<int>* idx, size_t* count) {
  std::ifstream in(path);
  if (!in.is_open()) return SBD_ERR_IO;
  std::string line, header;
  bool in_header = true;
  size_t cnt = 0;
  while (std::getline(in, line)) {
    if (line.empty() || line[0] == '!') continue;
    if (in_header) {
      if (line.find("&END") != std::string::npos) {
        in_header = false;
        continue;
      }
      const size_t pos = line.find("&FCI");
      if (pos != std::string::npos) line.erase(pos, 4);
      header += line;
    } else {
      std::stringstream ss(line);
      double v;
      int i, j, k, l;
      if (!(ss >> v >> i >> j >> k >> l)) continue;
      if (vals) {
        vals->push_back(v);
        idx->push_back(i);
        idx->push_back(j);
        idx->push_back(k);
        idx->push_back(l);
      }
      cnt++;
    }
  }
  std::map<std::string, std::string> kv;
  std::stringstream hs(header);
  std::string tok;
  while (std::getline(hs, tok, ',')) {
    const size_t eq = tok.find('=');
    if (eq == std::string::npos) continue;
    std::string k = tok.substr(0, eq), v = tok.substr(eq + 1);
    auto trim = [](std::string& s) {
      s.erase(0, s.find_first_not_of(" \t"));
      s.erase(s.find_last_not_of(" \t") + 1);
    };
    trim(k);
    trim(v);
    kv[k] = v;
  }
  if (norb) *norb = kv.count("NORB") ? std::atoi(kv["NORB"].c_str()) : 0;
  if (nelec) *nelec = kv.count("NELEC") ? std::atoi(kv["NELEC"].c_str()) : 0;
  if (count) *count = cnt;
  return SBD_OK;
}
}  // namespace

int sbd_fcidump_read(const char* path, int* norb, int* nelec, size_t* n_entries) {
  if (!path) return SBD_ERR_ARG;
  return fcidump_parse(path, norb, nelec, nullptr, nullptr, n_entries);
}

int sbd_fcidump_entries(const char* path, double* values, int* idx, size_t n_entries) {
  if (!path || !values || !idx) return SBD_ERR_ARG;
  std::vector<double> v;
  std::vector<int> ix;
  size_t cnt = 0;
  int rc = fcidump_parse(path, nullptr, nullptr, &v, &ix, &cnt);
  if (rc) return rc;
  if (cnt != n_entries) return SBD_ERR_ARG;
  std::memcpy(values, v.data(), sizeof(double) * cnt);
  std::memcpy(idx, ix.data(), sizeof(int) * 4 * cnt);
  return SBD_OK;
}

sbd_ctx* sbd_ctx_create_fcidump(int norb, size_t n_entries, const double* values, const int* idx,
                                const uint64_t* a, size_t na, const uint64_t* b, size_t nb,
                                int* status) {
  if (norb < 1 || norb > 64 || (n_entries && (!values || !idx))) {
    if (status) *status = SBD_ERR_ARG;
    return nullptr;
  }
  std::string msg;
  if (!check_entries(norb, n_entries, idx, msg)) {
    if (status) *status = SBD_ERR_ARG;
    return nullptr;
  }
  double I0;
  std::vector<double> I1, eri;
  setup_integrals(norb, n_entries, values, idx, I0, I1, eri);
  return create_common(norb, I0, std::move(I1), std::move(eri), a, na, b, nb, status);
}

sbd_ctx* sbd_ctx_create(int norb, double I0, const double* h1, const double* eri_packed,
                        const uint64_t* a, size_t na, const uint64_t* b, size_t nb, int* status) {
  if (norb < 1 || norb > 64 || !h1 || !eri_packed) {
    if (status) *status = SBD_ERR_ARG;
    return nullptr;
  }
  const int L = norb, L2 = 2 * norb;
  const int64_t B = (int64_t)L * (L + 1) / 2;
  std::vector<double> I1((size_t)L2 * L2, 0.0), eri(eri_packed, eri_packed + B * (B + 1) / 2);
  for (int p = 0; p < L; p++)
    for (int q = 0; q < L; q++) {
      I1[(size_t)(2 * p) * L2 + 2 * q] = h1[(size_t)p * L + q];
      I1[(size_t)(2 * p + 1) * L2 + 2 * q + 1] = h1[(size_t)p * L + q];
    }
  return create_common(norb, I0, std::move(I1), std::move(eri), a, na, b, nb, status);
}

void sbd_ctx_destroy(sbd_ctx* ctx) {
  if (!ctx) return;
  void* ptrs[] = {ctx->d_a,       ctx->d_b,       ctx->d_I1,      ctx->d_eri,     ctx->d_J,
                  ctx->d_K,       ctx->d_G,       ctx->d_Vab,     ctx->ta.s_off,  ctx->ta.d_off,  ctx->ta.s_rec,
                  ctx->ta.d_rec,  ctx->tb.s_off,  ctx->tb.d_off,  ctx->tb.s_rec,  ctx->tb.d_rec,
                  ctx->d_partials, ctx->d_results, ctx->ta.with_s, ctx->tb.with_s,
                  ctx->ta.ell64, ctx->ta.ell32, ctx->ta.ell_off, ctx->ta.ell_ket,
                  ctx->tb.ell64, ctx->tb.ell32, ctx->tb.ell_off, ctx->tb.ell_ket, ctx->d_sdot,
                  ctx->ta.ell_pos, ctx->tb.ell_pos, ctx->ta.col_rec, ctx->tb.col_rec,
                  ctx->ta.s_bra,   ctx->tb.s_bra,   ctx->d_Ea,       ctx->d_Eb,
                  ctx->d_ab_sell};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (ctx->h_results) (void)hipHostFree(ctx->h_results);
  delete ctx;
}

int sbd_set_sigma_kernel(sbd_ctx* ctx, int kernel) {
  if (!ctx || kernel < 0 || kernel > 4) return SBD_ERR_ARG;
  ctx->sigma_kernel = kernel;
  return SBD_OK;
}

int sbd_set_sigma_slab(sbd_ctx* ctx, int cols) {
  if (!ctx || (cols != 256 && cols != 512 && cols != 1024)) return SBD_ERR_ARG;
  ctx->stream_cols = cols;
  return SBD_OK;
}

int sbd_ctx_dims(const sbd_ctx* ctx, uint64_t* out) {
  if (!ctx || !out) return SBD_ERR_ARG;
  out[0] = ctx->na;
  out[1] = ctx->nb;
  out[2] = ctx->N;
  out[3] = (uint64_t)ctx->norb;
  return SBD_OK;
}

int sbd_make_helpers(sbd_ctx* ctx) {
  if (!ctx) return SBD_ERR_ARG;
  if (ctx->tables_built) return SBD_OK;
  int rc = build_tables(ctx, nullptr);
  if (!rc) rc = build_single_cache(ctx, nullptr);
  if (!rc) rc = build_abtable(ctx, nullptr);
  if (!rc && hipStreamSynchronize(nullptr) != hipSuccess) rc = SBD_ERR_DEVICE;
  if (rc) return rc;
  ctx->tables_built = true;
  return SBD_OK;
}

int sbd_table_counts(const sbd_ctx* ctx, uint64_t* out) {
  if (!ctx || !out) return SBD_ERR_ARG;
  if (!ctx->tables_built) return SBD_ERR_STATE;
  out[0] = ctx->ta.n_s;
  out[1] = ctx->ta.n_d;
  out[2] = ctx->tb.n_s;
  out[3] = ctx->tb.n_d;
  return SBD_OK;
}

int sbd_table_get(const sbd_ctx* ctx, int which, int64_t* off, int64_t* ket, int32_t* cran) {
  if (!ctx || which < 0 || which > 3 || !off) return SBD_ERR_ARG;
  if (!ctx->tables_built) return SBD_ERR_STATE;
  const SpinTables& T = which < 2 ? ctx->ta : ctx->tb;
  const bool dbl = which % 2 == 1;
  const int spin = which < 2 ? 0 : 1;
  const uint64_t n = which < 2 ? ctx->na : ctx->nb;
  const uint64_t cnt = dbl ? T.n_d : T.n_s;
  std::vector<uint32_t> o(n + 1);
  std::vector<PairRec> r(cnt);
  if (hipMemcpy(o.data(), dbl ? T.d_off : T.s_off, 4 * (n + 1), hipMemcpyDeviceToHost) !=
      hipSuccess)
    return SBD_ERR_DEVICE;
  if (cnt && hipMemcpy(r.data(), dbl ? T.d_rec : T.s_rec, sizeof(PairRec) * cnt,
                       hipMemcpyDeviceToHost) != hipSuccess)
    return SBD_ERR_DEVICE;
  for (uint64_t i = 0; i <= n; i++) off[i] = o[i];
  const std::vector<uint64_t>& S = which < 2 ? ctx->a_host : ctx->b_host;
  for (uint64_t i = 0; i < n; i++) {
    for (uint64_t q = o[i]; q < o[i + 1]; q++) {
      if (ket) ket[q] = r[q].ket;
      if (!cran) continue;
      if (dbl) {
        // OrbitalDifference (determinants.h:406-449): bra-only, then ket-only, ascending
        uint64_t x = S[i] & ~S[r[q].ket], y = S[r[q].ket] & ~S[i];
        const int x0 = __builtin_ctzll(x), x1 = __builtin_ctzll(x & (x - 1));
        const int y0 = __builtin_ctzll(y), y1 = __builtin_ctzll(y & (y - 1));
        cran[4 * q + 0] = 2 * x0 + spin;
        cran[4 * q + 1] = 2 * x1 + spin;
        cran[4 * q + 2] = 2 * y0 + spin;
        cran[4 * q + 3] = 2 * y1 + spin;
      } else {
        cran[2 * q + 0] = 2 * code_x0(r[q].code) + spin;
        cran[2 * q + 1] = 2 * code_y0(r[q].code) + spin;
      }
    }
  }
  return SBD_OK;
}

int sbd_hdiag(sbd_ctx* ctx, double* hii, void* stream) {
  if (!ctx || !hii) return SBD_ERR_ARG;
  return launch_hdiag(ctx, hii, (hipStream_t)stream);
}

int sbd_hdiag_f32(sbd_ctx* ctx, float* hii, void* stream) {
  if (!ctx || !hii) return SBD_ERR_ARG;
  return launch_hdiag_f32(ctx, hii, (hipStream_t)stream);
}

int sbd_sigma(sbd_ctx* ctx, const double* hii, const double* Wk, double* Wb, int accumulate,
              void* stream) {
  if (!ctx || !hii || !Wk || !Wb) return SBD_ERR_ARG;
  return launch_sigma<double>(ctx, hii, Wk, Wb, accumulate, (hipStream_t)stream);
}

int sbd_sigma_f32(sbd_ctx* ctx, const float* hii, const float* Wk, float* Wb, int accumulate,
                  void* stream) {
  if (!ctx || !hii || !Wk || !Wb) return SBD_ERR_ARG;
  return launch_sigma<float>(ctx, hii, Wk, Wb, accumulate, (hipStream_t)stream);
}

int sbd_inner(sbd_ctx* ctx, const double* x, const double* y, uint64_t n, double* out,
              void* stream) {
  if (!ctx || !x || !y || !out) return SBD_ERR_ARG;
  const double* X[1] = {x};
  return vec_multidot<double>(ctx, X, 1, y, n, out, (hipStream_t)stream);
}

int sbd_davidson(sbd_ctx* ctx, const double* hii, double* W, const sbd_options* opt,
                 double* energy, double* hist, sbd_report* rep) {
  if (!ctx || !hii || !W || !opt) return SBD_ERR_ARG;
  if (!ctx->tables_built) return set_err(ctx, SBD_ERR_STATE, "davidson before sbd_make_helpers");
  std::vector<double> h;
  if (rep) std::memset(rep, 0, sizeof(sbd_report));
  const double t0 = now_s();
  int rc = davidson_core<double>(ctx, hii, W, opt->max_it, opt->max_nb, opt->eps, opt->max_time,
                                 opt->verbose, energy, hist ? &h : nullptr, rep, t0);
  if (hist) {
    const size_t cap = (size_t)4 * opt->max_it * opt->max_nb;
    std::memcpy(hist, h.data(), sizeof(double) * std::min(cap, h.size()));
  }
  if (rep) rep->t_davidson += now_s() - t0;
  return rc;
}

int sbd_lanczos(sbd_ctx* ctx, const double* hii, double* W, const sbd_options* opt,
                double* energy, double* hist, sbd_report* rep) {
  if (!ctx || !hii || !W || !opt) return SBD_ERR_ARG;
  if (!ctx->tables_built) return set_err(ctx, SBD_ERR_STATE, "lanczos before sbd_make_helpers");
  std::vector<double> h;
  if (rep) std::memset(rep, 0, sizeof(sbd_report));
  const double t0 = now_s();
  int rc = lanczos_core(ctx, hii, W, opt->max_it, opt->max_nb, opt->eps, opt->verbose, energy,
                        hist ? &h : nullptr, rep, t0);
  if (hist) {
    const size_t cap = (size_t)4 * opt->max_it * opt->max_nb;
    std::memcpy(hist, h.data(), sizeof(double) * std::min(cap, h.size()));
  }
  if (rep) rep->t_davidson += now_s() - t0;
  return rc;
}

int sbd_diag_footprint(const sbd_ctx* ctx, const sbd_options* opt, uint64_t* bytes) {
  if (!ctx || !opt || !bytes) return SBD_ERR_ARG;
  const uint64_t N = ctx->N;
  const uint64_t nb = (uint64_t)opt->max_nb;
  uint64_t v = (2 * nb + 3) * N * 8;  // C, HC, spare, W, hii
  if (opt->precision == 1) v = (2 * nb + 3) * N * 4 + (2 * nb + 3) * N * 8;
  const uint64_t L = (uint64_t)ctx->norb;
  const uint64_t ints = 8 * (4 * L * L + ctx->eri.size() + 2 * L * L + 4 * L * L * L);
  const uint64_t strs = 8 * (ctx->na + ctx->nb);
  // tables: estimated from counts when built, else a pessimistic bound of
  // C(n,1)C(L-n,1) + C(n,2)C(L-n,2) records per string
  uint64_t tab = 0;
  if (ctx->tables_built) {
    tab = 16 * (ctx->ta.n_s + ctx->ta.n_d + ctx->tb.n_s + ctx->tb.n_d) + 8 * (ctx->na + ctx->nb + 2);
  } else {
    auto recs = [&](const std::vector<uint64_t>& s) {
      const uint64_t n = s.empty() ? 0 : (uint64_t)__builtin_popcountll(s[0]);
      const uint64_t v1 = n * (L - n), v2 = n * (n - 1) / 2 * (L - n) * (L - n - 1) / 2;
      const uint64_t cap = s.size();
      return s.size() * std::min<uint64_t>(v1 + v2, cap);
    };
    tab = 16 * (recs(ctx->a_host) + recs(ctx->b_host)) + 8 * (ctx->na + ctx->nb + 2);
  }
  *bytes = v + ints + strs + tab;
  return SBD_OK;
}

int sbd_diag(sbd_ctx* ctx, const sbd_options* opt, double* energy, double* W_out, double* hist,
             int hist_cap, sbd_report* rep) {
  return sbd_diag_from(ctx, opt, nullptr, energy, W_out, hist, hist_cap, rep);
}

static int diag_impl(sbd_ctx* ctx, const sbd_options* opt, const double* W_init, int64_t unit,
                     double* energy, double* W_out, double* hist, int hist_cap, sbd_report* rep);

int sbd_diag_from(sbd_ctx* ctx, const sbd_options* opt, const double* W_init, double* energy,
                  double* W_out, double* hist, int hist_cap, sbd_report* rep) {
  return diag_impl(ctx, opt, W_init, -1, energy, W_out, hist, hist_cap, rep);
}

int sbd_diag_from_det(sbd_ctx* ctx, const sbd_options* opt, uint64_t a_string, uint64_t b_string,
                      double* energy, double* W_out, double* hist, int hist_cap,
                      sbd_report* rep) {
  if (!ctx || !opt) return SBD_ERR_ARG;
  // unique_index of BasisInitVectorFromDeterminants (davidson.h:97-110): the
  // strings of a context are sorted and unique, so a match is unique
  auto find = [](const std::vector<uint64_t>& v, uint64_t x) -> int64_t {
    auto it = std::lower_bound(v.begin(), v.end(), x);
    return (it != v.end() && *it == x) ? (int64_t)(it - v.begin()) : -1;
  };
  const int64_t ia = find(ctx->a_host, a_string), ib = find(ctx->b_host, b_string);
  if (ia < 0) return set_err(ctx, SBD_ERR_ARG, "initial alpha determinant is not present in the basis");
  if (ib < 0) return set_err(ctx, SBD_ERR_ARG, "initial beta determinant is not present in the basis");
  return diag_impl(ctx, opt, nullptr, ia * (int64_t)ctx->nb + ib, energy, W_out, hist, hist_cap,
                   rep);
}

static int diag_impl(sbd_ctx* ctx, const sbd_options* opt, const double* W_init, int64_t unit,
                     double* energy, double* W_out, double* hist, int hist_cap, sbd_report* rep) {
  if (!ctx || !opt) return SBD_ERR_ARG;
  std::vector<double> hv;
  std::vector<double>* H = hist ? &hv : nullptr;
  // methods 0/1 Davidson, 2/3 Lanczos (sbdiag.h:396-468, 527-600): 1 and 3 are
  // the reference's stored-Hamiltonian forms of 0 and 2, whose iterates are the
  // same; here H is never stored
  if (opt->method < 0 || opt->method > 3)
    return set_err(ctx, SBD_ERR_ARG, "method must be 0..3 (0/1 Davidson, 2/3 Lanczos)");
  const bool lanczos = opt->method >= 2;
  if (lanczos && opt->precision != 0)
    return set_err(ctx, SBD_ERR_ARG, "precision 1 applies to the Davidson methods (0/1)");
  if (opt->init != 0 && opt->init != 1) return set_err(ctx, SBD_ERR_ARG, "init must be 0 or 1");
  sbd_report local{};
  sbd_report* R = rep ? rep : &local;
  std::memset(R, 0, sizeof(sbd_report));
  const uint64_t N = ctx->N;
  hipStream_t st = ctx->stream;
  double t0 = now_s();
  int rc = sbd_make_helpers(ctx);
  if (rc) return rc;
  R->t_tables = now_s() - t0;
  // Device memory is held to what each phase needs: the FP32 phase holds its
  // vectors and FP32 diagonal (plus the FP64 diagonal when the full FP32
  // subspace still fits beside it), and frees them (but W) before the FP64
  // phase; the final <W, H W> takes its H W buffer once the subspace vectors
  // are gone.
  DevBuf<double> W, hii;
  double E = 0.0;
  t0 = now_s();
  const double t_solve0 = t0;
  if (opt->precision == 1) {
    // FP32 Davidson (FP32 vectors and sigma; FP64 reductions, Ritz and GS),
    // then the FP64 correction sweep from its vector.
    DevBuf<float> W32;
    {
      DevBuf<float> hii32;
      if (hipMalloc(&hii32.p, 4 * N) != hipSuccess || hipMalloc(&W32.p, 4 * N) != hipSuccess)
        return set_err(ctx, SBD_ERR_MEMORY, "device allocation failed");
      rc = start_vector<float>(ctx, opt, W_init, W32.p, unit);
      if (rc) return rc;
      const double th = now_s();
      // The FP32 diagonal is the FP64 one rounded once (k_hdiag forms it in
      // FP64 either way), so when HBM holds the FP64 diagonal beside the full
      // FP32 subspace it is made once here, rounded, and kept for the FP64
      // phase; otherwise each phase makes its own.
      size_t fr = 0, tot = 0;
      bool keep64 = hipMemGetInfo(&fr, &tot) == hipSuccess &&
                    (double)fr >= 8.0 * N + 2.0 * opt->max_nb * 4.0 * N + 0.02 * (double)tot;
      if (keep64 && hipMalloc(&hii.p, 8 * N) != hipSuccess) {
        (void)hipGetLastError();  // not an error: the FP32 diagonal is made on its own
        hii.p = nullptr;
        keep64 = false;
      }
      if (keep64) {
        rc = launch_hdiag(ctx, hii.p, st);
        if (!rc) rc = vec_convert<double, float>(ctx, hii.p, hii32.p, N, st);
      } else {
        rc = launch_hdiag_f32(ctx, hii32.p, st);
      }
      if (rc) return rc;
      SBD_HIP_TRY(ctx, hipStreamSynchronize(st));
      R->t_hdiag += now_s() - th;
      const double e32 = std::max(opt->fp32_eps, opt->eps);
      rc = davidson_core<float>(ctx, hii32.p, W32.p, opt->max_it, opt->max_nb, e32, opt->max_time,
                                opt->verbose, &E, H, R, t_solve0);
      if (rc) return rc;
      R->steps_fp32 = R->steps;
    }
    if (hipMalloc(&W.p, 8 * N) != hipSuccess)
      return set_err(ctx, SBD_ERR_MEMORY, "device allocation failed");
    rc = vec_convert<float, double>(ctx, W32.p, W.p, N, st);
    if (rc) return rc;
    SBD_HIP_TRY(ctx, hipStreamSynchronize(st));
    (void)hipFree(W32.p);
    W32.p = nullptr;
    // renormalise in FP64: the FP32 vector's norm is 1 only to FP32 accuracy,
    // and the Davidson start vector is taken as normalised
    const double* Wc[1] = {W.p};
    double n2 = 0.0;
    rc = vec_multidot<double>(ctx, Wc, 1, W.p, N, &n2, st);
    if (!rc && n2 > 0.0) rc = vec_scale<double>(ctx, W.p, 1.0 / std::sqrt(n2), N, st);
    if (rc) return rc;
  } else {
    if (hipMalloc(&W.p, 8 * N) != hipSuccess)
      return set_err(ctx, SBD_ERR_MEMORY, "device allocation failed");
    rc = start_vector<double>(ctx, opt, W_init, W.p, unit);
    if (rc) return rc;
  }
  if (!hii.p) {
    if (hipMalloc(&hii.p, 8 * N) != hipSuccess)
      return set_err(ctx, SBD_ERR_MEMORY, "device allocation failed");
    const double th = now_s();
    rc = launch_hdiag(ctx, hii.p, st);
    if (rc) return rc;
    SBD_HIP_TRY(ctx, hipStreamSynchronize(st));
    R->t_hdiag += now_s() - th;
  }
  if (lanczos)
    rc = lanczos_core(ctx, hii.p, W.p, opt->max_it, opt->max_nb, opt->eps, opt->verbose, &E, H,
                      R, t_solve0);
  else
    rc = davidson_core<double>(ctx, hii.p, W.p, opt->max_it, opt->max_nb, opt->eps,
                               opt->max_time, opt->verbose, &E, H, R, t_solve0);
  if (rc) return rc;
  R->t_davidson = now_s() - t_solve0;
  if (hist && hist_cap > 0) {
    const size_t n = std::min((size_t)4 * hist_cap, hv.size());
    std::memcpy(hist, hv.data(), sizeof(double) * n);
  }
  // E = <W, H W> with a fresh sigma (sbdiag.h:488-526); the sigma's fused
  // fixed-order <T, W> sum is that inner product
  double e = 0.0;
  {
    DevBuf<double> HW;
    if (hipMalloc(&HW.p, 8 * N) != hipSuccess)
      return set_err(ctx, SBD_ERR_MEMORY, "device allocation failed");
    rc = launch_sigma<double>(ctx, hii.p, W.p, HW.p, 0, st, &e);
    if (rc) return rc;
  }
  if (opt->verbose) {
    printf(" Energy = %.16g\n", e);
    fflush(stdout);
  }
  if (energy) *energy = e;
  if (W_out) {
    SBD_HIP_TRY(ctx, hipMemcpyAsync(W_out, W.p, 8 * N, hipMemcpyDeviceToHost, st));
    SBD_HIP_TRY(ctx, hipStreamSynchronize(st));
  }
  return SBD_OK;
}

}  // extern "C"
