// Block Davidson for the lowest few states (Davidson-Liu with thick restart).
//
// The reference's TPB Davidson (chemistry/tpb/davidson.h:707-970) finds the
// lowest root only.  This solver keeps a subspace of up to max_sub vectors,
// takes the nroots lowest Ritz pairs each iteration, adds one preconditioned
// residual (r / (theta - d), the reference's preconditioner with its 1e-12
// guard) per unconverged root, orthogonalised twice by classical Gram-Schmidt,
// and, when the subspace is full, restarts GD+k style from the span of the
// current Ritz vectors and those of the previous iteration (a plain restart
// to the R Ritz vectors stalls the higher roots; keeping the previous ones
// also saves a quarter of the sigma applications over keeping 2R).  Reductions,
// the projected eigenproblem and the Gram-Schmidt coefficients are FP64; the
// vectors are VT (FP64, or FP32 for the mixed-precision mode, which finishes
// with an FP64 pass from the converted vectors).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "davidson_job.hpp"
#include "sbd_internal.hpp"

using namespace sbdx;

namespace {

double clock_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

template <typename VT>
struct Buf {
  VT* p = nullptr;
  Buf() = default;
  Buf(const Buf&) = delete;
  Buf(Buf&& o) noexcept : p(o.p) { o.p = nullptr; }
  ~Buf() {
    if (p) (void)hipFree(p);
  }
};

// x[i] = uniform_pm1(elem_hash(seed, i)): the elements of the seeded random
// vector of sbd_diag's init = 1 (before normalisation), each computed on its own
template <typename VT>
__global__ void k_random_vec(VT* __restrict__ x, uint64_t n, uint64_t seed) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    x[i] = (VT)uniform_pm1(elem_hash(seed, i));
}

template <typename VT>
__global__ void k_fill(VT* __restrict__ x, uint64_t n, VT v) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    x[i] = v;
}

// per-block minimum of d over indices not in the excluded list (ties: lowest index)
template <typename VT>
__global__ __launch_bounds__(256) void k_argmin(const VT* __restrict__ d, uint64_t n,
                                                const uint64_t* __restrict__ excl, int nex,
                                                double* __restrict__ bval,
                                                uint64_t* __restrict__ bidx) {
  __shared__ double sv[256];
  __shared__ uint64_t si[256];
  double best = INFINITY;
  uint64_t bi = ~0ULL;
  for (uint64_t i = blockIdx.x * 256ULL + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
    bool skip = false;
    for (int e = 0; e < nex; e++) skip |= (excl[e] == i);
    const double v = (double)d[i];
    if (!skip && (v < best || (v == best && i < bi))) {
      best = v;
      bi = i;
    }
  }
  sv[threadIdx.x] = best;
  si[threadIdx.x] = bi;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      const double o = sv[threadIdx.x + w];
      const uint64_t oi = si[threadIdx.x + w];
      if (o < sv[threadIdx.x] || (o == sv[threadIdx.x] && oi < si[threadIdx.x])) {
        sv[threadIdx.x] = o;
        si[threadIdx.x] = oi;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    bval[blockIdx.x] = sv[0];
    bidx[blockIdx.x] = si[0];
  }
}

template <typename VT>
std::vector<const VT*> cptr(const std::vector<VT*>& v, int k) {
  return std::vector<const VT*>(v.begin(), v.begin() + k);
}

}  // namespace

namespace sbdx {

// Indices of the nroots smallest entries of d (ascending value, then index).
template <typename VT>
int lowest_indices(sbd_ctx* ctx, const VT* d, uint64_t n, int nroots, std::vector<uint64_t>& out) {
  const int grid = 1024;
  hipStream_t st = ctx->stream;
  double* bval = nullptr;
  uint64_t *bidx = nullptr, *excl = nullptr;
  SBD_HIP_TRY(ctx, hipMalloc(&bval, sizeof(double) * grid));
  SBD_HIP_TRY(ctx, hipMalloc(&bidx, sizeof(uint64_t) * grid));
  SBD_HIP_TRY(ctx, hipMalloc(&excl, sizeof(uint64_t) * (nroots + 1)));
  out.clear();
  std::vector<double> hv(grid);
  std::vector<uint64_t> hi(grid);
  int rc = SBD_OK;
  for (int r = 0; r < nroots && !rc; r++) {
    if (!out.empty() && hipMemcpyAsync(excl, out.data(), sizeof(uint64_t) * out.size(),
                                       hipMemcpyHostToDevice, st) != hipSuccess) {
      rc = SBD_ERR_DEVICE;
      break;
    }
    hipLaunchKernelGGL(k_argmin<VT>, dim3(grid), dim3(256), 0, st, d, n, excl,
                       (int)out.size(), bval, bidx);
    if (hipMemcpyAsync(hv.data(), bval, sizeof(double) * grid, hipMemcpyDeviceToHost, st) !=
            hipSuccess ||
        hipMemcpyAsync(hi.data(), bidx, sizeof(uint64_t) * grid, hipMemcpyDeviceToHost, st) !=
            hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
      rc = SBD_ERR_DEVICE;
      break;
    }
    double b = INFINITY;
    uint64_t bi = ~0ULL;
    for (int g = 0; g < grid; g++)
      if (hv[g] < b || (hv[g] == b && hi[g] < bi)) {
        b = hv[g];
        bi = hi[g];
      }
    if (bi == ~0ULL) break;
    out.push_back(bi);
  }
  (void)hipFree(bval);
  (void)hipFree(bidx);
  (void)hipFree(excl);
  return rc;
}

// Orthonormalise t against C[0..m) (two classical Gram-Schmidt passes);
// returns false (and leaves t unusable) when t lies in the span.
template <typename VT>
static int orthonormalise(sbd_ctx* ctx, const std::vector<VT*>& C, int m, VT* t, uint64_t N,
                          bool* ok) {
  hipStream_t st = ctx->stream;
  *ok = false;
  std::vector<double> o(kMaxVecs + 1, 0.0);
  double n2 = 0.0;
  const VT* tp[1] = {t};
  int rc = vec_multidot<VT>(ctx, tp, 1, t, N, &n2, st);
  if (rc) return rc;
  if (!(n2 > 0.0)) return SBD_OK;
  for (int pass = 0; pass < 2; pass++) {
    if (m > 0) {
      const auto Cc = cptr(C, m);
      rc = vec_multidot<VT>(ctx, Cc.data(), m, t, N, o.data(), st);
      if (rc) return rc;
      double p2 = n2;
      for (int j = 0; j < m; j++) p2 -= o[j] * o[j];
      if (!(p2 > 1e-20 * n2)) return SBD_OK;  // nothing left outside the span
      double nn = 0.0;
      rc = vec_gs_update<VT>(ctx, Cc.data(), m, o.data(), t, 1.0, N, &nn, st);
      if (rc) return rc;
      n2 = nn;
    }
  }
  if (!(n2 > 0.0)) return SBD_OK;
  rc = vec_scale<VT>(ctx, t, 1.0 / std::sqrt(n2), N, st);
  if (rc) return rc;
  *ok = true;
  return SBD_OK;
}

// X: nroots device vectors (start vectors in, normalised Ritz vectors out).
template <typename VT>
int block_davidson(sbd_ctx* ctx, const VT* dii, int nroots, std::vector<VT*>& X, int max_it,
                   int max_sub, double eps, double max_time, double t_start, int verbose,
                   std::vector<double>& theta_out, std::vector<double>& res_out, int* steps_out,
                   std::vector<double>* hist) {
  const uint64_t N = ctx->N;
  hipStream_t st = ctx->stream;
  const int R = nroots;
  const int M = std::min(std::max(max_sub, 3 * R), kMaxVecs);
  std::vector<Buf<VT>> Cb(M), HCb(M), Yb(4 * R);
  for (int k = 0; k < M; k++)
    if (hipMalloc(&Cb[k].p, sizeof(VT) * N) != hipSuccess ||
        hipMalloc(&HCb[k].p, sizeof(VT) * N) != hipSuccess)
      return SBD_ERR_MEMORY;
  for (auto& y : Yb)
    if (hipMalloc(&y.p, sizeof(VT) * N) != hipSuccess) return SBD_ERR_MEMORY;
  std::vector<VT*> C(M), HC(M);
  for (int k = 0; k < M; k++) {
    C[k] = Cb[k].p;
    HC[k] = HCb[k].p;
  }
  std::vector<double> H((size_t)M * M, 0.0), U((size_t)M * M), th(M), o(2 * kMaxVecs + 3);
  // start subspace: the given vectors, orthonormalised
  int m = 0;
  for (int r = 0; r < R; r++) {
    SBD_HIP_TRY(ctx, hipMemcpyAsync(C[m], X[r], sizeof(VT) * N, hipMemcpyDeviceToDevice, st));
    bool ok = false;
    int rc = orthonormalise<VT>(ctx, C, m, C[m], N, &ok);
    if (rc) return rc;
    if (ok) m++;
  }
  if (m < R) return SBD_ERR_ARG;
  // sigma of the start vectors and their projected matrix
  auto add_columns = [&](int j0, int j1) -> int {
    for (int j = j0; j < j1; j++) {
      int rc = launch_sigma<VT>(ctx, dii, C[j], HC[j], 0, st);
      if (rc) return rc;
    }
    for (int j = j0; j < j1; j++) {
      const auto Cc = cptr(C, j1);
      int rc = vec_multidot<VT>(ctx, Cc.data(), j1, HC[j], N, o.data(), st);
      if (rc) return rc;
      for (int i = 0; i < j1; i++) {
        H[i + (size_t)M * j] = o[i];
        H[j + (size_t)M * i] = o[i];
      }
    }
    return SBD_OK;
  };
  int rc = add_columns(0, m);
  if (rc) return rc;
  std::vector<double> res(R, 0.0), Uprev((size_t)M * R, 0.0);
  int steps = 0, mprev = 0;
  bool done = false;
  auto rayleigh_ritz = [&]() {
    for (int j = 0; j < m; j++)
      for (int i = 0; i < m; i++) U[i + (size_t)M * j] = H[i + (size_t)M * j];
    syev_jacobi(m, U.data(), M, th.data());
  };
  for (int it = 0; it < max_it && !done; it++) {
    rayleigh_ritz();
    if (m == M) {
      // restart basis: orthonormal coordinates (two modified Gram-Schmidt
      // passes, FP64 on the host) spanning the R Ritz vectors and the R Ritz
      // vectors of the previous iteration, whose coordinates in today's basis
      // are yesterday's padded with zeros (the basis only grew since)
      std::vector<double> Q;
      int K = 0;
      auto push = [&](std::vector<double> v) {
        for (int pass = 0; pass < 2; pass++)
          for (int k = 0; k < K; k++) {
            double d = 0.0;
            for (int i = 0; i < m; i++) d += Q[(size_t)k * m + i] * v[i];
            for (int i = 0; i < m; i++) v[i] -= d * Q[(size_t)k * m + i];
          }
        double n2 = 0.0;
        for (int i = 0; i < m; i++) n2 += v[i] * v[i];
        if (!(n2 > 1e-8)) return;  // (nearly) inside the span already
        const double s = 1.0 / std::sqrt(n2);
        for (int i = 0; i < m; i++) Q.push_back(v[i] * s);
        K++;
      };
      for (int r = 0; r < R; r++) {
        std::vector<double> v(m);
        for (int i = 0; i < m; i++) v[i] = U[i + (size_t)M * r];
        push(v);
      }
      for (int r = 0; r < R && mprev > 0; r++) {
        std::vector<double> v(m, 0.0);
        for (int i = 0; i < mprev; i++) v[i] = Uprev[i + (size_t)M * r];
        push(v);
      }
      const auto Cc = cptr(C, m), HCc = cptr(HC, m);
      for (int k = 0; k < K; k++) {
        double n2 = 0.0;
        rc = vec_lincomb<VT>(ctx, Cc.data(), m, &Q[(size_t)k * m], 1.0, Yb[k].p, N, &n2, st);
        if (!rc)
          rc = vec_lincomb<VT>(ctx, HCc.data(), m, &Q[(size_t)k * m], 1.0, Yb[2 * R + k].p, N,
                               &n2, st);
        if (rc) return rc;
      }
      for (int k = 0; k < K; k++) {
        std::swap(Cb[k].p, Yb[k].p);
        std::swap(HCb[k].p, Yb[2 * R + k].p);
        C[k] = Cb[k].p;
        HC[k] = HCb[k].p;
      }
      m = K;
      for (int j = 0; j < m; j++) {
        const auto Cm = cptr(C, m);
        rc = vec_multidot<VT>(ctx, Cm.data(), m, HC[j], N, o.data(), st);
        if (rc) return rc;
        for (int i = 0; i < m; i++) {
          H[i + (size_t)M * j] = o[i];
          H[j + (size_t)M * i] = o[i];
        }
      }
      rayleigh_ritz();
    }
    // this iteration's Ritz coordinates, for the next restart
    for (int r = 0; r < R; r++)
      for (int i = 0; i < m; i++) Uprev[i + (size_t)M * r] = U[i + (size_t)M * r];
    mprev = m;
    steps++;
    // residuals of the R lowest Ritz pairs; a correction for each unconverged one
    int added = 0;
    bool all = true;
    for (int r = 0; r < R; r++) {
      std::vector<double> u(m);
      for (int i = 0; i < m; i++) u[i] = U[i + (size_t)M * r];
      VT* t = (m + added < M) ? C[m + added] : Yb[0].p;
      const auto Cc = cptr(C, m), HCc = cptr(HC, m);
      rc = vec_ritz_precond<VT>(ctx, Cc.data(), HCc.data(), m, u.data(), th[r], dii, t, N,
                                o.data(), st);
      if (rc) return rc;
      res[r] = std::sqrt(o[0]);
      if (res[r] < eps) continue;
      all = false;
      if (m + added >= M) continue;  // full: the next iteration restarts
      bool ok = false;
      rc = orthonormalise<VT>(ctx, C, m + added, t, N, &ok);
      if (rc) return rc;
      if (ok) added++;
    }
    double rmax = 0.0;
    for (double x : res) rmax = std::max(rmax, x);
    if (hist) {
      hist->push_back(th[0]);
      hist->push_back(rmax);
      hist->push_back(clock_s() - t_start);
      hist->push_back(sizeof(VT) == 4 ? 32.0 : 64.0);
    }
    if (verbose) {
      printf(" Block Davidson iteration %d (max tol=%g):", it, rmax);
      for (int r = 0; r < R; r++) printf(" %.16g", th[r]);
      printf("\n");
      fflush(stdout);
    }
    if (all) {
      done = true;
      break;
    }
    if (added == 0) break;  // every correction lies in the subspace: stagnated
    if (clock_s() - t_start > max_time) break;
    rc = add_columns(m, m + added);
    if (rc) return rc;
    m += added;
  }
  // Ritz vectors of the final subspace
  for (int j = 0; j < m; j++)
    for (int i = 0; i < m; i++) U[i + (size_t)M * j] = H[i + (size_t)M * j];
  syev_jacobi(m, U.data(), M, th.data());
  for (int r = 0; r < R; r++) {
    std::vector<double> u(m);
    for (int i = 0; i < m; i++) u[i] = U[i + (size_t)M * r];
    const auto Cc = cptr(C, m);
    double n2 = 0.0;
    rc = vec_lincomb<VT>(ctx, Cc.data(), m, u.data(), 1.0, X[r], N, &n2, st);
    if (!rc) rc = vec_scale<VT>(ctx, X[r], 1.0 / std::sqrt(n2), N, st);
    if (rc) return rc;
  }
  if (hipStreamSynchronize(st) != hipSuccess) return SBD_ERR_DEVICE;
  theta_out.assign(th.begin(), th.begin() + R);
  res_out = res;
  *steps_out = steps;
  return done ? SBD_OK : SBD_ERR_NOT_CONVERGED;
}

}  // namespace sbdx

extern "C" {

int sbd_diag_roots(sbd_ctx* ctx, const sbd_options* opt, int nroots, double* energies,
                   double* residuals, double* W_out, double* hist, int hist_cap,
                   sbd_report* rep) {
  if (!ctx || !opt || !energies || nroots < 1 || nroots > kMaxVecs / 3) return SBD_ERR_ARG;
  if (opt->init != 0 && opt->init != 1) return SBD_ERR_ARG;
  sbd_report local{};
  sbd_report* R = rep ? rep : &local;
  std::memset(R, 0, sizeof(sbd_report));
  const double t0 = clock_s();
  int rc = sbd_make_helpers(ctx);
  if (rc) return rc;
  R->t_tables = clock_s() - t0;
  const uint64_t N = ctx->N;
  hipStream_t st = ctx->stream;
  if ((uint64_t)nroots > N) return SBD_ERR_ARG;
  Buf<double> hii;
  std::vector<Buf<double>> Xb(nroots);
  if (hipMalloc(&hii.p, 8 * N) != hipSuccess) return SBD_ERR_MEMORY;
  for (auto& x : Xb)
    if (hipMalloc(&x.p, 8 * N) != hipSuccess) return SBD_ERR_MEMORY;
  const double t1 = clock_s();
  rc = launch_hdiag(ctx, hii.p, st);
  if (rc) return rc;
  if (hipStreamSynchronize(st) != hipSuccess) return SBD_ERR_DEVICE;
  R->t_hdiag = clock_s() - t1;
  // start vectors: unit vectors on the lowest diagonal elements (init 0) or
  // the seeded random stream with seed, seed + 1, ... (init 1)
  const unsigned g = (unsigned)std::min<uint64_t>((N + 255) / 256, 4096);
  if (opt->init == 0) {
    std::vector<uint64_t> idx;
    rc = lowest_indices<double>(ctx, hii.p, N, nroots, idx);
    if (rc) return rc;
    for (int r = 0; r < nroots; r++) {
      hipLaunchKernelGGL(k_fill<double>, dim3(g), dim3(256), 0, st, Xb[r].p, N, 0.0);
      const double one = 1.0;
      if (hipMemcpyAsync(Xb[r].p + idx[r], &one, 8, hipMemcpyHostToDevice, st) != hipSuccess ||
          hipStreamSynchronize(st) != hipSuccess)
        return SBD_ERR_DEVICE;
    }
  } else {
    for (int r = 0; r < nroots; r++)
      hipLaunchKernelGGL(k_random_vec<double>, dim3(g), dim3(256), 0, st, Xb[r].p, N,
                         opt->seed + (uint64_t)r);
  }
  std::vector<double> hv, th, res;
  std::vector<double>* H = hist ? &hv : nullptr;
  int steps = 0;
  const double t2 = clock_s();
  int conv_rc = SBD_OK;
  if (opt->precision == 1) {
    // FP32 block phase, then the FP64 block phase from its vectors
    Buf<float> hii32;
    std::vector<Buf<float>> X32(nroots);
    if (hipMalloc(&hii32.p, 4 * N) != hipSuccess) return SBD_ERR_MEMORY;
    for (auto& x : X32)
      if (hipMalloc(&x.p, 4 * N) != hipSuccess) return SBD_ERR_MEMORY;
    rc = vec_convert<double, float>(ctx, hii.p, hii32.p, N, st);
    for (int r = 0; r < nroots && !rc; r++)
      rc = vec_convert<double, float>(ctx, Xb[r].p, X32[r].p, N, st);
    if (rc) return rc;
    std::vector<float*> Xp(nroots);
    for (int r = 0; r < nroots; r++) Xp[r] = X32[r].p;
    const double e32 = std::max(opt->fp32_eps, opt->eps);
    rc = block_davidson<float>(ctx, hii32.p, nroots, Xp, opt->max_it, opt->max_nb, e32,
                               opt->max_time, t2, opt->verbose, th, res, &steps, H);
    if (rc && rc != SBD_ERR_NOT_CONVERGED) return rc;
    R->steps_fp32 = steps;
    for (int r = 0; r < nroots; r++) {
      rc = vec_convert<float, double>(ctx, X32[r].p, Xb[r].p, N, st);
      if (rc) return rc;
    }
  }
  std::vector<double*> Xp(nroots);
  for (int r = 0; r < nroots; r++) Xp[r] = Xb[r].p;
  int steps64 = 0;
  conv_rc = block_davidson<double>(ctx, hii.p, nroots, Xp, opt->max_it, opt->max_nb, opt->eps,
                                   opt->max_time, t2, opt->verbose, th, res, &steps64, H);
  if (conv_rc && conv_rc != SBD_ERR_NOT_CONVERGED) return conv_rc;
  R->steps = R->steps_fp32 + steps64;
  R->t_davidson = clock_s() - t2;
  // energies <X_r, H X_r> with fresh FP64 sigmas (as sbd_diag)
  Buf<double> HX;
  if (hipMalloc(&HX.p, 8 * N) != hipSuccess) return SBD_ERR_MEMORY;
  for (int r = 0; r < nroots; r++) {
    rc = launch_sigma<double>(ctx, hii.p, Xp[r], HX.p, 0, st);
    const double* Xc[1] = {Xp[r]};
    if (!rc) rc = vec_multidot<double>(ctx, Xc, 1, HX.p, N, &energies[r], st);
    if (rc) return rc;
    if (residuals) residuals[r] = res[r];
    if (W_out && (hipMemcpyAsync(W_out + (size_t)r * N, Xp[r], 8 * N, hipMemcpyDeviceToHost,
                                 st) != hipSuccess ||
                  hipStreamSynchronize(st) != hipSuccess))
      return SBD_ERR_DEVICE;
  }
  if (hist && hist_cap > 0) {
    const size_t n = std::min((size_t)4 * hist_cap, hv.size());
    std::memcpy(hist, hv.data(), sizeof(double) * n);
  }
  double rmax = 0.0;
  for (double x : res) rmax = std::max(rmax, x);
  R->converged = conv_rc == SBD_OK ? 1 : 0;
  R->residual = rmax;
  R->energy = energies[0];
  return SBD_OK;
}

}  // extern "C"
