// Davidson vector passes (reference: chemistry/tpb/davidson.h:775-922 and the
// GPU variant chemistry/basic/davidson_thrust.h:185-342, thrust_kernels.h).
//
// The reference makes one pass per dot product / AXPY (about 5k+9 passes per
// step for subspace size k).  Here each step is three fused passes:
//   multidot     : H[j, k] = <C_j, HC_k> for all j at once
//   ritz_precond : w = sum u_j C_j, r = -(sum u_j HC_j) + E0 w, the
//                  preconditioned t = r / (E0 - d), and in the same pass
//                  |r|^2, |w|^2, <C_j, t>, |t|^2
//   gs_update    : classical Gram-Schmidt t -= sum o_j C_j, scaled, |t|^2
// All sums are FP64 with per-thread Kahan compensation; every kernel runs a
// fixed grid and the block partials are summed in a fixed order, so results
// are deterministic run to run.
#include "sbd_internal.hpp"

namespace sbdx {

constexpr int kVecBlock = 256;
constexpr int kVecGrid = 1024;
constexpr int kMaxSums = 2 * kMaxVecs + 4;

template <typename VT>
struct VecPtrs {
  const VT* p[kMaxVecs];
};
struct Coefs {
  double c[kMaxVecs];
};

struct Kahan {
  double s = 0.0, c = 0.0;
  __device__ inline void add(double x) {
    const double y = x - c;
    const double t = s + y;
    c = (t - s) - y;
    s = t;
  }
};

// Block-reduce m per-thread values (value j held in v[j]) and write partials.
template <int MAXM>
__device__ inline void block_partials(double (&v)[MAXM], int m, double* __restrict__ part) {
  __shared__ double red[kVecBlock / 64][MAXM];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int j = 0; j < m; j++) {
    double x = v[j];
    for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
    if (lane == 0) red[wid][j] = x;
  }
  __syncthreads();
  if (threadIdx.x < m) {
    double s = 0.0;
    for (int w = 0; w < kVecBlock / 64; w++) s += red[w][threadIdx.x];
    part[blockIdx.x * m + threadIdx.x] = s;
  }
}

__global__ void k_finalize(const double* __restrict__ part, int nblk, int m,
                           double* __restrict__ out) {
  const int j = threadIdx.x;
  if (j >= m) return;
  Kahan k;
  for (int b = 0; b < nblk; b++) k.add(part[b * m + j]);
  out[j] = k.s;
}

template <typename VT, int MAXM>
__global__ __launch_bounds__(kVecBlock) void k_multidot(VecPtrs<VT> X, int m,
                                                        const VT* __restrict__ y, uint64_t n,
                                                        double* __restrict__ part) {
  Kahan acc[MAXM];
  for (uint64_t i = blockIdx.x * (uint64_t)kVecBlock + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * kVecBlock) {
    const double yi = (double)y[i];
#pragma unroll
    for (int j = 0; j < MAXM; j++)
      if (j < m) acc[j].add((double)X.p[j][i] * yi);
  }
  double v[MAXM];
#pragma unroll
  for (int j = 0; j < MAXM; j++) v[j] = acc[j].s;
  block_partials<MAXM>(v, m, part);
}

// sums: [0] |r|^2, [1] |w|^2, [2] |t|^2, [3 .. 3+k) <C_j, t>
template <typename VT, int MAXK>
__global__ __launch_bounds__(kVecBlock) void k_ritz_precond(VecPtrs<VT> C, VecPtrs<VT> HC, int k,
                                                            Coefs u, double E0,
                                                            const VT* __restrict__ dii,
                                                            VT* __restrict__ t, uint64_t n,
                                                            double* __restrict__ part) {
  constexpr double eps_reg = 1.0e-12;
  Kahan rr, ww, tt, ot[MAXK];
  for (uint64_t i = blockIdx.x * (uint64_t)kVecBlock + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * kVecBlock) {
    double c[MAXK];
#pragma unroll
    for (int j = 0; j < MAXK; j++) c[j] = j < k ? (double)C.p[j][i] : 0.0;
    double w = c[0] * u.c[0];
    double r = (double)HC.p[0][i] * (-1.0 * u.c[0]);
#pragma unroll
    for (int j = 1; j < MAXK; j++) {
      if (j < k) {
        w += c[j] * u.c[j];
        r += (double)HC.p[j][i] * (-1.0 * u.c[j]);
      }
    }
    r += E0 * w;
    const double den = E0 - (double)dii[i];
    double tv = fabs(den) > eps_reg ? r / den : r / (den - eps_reg);
    const VT ts = (VT)tv;
    t[i] = ts;
    tv = (double)ts;
    rr.add(r * r);
    ww.add(w * w);
    tt.add(tv * tv);
#pragma unroll
    for (int j = 0; j < MAXK; j++)
      if (j < k) ot[j].add(c[j] * tv);
  }
  double v[MAXK + 3];
  v[0] = rr.s;
  v[1] = ww.s;
  v[2] = tt.s;
#pragma unroll
  for (int j = 0; j < MAXK; j++) v[3 + j] = ot[j].s;
  block_partials<MAXK + 3>(v, k + 3, part);
}

template <typename VT, int MAXK>
__global__ __launch_bounds__(kVecBlock) void k_gs_update(VecPtrs<VT> C, int k, Coefs o,
                                                         VT* __restrict__ t, double scale,
                                                         uint64_t n, double* __restrict__ part) {
  Kahan tt;
  for (uint64_t i = blockIdx.x * (uint64_t)kVecBlock + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * kVecBlock) {
    double x = (double)t[i];
#pragma unroll
    for (int j = 0; j < MAXK; j++)
      if (j < k) x -= o.c[j] * (double)C.p[j][i];
    const VT xs = (VT)(x * scale);
    t[i] = xs;
    tt.add((double)xs * (double)xs);
  }
  double v[1] = {tt.s};
  block_partials<1>(v, 1, part);
}

template <typename VT, int MAXK>
__global__ __launch_bounds__(kVecBlock) void k_lincomb(VecPtrs<VT> C, int k, Coefs u,
                                                       VT* __restrict__ out, uint64_t n,
                                                       double* __restrict__ part) {
  Kahan tt;
  for (uint64_t i = blockIdx.x * (uint64_t)kVecBlock + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * kVecBlock) {
    double x = (double)C.p[0][i] * u.c[0];
#pragma unroll
    for (int j = 1; j < MAXK; j++)
      if (j < k) x += (double)C.p[j][i] * u.c[j];
    const VT xs = (VT)x;
    out[i] = xs;
    tt.add((double)xs * (double)xs);
  }
  double v[1] = {tt.s};
  block_partials<1>(v, 1, part);
}

template <typename VT>
__global__ void k_scale(VT* __restrict__ x, double s, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    x[i] = (VT)((double)x[i] * s);
}

template <typename VS, typename VD>
__global__ void k_convert(const VS* __restrict__ x, VD* __restrict__ y, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    y[i] = (VD)x[i];
}

// ---------------------------------------------------------------- launchers

static int ensure_scratch(sbd_ctx* ctx) {
  const size_t need = (size_t)kVecGrid * kMaxSums;
  if (ctx->partial_cap >= need) return SBD_OK;
  if (ctx->d_partials) (void)hipFree(ctx->d_partials);
  if (ctx->d_results) (void)hipFree(ctx->d_results);
  if (ctx->h_results) (void)hipHostFree(ctx->h_results);
  SBD_HIP_TRY(ctx, hipMalloc(&ctx->d_partials, sizeof(double) * need));
  SBD_HIP_TRY(ctx, hipMalloc(&ctx->d_results, sizeof(double) * kMaxSums));
  SBD_HIP_TRY(ctx, hipHostMalloc(&ctx->h_results, sizeof(double) * kMaxSums));
  ctx->partial_cap = need;
  return SBD_OK;
}

static int grid_for(uint64_t n) {
  uint64_t g = (n + kVecBlock - 1) / kVecBlock;
  if (g > (uint64_t)kVecGrid) g = kVecGrid;
  return g < 1 ? 1 : (int)g;
}

static int finish(sbd_ctx* ctx, int grid, int m, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(64 * ((m + 63) / 64)), 0, st, ctx->d_partials, grid,
                     m, ctx->d_results);
  SBD_HIP_TRY(ctx, hipGetLastError());
  SBD_HIP_TRY(ctx, hipMemcpyAsync(ctx->h_results, ctx->d_results, sizeof(double) * m,
                                  hipMemcpyDeviceToHost, st));
  SBD_HIP_TRY(ctx, hipStreamSynchronize(st));
  for (int j = 0; j < m; j++) out[j] = ctx->h_results[j];
  return SBD_OK;
}

template <typename VT>
static VecPtrs<VT> pack(const VT* const* X, int m) {
  VecPtrs<VT> P{};
  for (int j = 0; j < m; j++) P.p[j] = X[j];
  return P;
}

static Coefs packc(const double* c, int m) {
  Coefs C{};
  for (int j = 0; j < m; j++) C.c[j] = c[j];
  return C;
}

// dispatch a kernel template over compile-time sizes 4/8/16/24/40
#define SBD_DISPATCH_K(k, KERN, ...)                                                       \
  do {                                                                                     \
    if ((k) <= 4) hipLaunchKernelGGL((KERN<VT, 4>), __VA_ARGS__);                          \
    else if ((k) <= 8) hipLaunchKernelGGL((KERN<VT, 8>), __VA_ARGS__);                     \
    else if ((k) <= 16) hipLaunchKernelGGL((KERN<VT, 16>), __VA_ARGS__);                   \
    else if ((k) <= 24) hipLaunchKernelGGL((KERN<VT, 24>), __VA_ARGS__);                   \
    else hipLaunchKernelGGL((KERN<VT, kMaxVecs>), __VA_ARGS__);                            \
  } while (0)

template <typename VT>
int vec_multidot(sbd_ctx* ctx, const VT* const* X, int m, const VT* y, uint64_t n, double* out,
                 hipStream_t st) {
  if (m < 1 || m > kMaxVecs) return SBD_ERR_ARG;
  int rc = ensure_scratch(ctx);
  if (rc) return rc;
  const int g = grid_for(n);
  SBD_DISPATCH_K(m, k_multidot, dim3(g), dim3(kVecBlock), 0, st, pack(X, m), m, y, n,
                 ctx->d_partials);
  SBD_HIP_TRY(ctx, hipGetLastError());
  return finish(ctx, g, m, out, st);
}

template <typename VT>
int vec_ritz_precond(sbd_ctx* ctx, const VT* const* C, const VT* const* HC, int k, const double* u,
                     double E0, const VT* dii, VT* t, uint64_t n, double* out, hipStream_t st) {
  if (k < 1 || k > kMaxVecs) return SBD_ERR_ARG;
  int rc = ensure_scratch(ctx);
  if (rc) return rc;
  const int g = grid_for(n);
  SBD_DISPATCH_K(k, k_ritz_precond, dim3(g), dim3(kVecBlock), 0, st, pack(C, k), pack(HC, k), k,
                 packc(u, k), E0, dii, t, n, ctx->d_partials);
  SBD_HIP_TRY(ctx, hipGetLastError());
  return finish(ctx, g, k + 3, out, st);
}

template <typename VT>
int vec_gs_update(sbd_ctx* ctx, const VT* const* C, int k, const double* coef, VT* t, double scale,
                  uint64_t n, double* out_norm2, hipStream_t st) {
  if (k < 1 || k > kMaxVecs) return SBD_ERR_ARG;
  int rc = ensure_scratch(ctx);
  if (rc) return rc;
  const int g = grid_for(n);
  SBD_DISPATCH_K(k, k_gs_update, dim3(g), dim3(kVecBlock), 0, st, pack(C, k), k, packc(coef, k), t,
                 scale, n, ctx->d_partials);
  SBD_HIP_TRY(ctx, hipGetLastError());
  return finish(ctx, g, 1, out_norm2, st);
}

template <typename VT>
int vec_lincomb(sbd_ctx* ctx, const VT* const* C, int k, const double* u, VT* out, uint64_t n,
                double* out_norm2, hipStream_t st) {
  if (k < 1 || k > kMaxVecs) return SBD_ERR_ARG;
  int rc = ensure_scratch(ctx);
  if (rc) return rc;
  const int g = grid_for(n);
  SBD_DISPATCH_K(k, k_lincomb, dim3(g), dim3(kVecBlock), 0, st, pack(C, k), k, packc(u, k), out, n,
                 ctx->d_partials);
  SBD_HIP_TRY(ctx, hipGetLastError());
  return finish(ctx, g, 1, out_norm2, st);
}

template <typename VT>
int vec_scale(sbd_ctx* ctx, VT* x, double s, uint64_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_scale<VT>, dim3(grid_for(n) * 2), dim3(kVecBlock), 0, st, x, s, n);
  SBD_HIP_TRY(ctx, hipGetLastError());
  return SBD_OK;
}

template <typename VS, typename VD>
int vec_convert(sbd_ctx* ctx, const VS* x, VD* y, uint64_t n, hipStream_t st) {
  hipLaunchKernelGGL((k_convert<VS, VD>), dim3(grid_for(n) * 2), dim3(kVecBlock), 0, st, x, y, n);
  SBD_HIP_TRY(ctx, hipGetLastError());
  return SBD_OK;
}

#define SBD_INST(VT)                                                                          \
  template int vec_multidot<VT>(sbd_ctx*, const VT* const*, int, const VT*, uint64_t, double*, \
                                hipStream_t);                                                 \
  template int vec_ritz_precond<VT>(sbd_ctx*, const VT* const*, const VT* const*, int,        \
                                    const double*, double, const VT*, VT*, uint64_t, double*,  \
                                    hipStream_t);                                             \
  template int vec_gs_update<VT>(sbd_ctx*, const VT* const*, int, const double*, VT*, double,  \
                                 uint64_t, double*, hipStream_t);                             \
  template int vec_lincomb<VT>(sbd_ctx*, const VT* const*, int, const double*, VT*, uint64_t,  \
                               double*, hipStream_t);                                         \
  template int vec_scale<VT>(sbd_ctx*, VT*, double, uint64_t, hipStream_t);
SBD_INST(double)
SBD_INST(float)
template int vec_convert<float, double>(sbd_ctx*, const float*, double*, uint64_t, hipStream_t);
template int vec_convert<double, float>(sbd_ctx*, const double*, float*, uint64_t, hipStream_t);

}  // namespace sbdx
