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
#include <algorithm>
#include <cmath>
#include <initializer_list>
#include <mutex>

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

// The share of one pass that a workgroup computes: block bid of nblk (the
// launch grid for one vector pass, or one job's range inside a batch).
struct Grid {
  uint32_t bid, nblk;
};
__device__ inline Grid launch_grid() { return Grid{blockIdx.x, gridDim.x}; }

// Block-reduce m per-thread values (value j held in v[j]) and write partials.
template <int MAXM>
__device__ inline void block_partials(Grid g, double (&v)[MAXM], int m,
                                      double* __restrict__ part) {
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
    part[g.bid * m + threadIdx.x] = s;
  }
}

// One block per sum: thread t adds block partials t, t + 256, ... (compensated),
// then the 256 thread sums are combined in a fixed tree.
__global__ __launch_bounds__(kVecBlock) void k_finalize(const double* __restrict__ part, int nblk,
                                                        int m, double* __restrict__ out) {
  __shared__ double red[kVecBlock];
  const int j = blockIdx.x;
  Kahan k;
  for (int b = threadIdx.x; b < nblk; b += kVecBlock) k.add(part[b * m + j]);
  red[threadIdx.x] = k.s;
  __syncthreads();
  for (int w = kVecBlock / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[j] = red[0];
}

// Elements are read as 16-byte chunks (two doubles or four floats) when every
// vector of the call is 16-byte aligned; the remainder (and any call with an
// unaligned vector) goes element by element.  Each thread owns a fixed set of
// chunks, so the per-thread sums, and hence the results, are the same on
// every run.
template <typename VT, int W>
struct alignas(W * sizeof(VT)) Chunk {
  VT v[W];
};
template <int W, typename VT>
__device__ inline Chunk<VT, W> ldc(const VT* __restrict__ p, uint64_t i) {
  return *reinterpret_cast<const Chunk<VT, W>*>(p + i);
}
template <int W, typename VT>
__device__ inline void stc(VT* __restrict__ p, uint64_t i, const Chunk<VT, W>& c) {
  *reinterpret_cast<Chunk<VT, W>*>(p + i) = c;
}
template <typename VT>
constexpr int chunk_of() {
  return 16 / (int)sizeof(VT);  // 16-byte loads: two doubles or four floats
}

// Calls body.template operator()<W>(i) over this thread's share of [0, n):
// W-wide chunks at i = W * (tid + j * nthreads) while they fit (if packed),
// then single elements.
template <typename VT, typename F, int P = chunk_of<VT>()>
__device__ inline void for_chunks(Grid g, uint64_t n, bool packed, F&& body) {
  const uint64_t tid = g.bid * (uint64_t)kVecBlock + threadIdx.x;
  const uint64_t nth = (uint64_t)g.nblk * kVecBlock;
  const uint64_t np = packed ? n / P : 0;
  for (uint64_t q = tid; q < np; q += nth) body.template operator()<P>(q * P);
  for (uint64_t i = np * P + tid; i < n; i += nth) body.template operator()<1>(i);
}

template <typename VT, int MAXM>
__global__ __launch_bounds__(kVecBlock) void k_multidot(VecPtrs<VT> X, int m,
                                                        const VT* __restrict__ y, uint64_t n,
                                                        bool packed, double* __restrict__ part) {
  Kahan acc[MAXM];
  for_chunks<VT>(launch_grid(), n, packed, [&]<int W>(uint64_t i) {
    const Chunk<VT, W> yv = ldc<W>(y, i);
#pragma unroll
    for (int j = 0; j < MAXM; j++) {
      if (j < m) {
        const Chunk<VT, W> xv = ldc<W>(X.p[j], i);
#pragma unroll
        for (int e = 0; e < W; e++) acc[j].add((double)xv.v[e] * (double)yv.v[e]);
      }
    }
  });
  double v[MAXM];
#pragma unroll
  for (int j = 0; j < MAXM; j++) v[j] = acc[j].s;
  block_partials<MAXM>(launch_grid(), v, m, part);
}

// One pass of the Davidson step after the Ritz problem:
//   w = sum u_j C_j,  r = -(sum u_j HC_j) + E0 w,  t = r / (E0 - d)
// with the 1e-12 guard of the reference preconditioner (davidson.h:868-877).
// sums: [0] |r|^2, [1] |w|^2, [2] |t|^2, [3, 3+k) <C_j, t>, [3+MAXK, 3+MAXK+k)
// <HC_j, t> (t as stored); the last group gives the next projected-matrix
// column by symmetry of H, so no separate overlap pass follows the sigma.
template <typename VT, int MAXK, bool kHold>
__device__ inline void ritz_body(Grid g, const VecPtrs<VT>& C, const VecPtrs<VT>& HC, int k,
                                 const Coefs& u, double E0, const VT* __restrict__ dii,
                                 VT* __restrict__ t, uint64_t n, bool packed,
                                 double* __restrict__ part, bool wout) {
  constexpr double eps_reg = 1.0e-12;
  // norms compensated; the 2k overlaps as plain FP64 per-thread sums (the
  // register budget of 2k compensated sums would cut the waves in flight)
  Kahan rr, ww, tt;
  double ot[MAXK], oh[MAXK];
#pragma unroll
  for (int j = 0; j < MAXK; j++) ot[j] = oh[j] = 0.0;
  // kHold (FP64 single launches up to 12 vectors per side): all 2k chunks
  // are held in registers until t is known, so each is read once.
  // Otherwise (FP32, wider FP64 subspaces and the batched launches, where
  // holding them would spill): w and r are accumulated as the chunks arrive
  // and, once t is known, the chunks are taken again for the overlaps: the
  // first kStage pairs (C_j, HC_j) from LDS, where each thread parked them,
  // the others read again newest first (the lines most likely still in L2).
  // No chunk is held in registers across the division, so four waves fit per
  // SIMD (the 32 KiB of LDS allows four workgroups per CU as well).  Every
  // accumulator takes its terms in the same order in both forms, so the sums
  // are the same.
  constexpr int kStage = 4;
  __shared__ Chunk<VT, chunk_of<VT>()> stage[kHold ? 1 : 2 * kStage][kVecBlock];
  auto hold = [&]<int W>(uint64_t i) {
    double w[W], r[W];
    Chunk<VT, W> cv[MAXK], hv[MAXK];
#pragma unroll
    for (int j = 0; j < MAXK; j++) {
      if (j < k) {
        cv[j] = ldc<W>(C.p[j], i);
        hv[j] = ldc<W>(HC.p[j], i);
      }
    }
    const Chunk<VT, W> dv = ldc<W>(dii, i);
#pragma unroll
    for (int e = 0; e < W; e++) {
      w[e] = (double)cv[0].v[e] * u.c[0];
      r[e] = (double)hv[0].v[e] * (-1.0 * u.c[0]);
    }
#pragma unroll
    for (int j = 1; j < MAXK; j++) {
      if (j < k) {
#pragma unroll
        for (int e = 0; e < W; e++) {
          w[e] += (double)cv[j].v[e] * u.c[j];
          r[e] += (double)hv[j].v[e] * (-1.0 * u.c[j]);
        }
      }
    }
    Chunk<VT, W> ts;
#pragma unroll
    for (int e = 0; e < W; e++) {
      r[e] += E0 * w[e];
      const double den = E0 - (double)dv.v[e];
      const double tv = fabs(den) > eps_reg ? r[e] / den : r[e] / (den - eps_reg);
      ts.v[e] = (VT)tv;
    }
    stc<W>(t, i, ts);
#pragma unroll
    for (int e = 0; e < W; e++) {
      const double tv = (double)ts.v[e];
      rr.add(r[e] * r[e]);
      ww.add(w[e] * w[e]);
      tt.add(tv * tv);
#pragma unroll
      for (int j = 0; j < MAXK; j++) {
        if (j < k) {
          ot[j] += (double)cv[j].v[e] * tv;
          oh[j] += (double)hv[j].v[e] * tv;
        }
      }
    }
  };
  auto reread = [&]<int W>(uint64_t i) {
    double w[W], r[W];
#pragma unroll
    for (int e = 0; e < W; e++) w[e] = r[e] = 0.0;
#pragma unroll
    for (int j = 0; j < MAXK; j++) {
      if (j < k) {
        const Chunk<VT, W> cv = ldc<W>(C.p[j], i);
        const Chunk<VT, W> hv = ldc<W>(HC.p[j], i);
        if constexpr (W == chunk_of<VT>() && !kHold) {
          if (j < kStage) {
            stage[2 * j][threadIdx.x] = cv;
            stage[2 * j + 1][threadIdx.x] = hv;
          }
        }
#pragma unroll
        for (int e = 0; e < W; e++) {
          if (j == 0) {
            w[e] = (double)cv.v[e] * u.c[0];
            r[e] = (double)hv.v[e] * (-1.0 * u.c[0]);
          } else {
            w[e] += (double)cv.v[e] * u.c[j];
            r[e] += (double)hv.v[e] * (-1.0 * u.c[j]);
          }
        }
      }
    }
    const Chunk<VT, W> dv = ldc<W>(dii, i);
    Chunk<VT, W> ts;
#pragma unroll
    for (int e = 0; e < W; e++) {
      r[e] += E0 * w[e];
      const double den = E0 - (double)dv.v[e];
      const double tv = fabs(den) > eps_reg ? r[e] / den : r[e] / (den - eps_reg);
      ts.v[e] = (VT)tv;
    }
    stc<W>(t, i, ts);
    double tv[W];
#pragma unroll
    for (int e = 0; e < W; e++) {
      tv[e] = (double)ts.v[e];
      rr.add(r[e] * r[e]);
      ww.add(w[e] * w[e]);
      tt.add(tv[e] * tv[e]);
    }
    // the second reads must not be moved above the division
    uint64_t i2 = i;
    asm volatile("" : "+v"(i2));
#pragma unroll
    for (int j = MAXK - 1; j >= 0; j--) {
      if (j < k) {
        Chunk<VT, W> cv, hv;
        bool staged = false;
        if constexpr (W == chunk_of<VT>() && !kHold) {
          if (j < kStage) {
            cv = stage[2 * j][threadIdx.x];
            hv = stage[2 * j + 1][threadIdx.x];
            staged = true;
          }
        }
        if (!staged) {
          cv = ldc<W>(C.p[j], i2);
          hv = ldc<W>(HC.p[j], i2);
        }
#pragma unroll
        for (int e = 0; e < W; e++) {
          ot[j] += (double)cv.v[e] * tv[e];
          oh[j] += (double)hv.v[e] * tv[e];
        }
      }
    }
  };
  // wout (the last step of a restart cycle, where t has no slot in the
  // subspace): the pass stores the restart vector w = sum u_j C_j in t's place
  // and sums |w|^2 of the stored values in the place of |t|^2, exactly as
  // k_lincomb forms and sums it (same chunks, same order), with |r|^2 and
  // |w|^2 as in the full pass; the overlaps are not formed
  auto restart = [&]<int W>(uint64_t i) {
    double w[W], r[W];
#pragma unroll
    for (int j = 0; j < MAXK; j++) {
      if (j < k) {
        const Chunk<VT, W> cv = ldc<W>(C.p[j], i);
        const Chunk<VT, W> hv = ldc<W>(HC.p[j], i);
#pragma unroll
        for (int e = 0; e < W; e++) {
          if (j == 0) {
            w[e] = (double)cv.v[e] * u.c[0];
            r[e] = (double)hv.v[e] * (-1.0 * u.c[0]);
          } else {
            w[e] += (double)cv.v[e] * u.c[j];
            r[e] += (double)hv.v[e] * (-1.0 * u.c[j]);
          }
        }
      }
    }
    Chunk<VT, W> ws;
#pragma unroll
    for (int e = 0; e < W; e++) {
      r[e] += E0 * w[e];
      ws.v[e] = (VT)w[e];
    }
    stc<W>(t, i, ws);
#pragma unroll
    for (int e = 0; e < W; e++) {
      rr.add(r[e] * r[e]);
      ww.add(w[e] * w[e]);
      tt.add((double)ws.v[e] * (double)ws.v[e]);
    }
  };
  if (wout) for_chunks<VT>(g, n, packed, restart);
  else if constexpr (kHold) for_chunks<VT>(g, n, packed, hold);
  else for_chunks<VT>(g, n, packed, reread);
  double v[2 * MAXK + 3];
  v[0] = rr.s;
  v[1] = ww.s;
  v[2] = tt.s;
#pragma unroll
  for (int j = 0; j < MAXK; j++) {
    v[3 + j] = ot[j];
    v[3 + MAXK + j] = oh[j];
  }
  block_partials<2 * MAXK + 3>(g, v, 2 * MAXK + 3, part);
}

template <typename VT, int MAXK>
__global__ __launch_bounds__(kVecBlock) void k_ritz_precond(VecPtrs<VT> C, VecPtrs<VT> HC, int k,
                                                            Coefs u, double E0,
                                                            const VT* __restrict__ dii,
                                                            VT* __restrict__ t, uint64_t n,
                                                            bool packed,
                                                            double* __restrict__ part,
                                                            bool wout) {
  ritz_body<VT, MAXK, sizeof(VT) == 8 && MAXK <= 12>(launch_grid(), C, HC, k, u, E0, dii, t, n,
                                                      packed, part, wout);
}

// Classical Gram-Schmidt update t = scale * (t - sum o_j C_j); sums |t|^2.
template <typename VT, int MAXK>
__device__ inline void gs_body(Grid g, const VecPtrs<VT>& C, int k, const Coefs& o,
                               VT* __restrict__ t, double scale, uint64_t n, bool packed,
                               double* __restrict__ part) {
  Kahan tt;
  for_chunks<VT>(g, n, packed, [&]<int W>(uint64_t i) {
    const Chunk<VT, W> tv = ldc<W>(t, i);
    double x[W];
#pragma unroll
    for (int e = 0; e < W; e++) x[e] = (double)tv.v[e];
#pragma unroll
    for (int j = 0; j < MAXK; j++) {
      if (j < k) {
        const Chunk<VT, W> cv = ldc<W>(C.p[j], i);
#pragma unroll
        for (int e = 0; e < W; e++) x[e] -= o.c[j] * (double)cv.v[e];
      }
    }
    Chunk<VT, W> xs;
#pragma unroll
    for (int e = 0; e < W; e++) xs.v[e] = (VT)(x[e] * scale);
    stc<W>(t, i, xs);
#pragma unroll
    for (int e = 0; e < W; e++) tt.add((double)xs.v[e] * (double)xs.v[e]);
  });
  double v[1] = {tt.s};
  block_partials<1>(g, v, 1, part);
}

template <typename VT, int MAXK>
__global__ __launch_bounds__(kVecBlock) void k_gs_update(VecPtrs<VT> C, int k, Coefs o,
                                                         VT* __restrict__ t, double scale,
                                                         uint64_t n, bool packed,
                                                         double* __restrict__ part) {
  gs_body<VT, MAXK>(launch_grid(), C, k, o, t, scale, n, packed, part);
}

// ---------------------------------------------------------------- batches
// Several jobs' passes in one launch: job j owns workgroups [blk0, blk0 + nblk)
// with nblk = the grid it would get alone, so its sums are the same as in a
// launch of its own.

template <typename VT>
struct RitzJob {
  VecPtrs<VT> C, HC;
  Coefs u;
  const VT* dii;
  VT* t;
  uint64_t n;
  double E0;
  int k, packed, wout;
  uint32_t blk0, nblk;
  double* part;
};
template <typename VT>
struct GsJob {
  VecPtrs<VT> C;
  Coefs o;
  VT* t;
  uint64_t n;
  double scale;
  int k, packed;
  uint32_t blk0, nblk;
  double* part;
};
struct FinJob {
  const double* part;
  double* out;
  int nblk, m;
};

template <typename J>
__device__ inline const J& find_job(const J* __restrict__ jobs, int njobs) {
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (jobs[mid].blk0 <= blockIdx.x) lo = mid;
    else hi = mid - 1;
  }
  return jobs[lo];
}

template <typename VT, int MAXK>
__global__ __launch_bounds__(kVecBlock) void k_ritz_batch(const RitzJob<VT>* __restrict__ jobs,
                                                          int njobs) {
  const RitzJob<VT>& J = find_job(jobs, njobs);
  ritz_body<VT, MAXK, false>(Grid{blockIdx.x - J.blk0, J.nblk}, J.C, J.HC, J.k, J.u, J.E0, J.dii, J.t,
                             J.n, J.packed != 0, J.part, J.wout != 0);
}

template <typename VT, int MAXK>
__global__ __launch_bounds__(kVecBlock) void k_gs_batch(const GsJob<VT>* __restrict__ jobs,
                                                        int njobs) {
  const GsJob<VT>& J = find_job(jobs, njobs);
  gs_body<VT, MAXK>(Grid{blockIdx.x - J.blk0, J.nblk}, J.C, J.k, J.o, J.t, J.scale, J.n,
                    J.packed != 0, J.part);
}

// finalize for several jobs: block (j, s) sums job j's partials of sum s with
// the arithmetic of k_finalize
__global__ __launch_bounds__(kVecBlock) void k_finalize_batch(const FinJob* __restrict__ jobs,
                                                              int maxm) {
  __shared__ double red[kVecBlock];
  const FinJob& J = jobs[blockIdx.x / maxm];
  const int j = blockIdx.x % maxm;
  if (j >= J.m) return;  // uniform over the block
  Kahan k;
  for (int b = threadIdx.x; b < J.nblk; b += kVecBlock) k.add(J.part[b * J.m + j]);
  red[threadIdx.x] = k.s;
  __syncthreads();
  for (int w = kVecBlock / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) J.out[j] = red[0];
}

// out = scale * sum u_j C_j; sums |out|^2.
template <typename VT, int MAXK>
__global__ __launch_bounds__(kVecBlock) void k_lincomb(VecPtrs<VT> C, int k, Coefs u,
                                                       VT* __restrict__ out, double scale,
                                                       uint64_t n, bool packed,
                                                       double* __restrict__ part) {
  Kahan tt;
  for_chunks<VT>(launch_grid(), n, packed, [&]<int W>(uint64_t i) {
    double x[W];
    const Chunk<VT, W> c0 = ldc<W>(C.p[0], i);
#pragma unroll
    for (int e = 0; e < W; e++) x[e] = (double)c0.v[e] * u.c[0];
#pragma unroll
    for (int j = 1; j < MAXK; j++) {
      if (j < k) {
        const Chunk<VT, W> cv = ldc<W>(C.p[j], i);
#pragma unroll
        for (int e = 0; e < W; e++) x[e] += (double)cv.v[e] * u.c[j];
      }
    }
    Chunk<VT, W> xs;
#pragma unroll
    for (int e = 0; e < W; e++) xs.v[e] = (VT)(x[e] * scale);
    stc<W>(out, i, xs);
#pragma unroll
    for (int e = 0; e < W; e++) tt.add((double)xs.v[e] * (double)xs.v[e]);
  });
  double v[1] = {tt.s};
  block_partials<1>(launch_grid(), v, 1, part);
}

template <typename VT>
__global__ void k_scale(VT* __restrict__ x, double s, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    x[i] = (VT)((double)x[i] * s);
}

// Start vectors on the device (BasisInitVector, davidson.h:144-184).
// Randomize (dm_vector.h:101-191): element i is uniform_pm1(elem_hash(seed, i)),
// computed where it is needed, so the squared norm is summed (compensated, per
// thread, fixed-order) without storing the vector, and the vector is then
// written already scaled by 1 / norm.
__global__ __launch_bounds__(kVecBlock) void k_random_norm2(uint64_t n, uint64_t seed,
                                                            double* __restrict__ part) {
  Kahan k;
  const Grid g = launch_grid();
  for (uint64_t i = g.bid * (uint64_t)kVecBlock + threadIdx.x; i < n;
       i += (uint64_t)g.nblk * kVecBlock) {
    const double w = uniform_pm1(elem_hash(seed, i));
    k.add(w * w);
  }
  double v[1] = {k.s};
  block_partials<1>(g, v, 1, part);
}

template <typename VT>
__global__ void k_random_fill(VT* __restrict__ x, uint64_t n, uint64_t seed, double factor) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    x[i] = (VT)(uniform_pm1(elem_hash(seed, i)) * factor);
}

template <typename VT>
__global__ void k_fill(VT* __restrict__ x, uint64_t n, VT v) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    x[i] = v;
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
  hipLaunchKernelGGL(k_finalize, dim3(m), dim3(kVecBlock), 0, st, ctx->d_partials, grid,
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

// compile-time sizes the kernels are instantiated for
static int size_class(int k) {
  return k <= 4 ? 4 : k <= 8 ? 8 : k <= 12 ? 12 : k <= 16 ? 16 : k <= 24 ? 24 : kMaxVecs;
}
#define SBD_DISPATCH_K(k, KERN, ...)                                                       \
  do {                                                                                     \
    switch (size_class(k)) {                                                               \
      case 4: hipLaunchKernelGGL((KERN<VT, 4>), __VA_ARGS__); break;                       \
      case 8: hipLaunchKernelGGL((KERN<VT, 8>), __VA_ARGS__); break;                       \
      case 12: hipLaunchKernelGGL((KERN<VT, 12>), __VA_ARGS__); break;                     \
      case 16: hipLaunchKernelGGL((KERN<VT, 16>), __VA_ARGS__); break;                     \
      case 24: hipLaunchKernelGGL((KERN<VT, 24>), __VA_ARGS__); break;                     \
      default: hipLaunchKernelGGL((KERN<VT, kMaxVecs>), __VA_ARGS__); break;               \
    }                                                                                      \
  } while (0)

// true when every pointer is 16-byte aligned (chunked loads allowed)
template <typename VT>
static bool aligned16(std::initializer_list<const VT*> ps) {
  for (const VT* p : ps)
    if (reinterpret_cast<uintptr_t>(p) % 16) return false;
  return true;
}
template <typename VT>
static bool aligned16(const VT* const* X, int m) {
  for (int j = 0; j < m; j++)
    if (reinterpret_cast<uintptr_t>(X[j]) % 16) return false;
  return true;
}

template <typename VT>
int vec_multidot(sbd_ctx* ctx, const VT* const* X, int m, const VT* y, uint64_t n, double* out,
                 hipStream_t st) {
  if (m < 1 || m > kMaxVecs) return SBD_ERR_ARG;
  int rc = ensure_scratch(ctx);
  if (rc) return rc;
  const int g = grid_for(n);
  const bool packed = aligned16(X, m) && aligned16<VT>({y});
  SBD_DISPATCH_K(m, k_multidot, dim3(g), dim3(kVecBlock), 0, st, pack(X, m), m, y, n, packed,
                 ctx->d_partials);
  SBD_HIP_TRY(ctx, hipGetLastError());
  return finish(ctx, g, m, out, st);
}

template <typename VT>
int vec_ritz_precond(sbd_ctx* ctx, const VT* const* C, const VT* const* HC, int k, const double* u,
                     double E0, const VT* dii, VT* t, uint64_t n, double* out, hipStream_t st,
                     bool wout) {
  if (k < 1 || k > kMaxVecs) return SBD_ERR_ARG;
  int rc = ensure_scratch(ctx);
  if (rc) return rc;
  const int g = grid_for(n);
  const bool packed = aligned16(C, k) && aligned16(HC, k) && aligned16<VT>({dii, t});
  SBD_DISPATCH_K(k, k_ritz_precond, dim3(g), dim3(kVecBlock), 0, st, pack(C, k), pack(HC, k), k,
                 packc(u, k), E0, dii, t, n, packed, ctx->d_partials, wout);
  SBD_HIP_TRY(ctx, hipGetLastError());
  const int K = size_class(k);
  std::vector<double> res(2 * K + 3);
  rc = finish(ctx, g, 2 * K + 3, res.data(), st);
  if (rc) return rc;
  for (int j = 0; j < 3 + k; j++) out[j] = res[j];
  for (int j = 0; j < k; j++) out[3 + k + j] = res[3 + K + j];
  return SBD_OK;
}

template <typename VT>
int vec_gs_update(sbd_ctx* ctx, const VT* const* C, int k, const double* coef, VT* t, double scale,
                  uint64_t n, double* out_norm2, hipStream_t st) {
  if (k < 1 || k > kMaxVecs) return SBD_ERR_ARG;
  int rc = ensure_scratch(ctx);
  if (rc) return rc;
  const int g = grid_for(n);
  const bool packed = aligned16(C, k) && aligned16<VT>({t});
  SBD_DISPATCH_K(k, k_gs_update, dim3(g), dim3(kVecBlock), 0, st, pack(C, k), k, packc(coef, k), t,
                 scale, n, packed, ctx->d_partials);
  SBD_HIP_TRY(ctx, hipGetLastError());
  return finish(ctx, g, 1, out_norm2, st);
}

template <typename VT>
int vec_lincomb(sbd_ctx* ctx, const VT* const* C, int k, const double* u, double scale, VT* out,
                uint64_t n, double* out_norm2, hipStream_t st) {
  if (k < 1 || k > kMaxVecs) return SBD_ERR_ARG;
  int rc = ensure_scratch(ctx);
  if (rc) return rc;
  const int g = grid_for(n);
  const bool packed = aligned16(C, k) && aligned16<VT>({out});
  SBD_DISPATCH_K(k, k_lincomb, dim3(g), dim3(kVecBlock), 0, st, pack(C, k), k, packc(u, k), out,
                 scale, n, packed, ctx->d_partials);
  SBD_HIP_TRY(ctx, hipGetLastError());
  return finish(ctx, g, 1, out_norm2, st);
}

template <typename VT>
int vec_scale(sbd_ctx* ctx, VT* x, double s, uint64_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_scale<VT>, dim3(grid_for(n) * 2), dim3(kVecBlock), 0, st, x, s, n);
  SBD_HIP_TRY(ctx, hipGetLastError());
  return SBD_OK;
}

template <typename VT>
int vec_random(sbd_ctx* ctx, VT* x, uint64_t n, uint64_t seed, hipStream_t st) {
  int rc = ensure_scratch(ctx);
  if (rc) return rc;
  const int g = grid_for(n);
  hipLaunchKernelGGL(k_random_norm2, dim3(g), dim3(kVecBlock), 0, st, n, seed, ctx->d_partials);
  SBD_HIP_TRY(ctx, hipGetLastError());
  double n2 = 0.0;
  rc = finish(ctx, g, 1, &n2, st);
  if (rc) return rc;
  hipLaunchKernelGGL(k_random_fill<VT>, dim3(grid_for(n) * 2), dim3(kVecBlock), 0, st, x, n, seed,
                     1.0 / std::sqrt(n2));
  SBD_HIP_TRY(ctx, hipGetLastError());
  return SBD_OK;
}

template <typename VT>
int vec_fill(sbd_ctx* ctx, VT* x, uint64_t n, double v, hipStream_t st) {
  hipLaunchKernelGGL(k_fill<VT>, dim3(grid_for(n) * 2), dim3(kVecBlock), 0, st, x, n, (VT)v);
  SBD_HIP_TRY(ctx, hipGetLastError());
  return SBD_OK;
}

template <typename VS, typename VD>
int vec_convert(sbd_ctx* ctx, const VS* x, VD* y, uint64_t n, hipStream_t st) {
  hipLaunchKernelGGL((k_convert<VS, VD>), dim3(grid_for(n) * 2), dim3(kVecBlock), 0, st, x, y, n);
  SBD_HIP_TRY(ctx, hipGetLastError());
  return SBD_OK;
}

// ---------------------------------------------------------------- batch launchers

int batch_reserve(BatchScratch& sc, size_t job_bytes, size_t parts, size_t outs) {
  auto grow = [](void** p, size_t* cap, size_t need, bool pinned) -> bool {
    if (*cap >= need) return true;
    if (*p) {
      if (pinned) (void)hipHostFree(*p);
      else (void)hipFree(*p);
    }
    *p = nullptr;
    *cap = 0;
    const size_t n = need + need / 2 + 256;
    if ((pinned ? hipHostMalloc(p, n) : hipMalloc(p, n)) != hipSuccess) return false;
    *cap = n;
    return true;
  };
  if (!grow(&sc.d_jobs, &sc.jobs_cap, job_bytes, false) ||
      !grow((void**)&sc.d_part, &sc.part_cap, parts * sizeof(double), false) ||
      !grow((void**)&sc.d_out, &sc.out_cap, outs * sizeof(double), false) ||
      !grow((void**)&sc.h_out, &sc.hout_cap, outs * sizeof(double), true) ||
      !grow(&sc.d_fin, &sc.fin_cap, job_bytes, false))
    return SBD_ERR_MEMORY;
  return SBD_OK;
}

void batch_free(BatchScratch& sc) {
  if (sc.d_jobs) (void)hipFree(sc.d_jobs);
  if (sc.d_part) (void)hipFree(sc.d_part);
  if (sc.d_out) (void)hipFree(sc.d_out);
  if (sc.h_out) (void)hipHostFree(sc.h_out);
  if (sc.d_fin) (void)hipFree(sc.d_fin);
  sc = BatchScratch{};
}

// partials of every job, then their fixed-order sums, copied to the host
static int batch_finish(BatchScratch& sc, const std::vector<FinJob>& fin, int maxm,
                        size_t nout, hipStream_t st) {
  if (hipMemcpyAsync(sc.d_fin, fin.data(), sizeof(FinJob) * fin.size(), hipMemcpyHostToDevice,
                     st) != hipSuccess)
    return SBD_ERR_DEVICE;
  hipLaunchKernelGGL(k_finalize_batch, dim3((unsigned)(fin.size() * maxm)), dim3(kVecBlock), 0, st,
                     (const FinJob*)sc.d_fin, maxm);
  if (hipGetLastError() != hipSuccess) return SBD_ERR_DEVICE;
  if (hipMemcpyAsync(sc.h_out, sc.d_out, sizeof(double) * nout, hipMemcpyDeviceToHost, st) !=
      hipSuccess)
    return SBD_ERR_DEVICE;
  if (hipStreamSynchronize(st) != hipSuccess) return SBD_ERR_DEVICE;
  return SBD_OK;
}

template <typename VT>
int vec_ritz_batch(BatchScratch& sc, const RitzArgs<VT>* a, int njobs, double* const* out,
                   hipStream_t st) {
  if (njobs < 1) return SBD_OK;
  int maxk = 1;
  for (int j = 0; j < njobs; j++) {
    if (a[j].k < 1 || a[j].k > kMaxVecs) return SBD_ERR_ARG;
    maxk = std::max(maxk, a[j].k);
  }
  const int K = size_class(maxk), m = 2 * K + 3;
  std::vector<RitzJob<VT>> jobs(njobs);
  std::vector<FinJob> fin(njobs);
  size_t parts = 0;
  uint32_t blk = 0;
  for (int j = 0; j < njobs; j++) {
    const RitzArgs<VT>& x = a[j];
    RitzJob<VT>& J = jobs[j];
    J.C = pack(x.C, x.k);
    J.HC = pack(x.HC, x.k);
    J.u = packc(x.u, x.k);
    J.dii = x.dii;
    J.t = x.t;
    J.n = x.n;
    J.E0 = x.E0;
    J.k = x.k;
    J.packed = aligned16(x.C, x.k) && aligned16(x.HC, x.k) && aligned16<VT>({x.dii, x.t});
    J.wout = x.wout ? 1 : 0;
    J.nblk = (uint32_t)grid_for(x.n);
    J.blk0 = blk;
    blk += J.nblk;
    parts += (size_t)J.nblk * m;
  }
  int rc = batch_reserve(sc, std::max(sizeof(RitzJob<VT>), sizeof(FinJob)) * njobs, parts,
                         (size_t)njobs * m);
  if (rc) return rc;
  size_t off = 0;
  for (int j = 0; j < njobs; j++) {
    jobs[j].part = sc.d_part + off;
    fin[j] = FinJob{sc.d_part + off, sc.d_out + (size_t)j * m, (int)jobs[j].nblk, m};
    off += (size_t)jobs[j].nblk * m;
  }
  if (hipMemcpyAsync(sc.d_jobs, jobs.data(), sizeof(RitzJob<VT>) * njobs, hipMemcpyHostToDevice,
                     st) != hipSuccess)
    return SBD_ERR_DEVICE;
  const RitzJob<VT>* dj = (const RitzJob<VT>*)sc.d_jobs;
  switch (K) {
    case 4: hipLaunchKernelGGL((k_ritz_batch<VT, 4>), dim3(blk), dim3(kVecBlock), 0, st, dj, njobs); break;
    case 8: hipLaunchKernelGGL((k_ritz_batch<VT, 8>), dim3(blk), dim3(kVecBlock), 0, st, dj, njobs); break;
    case 12: hipLaunchKernelGGL((k_ritz_batch<VT, 12>), dim3(blk), dim3(kVecBlock), 0, st, dj, njobs); break;
    case 16: hipLaunchKernelGGL((k_ritz_batch<VT, 16>), dim3(blk), dim3(kVecBlock), 0, st, dj, njobs); break;
    case 24: hipLaunchKernelGGL((k_ritz_batch<VT, 24>), dim3(blk), dim3(kVecBlock), 0, st, dj, njobs); break;
    default: hipLaunchKernelGGL((k_ritz_batch<VT, kMaxVecs>), dim3(blk), dim3(kVecBlock), 0, st, dj, njobs); break;
  }
  if (hipGetLastError() != hipSuccess) return SBD_ERR_DEVICE;
  rc = batch_finish(sc, fin, m, (size_t)njobs * m, st);
  if (rc) return rc;
  for (int j = 0; j < njobs; j++) {
    const double* r = sc.h_out + (size_t)j * m;
    const int k = a[j].k;
    for (int q = 0; q < 3 + k; q++) out[j][q] = r[q];
    for (int q = 0; q < k; q++) out[j][3 + k + q] = r[3 + K + q];
  }
  return SBD_OK;
}

template <typename VT>
int vec_gs_batch(BatchScratch& sc, const GsArgs<VT>* a, int njobs, double* out_norm2,
                 hipStream_t st) {
  if (njobs < 1) return SBD_OK;
  int maxk = 1;
  for (int j = 0; j < njobs; j++) {
    if (a[j].k < 1 || a[j].k > kMaxVecs) return SBD_ERR_ARG;
    maxk = std::max(maxk, a[j].k);
  }
  const int K = size_class(maxk);
  std::vector<GsJob<VT>> jobs(njobs);
  std::vector<FinJob> fin(njobs);
  uint32_t blk = 0;
  for (int j = 0; j < njobs; j++) {
    const GsArgs<VT>& x = a[j];
    GsJob<VT>& J = jobs[j];
    J.C = pack(x.C, x.k);
    J.o = packc(x.o, x.k);
    J.t = x.t;
    J.n = x.n;
    J.scale = x.scale;
    J.k = x.k;
    J.packed = aligned16(x.C, x.k) && aligned16<VT>({x.t});
    J.nblk = (uint32_t)grid_for(x.n);
    J.blk0 = blk;
    blk += J.nblk;
  }
  int rc = batch_reserve(sc, std::max(sizeof(GsJob<VT>), sizeof(FinJob)) * njobs, blk, njobs);
  if (rc) return rc;
  size_t off = 0;
  for (int j = 0; j < njobs; j++) {
    jobs[j].part = sc.d_part + off;
    fin[j] = FinJob{sc.d_part + off, sc.d_out + j, (int)jobs[j].nblk, 1};
    off += jobs[j].nblk;
  }
  if (hipMemcpyAsync(sc.d_jobs, jobs.data(), sizeof(GsJob<VT>) * njobs, hipMemcpyHostToDevice,
                     st) != hipSuccess)
    return SBD_ERR_DEVICE;
  const GsJob<VT>* dj = (const GsJob<VT>*)sc.d_jobs;
  switch (K) {
    case 4: hipLaunchKernelGGL((k_gs_batch<VT, 4>), dim3(blk), dim3(kVecBlock), 0, st, dj, njobs); break;
    case 8: hipLaunchKernelGGL((k_gs_batch<VT, 8>), dim3(blk), dim3(kVecBlock), 0, st, dj, njobs); break;
    case 12: hipLaunchKernelGGL((k_gs_batch<VT, 12>), dim3(blk), dim3(kVecBlock), 0, st, dj, njobs); break;
    case 16: hipLaunchKernelGGL((k_gs_batch<VT, 16>), dim3(blk), dim3(kVecBlock), 0, st, dj, njobs); break;
    case 24: hipLaunchKernelGGL((k_gs_batch<VT, 24>), dim3(blk), dim3(kVecBlock), 0, st, dj, njobs); break;
    default: hipLaunchKernelGGL((k_gs_batch<VT, kMaxVecs>), dim3(blk), dim3(kVecBlock), 0, st, dj, njobs); break;
  }
  if (hipGetLastError() != hipSuccess) return SBD_ERR_DEVICE;
  rc = batch_finish(sc, fin, 1, njobs, st);
  if (rc) return rc;
  for (int j = 0; j < njobs; j++) out_norm2[j] = sc.h_out[j];
  return SBD_OK;
}

template int vec_ritz_batch<double>(BatchScratch&, const RitzArgs<double>*, int, double* const*,
                                    hipStream_t);
template int vec_gs_batch<double>(BatchScratch&, const GsArgs<double>*, int, double*,
                                  hipStream_t);

#define SBD_INST(VT)                                                                          \
  template int vec_multidot<VT>(sbd_ctx*, const VT* const*, int, const VT*, uint64_t, double*, \
                                hipStream_t);                                                 \
  template int vec_ritz_precond<VT>(sbd_ctx*, const VT* const*, const VT* const*, int,        \
                                    const double*, double, const VT*, VT*, uint64_t, double*,  \
                                    hipStream_t, bool);                                        \
  template int vec_gs_update<VT>(sbd_ctx*, const VT* const*, int, const double*, VT*, double,  \
                                 uint64_t, double*, hipStream_t);                             \
  template int vec_lincomb<VT>(sbd_ctx*, const VT* const*, int, const double*, double, VT*,    \
                               uint64_t, double*, hipStream_t);                               \
  template int vec_scale<VT>(sbd_ctx*, VT*, double, uint64_t, hipStream_t);                    \
  template int vec_random<VT>(sbd_ctx*, VT*, uint64_t, uint64_t, hipStream_t);                 \
  template int vec_fill<VT>(sbd_ctx*, VT*, uint64_t, double, hipStream_t);
SBD_INST(double)
SBD_INST(float)
template int vec_convert<float, double>(sbd_ctx*, const float*, double*, uint64_t, hipStream_t);
template int vec_convert<double, float>(sbd_ctx*, const double*, float*, uint64_t, hipStream_t);

}  // namespace sbdx

// ------------------------------------------- inner product without allocation
//
// sbd_inner_noalloc: the same kernels and grid as sbd_inner (k_multidot for
// one vector, then k_finalize), with the block partials and the result slot
// in a fixed device array instead of the context's lazily allocated buffers,
// for callers that cannot allocate device memory.  Every partial the
// finalize step reads is written by the same call first, so nothing needs
// clearing between calls.  All calls on one device share the array: calls
// that overlap on one device (other host threads, other streams) race.

namespace {

constexpr size_t kNoAllocDoubles = (size_t)sbdx::kVecGrid + 64;  // partials + result, padded
static __device__ __align__(256) double scratch[kNoAllocDoubles];

constexpr int kNoAllocDevices = 64;
std::mutex g_noalloc_mu;
double* g_noalloc_ptr[kNoAllocDevices] = {};  // device address of scratch, per device

// The current device's copy of scratch, looked up once per device with that
// device current (the array's host-side name is not a device address).
int noalloc_scratch(sbd_ctx* ctx, double** out) {
  int dev = -1;
  SBD_HIP_TRY(ctx, hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoAllocDevices)
    return sbdx::set_err(ctx, SBD_ERR_DEVICE, "sbd_inner_noalloc: device index beyond 63");
  std::lock_guard<std::mutex> g(g_noalloc_mu);
  if (!g_noalloc_ptr[dev]) {
    void* p = nullptr;
    SBD_HIP_TRY(ctx, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return sbdx::set_err(ctx, SBD_ERR_DEVICE, "sbd_inner_noalloc: no scratch address");
    g_noalloc_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_noalloc_ptr[dev];
  return SBD_OK;
}

}  // namespace

extern "C" int sbd_inner_noalloc(sbd_ctx* ctx, const double* x, const double* y, uint64_t n,
                                 double* out, void* stream) {
  using namespace sbdx;
  if (!ctx || !x || !y || !out) return SBD_ERR_ARG;
  const hipStream_t st = (hipStream_t)stream;
  const int g = grid_for(n);
  if ((size_t)g + 1 > kNoAllocDoubles)
    return set_err(ctx, SBD_ERR_ARG, "sbd_inner_noalloc: reduction exceeds the fixed scratch");
  double* part = nullptr;
  int rc = noalloc_scratch(ctx, &part);
  if (rc) return rc;
  double* res = part + g;
  const double* X[1] = {x};
  const bool packed = aligned16(X, 1) && aligned16<double>({y});
  hipLaunchKernelGGL((k_multidot<double, 4>), dim3(g), dim3(kVecBlock), 0, st, pack(X, 1), 1, y,
                     n, packed, part);
  SBD_HIP_TRY(ctx, hipGetLastError());
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kVecBlock), 0, st, part, g, 1, res);
  SBD_HIP_TRY(ctx, hipGetLastError());
  SBD_HIP_TRY(ctx, hipMemcpyAsync(out, res, sizeof(double), hipMemcpyDeviceToHost, st));
  SBD_HIP_TRY(ctx, hipStreamSynchronize(st));
  return SBD_OK;
}
