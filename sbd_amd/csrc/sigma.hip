// Diagonal and sigma (H * v) kernels for the tensor-product basis.
//
// Reference behaviour: makeQChamDiagTerms (chemistry/tpb/qcham.h:290-340) and
// the CPU mult (chemistry/tpb/mult.h:625-881), whose per-element summation
// order is followed exactly: diagonal, alpha singles, alpha doubles, beta
// singles, beta doubles, alpha-single x beta-single.  Matrix elements are the
// Slater-Condon rules of chemistry/basic/determinants.h:451-518 evaluated with
// the same operations, so FP64 results are bit-identical to the CPU path.
//
// Sigma layout: one workgroup per bra alpha row ia; threads stride over the
// beta columns ib so every global access to T, hii and W is coalesced along
// ib.  The bra row T[ia, :] is staged in LDS (when it fits) because all beta
// excitations of (ia, ib) land in that row.  Alpha-pair records are
// workgroup-uniform (scalar loads); beta-pair records are per column.  There
// are no atomics: each output element has one writer.
#include "sbd_internal.hpp"

namespace sbdx {

// ---------------------------------------------------------------- diagonal

constexpr int kDiagBlock = 256;

// ZeroExcite (determinants.h:451-474): for occupied spin orbitals i < j in
// ascending order: += I1(i,i); += Direct(i,j); -= Exchange(i,j) if same spin;
// finally + I0.  Direct/Exchange are looked up as L x L spatial tables, which
// hold the same values as the reference's 2L x 2L matrices.
template <typename OT>
__global__ __launch_bounds__(kDiagBlock) void k_hdiag(const uint64_t* __restrict__ a,
                                                      const uint64_t* __restrict__ b, uint64_t nb,
                                                      uint64_t N, int L, double I0,
                                                      const double* __restrict__ I1,
                                                      const double* __restrict__ J,
                                                      const double* __restrict__ K,
                                                      OT* __restrict__ hii) {
  extern __shared__ double lds[];
  double* sJ = lds;
  double* sK = lds + L * L;
  double* sd = lds + 2 * L * L;
  for (int t = threadIdx.x; t < L * L; t += blockDim.x) {
    sJ[t] = J[t];
    sK[t] = K[t];
  }
  for (int t = threadIdx.x; t < 2 * L; t += blockDim.x) sd[t] = I1[t * 2 * L + t];
  __syncthreads();
  for (uint64_t idx = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; idx < N;
       idx += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t ia = idx / nb, ib = idx - ia * nb;
    const uint64_t A = a[ia], B = b[ib];
    uint64_t D[2];
    D[0] = spread32(A) | (spread32(B) << 1);
    D[1] = spread32(A >> 32) | (spread32(B >> 32) << 1);
    double e = 0.0;
    for (int wi = 0; wi < 2; wi++) {
      uint64_t mi = D[wi];
      while (mi) {
        const int i = __ffsll((unsigned long long)mi) - 1 + 64 * wi;
        mi &= mi - 1;
        e += sd[i];
        const int pi = i >> 1;
        // j > i: rest of this word, then the next word
        uint64_t mj = mi;
        for (int wj = wi; wj < 2; wj++) {
          if (wj != wi) mj = D[wj];
          while (mj) {
            const int j = __ffsll((unsigned long long)mj) - 1 + 64 * wj;
            mj &= mj - 1;
            const int pj = j >> 1;
            e += sJ[pi * L + pj];
            if ((i & 1) == (j & 1)) e -= sK[pi * L + pj];
          }
        }
      }
    }
    hii[idx] = (OT)(e + I0);
  }
}

template <typename OT>
static int hdiag_impl(sbd_ctx* ctx, OT* hii, hipStream_t st) {
  const int L = ctx->norb;
  const size_t lds = sizeof(double) * (2 * L * L + 2 * L);
  uint64_t blocks = (ctx->N + kDiagBlock - 1) / kDiagBlock;
  if (blocks > (uint64_t)ctx->cu_count * 64) blocks = (uint64_t)ctx->cu_count * 64;
  if (blocks == 0) return SBD_OK;
  hipLaunchKernelGGL(k_hdiag<OT>, dim3((unsigned)blocks), dim3(kDiagBlock), lds, st, ctx->d_a,
                     ctx->d_b, ctx->nb, ctx->N, L, ctx->I0, ctx->d_I1, ctx->d_J, ctx->d_K, hii);
  SBD_HIP_TRY(ctx, hipGetLastError());
  return SBD_OK;
}

int launch_hdiag(sbd_ctx* ctx, double* hii, hipStream_t st) { return hdiag_impl(ctx, hii, st); }
int launch_hdiag_f32(sbd_ctx* ctx, float* hii, hipStream_t st) { return hdiag_impl(ctx, hii, st); }

// ---------------------------------------------------------------- sigma

template <typename VT>
struct SigmaArgs {
  const uint64_t* a;
  const uint64_t* b;
  uint64_t na, nb;
  int L;
  const VT* hii;
  const VT* T;
  VT* W;
  int accumulate;
  const uint32_t* as_off;
  const PairRec* as_rec;
  const uint32_t* ad_off;
  const PairRec* ad_rec;
  const uint32_t* bs_off;
  const PairRec* bs_rec;
  const uint32_t* bd_off;
  const PairRec* bd_rec;
  const double* G;   // [2][L][L][2L]
  const double* I2;  // packed ERI
};

// OneExcite energy (before the sign): val + sum over occupied spin orbitals j
// of the bra, ascending, of G[j].
__device__ inline double one_excite_sum(double val, const double* __restrict__ g, uint64_t D0,
                                        uint64_t D1) {
  double e = val;
  while (D0) {
    e += g[__ffsll((unsigned long long)D0) - 1];
    D0 &= D0 - 1;
  }
  while (D1) {
    e += g[__ffsll((unsigned long long)D1) - 1 + 64];
    D1 &= D1 - 1;
  }
  return e;
}

// number of determinant bits strictly between spin orbitals lo < hi
__device__ inline int popc_between(uint64_t D0, uint64_t D1, int lo, int hi) {
  const int s = lo + 1, e = hi;
  int c = 0;
  if (s < 64) c += __popcll(D0 & bits_range(s, e < 64 ? e : 64));
  if (e > 64) c += __popcll(D1 & bits_range(s > 64 ? s - 64 : 0, e - 64));
  return c;
}

// TwoExcite (determinants.h:500-518) on the full determinant.
__device__ inline double two_excite_full(uint64_t D0, uint64_t D1, int i, int j, int a, int b,
                                         const double* __restrict__ I2) {
  const int I = i < j ? i : j, J = i < j ? j : i;
  const int A = a < b ? a : b, B = a < b ? b : a;
  int par = popc_between(D0, D1, I < A ? I : A, I < A ? A : I);
  par += popc_between(D0, D1, J < B ? J : B, J < B ? B : J);
  double sgn = (par & 1) ? -1.0 : 1.0;
  if (A > J || B < I) sgn *= -1.0;
  return sgn * (eri_value(I2, A, I, B, J) - eri_value(I2, A, J, B, I));
}

constexpr int kSigBlock = 512;
constexpr int kBatch = 8;

template <typename VT, typename AT, bool kRowLds>
__global__ __launch_bounds__(kSigBlock) void k_sigma_rows(const SigmaArgs<VT> p) {
  extern __shared__ unsigned char lds_raw[];
  VT* trow = reinterpret_cast<VT*>(lds_raw);
  const uint64_t nb = p.nb;
  const int L = p.L;
  const int64_t GL = (int64_t)L * L * 2 * L;  // one spin's G block
  for (uint64_t ia = blockIdx.x; ia < p.na; ia += gridDim.x) {
    const VT* __restrict__ Trow = p.T + ia * nb;
    if (kRowLds) {
      __syncthreads();
      for (uint64_t ib = threadIdx.x; ib < nb; ib += blockDim.x) trow[ib] = Trow[ib];
      __syncthreads();
    }
    const uint64_t A = p.a[ia];
    const uint64_t A0 = spread32(A), A1 = spread32(A >> 32);
    const uint32_t as0 = p.as_off[ia], as1 = p.as_off[ia + 1];
    const uint32_t ad0 = p.ad_off[ia], ad1 = p.ad_off[ia + 1];
    for (uint64_t ib = threadIdx.x; ib < nb; ib += blockDim.x) {
      const uint64_t B = p.b[ib];
      const uint64_t D0 = A0 | (spread32(B) << 1);
      const uint64_t D1 = A1 | (spread32(B >> 32) << 1);
      const uint64_t idx = ia * nb + ib;
      AT acc = p.accumulate ? (AT)p.W[idx] : (AT)0;
      const VT tself = kRowLds ? trow[ib] : Trow[ib];
      acc += (AT)p.hii[idx] * (AT)tself;

      // alpha singles (task type 2, OneExcite)
      for (uint32_t q = as0; q < as1; q++) {
        const PairRec r = p.as_rec[q];
        const int x = code_x0(r.code), y = code_y0(r.code);
        const double* g = p.G + ((int64_t)x * L + y) * (2 * L);
        double e = one_excite_sum(r.val, g, D0, D1);
        const int par = code_par(r.code) ^ (__popcll(B & rec_other_mask(0, r.code, false)) & 1);
        if (par) e = -e;
        acc += (AT)e * (AT)p.T[(uint64_t)r.ket * nb + ib];
      }
      // alpha doubles (task type 2, TwoExcite): loads batched ahead of the
      // in-order accumulation
      for (uint32_t q0 = ad0; q0 < ad1; q0 += kBatch) {
        VT tv[kBatch];
        double ev[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; u++) {
          if (q0 + u < ad1) {
            const PairRec r = p.ad_rec[q0 + u];
            tv[u] = p.T[(uint64_t)r.ket * nb + ib];
            ev[u] = (__popcll(B & rec_other_mask(0, r.code, true)) & 1) ? -r.val : r.val;
          }
        }
#pragma unroll
        for (int u = 0; u < kBatch; u++)
          if (q0 + u < ad1) acc += (AT)ev[u] * (AT)tv[u];
      }
      // beta singles (task type 1, OneExcite)
      const uint32_t bs0 = p.bs_off[ib], bs1 = p.bs_off[ib + 1];
      for (uint32_t q = bs0; q < bs1; q++) {
        const PairRec r = p.bs_rec[q];
        const int x = code_x0(r.code), y = code_y0(r.code);
        const double* g = p.G + GL + ((int64_t)x * L + y) * (2 * L);
        double e = one_excite_sum(r.val, g, D0, D1);
        const int par = code_par(r.code) ^ (__popcll(A & rec_other_mask(1, r.code, false)) & 1);
        if (par) e = -e;
        const VT t = kRowLds ? trow[r.ket] : Trow[r.ket];
        acc += (AT)e * (AT)t;
      }
      // beta doubles (task type 1, TwoExcite)
      const uint32_t bd0 = p.bd_off[ib], bd1 = p.bd_off[ib + 1];
      for (uint32_t q0 = bd0; q0 < bd1; q0 += kBatch) {
        VT tv[kBatch];
        double ev[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; u++) {
          if (q0 + u < bd1) {
            const PairRec r = p.bd_rec[q0 + u];
            tv[u] = kRowLds ? trow[r.ket] : Trow[r.ket];
            ev[u] = (__popcll(A & rec_other_mask(1, r.code, true)) & 1) ? -r.val : r.val;
          }
        }
#pragma unroll
        for (int u = 0; u < kBatch; u++)
          if (q0 + u < bd1) acc += (AT)ev[u] * (AT)tv[u];
      }
      // alpha single x beta single (task type 0, TwoExcite(D, ax, bx, ay, by))
      if (bs1 > bs0) {
        for (uint32_t q = as0; q < as1; q++) {
          const PairRec ra = p.as_rec[q];
          const int ai = 2 * code_x0(ra.code), aa = 2 * code_y0(ra.code);
          const VT* __restrict__ Tk = p.T + (uint64_t)ra.ket * nb;
          for (uint32_t k = bs0; k < bs1; k++) {
            const PairRec rb = p.bs_rec[k];
            const int bi = 2 * code_x0(rb.code) + 1, ba = 2 * code_y0(rb.code) + 1;
            const double e = two_excite_full(D0, D1, ai, bi, aa, ba, p.I2);
            acc += (AT)e * (AT)Tk[rb.ket];
          }
        }
      }
      p.W[idx] = (VT)acc;
    }
  }
}

template <typename VT>
int launch_sigma(sbd_ctx* ctx, const VT* hii, const VT* T, VT* W, int accumulate, hipStream_t st) {
  if (!ctx->tables_built) {
    ctx->last_error = "sigma called before sbd_make_helpers";
    return SBD_ERR_STATE;
  }
  SigmaArgs<VT> p;
  p.a = ctx->d_a;
  p.b = ctx->d_b;
  p.na = ctx->na;
  p.nb = ctx->nb;
  p.L = ctx->norb;
  p.hii = hii;
  p.T = T;
  p.W = W;
  p.accumulate = accumulate;
  p.as_off = ctx->ta.s_off;
  p.as_rec = ctx->ta.s_rec;
  p.ad_off = ctx->ta.d_off;
  p.ad_rec = ctx->ta.d_rec;
  p.bs_off = ctx->tb.s_off;
  p.bs_rec = ctx->tb.s_rec;
  p.bd_off = ctx->tb.d_off;
  p.bd_rec = ctx->tb.d_rec;
  p.G = ctx->d_G;
  p.I2 = ctx->d_eri;
  if (ctx->na == 0 || ctx->nb == 0) return SBD_OK;
  const size_t row_bytes = sizeof(VT) * ctx->nb;
  unsigned grid = (unsigned)(ctx->na < 0x7FFFFFFFULL ? ctx->na : 0x7FFFFFFFULL);
  if (row_bytes <= (size_t)ctx->lds_max) {
    hipLaunchKernelGGL((k_sigma_rows<VT, VT, true>), dim3(grid), dim3(kSigBlock), row_bytes, st, p);
  } else {
    hipLaunchKernelGGL((k_sigma_rows<VT, VT, false>), dim3(grid), dim3(kSigBlock), 0, st, p);
  }
  SBD_HIP_TRY(ctx, hipGetLastError());
  return SBD_OK;
}

template int launch_sigma<double>(sbd_ctx*, const double*, const double*, double*, int,
                                  hipStream_t);
template int launch_sigma<float>(sbd_ctx*, const float*, const float*, float*, int, hipStream_t);

int sigma_set_lds_limit(sbd_ctx* ctx) {
  // allow the row-staged variants to use the device's full LDS
  int lds = 65536;
  SBD_HIP_TRY(ctx, hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock,
                                         ctx->device));
  (void)hipFuncSetAttribute((const void*)k_sigma_rows<double, double, true>,
                      hipFuncAttributeMaxDynamicSharedMemorySize, lds);
  (void)hipFuncSetAttribute((const void*)k_sigma_rows<float, float, true>,
                      hipFuncAttributeMaxDynamicSharedMemorySize, lds);
  (void)hipGetLastError();
  ctx->lds_max = lds;
  return SBD_OK;
}

}  // namespace sbdx
