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
#include <algorithm>
#include <type_traits>
#include <vector>

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
  const EllSlot64* bd_ell64;  // beta doubles, sliced-ELL order (FP64 slots)
  const EllSlot32* bd_ell32;  // the same slots for the FP32 sigma
  const uint32_t* bd_ell_ket; // kets of the slots (unpacked tables only)
  const uint32_t* bd_ell_off;
  const EllPos* bd_ell_pos;    // column and its doubles count at each block-sorted position
  const ColRec* b_rec;         // beta string and single-excitation range per column
  const double* Ea;  // alpha single elements [q][ib] (or null)
  const double* Eb;  // beta single elements [ia][q] (or null)
  uint32_t Sb;       // number of beta single records
  const double* G;    // [2][L][L][2L]
  const double* I2;   // packed ERI
  const double* Vab;  // [L][L][L][L] alpha-single x beta-single integral parts
  int skip_ab = 0;    // leave the alpha-single x beta-single terms to k_sigma_ab
};

template <typename VT>
struct SigmaJob {
  SigmaArgs<VT> p;
  uint32_t blk0, ngrp;  // first workgroup and number of row groups
  double* dot_part;     // ngrp * 16 partials of <T, W>
};

// OneExcite energy (before the sign): val + sum over occupied spin orbitals j
// of the bra, ascending, of G[j].  The gathers are issued kGBatch at a time and
// then added in ascending order, so the result is the sequential sum.
constexpr int kGBatch = 8;
__device__ inline double one_excite_sum(double val, const double* __restrict__ g, uint64_t D0,
                                        uint64_t D1) {
  double e = val;
#pragma unroll
  for (int w = 0; w < 2; w++) {
    uint64_t x = w ? D1 : D0;
    while (x) {
      double gv[kGBatch];
      int cnt = 0;
#pragma unroll
      for (int k = 0; k < kGBatch; k++) {
        if (x) {
          gv[k] = g[__ffsll((unsigned long long)x) - 1 + 64 * w];
          x &= x - 1;
          cnt = k + 1;
        }
      }
#pragma unroll
      for (int k = 0; k < kGBatch; k++)
        if (k < cnt) e += gv[k];
    }
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

// Alpha-single x beta-single element TwoExcite(D, 2xa, 2xb+1, 2ya, 2yb+1)
// (determinants.h:500-518): the parity of D over (I, A) and (J, B) times the
// tabulated integral part (which carries the order-dependent sign): the same
// operations as TwoExcite, so the same value bit for bit.
__device__ inline double ab_elem(uint64_t D0, uint64_t D1, int xa, int ya, int xb, int yb, int L,
                                 const double* __restrict__ Vab) {
  const int i = 2 * xa, j = 2 * xb + 1, a = 2 * ya, b = 2 * yb + 1;
  const int I = i < j ? i : j, J = i < j ? j : i;
  const int A = a < b ? a : b, B = a < b ? b : a;
  int par = popc_between(D0, D1, I < A ? I : A, I < A ? A : I);
  par += popc_between(D0, D1, J < B ? J : B, J < B ? B : J);
  const double v = Vab[(((int64_t)xa * L + ya) * L + xb) * L + yb];
  const double sgn = (par & 1) ? -1.0 : 1.0;
  return sgn * v;
}

// Single-excitation elements with their full sign (OneExcite, determinants.h:
// 476-498, on the determinant (alpha string A, beta string B)).
__device__ inline double alpha_single_elem(const PairRec& r, uint64_t A, uint64_t B, int L,
                                           const double* __restrict__ G) {
  const uint64_t D0 = spread32(A) | (spread32(B) << 1);
  const uint64_t D1 = spread32(A >> 32) | (spread32(B >> 32) << 1);
  const double* g = G + ((int64_t)code_x0(r.code) * L + code_y0(r.code)) * (2 * L);
  double e = one_excite_sum(r.val, g, D0, D1);
  if (code_par(r.code) ^ (__popcll(B & beta_mask_single(r.code)) & 1)) e = -e;
  return e;
}
__device__ inline double beta_single_elem(const PairRec& r, uint64_t A, uint64_t B, int L,
                                          const double* __restrict__ G) {
  const uint64_t D0 = spread32(A) | (spread32(B) << 1);
  const uint64_t D1 = spread32(A >> 32) | (spread32(B >> 32) << 1);
  const double* g =
      G + (int64_t)L * L * 2 * L + ((int64_t)code_x0(r.code) * L + code_y0(r.code)) * (2 * L);
  double e = one_excite_sum(r.val, g, D0, D1);
  if (code_par(r.code) ^ alpha_par_single(prefix_parity(A), r.code)) e = -e;
  return e;
}

// Ea[q * nb + ib] for every alpha single record q and beta string ib.
__global__ void k_single_alpha(const uint64_t* __restrict__ a, const uint64_t* __restrict__ b,
                               uint64_t nb, uint64_t total, int L,
                               const PairRec* __restrict__ rec, const uint32_t* __restrict__ bra,
                               const double* __restrict__ G, double* __restrict__ Ea) {
  for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < total;
       t += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t q = t / nb, ib = t - q * nb;
    Ea[t] = alpha_single_elem(rec[q], a[bra[q]], b[ib], L, G);
  }
}

// Eb[ia * Sb + q] for every alpha string ia and beta single record q.
__global__ void k_single_beta(const uint64_t* __restrict__ a, const uint64_t* __restrict__ b,
                              uint64_t Sb, uint64_t total, int L, const PairRec* __restrict__ rec,
                              const uint32_t* __restrict__ bra, const double* __restrict__ G,
                              double* __restrict__ Eb) {
  for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < total;
       t += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t ia = t / Sb, q = t - ia * Sb;
    Eb[t] = beta_single_elem(rec[q], a[ia], b[bra[q]], L, G);
  }
}

constexpr int kSigBlock = kColBlock;  // 16 waves: one workgroup per CU when the row fills LDS
constexpr int kHalf = kSigBlock / 2;  // threads of each role (8 waves)
constexpr int kSigCols = kColBlock / kHalf;  // columns (positions) a thread carries per block
constexpr unsigned kSigMaxGrid = 65536;  // rows beyond this are taken grid-stride
constexpr int kBatch = 8;        // alpha pairs whose ket loads are issued together
constexpr int kABatch = 4;       // beta singles per alpha single fetched together (alpha-beta)
// beta double slots of one position fetched together: fewer when the partial
// sums of several rows share the registers
template <typename VT, int kRows>
constexpr int kBBatch = sizeof(VT) * kRows >= 16 ? 2 : sizeof(VT) * kRows >= 8 ? 5 : 8;

// Bra row staged in LDS.  A thread issues up to kStageLoads loads before its
// stores, so a 20,000-element row takes two memory round trips, not five.
constexpr int kStageLoads = 16;
template <typename VT>
__device__ inline void stage_row(VT* __restrict__ trow, const VT* __restrict__ Trow, uint32_t nb) {
  uint32_t j = threadIdx.x;
  for (; j + (kStageLoads - 1) * kSigBlock < nb; j += kStageLoads * kSigBlock) {
    VT v[kStageLoads];
#pragma unroll
    for (int u = 0; u < kStageLoads; u++) v[u] = Trow[j + u * kSigBlock];
#pragma unroll
    for (int u = 0; u < kStageLoads; u++) trow[j + u * kSigBlock] = v[u];
  }
  if (j < nb) {
    VT v[kStageLoads];
#pragma unroll
    for (int u = 0; u < kStageLoads; u++)
      if (j + u * kSigBlock < nb) v[u] = Trow[j + u * kSigBlock];
#pragma unroll
    for (int u = 0; u < kStageLoads; u++)
      if (j + u * kSigBlock < nb) trow[j + u * kSigBlock] = v[u];
  }
}

// v, or -v when the low bit of par is set (the sign bit flipped: the same
// value as negation)
__device__ inline double flip_sign(double v, uint32_t par) {
  return __hiloint2double(__double2hiint(v) ^ (int)(par << 31), __double2loint(v));
}
__device__ inline float flip_sign(float v, uint32_t par) {
  return __int_as_float(__float_as_int(v) ^ (int)(par << 31));
}

// Beta-double ELL slot as the kernel of element type VT reads it: value, alpha
// sign mask and ket (from the slot's upper bits when packed).
template <typename VT>
struct EllOf;
template <>
struct EllOf<double> {
  template <bool kPk>
  __device__ static inline void load(const SigmaArgs<double>& p, uint32_t i, uint32_t& ket,
                                     double& val, uint64_t& mask) {
    const EllSlot64 s = p.bd_ell64[i];
    val = s.val;
    mask = s.mk;
    ket = kPk ? (uint32_t)(s.mk >> kEllKetShift) : p.bd_ell_ket[i];
  }
};
template <>
struct EllOf<float> {
  template <bool kPk>
  __device__ static inline void load(const SigmaArgs<float>& p, uint32_t i, uint32_t& ket,
                                     float& val, uint64_t& mask) {
    const EllSlot32 s = p.bd_ell32[i];
    val = s.val;
    mask = ((uint64_t)s.mhi << 32) | s.mlo;
    ket = kPk ? (s.mhi >> (kEllKetShift - 32)) : p.bd_ell_ket[i];
  }
};

// Phase 1 of one 1024-column block for the rows of a group: thread t (of the
// first half) owns columns blk + t + 512 j (records cr[j], loaded ahead):
// diagonal, alpha singles, alpha doubles and beta singles, in that order for
// every row; the partial sums are stored to W.  The alpha records of the
// group's rows are contiguous in the CSR, so they are walked as one list (each
// record adds to its own row's sum, rows in order): the ket loads of all rows
// are batched together.  Every access to T, hii and W is a coalesced stream
// along the columns; the alpha records are workgroup-uniform.  Columns past
// the end are clamped to the last one (loads stay in bounds; nothing is
// stored).
template <typename VT, bool kRowLds, bool kCache, int kRows>
__device__ inline void sigma_phase1(const SigmaArgs<VT>& p, const VT* __restrict__ trows,
                                    const uint64_t* A, uint64_t ia0, int nr, uint32_t blk,
                                    uint32_t t, const ColRec* cr) {
  using AT = VT;
  constexpr int C = kSigCols;
  const uint32_t nb = (uint32_t)p.nb;
  const int L = p.L;
  uint32_t col[C];
  bool ok[C];
#pragma unroll
  for (int j = 0; j < C; j++) {
    const uint32_t c = blk + j * kHalf + t;
    ok[j] = c < nb;
    col[j] = ok[j] ? c : nb - 1;
  }
  // record ranges: row r's records end at as_end[r] / ad_end[r]
  uint32_t as_end[kRows], ad_end[kRows];
#pragma unroll
  for (int r = 0; r < kRows; r++) {
    const uint64_t ia = ia0 + (r < nr ? r : nr - 1);
    as_end[r] = p.as_off[ia + 1];
    ad_end[r] = p.ad_off[ia + 1];
  }
  AT acc[kRows][C];
#pragma unroll
  for (int r = 0; r < kRows; r++) {
    if (r >= nr) continue;
    const uint64_t ia = ia0 + r;
    const VT* __restrict__ Tsrc = kRowLds ? trows + (size_t)r * nb : p.T + ia * nb;
#pragma unroll
    for (int j = 0; j < C; j++) {
      const VT hv = p.hii[ia * nb + col[j]];
      const VT ts = Tsrc[col[j]];
      acc[r][j] = p.accumulate ? (AT)p.W[ia * nb + col[j]] : (AT)0;
      acc[r][j] += (AT)hv * (AT)ts;
    }
  }
  // alpha singles (task type 2, OneExcite), rows in order
  {
    const uint32_t q1 = as_end[nr - 1];
    int r = 0;
    for (uint32_t q = p.as_off[ia0]; q < q1; q++) {
      while (q >= as_end[r]) r++;  // uniform
      const PairRec rc = p.as_rec[q];
      const VT* __restrict__ Tk = p.T + (uint64_t)rc.ket * nb;
#pragma unroll
      for (int rr = 0; rr < kRows; rr++) {
        if (rr != r) continue;
#pragma unroll
        for (int j = 0; j < C; j++) {
          const VT tk = Tk[col[j]];
          const double e = kCache ? p.Ea[(uint64_t)q * nb + col[j]]
                                  : alpha_single_elem(rc, A[rr], cr[j].B, L, p.G);
          acc[rr][j] += (AT)e * (AT)tk;
        }
      }
    }
  }
  // alpha doubles (task type 2, TwoExcite): ket loads batched ahead over the
  // records of all rows
  {
    const uint32_t q1 = ad_end[nr - 1];
    for (uint32_t q0 = p.ad_off[ia0]; q0 < q1; q0 += kBatch) {
      VT tv[kBatch][C];
      double vv[kBatch];
      uint64_t mm[kBatch];
#pragma unroll
      for (int u = 0; u < kBatch; u++) {
        if (q0 + u < q1) {
          const PairRec rc = p.ad_rec[q0 + u];
          const VT* __restrict__ Tk = p.T + (uint64_t)rc.ket * nb;
#pragma unroll
          for (int j = 0; j < C; j++) tv[u][j] = Tk[col[j]];
          vv[u] = rc.val;
          mm[u] = beta_mask_double(rc.code);
        }
      }
#pragma unroll
      for (int u = 0; u < kBatch; u++) {
        if (q0 + u < q1) {
          int r = 0;
#pragma unroll
          for (int rr = 0; rr + 1 < kRows; rr++) r += q0 + u >= ad_end[rr] ? 1 : 0;
#pragma unroll
          for (int j = 0; j < C; j++) {
            const double e = flip_sign(vv[u], (uint32_t)__popcll(cr[j].B & mm[u]));
            const AT x = (AT)e * (AT)tv[u][j];
#pragma unroll
            for (int rr = 0; rr < kRows; rr++)
              if (rr == r) acc[rr][j] += x;
          }
        }
      }
    }
  }
  // beta singles (task type 1, OneExcite): rare, per column
#pragma unroll
  for (int r = 0; r < kRows; r++) {
    if (r >= nr) continue;
    const uint64_t ia = ia0 + r;
    const VT* __restrict__ Tsrc = kRowLds ? trows + (size_t)r * nb : p.T + ia * nb;
#pragma unroll
    for (int j = 0; j < C; j++) {
      const uint32_t s1 = ok[j] ? cr[j].s1 : cr[j].s0;
      for (uint32_t q = cr[j].s0; q < s1; q++) {
        const PairRec rc = p.bs_rec[q];
        const double e = kCache ? p.Eb[ia * (uint64_t)p.Sb + q]
                                : beta_single_elem(rc, A[r], cr[j].B, L, p.G);
        acc[r][j] += (AT)e * (AT)Tsrc[rc.ket];
      }
    }
#pragma unroll
    for (int j = 0; j < C; j++)
      if (ok[j]) p.W[ia * nb + col[j]] = (VT)acc[r][j];
  }
}

// Phase 2 of one 1024-column block: thread t (of the second half) owns the
// columns col[j] at block-sorted positions blk + t + 512 j (the block's columns
// in descending order of their beta-double count; columns loaded ahead),
// reloads their partial sums and adds the beta doubles and the alpha-single x
// beta-single terms of every row of the group.  Each wavefront takes 64
// positions of nearly equal count, so its ELL slices carry little padding;
// the slots of a slice sit at fixed strides from one base and are fetched
// together with the partial sums (their addresses depend on nothing loaded),
// once for all rows of the group.  Returns the thread's share of <T, W> when
// `dot` is set.
template <typename VT, bool kRowLds, int kRows, bool kPk>
__device__ inline double sigma_phase2(const SigmaArgs<VT>& p, const VT* __restrict__ trows,
                                      const uint64_t* A, uint64_t ia0, int nr, uint32_t blk,
                                      uint32_t t, const EllPos* pc, bool dot) {
  using AT = VT;
  constexpr int C = kSigCols;
  constexpr int BB = kBBatch<VT, kRows>;
  const uint32_t nb = (uint32_t)p.nb;
  const int L = p.L;
  double tw = 0.0;
  bool okp[C];
  uint32_t col[C], i0[C], w[C];
#pragma unroll
  for (int j = 0; j < C; j++) {
    const uint32_t pos = blk + j * kHalf + t;
    okp[j] = pos < nb;
    col[j] = pc[j].col;
    // the slice's offset and width are uniform over the wavefront (a position
    // past the end is clamped: its slice is then the last one)
    const uint32_t sl = __builtin_amdgcn_readfirstlane((okp[j] ? pos : nb - 1) / kEllSlice);
    const uint32_t o0 = p.bd_ell_off[sl];
    w[j] = (p.bd_ell_off[sl + 1] - o0) / kEllSlice;
    i0[j] = o0 + pos % kEllSlice;
  }
  AT acc[kRows][C];
#pragma unroll
  for (int r = 0; r < kRows; r++)
#pragma unroll
    for (int j = 0; j < C; j++)
      acc[r][j] = (r < nr && okp[j]) ? (AT)p.W[(ia0 + r) * nb + col[j]] : (AT)0;
  uint32_t wmax = 0;
#pragma unroll
  for (int j = 0; j < C; j++) wmax = w[j] > wmax ? w[j] : wmax;
  // beta doubles (task type 1, TwoExcite): slot k of position j at
  // i0[j] + 64 k for k below the slice width (past the position's own count
  // the slice holds zero slots).  A batch issues all its loads together;
  // slots past the width read the zero slot after the last slice, so every
  // product is added without a condition (an exact zero for those).
  const uint32_t zs = p.bd_ell_off[(nb + kEllSlice - 1) / kEllSlice];
  for (uint32_t k0 = 0; k0 < wmax; k0 += BB) {
    uint32_t kk[C][BB];
    VT vv[C][BB];
    uint64_t mm[C][BB];
#pragma unroll
    for (int j = 0; j < C; j++)
#pragma unroll
      for (int u = 0; u < BB; u++)
        EllOf<VT>::template load<kPk>(p, k0 + u < w[j] ? i0[j] + (k0 + u) * kEllSlice : zs,
                                      kk[j][u], vv[j][u], mm[j][u]);
#pragma unroll
    for (int r = 0; r < kRows; r++) {
      if (r >= nr) continue;
      const VT* __restrict__ Tsrc = kRowLds ? trows + (size_t)r * nb : p.T + (ia0 + r) * nb;
#pragma unroll
      for (int j = 0; j < C; j++) {
        VT tt[BB];
#pragma unroll
        for (int u = 0; u < BB; u++) tt[u] = Tsrc[kk[j][u]];
#pragma unroll
        for (int u = 0; u < BB; u++) {
          const VT e = flip_sign(vv[j][u], (uint32_t)__popcll(A[r] & mm[j][u]));
          acc[r][j] += e * tt[u];
        }
      }
    }
  }
#pragma unroll
  for (int r = 0; r < kRows; r++) {
    if (r >= nr) continue;
    const uint64_t ia = ia0 + r;
    const uint32_t as0 = p.as_off[ia], as1 = p.as_off[ia + 1];
#pragma unroll
    for (int j = 0; j < C; j++) {
      if (!okp[j]) continue;
      // alpha single x beta single (task type 0, TwoExcite(D, ax, bx, ay, by)):
      // only rows with alpha singles (uniform) look at the column's singles
      if (!p.skip_ab && as1 > as0) {
        const ColRec cb = p.b_rec[col[j]];
        if (cb.s1 > cb.s0) {
          const uint64_t D0 = spread32(A[r]) | (spread32(cb.B) << 1);
          const uint64_t D1 = spread32(A[r] >> 32) | (spread32(cb.B >> 32) << 1);
          for (uint32_t q = as0; q < as1; q++) {
            const PairRec ra = p.as_rec[q];
            const int xa = code_x0(ra.code), ya = code_y0(ra.code);
            const VT* __restrict__ Tk = p.T + (uint64_t)ra.ket * nb;
            // beta records kABatch at a time: their ket and integral loads are
            // issued together, the sums are added in order (slots past the
            // end repeat the last record and are discarded)
            for (uint32_t k0 = cb.s0; k0 < cb.s1; k0 += kABatch) {
              uint32_t code[kABatch];
              VT tv[kABatch];
#pragma unroll
              for (int u = 0; u < kABatch; u++) {
                const PairRec rb = p.bs_rec[k0 + u < cb.s1 ? k0 + u : cb.s1 - 1];
                code[u] = rb.code;
                tv[u] = Tk[rb.ket];
              }
#pragma unroll
              for (int u = 0; u < kABatch; u++) {
                const double e =
                    ab_elem(D0, D1, xa, ya, code_x0(code[u]), code_y0(code[u]), L, p.Vab);
                const AT nv = acc[r][j] + (AT)e * (AT)tv[u];
                acc[r][j] = (k0 + u < cb.s1) ? nv : acc[r][j];
              }
            }
          }
        }
      }
      p.W[ia * nb + col[j]] = (VT)acc[r][j];
      if (dot) {
        const VT tc = kRowLds ? trows[(size_t)r * nb + col[j]] : p.T[ia * nb + col[j]];
        tw += (double)tc * (double)(VT)acc[r][j];
      }
    }
  }
  return tw;
}

// One workgroup per kRows consecutive bra alpha rows (grid-stride over row
// groups); the beta columns are taken in blocks of 1024, each in two phases
// (phase 1 and phase 2 above).  The two halves of the workgroup work as a
// pipeline: while the first eight wavefronts run phase 1 of block g (streams
// from HBM), the other eight run phase 2 of block g - 1 (ELL slots, mostly from
// L2, and gathers from the LDS rows); one barrier per block hands the partial
// sums over through W.  Each half loads the per-column (per-position) records
// of its next block before the barrier, so a phase starts with every address
// it needs.  Each element is accumulated in the reference order: diagonal,
// alpha singles, alpha doubles, beta singles (phase 1), beta doubles and
// alpha-single x beta-single terms (phase 2).  The bra rows are staged in LDS
// (when they fit) because every beta excitation of (ia, ib) lands in them.
// Every pointer is a restrict-qualified kernel argument: W aliases no input,
// so the workgroup-uniform alpha records are fetched with scalar loads.
// Row groups bid, bid + ngrp, ... of one sigma (the workgroup's share);
// dot_part (optional) receives this workgroup's per-wavefront <T, W> sums at
// [bid * 16 + wave].
template <typename VT, bool kRowLds, bool kCache, int kRows, bool kPk>
__device__ inline void sigma_rows_body(const SigmaArgs<VT>& p, uint32_t bid, uint32_t ngrp,
                                       double* __restrict__ dot_part) {
  static_assert(kRowLds || kRows == 1, "row groups need the rows in LDS");
  static_assert(kSigCols * kHalf == kColBlock, "each half covers one block");
  constexpr int C = kSigCols;
  extern __shared__ unsigned char lds_raw[];
  VT* trows = reinterpret_cast<VT*>(lds_raw);
  const uint32_t nb = (uint32_t)p.nb;
  const uint32_t nblk = (nb + kColBlock - 1) / kColBlock;
  const bool streamer = threadIdx.x < kHalf;  // uniform over each wavefront
  const uint32_t t = streamer ? threadIdx.x : threadIdx.x - kHalf;
  double tw = 0.0;  // sum of T * W over this thread's outputs (when dot_part is set)
  for (uint64_t ia0 = (uint64_t)bid * kRows; ia0 < p.na; ia0 += (uint64_t)ngrp * kRows) {
    const int nr = (int)(p.na - ia0 < (uint64_t)kRows ? p.na - ia0 : kRows);  // uniform
    __syncthreads();
    if (kRowLds) {
#pragma unroll
      for (int r = 0; r < kRows; r++)
        if (r < nr) stage_row<VT>(trows + (size_t)r * nb, p.T + (ia0 + r) * nb, nb);
      __syncthreads();
    }
    uint64_t A[kRows];
#pragma unroll
    for (int r = 0; r < kRows; r++) A[r] = r < nr ? p.a[ia0 + r] : 0;
    // records of block 0 (columns for phase 1, block-sorted columns for phase 2)
    ColRec cr[C];
    EllPos pc[C];
#pragma unroll
    for (int j = 0; j < C; j++) {
      const uint32_t i = j * kHalf + t < nb ? j * kHalf + t : nb - 1;
      if (streamer) cr[j] = p.b_rec[i];
      else pc[j] = p.bd_ell_pos[i];
    }
    for (uint32_t g = 0; g <= nblk; g++) {
      if (streamer) {
        if (g < nblk && g * kColBlock + t < nb) {
          sigma_phase1<VT, kRowLds, kCache, kRows>(p, trows, A, ia0, nr, g * kColBlock, t, cr);
          if (g + 1 < nblk) {
#pragma unroll
            for (int j = 0; j < C; j++) {
              const uint32_t c = (g + 1) * kColBlock + j * kHalf + t;
              cr[j] = p.b_rec[c < nb ? c : nb - 1];
            }
          }
        }
      } else if (g > 0 && (g - 1) * kColBlock + t < nb) {
        tw += sigma_phase2<VT, kRowLds, kRows, kPk>(p, trows, A, ia0, nr, (g - 1) * kColBlock,
                                                    t, pc, dot_part != nullptr);
        if (g < nblk) {
#pragma unroll
          for (int j = 0; j < C; j++) {
            const uint32_t i = g * kColBlock + j * kHalf + t;
            pc[j] = p.bd_ell_pos[i < nb ? i : nb - 1];
          }
        }
      }
      __syncthreads();
    }
  }
  if (dot_part) {
    for (int o = 32; o > 0; o >>= 1) tw += __shfl_down(tw, o, 64);
    if ((threadIdx.x & 63) == 0) dot_part[bid * (kSigBlock / 64) + threadIdx.x / 64] = tw;
  }
}

template <typename VT, bool kRowLds, bool kCache, int kRows, bool kPk>
__global__ __launch_bounds__(kSigBlock) void k_sigma_rows(
    const uint64_t* __restrict__ a_, const uint64_t* __restrict__ b_, uint64_t na_, uint64_t nb_,
    int L_, const VT* __restrict__ hii_, const VT* __restrict__ T_, VT* __restrict__ W_,
    int accumulate_, const uint32_t* __restrict__ as_off_, const PairRec* __restrict__ as_rec_,
    const uint32_t* __restrict__ ad_off_, const PairRec* __restrict__ ad_rec_,
    const uint32_t* __restrict__ bs_off_, const PairRec* __restrict__ bs_rec_,
    const uint32_t* __restrict__ bd_off_, const EllSlot64* __restrict__ bd_ell64_,
    const EllSlot32* __restrict__ bd_ell32_, const uint32_t* __restrict__ bd_ell_ket_,
    const uint32_t* __restrict__ bd_ell_off_, const EllPos* __restrict__ bd_ell_pos_,
    const ColRec* __restrict__ b_rec_, const double* __restrict__ Ea_, const double* __restrict__ Eb_, uint32_t Sb_,
    const double* __restrict__ G_, const double* __restrict__ I2_,
    const double* __restrict__ Vab_, int skip_ab_, double* __restrict__ dot_part) {
  const SigmaArgs<VT> p{a_,       b_,          na_,         nb_,         L_,        hii_,
                        T_,       W_,          accumulate_, as_off_,     as_rec_,   ad_off_,
                        ad_rec_,  bs_off_,     bs_rec_,     bd_off_,     bd_ell64_, bd_ell32_,
                        bd_ell_ket_, bd_ell_off_, bd_ell_pos_, b_rec_,  Ea_,       Eb_,
                        Sb_,      G_,          I2_,         Vab_,      skip_ab_};
  sigma_rows_body<VT, kRowLds, kCache, kRows, kPk>(p, blockIdx.x, gridDim.x, dot_part);
}

// Several sigmas in one launch (the serving engine's batch): job j owns
// workgroups [blk0, blk0 + ngrp) and each of them one row group, so a job's
// results, including its <T, W> partials, do not depend on the other jobs.
// Jobs here have packed tables, single-excitation caches and rows in LDS.
template <typename VT, int kRows>
__global__ __launch_bounds__(kSigBlock) void k_sigma_batch(const SigmaJob<VT>* __restrict__ jobs,
                                                           int njobs) {
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (jobs[mid].blk0 <= blockIdx.x) lo = mid;
    else hi = mid - 1;
  }
  // the job's arguments as a uniform local copy (kept in scalar registers like
  // a kernel's own arguments; read through a reference into the job table they
  // were held in vector registers and spilled)
  lo = __builtin_amdgcn_readfirstlane(lo);
  const SigmaArgs<VT> p = jobs[lo].p;
  const uint32_t blk0 = jobs[lo].blk0, ngrp = jobs[lo].ngrp;
  double* const dot_part = jobs[lo].dot_part;
  sigma_rows_body<VT, true, true, kRows, true>(p, blockIdx.x - blk0, ngrp, dot_part);
}

// ---------------------------------------------------------------- split sigma
//
// The row kernel sweeps whole rows, so the alpha ket rows T[ja, :] that a bra
// row reads come back from HBM once per pair: seven times per element at the
// bench's size, with no reuse in between (the alpha connectivity of random
// strings has no locality).  The split sigma cuts the same sum at the same
// point into two kernels:
//
//   k_sigma_stream  diagonal, alpha singles, alpha doubles, for tiles of one
//                   bra row x kStreamBlock * J consecutive columns, taken
//                   slab-major (every row of one column slab, then the next
//                   slab).  The rows of T touched while a slab is in flight
//                   are that slab's columns only (nα x slab width elements),
//                   so the seven ket reads of an element find their lines in
//                   the memory-side cache instead of HBM;
//   k_sigma_bd      beta singles, beta doubles and alpha-single x beta-single
//                   terms (sigma_phase2 with all sixteen wavefronts), bra rows
//                   staged in LDS, reading the partial sums the first kernel
//                   stored in W.
//
// Per element the terms are added in the same order with the same operations
// as in k_sigma_rows, so the results are the same bits.
constexpr int kStreamBlock = 256;

// tiles t0, t0 + stride, ... below ntiles of one sigma
template <typename VT, bool kCache, int J>
__device__ inline void sigma_stream_body(const SigmaArgs<VT>& p, uint64_t t0, uint64_t stride,
                                         uint64_t ntiles) {
  using AT = VT;
  constexpr int SW = kStreamBlock * J;
  const uint32_t nb = (uint32_t)p.nb;
  const int L = p.L;
  for (uint64_t tile = t0; tile < ntiles; tile += stride) {
    const uint64_t slab = tile / p.na;
    const uint64_t ia = tile - slab * p.na;
    const uint32_t base = (uint32_t)slab * SW;
    const uint64_t A = p.a[ia];
    uint32_t col[J];
    bool ok[J];
    uint64_t B[J];
    AT acc[J];
#pragma unroll
    for (int j = 0; j < J; j++) {
      const uint32_t c = base + threadIdx.x + j * kStreamBlock;
      ok[j] = c < nb;
      col[j] = ok[j] ? c : nb - 1;
    }
#pragma unroll
    for (int j = 0; j < J; j++) {
      B[j] = p.b[col[j]];
      const VT hv = p.hii[ia * nb + col[j]];
      const VT ts = p.T[ia * nb + col[j]];
      acc[j] = p.accumulate ? (AT)p.W[ia * nb + col[j]] : (AT)0;
      acc[j] += (AT)hv * (AT)ts;
    }
    // alpha singles (task type 2, OneExcite)
    const uint32_t as1 = p.as_off[ia + 1];
    for (uint32_t q = p.as_off[ia]; q < as1; q++) {
      const PairRec rc = p.as_rec[q];
      const VT* __restrict__ Tk = p.T + (uint64_t)rc.ket * nb;
#pragma unroll
      for (int j = 0; j < J; j++) {
        const VT tk = Tk[col[j]];
        const double e = kCache ? p.Ea[(uint64_t)q * nb + col[j]]
                                : alpha_single_elem(rc, A, B[j], L, p.G);
        acc[j] += (AT)e * (AT)tk;
      }
    }
    // alpha doubles (task type 2, TwoExcite), ket loads kBatch records deep
    const uint32_t q1 = p.ad_off[ia + 1];
    for (uint32_t q0 = p.ad_off[ia]; q0 < q1; q0 += kBatch) {
      VT tv[kBatch][J];
      double vv[kBatch];
      uint64_t mm[kBatch];
#pragma unroll
      for (int u = 0; u < kBatch; u++) {
        if (q0 + u < q1) {
          const PairRec rc = p.ad_rec[q0 + u];
          const VT* __restrict__ Tk = p.T + (uint64_t)rc.ket * nb;
#pragma unroll
          for (int j = 0; j < J; j++) tv[u][j] = Tk[col[j]];
          vv[u] = rc.val;
          mm[u] = beta_mask_double(rc.code);
        }
      }
#pragma unroll
      for (int u = 0; u < kBatch; u++) {
        if (q0 + u < q1) {
#pragma unroll
          for (int j = 0; j < J; j++) {
            const double e = flip_sign(vv[u], (uint32_t)__popcll(B[j] & mm[u]));
            acc[j] += (AT)e * (AT)tv[u][j];
          }
        }
      }
    }
#pragma unroll
    for (int j = 0; j < J; j++)
      if (ok[j]) p.W[ia * nb + col[j]] = (VT)acc[j];
  }
}

template <typename VT, bool kCache, int J>
__global__ __launch_bounds__(kStreamBlock) void k_sigma_stream(const SigmaArgs<VT> p,
                                                               uint64_t ntiles) {
  sigma_stream_body<VT, kCache, J>(p, blockIdx.x, gridDim.x, ntiles);
}

// Several sigmas' streaming parts in one launch (the serving engine's batch):
// job j owns workgroups [blk0, blk0 + ngrp) and takes its tiles grid-stride
// over them.
template <typename VT>
struct StreamJob {
  SigmaArgs<VT> p;
  uint32_t blk0, ngrp;
  uint64_t ntiles;
};
template <typename VT, int J>
__global__ __launch_bounds__(kStreamBlock) void k_sigma_stream_batch(
    const StreamJob<VT>* __restrict__ jobs, int njobs) {
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (jobs[mid].blk0 <= blockIdx.x) lo = mid;
    else hi = mid - 1;
  }
  const StreamJob<VT>& Jb = jobs[lo];
  sigma_stream_body<VT, true, J>(Jb.p, blockIdx.x - Jb.blk0, Jb.ngrp, Jb.ntiles);
}

// Second kernel of the split sigma: one workgroup per group of kRows bra rows
// (grid-stride), the rows staged in LDS, then blocks of 2048 block-sorted
// positions, thread t taking positions blk + t + 1024 j (j < 2; their records
// loaded one block ahead).  Per position and row a block adds to the partial
// sum from W the beta singles (records from L2, elements from the cache, T
// from LDS), the beta doubles from the ELL slices BB slots at a time (the
// first batch is issued before the singles, so the two latencies overlap) and
// the alpha-single x beta-single terms, then stores the element.  The
// strings with single excitations come first in each block's order, so few
// wavefronts wait for them.  dot_part (optional): per-wavefront <T, W> sums
// at [blockIdx * 16 + wave].
constexpr int kBdCols = 2;
constexpr uint32_t kBdSpan = kBdCols * kSigBlock;  // positions per step

template <typename VT, int kRows, bool kPk, bool kCache>
__device__ inline double bd_block(const SigmaArgs<VT>& p, const VT* __restrict__ trows,
                                  const uint64_t* A, uint64_t ia0, int nr, uint32_t blk,
                                  uint32_t t, const EllPos* pc, bool dot, uint32_t zs,
                                  uint32_t as0, uint32_t as1) {
  using AT = VT;
  constexpr int C = kBdCols;
  constexpr int BB = kBBatch<VT, kRows>;
  const uint32_t nb = (uint32_t)p.nb;
  const int L = p.L;
  double tw = 0.0;
  bool okp[C];
  uint32_t i0[C], w[C];
#pragma unroll
  for (int j = 0; j < C; j++) {
    const uint32_t pos = blk + j * kSigBlock + t;
    okp[j] = pos < nb;
    const uint32_t sl = __builtin_amdgcn_readfirstlane((okp[j] ? pos : nb - 1) / kEllSlice);
    const uint32_t o0 = p.bd_ell_off[sl];
    w[j] = (p.bd_ell_off[sl + 1] - o0) / kEllSlice;
    i0[j] = o0 + pos % kEllSlice;
  }
  uint32_t wmax = 0;
#pragma unroll
  for (int j = 0; j < C; j++) wmax = w[j] > wmax ? w[j] : wmax;
  // positions without beta terms keep their partial sums as they are in W
  bool has[C];
#pragma unroll
  for (int j = 0; j < C; j++) has[j] = okp[j] && (pc[j].cnt != 0 || pc[j].s1 > pc[j].s0);
  AT acc[kRows][C];
#pragma unroll
  for (int r = 0; r < kRows; r++)
#pragma unroll
    for (int j = 0; j < C; j++)
      acc[r][j] = (r < nr && has[j]) ? (AT)p.W[(ia0 + r) * nb + pc[j].col] : (AT)0;
  // first batch of slots: issued now, added after the singles
  uint32_t kk[C][BB];
  VT vv[C][BB];
  uint64_t mm[C][BB];
#pragma unroll
  for (int j = 0; j < C; j++)
#pragma unroll
    for (int u = 0; u < BB; u++)
      EllOf<VT>::template load<kPk>(p, (uint32_t)u < w[j] ? i0[j] + u * kEllSlice : zs, kk[j][u],
                                    vv[j][u], mm[j][u]);
  // beta singles (task type 1, OneExcite)
#pragma unroll
  for (int j = 0; j < C; j++) {
    const uint32_t s1 = okp[j] ? pc[j].s1 : pc[j].s0;
    for (uint32_t q = pc[j].s0; q < s1; q++) {
      const PairRec rc = p.bs_rec[q];
#pragma unroll
      for (int r = 0; r < kRows; r++) {
        if (r >= nr) continue;
        const double e = kCache ? p.Eb[(ia0 + r) * (uint64_t)p.Sb + q]
                                : beta_single_elem(rc, A[r], p.b[pc[j].col], L, p.G);
        acc[r][j] += (AT)e * (AT)trows[(size_t)r * nb + rc.ket];
      }
    }
  }
  // beta doubles (task type 1, TwoExcite), in slot order; slots past a
  // slice's width read the zero slot
  for (uint32_t k0 = 0; k0 < wmax; k0 += BB) {
    if (k0 > 0) {
#pragma unroll
      for (int j = 0; j < C; j++)
#pragma unroll
        for (int u = 0; u < BB; u++)
          EllOf<VT>::template load<kPk>(p, k0 + u < w[j] ? i0[j] + (k0 + u) * kEllSlice : zs,
                                        kk[j][u], vv[j][u], mm[j][u]);
    }
#pragma unroll
    for (int r = 0; r < kRows; r++) {
      if (r >= nr) continue;
#pragma unroll
      for (int j = 0; j < C; j++) {
        VT tt[BB];
#pragma unroll
        for (int u = 0; u < BB; u++) tt[u] = trows[(size_t)r * nb + kk[j][u]];
#pragma unroll
        for (int u = 0; u < BB; u++) {
          const VT e = flip_sign(vv[j][u], (uint32_t)__popcll(A[r] & mm[j][u]));
          acc[r][j] += e * tt[u];
        }
      }
    }
  }
#pragma unroll
  for (int r = 0; r < kRows; r++) {
    if (r >= nr) continue;
    const uint64_t ia = ia0 + r;
    const uint32_t a0 = __shfl(as0, r, 64), a1 = __shfl(as1, r, 64);
#pragma unroll
    for (int j = 0; j < C; j++) {
      if (!okp[j]) continue;
      // alpha single x beta single (task type 0): rows with alpha singles
      // (uniform) and strings with beta singles
      if (!p.skip_ab && a1 > a0 && pc[j].s1 > pc[j].s0) {
        const uint64_t B = p.b[pc[j].col];
        const uint64_t D0 = spread32(A[r]) | (spread32(B) << 1);
        const uint64_t D1 = spread32(A[r] >> 32) | (spread32(B >> 32) << 1);
        for (uint32_t q = a0; q < a1; q++) {
          const PairRec ra = p.as_rec[q];
          const int xa = code_x0(ra.code), ya = code_y0(ra.code);
          const VT* __restrict__ Tk = p.T + (uint64_t)ra.ket * nb;
          for (uint32_t k0 = pc[j].s0; k0 < pc[j].s1; k0 += kABatch) {
            uint32_t code[kABatch];
            VT tv[kABatch];
#pragma unroll
            for (int u = 0; u < kABatch; u++) {
              const PairRec rb = p.bs_rec[k0 + u < pc[j].s1 ? k0 + u : pc[j].s1 - 1];
              code[u] = rb.code;
              tv[u] = Tk[rb.ket];
            }
#pragma unroll
            for (int u = 0; u < kABatch; u++) {
              const double e =
                  ab_elem(D0, D1, xa, ya, code_x0(code[u]), code_y0(code[u]), L, p.Vab);
              const AT nv = acc[r][j] + (AT)e * (AT)tv[u];
              acc[r][j] = (k0 + u < pc[j].s1) ? nv : acc[r][j];
            }
          }
        }
      }
      if (has[j]) p.W[ia * nb + pc[j].col] = (VT)acc[r][j];
      else if (dot) acc[r][j] = (AT)p.W[ia * nb + pc[j].col];
      if (dot) tw += (double)trows[(size_t)r * nb + pc[j].col] * (double)(VT)acc[r][j];
    }
  }
  return tw;
}

// row groups bid, bid + ngrp, ... of one sigma; dot_part (optional) receives
// this workgroup's per-wavefront <T, W> sums at [bid * 16 + wave]
template <typename VT, int kRows, bool kPk, bool kCache>
__device__ inline void sigma_bd_body(const SigmaArgs<VT>& p, uint32_t bid, uint32_t ngrp,
                                     double* __restrict__ dot_part) {
  constexpr int C = kBdCols;
  extern __shared__ unsigned char lds_raw[];
  VT* trows = reinterpret_cast<VT*>(lds_raw);
  const uint32_t nb = (uint32_t)p.nb;
  const uint32_t t = threadIdx.x;
  const uint32_t zs = p.bd_ell_off[(nb + kEllSlice - 1) / kEllSlice];
  double tw = 0.0;
  for (uint64_t ia0 = (uint64_t)bid * kRows; ia0 < p.na; ia0 += (uint64_t)ngrp * kRows) {
    const int nr = (int)(p.na - ia0 < (uint64_t)kRows ? p.na - ia0 : kRows);  // uniform
    __syncthreads();  // the previous group's rows are no longer read
#pragma unroll
    for (int r = 0; r < kRows; r++)
      if (r < nr) stage_row<VT>(trows + (size_t)r * nb, p.T + (ia0 + r) * nb, nb);
    uint64_t A[kRows];
#pragma unroll
    for (int r = 0; r < kRows; r++) A[r] = r < nr ? p.a[ia0 + r] : 0;
    // alpha single ranges of the rows: lane r holds row r's (read with __shfl)
    const uint32_t lr = (uint32_t)(threadIdx.x & 63) < (uint32_t)nr ? (threadIdx.x & 63) : 0;
    const uint32_t as0 = p.as_off[ia0 + lr], as1 = p.as_off[ia0 + lr + 1];
    EllPos pc[C];
#pragma unroll
    for (int j = 0; j < C; j++) {
      const uint32_t i = j * kSigBlock + t;
      pc[j] = p.bd_ell_pos[i < nb ? i : nb - 1];
    }
    __syncthreads();
    for (uint32_t blk = 0; blk < nb; blk += kBdSpan) {
      tw += bd_block<VT, kRows, kPk, kCache>(p, trows, A, ia0, nr, blk, t, pc,
                                             dot_part != nullptr, zs, as0, as1);
      if (blk + kBdSpan < nb) {
#pragma unroll
        for (int j = 0; j < C; j++) {
          const uint32_t i = blk + kBdSpan + j * kSigBlock + t;
          pc[j] = p.bd_ell_pos[i < nb ? i : nb - 1];
        }
      }
    }
  }
  if (dot_part) {
    for (int o = 32; o > 0; o >>= 1) tw += __shfl_down(tw, o, 64);
    if ((threadIdx.x & 63) == 0) dot_part[bid * (kSigBlock / 64) + threadIdx.x / 64] = tw;
  }
}

template <typename VT, int kRows, bool kPk, bool kCache>
__global__ __launch_bounds__(kSigBlock) void k_sigma_bd(const SigmaArgs<VT> p,
                                                       double* __restrict__ dot_part) {
  sigma_bd_body<VT, kRows, kPk, kCache>(p, blockIdx.x, gridDim.x, dot_part);
}

// Several sigmas' second parts in one launch: job j owns workgroups
// [blk0, blk0 + ngrp), the grid its sigma has alone, so its results and
// <T, W> partials are those of its own launch.  Packed, cached tables.
template <typename VT, int kRows>
__global__ __launch_bounds__(kSigBlock) void k_sigma_bd_batch(const SigmaJob<VT>* __restrict__ jobs,
                                                             int njobs) {
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (jobs[mid].blk0 <= blockIdx.x) lo = mid;
    else hi = mid - 1;
  }
  // a uniform local copy of the job's arguments, as in k_sigma_batch
  lo = __builtin_amdgcn_readfirstlane(lo);
  const SigmaArgs<VT> p = jobs[lo].p;
  const uint32_t blk0 = jobs[lo].blk0, ngrp = jobs[lo].ngrp;
  double* const dot_part = jobs[lo].dot_part;
  sigma_bd_body<VT, kRows, true, true>(p, blockIdx.x - blk0, ngrp, dot_part);
}

// ---------------------------------------------------------------- dense alpha x beta
//
// In densely connected bases (tens of single excitations per string, as in
// the reference's molecular inputs) the alpha-single x beta-single terms are
// most of the sigma: each costs the row kernel three dependent global loads
// (the beta record, a T element of the alpha ket row and an entry of the L^4
// integral table).  k_sigma_ab adds them after the row kernel (which skips
// them), in the same per-element order, with the data a term needs on chip:
//
//   * a workgroup owns a block of kAbCols columns (one per thread) and a
//     set of rows; the beta single records of its columns are a contiguous
//     range of the CSR and are staged in LDS once ({ket, code}, 8 bytes);
//   * for each row ia and each alpha single (ia -> ja, orbitals xa -> ya) the
//     L x L slice (ya xa | yb xb) of the packed ERI is staged in LDS;
//   * the term's sign is the fermionic sign of the alpha single on D followed
//     by the beta single on the result, which equals the reference's
//     TwoExcite parity x order sign x (the sign with which the surviving
//     direct or exchange integral enters): a per-row bit, a per-column bit,
//     the record's same-spin parity bit and two prefix-parity bits of the
//     ket alpha string A' (alpha orbitals in (xb, yb]).  The term is then
//     that sign applied to the integral, times T[ja, jb]: the same value as
//     the reference's, so the FP64 result is unchanged bit for bit.
#ifndef SBD_AB_COLS
#define SBD_AB_COLS 256
#endif
#ifndef SBD_AB_BATCH
#define SBD_AB_BATCH 16  // 4 / 8 / 16 on the reference's H2O 1e-7 FP64: 12.86 / 11.38 / 11.25 s
#endif
constexpr int kAbCols = SBD_AB_COLS;    // columns (threads) per workgroup
constexpr int kAbKetBits = 19;          // beta strings per basis for the packed records
constexpr int kAbBatch = SBD_AB_BATCH;  // records whose T elements are fetched together

// A beta single record packed for LDS: ket | xb << 19 | yb << 25 | parity << 31.
__device__ inline uint32_t ab_pack(const PairRec& r) {
  return r.ket | ((uint32_t)code_x0(r.code) << kAbKetBits) |
         ((uint32_t)code_y0(r.code) << (kAbKetBits + 6)) | ((uint32_t)code_par(r.code) << 31);
}

// Workgroup blockIdx = cb + ncb * rg: column block cb, rows rg, rg + nrg, ...
template <typename VT>
__global__ __launch_bounds__(kAbCols) void k_sigma_ab(SigmaArgs<VT> p, uint32_t ncb,
                                                     uint32_t nrg) {
  using AT = VT;
  extern __shared__ unsigned char lds_raw[];
  const int L = p.L;
  double* vsl = reinterpret_cast<double*>(lds_raw);  // L x L slice
  uint32_t* recs = reinterpret_cast<uint32_t*>(lds_raw + sizeof(double) * L * L);
  const uint32_t nb = (uint32_t)p.nb;
  const uint32_t cb = blockIdx.x % ncb, rg = blockIdx.x / ncb;
  const uint32_t c0 = cb * kAbCols;
  const uint32_t c = c0 + threadIdx.x;
  const bool ok = c < nb;
  const uint32_t cend = c0 + kAbCols < nb ? c0 + kAbCols : nb;
  // this block's beta single records (a contiguous range of the CSR)
  const uint32_t r0 = p.bs_off[c0], r1 = p.bs_off[cend];
  for (uint32_t r = threadIdx.x; r < r1 - r0; r += kAbCols) recs[r] = ab_pack(p.bs_rec[r0 + r]);
  const uint32_t q0 = ok ? p.bs_off[c] - r0 : 0, q1 = ok ? p.bs_off[c + 1] - r0 : 0;
  const uint64_t B = ok ? p.b[c] : 0;
  for (uint64_t ia = rg; ia < p.na; ia += nrg) {
    const uint32_t as0 = p.as_off[ia], as1 = p.as_off[ia + 1];
    if (as0 == as1) continue;  // uniform: no alpha singles in this row
    const uint64_t A = p.a[ia];
    AT acc = (ok && q1 > q0) ? (AT)p.W[ia * nb + c] : (AT)0;
    for (uint32_t qa = as0; qa < as1; qa++) {
      const PairRec ra = p.as_rec[qa];
      const int xa = code_x0(ra.code), ya = code_y0(ra.code);
      __syncthreads();  // the previous slice is no longer read (and, first, the records are in)
      for (int e = threadIdx.x; e < L * L; e += kAbCols)
        vsl[e] = p.I2[eri_index(ya, xa, e % L, e / L)];  // (ya xa | yb xb) at e = xb L + yb
      __syncthreads();
      if (q1 == q0) continue;
      const int lo = xa < ya ? xa : ya, hi = xa < ya ? ya : xa;
      const uint64_t PA1 = prefix_parity(A ^ (1ULL << xa) ^ (1ULL << ya));
      const uint32_t s_col =
          parity64(A & bits_range(lo + 1, hi)) ^ parity64(B & bits_range(lo, hi));
      const VT* __restrict__ Tk = p.T + (uint64_t)ra.ket * nb;
      for (uint32_t k0 = q0; k0 < q1; k0 += kAbBatch) {
        uint32_t rb[kAbBatch];
        VT tv[kAbBatch];
#pragma unroll
        for (int u = 0; u < kAbBatch; u++) {
          rb[u] = recs[k0 + u < q1 ? k0 + u : q1 - 1];
          tv[u] = Tk[rb[u] & ((1u << kAbKetBits) - 1)];
        }
#pragma unroll
        for (int u = 0; u < kAbBatch; u++) {
          const int xb = (rb[u] >> kAbKetBits) & 63, yb = (rb[u] >> (kAbKetBits + 6)) & 63;
          const uint32_t sg = s_col ^ (rb[u] >> 31) ^ (uint32_t)((PA1 >> xb) ^ (PA1 >> yb));
          const double e = flip_sign(vsl[xb * L + yb], sg & 1u);
          const AT nv = acc + (AT)e * (AT)tv[u];
          acc = (k0 + u < q1) ? nv : acc;
        }
      }
    }
    if (ok && q1 > q0) p.W[ia * nb + c] = (VT)acc;
  }
}

// Row-wise form of the dense pass (bases of up to kAbRowBlock * 40 beta
// strings whose row and integral slice fit LDS): one 1024-thread workgroup per
// bra row (grid-stride over rows), thread t owning columns t + 1024 j (j < C)
// with their sums kept across all of the row's alpha singles (in registers up
// to 8 columns, in private memory from 16 on).  For each alpha single the ket row
// T[ja, :] and the integral slice are staged in LDS; with two buffers (when
// they fit) the next ones are loaded into registers while the current ones are
// used and stored after, one barrier per alpha single.  The beta records come
// from a column-major packed copy (record k of column c at k nb + c), so a
// wavefront's record loads are coalesced; every gather of the inner loop is
// from LDS.  Same terms, same order, same operations as k_sigma_ab.
constexpr int kAbRowBlock = 1024;
// records whose LDS operands are fetched together (4 / 8 / 12 / 16 measured on the
// reference's H2O 1e-6: 1,692 / 1,563 / 1,581 / 1,554 ms per FP64 sigma)
constexpr int kAbRowBatch = 16;

template <typename VT, int C, bool kDbuf>
__global__ __launch_bounds__(kAbRowBlock) void k_sigma_ab_rows(SigmaArgs<VT> p,
                                                              const uint32_t* __restrict__ sell) {
  using AT = VT;
  extern __shared__ unsigned char lds_raw[];
  const int L = p.L;
  const int LL = L * L;
  const uint32_t nb = (uint32_t)p.nb;
  const size_t row_bytes = ((size_t)nb * sizeof(VT) + 15) & ~(size_t)15;
  const size_t buf_bytes = row_bytes + sizeof(double) * LL;
  auto trow = [&](int b) { return reinterpret_cast<VT*>(lds_raw + b * buf_bytes); };
  auto vslc = [&](int b) { return reinterpret_cast<double*>(lds_raw + b * buf_bytes + row_bytes); };
  constexpr int kV = 4;  // slice elements a thread stages (L*L <= 4096)
  // this thread's columns: record counts and strings (the same for every row)
  uint32_t cnt[C];
  uint64_t B[C];
#pragma unroll
  for (int j = 0; j < C; j++) {
    const uint32_t c = threadIdx.x + j * kAbRowBlock;
    cnt[j] = c < nb ? p.bs_off[c + 1] - p.bs_off[c] : 0;
    B[j] = c < nb ? p.b[c] : 0;
  }
  auto stage = [&](int b, uint32_t qa) {  // synchronous staging of alpha single qa
    const PairRec ra = p.as_rec[qa];
    const int xa = code_x0(ra.code), ya = code_y0(ra.code);
    const VT* __restrict__ Tk = p.T + (uint64_t)ra.ket * nb;
    VT* tr = trow(b);
    double* vs = vslc(b);
    VT v[C];
#pragma unroll
    for (int j = 0; j < C; j++) {
      const uint32_t c = threadIdx.x + j * kAbRowBlock;
      v[j] = c < nb ? Tk[c] : (VT)0;
    }
#pragma unroll
    for (int j = 0; j < C; j++) {
      const uint32_t c = threadIdx.x + j * kAbRowBlock;
      if (c < nb) tr[c] = v[j];
    }
    for (int e = threadIdx.x; e < LL; e += kAbRowBlock) vs[e] = p.I2[eri_index(ya, xa, e % L, e / L)];
  };
  for (uint64_t ia = blockIdx.x; ia < p.na; ia += gridDim.x) {
    const uint32_t as0 = p.as_off[ia], as1 = p.as_off[ia + 1];
    if (as0 == as1) continue;  // uniform: no alpha singles in this row
    const uint64_t A = p.a[ia];
    AT acc[C];
#pragma unroll
    for (int j = 0; j < C; j++)
      acc[j] = cnt[j] ? (AT)p.W[ia * nb + threadIdx.x + j * kAbRowBlock] : (AT)0;
    __syncthreads();  // the buffers of the previous row are no longer read
    stage(0, as0);
    __syncthreads();
    for (uint32_t qa = as0; qa < as1; qa++) {
      const int cur = kDbuf ? (int)((qa - as0) & 1) : 0;
      // next alpha single's ket row and slice, into registers
      VT nt[kDbuf ? C : 1];
      double nv[kDbuf ? kV : 1];
      const bool more = kDbuf && qa + 1 < as1;  // uniform
      if (more) {
        const PairRec rn = p.as_rec[qa + 1];
        const int xn = code_x0(rn.code), yn = code_y0(rn.code);
        const VT* __restrict__ Tn = p.T + (uint64_t)rn.ket * nb;
#pragma unroll
        for (int j = 0; j < (kDbuf ? C : 1); j++) {
          const uint32_t c = threadIdx.x + j * kAbRowBlock;
          nt[j] = c < nb ? Tn[c] : (VT)0;
        }
#pragma unroll
        for (int i = 0; i < (kDbuf ? kV : 1); i++) {
          const int e = threadIdx.x + i * kAbRowBlock;
          nv[i] = e < LL ? p.I2[eri_index(yn, xn, e % L, e / L)] : 0.0;
        }
      }
      const PairRec ra = p.as_rec[qa];
      const int xa = code_x0(ra.code), ya = code_y0(ra.code);
      const int lo = xa < ya ? xa : ya, hi = xa < ya ? ya : xa;
      const uint64_t PA1 = prefix_parity(A ^ (1ULL << xa) ^ (1ULL << ya));
      const uint32_t sA = parity64(A & bits_range(lo + 1, hi));
      const VT* tr = trow(cur);
      const double* vs = vslc(cur);
      // unrolled up to 8 columns per thread; from 16 on the unrolled copies
      // spill (412 B/lane FP64 at 16), so the loop stays rolled and the sums are
      // indexed by the uniform column counter (H2O 1e-6: 1.62 -> 1.56 s per sigma)
#pragma unroll(C > 8 ? 1 : C)
      for (int j = 0; j < C; j++) {
        if (!cnt[j]) continue;
        const uint32_t c = threadIdx.x + j * kAbRowBlock;
        const uint32_t s_col = sA ^ parity64(B[j] & bits_range(lo, hi));
        for (uint32_t k0 = 0; k0 < cnt[j]; k0 += kAbRowBatch) {
          uint32_t rb[kAbRowBatch];
#pragma unroll
          for (int u = 0; u < kAbRowBatch; u++)
            rb[u] = sell[(uint64_t)(k0 + u < cnt[j] ? k0 + u : cnt[j] - 1) * nb + c];
#pragma unroll
          for (int u = 0; u < kAbRowBatch; u++) {
            const int xb = (rb[u] >> kAbKetBits) & 63, yb = (rb[u] >> (kAbKetBits + 6)) & 63;
            const uint32_t sg = s_col ^ (rb[u] >> 31) ^ (uint32_t)((PA1 >> xb) ^ (PA1 >> yb));
            const double e = flip_sign(vs[xb * L + yb], sg & 1u);
            const AT nv2 = acc[j] + (AT)e * (AT)tr[rb[u] & ((1u << kAbKetBits) - 1)];
            acc[j] = (k0 + u < cnt[j]) ? nv2 : acc[j];
          }
        }
      }
      if (kDbuf) {
        if (more) {
          VT* tn = trow(cur ^ 1);
          double* vn = vslc(cur ^ 1);
#pragma unroll
          for (int j = 0; j < (kDbuf ? C : 1); j++) {
            const uint32_t c = threadIdx.x + j * kAbRowBlock;
            if (c < nb) tn[c] = nt[j];
          }
#pragma unroll
          for (int i = 0; i < (kDbuf ? kV : 1); i++) {
            const int e = threadIdx.x + i * kAbRowBlock;
            if (e < LL) vn[e] = nv[i];
          }
        }
        __syncthreads();
      } else if (qa + 1 < as1) {
        __syncthreads();
        stage(0, qa + 1);
        __syncthreads();
      }
    }
#pragma unroll
    for (int j = 0; j < C; j++)
      if (cnt[j]) p.W[ia * nb + threadIdx.x + j * kAbRowBlock] = (VT)acc[j];
  }
}

// Column-major packed copy of the beta single records for k_sigma_ab_rows:
// record k of column c (k < its count) at k nb + c.
__global__ void k_ab_sell(uint32_t nb, const uint32_t* __restrict__ off,
                          const PairRec* __restrict__ rec, uint32_t* __restrict__ sell) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nb) return;
  for (uint32_t k = 0; k < off[c + 1] - off[c]; k++) sell[(uint64_t)k * nb + c] = ab_pack(rec[off[c] + k]);
}

// Sparse bases (a few off-diagonal terms per determinant): one element per
// thread, no row staging, every wavefront streaming.  Workgroup tiles are
// 1024 consecutive columns of one row; thread t owns columns base + t + 256 j,
// j = 0..3.  Each element is summed in the order of k_sigma_rows
// (diagonal, alpha singles, alpha doubles, beta singles, beta doubles from the
// CSR in the order of the ELL slots, alpha-single x beta-single terms) with the
// same operations, so the results are the same bits.
constexpr int kElemBlock = 256;
constexpr double kSparseTerms = 2.0;  // off-diagonal pairs per alpha string + per beta string
// the same sum at which rows too wide for LDS take the row kernel (L2 gathers)
constexpr double kDenseRowTerms = 64.0;
constexpr int kElemCols = 4;
constexpr int kElemTile = kElemBlock * kElemCols;

template <typename VT, bool kCache>
__global__ __launch_bounds__(kElemBlock) void k_sigma_elems(const SigmaArgs<VT> p,
                                                            const PairRec* __restrict__ bd_rec,
                                                            uint64_t ntiles,
                                                            double* __restrict__ dot_part) {
  using AT = VT;
  const uint32_t nb = (uint32_t)p.nb;
  const int L = p.L;
  const uint64_t tpr = (nb + kElemTile - 1) / kElemTile;  // tiles per row
  double tw = 0.0;
  for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const uint64_t ia = tile / tpr;
    const uint32_t base = (uint32_t)(tile - ia * tpr) * kElemTile;
    const uint64_t A = p.a[ia];
    const VT* __restrict__ Trow = p.T + ia * nb;
    uint32_t col[kElemCols];
    bool ok[kElemCols];
    ColRec cb[kElemCols];
    AT acc[kElemCols];
#pragma unroll
    for (int j = 0; j < kElemCols; j++) {
      const uint32_t c = base + threadIdx.x + j * kElemBlock;
      ok[j] = c < nb;
      col[j] = ok[j] ? c : nb - 1;
      cb[j] = p.b_rec[col[j]];
      const VT hv = p.hii[ia * nb + col[j]];
      const VT ts = Trow[col[j]];
      acc[j] = p.accumulate ? (AT)p.W[ia * nb + col[j]] : (AT)0;
      acc[j] += (AT)hv * (AT)ts;
    }
    // alpha singles (uniform over the workgroup)
    const uint32_t as0 = p.as_off[ia], as1 = p.as_off[ia + 1];
    for (uint32_t q = as0; q < as1; q++) {
      const PairRec rc = p.as_rec[q];
      const VT* __restrict__ Tk = p.T + (uint64_t)rc.ket * nb;
#pragma unroll
      for (int j = 0; j < kElemCols; j++) {
        const VT tk = Tk[col[j]];
        const double e = kCache ? p.Ea[(uint64_t)q * nb + col[j]]
                                : alpha_single_elem(rc, A, cb[j].B, L, p.G);
        acc[j] += (AT)e * (AT)tk;
      }
    }
    // alpha doubles
    for (uint32_t q = p.ad_off[ia]; q < p.ad_off[ia + 1]; q++) {
      const PairRec rc = p.ad_rec[q];
      const VT* __restrict__ Tk = p.T + (uint64_t)rc.ket * nb;
      const uint64_t mm = beta_mask_double(rc.code);
#pragma unroll
      for (int j = 0; j < kElemCols; j++) {
        const VT tv = Tk[col[j]];
        const double e = flip_sign(rc.val, (uint32_t)__popcll(cb[j].B & mm));
        acc[j] += (AT)e * (AT)tv;
      }
    }
#pragma unroll
    for (int j = 0; j < kElemCols; j++) {
      // beta singles
      for (uint32_t q = cb[j].s0; q < cb[j].s1; q++) {
        const PairRec rc = p.bs_rec[q];
        const double e = kCache ? p.Eb[ia * (uint64_t)p.Sb + q]
                                : beta_single_elem(rc, A, cb[j].B, L, p.G);
        acc[j] += (AT)e * (AT)Trow[rc.ket];
      }
      // beta doubles: the value as the ELL slot of this precision holds it
      for (uint32_t q = p.bd_off[col[j]]; q < p.bd_off[col[j] + 1]; q++) {
        const PairRec rc = bd_rec[q];
        const uint64_t mask = bits_range(code_f(rc.code, 0) + 1, code_f(rc.code, 1) + 1) ^
                              bits_range(code_f(rc.code, 2) + 1, code_f(rc.code, 3) + 1);
        const VT e = flip_sign((VT)rc.val, (uint32_t)__popcll(A & mask));
        acc[j] += e * Trow[rc.ket];
      }
      // alpha single x beta single
      if (as1 > as0 && cb[j].s1 > cb[j].s0) {
        const uint64_t D0 = spread32(A) | (spread32(cb[j].B) << 1);
        const uint64_t D1 = spread32(A >> 32) | (spread32(cb[j].B >> 32) << 1);
        for (uint32_t q = as0; q < as1; q++) {
          const PairRec ra = p.as_rec[q];
          const int xa = code_x0(ra.code), ya = code_y0(ra.code);
          const VT* __restrict__ Tk = p.T + (uint64_t)ra.ket * nb;
          for (uint32_t k = cb[j].s0; k < cb[j].s1; k++) {
            const PairRec rb = p.bs_rec[k];
            const double e = ab_elem(D0, D1, xa, ya, code_x0(rb.code), code_y0(rb.code), L, p.Vab);
            acc[j] = acc[j] + (AT)e * (AT)Tk[rb.ket];
          }
        }
      }
      if (ok[j]) {
        p.W[ia * nb + col[j]] = (VT)acc[j];
        if (dot_part) tw += (double)Trow[col[j]] * (double)(VT)acc[j];
      }
    }
  }
  if (dot_part) {
    // per-wavefront partial, fixed order of lanes
    for (int o = 32; o > 0; o >>= 1) tw += __shfl_xor(tw, o, 64);
    if ((threadIdx.x & 63) == 0)
      dot_part[(uint64_t)blockIdx.x * (kElemBlock / 64) + (threadIdx.x >> 6)] = tw;
  }
}

// Sum of n partials in a fixed order: thread t adds partials t, t + 1024, ...
// (compensated), then the 1024 thread sums are combined in a fixed tree.  The
// loads are issued eight at a time ahead of the additions, which keep their
// order.
constexpr int kSumBlock = 1024;
__device__ inline void sum_fixed_block(const double* __restrict__ part, uint64_t n,
                                       double* __restrict__ out) {
  __shared__ double red[kSumBlock];
  constexpr int U = 8;
  double s = 0.0, c = 0.0;
  for (uint64_t i0 = threadIdx.x; i0 < n; i0 += (uint64_t)U * kSumBlock) {
    double v[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const uint64_t i = i0 + (uint64_t)u * kSumBlock;
      v[u] = i < n ? part[i] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (i0 + (uint64_t)u * kSumBlock < n) {
        const double y = v[u] - c;
        const double t = s + y;
        c = (t - s) - y;
        s = t;
      }
    }
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = kSumBlock / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}
__global__ __launch_bounds__(kSumBlock) void k_sum_fixed(const double* __restrict__ part,
                                                         uint64_t n, double* __restrict__ out) {
  sum_fixed_block(part, n, out);
}

int build_single_cache(sbd_ctx* ctx, hipStream_t st) {
  const uint64_t Sa = ctx->ta.n_s, Sb = ctx->tb.n_s;
  const uint64_t bytes = 8 * (Sa * ctx->nb + ctx->na * Sb);
  if (bytes == 0 || bytes > ctx->single_cache_budget) return SBD_OK;  // evaluate on the fly
  if (Sa) {
    SBD_HIP_TRY(ctx, hipMalloc(&ctx->d_Ea, 8 * Sa * ctx->nb));
    const uint64_t total = Sa * ctx->nb;
    const unsigned grid = (unsigned)std::min<uint64_t>((total + 255) / 256, 65535);
    hipLaunchKernelGGL(k_single_alpha, dim3(grid), dim3(256), 0, st, ctx->d_a, ctx->d_b, ctx->nb,
                       total, ctx->norb, ctx->ta.s_rec, ctx->ta.s_bra, ctx->d_G, ctx->d_Ea);
    SBD_HIP_TRY(ctx, hipGetLastError());
  }
  if (Sb) {
    SBD_HIP_TRY(ctx, hipMalloc(&ctx->d_Eb, 8 * ctx->na * Sb));
    const uint64_t total = ctx->na * Sb;
    const unsigned grid = (unsigned)std::min<uint64_t>((total + 255) / 256, 65535);
    hipLaunchKernelGGL(k_single_beta, dim3(grid), dim3(256), 0, st, ctx->d_a, ctx->d_b, Sb,
                       total, ctx->norb, ctx->tb.s_rec, ctx->tb.s_bra, ctx->d_G, ctx->d_Eb);
    SBD_HIP_TRY(ctx, hipGetLastError());
  }
  SBD_HIP_TRY(ctx, hipStreamSynchronize(st));
  return SBD_OK;
}

struct SumJob {
  const double* part;
  uint64_t n;
  double* out;
};
// k_sum_fixed for several jobs, one block each (same arithmetic)
__global__ __launch_bounds__(kSumBlock) void k_sum_fixed_batch(const SumJob* __restrict__ jobs) {
  const SumJob& J = jobs[blockIdx.x];
  sum_fixed_block(J.part, J.n, J.out);
}

template <typename VT>
static SigmaArgs<VT> sigma_args(const sbd_ctx* ctx, const VT* hii, const VT* T, VT* W,
                                int accumulate) {
  const SpinTables& A = ctx->ta;
  const SpinTables& B = ctx->tb;
  return SigmaArgs<VT>{ctx->d_a,  ctx->d_b,   ctx->na,      ctx->nb,   ctx->norb, hii,
                       T,         W,          accumulate,   A.s_off,   A.s_rec,   A.d_off,
                       A.d_rec,   B.s_off,    B.s_rec,      B.d_off,   B.ell64,   B.ell32,
                       B.ell_ket, B.ell_off,  B.ell_pos,    B.col_rec, ctx->d_Ea, ctx->d_Eb,
                       (uint32_t)B.n_s, ctx->d_G, ctx->d_eri, ctx->d_Vab};
}

// rows per workgroup for one context (a property of the context alone, so a
// job computes the same way in any batch): the most rows that fit LDS
// together and that the basis has, up to 4 FP64 rows (whose per-row state
// still fits the registers) or 8 FP32 rows
static int sigma_rows_for(const sbd_ctx* ctx, size_t elem) {
  const size_t row_bytes = elem * ctx->nb;
  int r = elem >= 8 ? 4 : 8;
  while (r > 1 && (r * row_bytes > (size_t)ctx->lds_max || ctx->na < (uint64_t)r)) r /= 2;
  return r;
}

// Which kernel a context's sigma takes.  One rule serves single launches and
// the serving engine's batches, so a solve computes the same way alone and in
// a batch: the per-element kernel for sparse bases (fewer than kSparseTerms
// off-diagonal pairs per alpha string plus per beta string; fewer than
// kDenseRowTerms when a row does not fit LDS), else the split form or the row
// kernel (sigma_use_split); sbd_set_sigma_kernel can force any of them.
static bool sigma_use_elems(const sbd_ctx* ctx, size_t elem) {
  if (ctx->sigma_kernel == 1 || ctx->sigma_kernel == 3 || ctx->sigma_kernel == 4) return false;
  if (ctx->sigma_kernel == 2) return true;
  const SpinTables& A = ctx->ta;
  const SpinTables& B = ctx->tb;
  const double terms = (double)(A.n_s + A.n_d) / (double)ctx->na +
                       (double)(B.n_s + B.n_d) / (double)ctx->nb;
  // rows that do not fit LDS: the row kernel gathers every beta excitation's
  // T element from L2.  On nearly sparse bases that is far slower than this
  // kernel (17x on 60,480 strings of 58 orbitals with the HF singles, about
  // 4 pairs per string); on dense ones far faster (the reference's H2O 1e-7,
  // 25,062 strings, about 2,500 pairs per alpha plus beta string: 11.1 s
  // against 71.3 s per FP64 sigma)
  if (elem * ctx->nb > (size_t)ctx->lds_max) return terms < kDenseRowTerms;
  return terms < kSparseTerms;
}

static bool sigma_cached(const sbd_ctx* ctx) {
  return (ctx->ta.n_s == 0 || ctx->d_Ea) && (ctx->tb.n_s == 0 || ctx->d_Eb);
}

constexpr double kSplitDenseDoubles = 64.0;  // beta doubles per string

// Whether a context outside the per-element rule takes the split form
// (k_sigma_stream + k_sigma_bd) rather than the row kernel.  Both need the bra
// rows in LDS.  By default the split form is taken when two or more rows fit
// LDS together (measured faster in every such case: FP32 at 20,000 beta
// strings, FCI norb 12, the reference's H2O) or the beta doubles are dense.
// Knob 4 forces it where it fits, 1 and 3 never take it.
template <typename VT>
static bool sigma_use_split(const sbd_ctx* ctx) {
  if (sizeof(VT) * ctx->nb > (size_t)ctx->lds_max) return false;
  if (ctx->sigma_kernel == 4) return true;
  if (ctx->sigma_kernel != 0) return false;
  // one row per workgroup: the row kernel's overlapped phases win on sparse
  // beta doubles (7 per string: 11.19 vs 11.78 ms), the split form on dense
  // ones (906 per string, the reference's H2O 1e-6: 1622 vs 1678 ms)
  return sigma_rows_for(ctx, sizeof(VT)) >= 2 ||
         (double)ctx->tb.n_d / (double)ctx->nb >= kSplitDenseDoubles;
}

// The row kernel leaves the alpha-single x beta-single terms to k_sigma_ab
// when the basis has at least kAbDense of them per determinant, its beta
// strings fit the packed records, and a column block's records and the
// integral slice fit LDS (decided once per context; knob 3 keeps them
// inline).
constexpr double kAbDense = 4.0;
static bool sigma_ab_pass(sbd_ctx* ctx) {
  if (ctx->sigma_kernel == 3) return false;
  if (ctx->ab_pass < 0) {
    ctx->ab_pass = 0;
    const SpinTables& A = ctx->ta;
    const SpinTables& B = ctx->tb;
    const double terms = ((double)A.n_s / (double)ctx->na) * ((double)B.n_s / (double)ctx->nb);
    if (terms >= kAbDense && ctx->nb <= (1ULL << kAbKetBits) &&
        B.s_off_host.size() == ctx->nb + 1) {
      uint32_t mx = 0;
      for (uint64_t c0 = 0; c0 < ctx->nb; c0 += kAbCols) {
        const uint64_t c1 = std::min<uint64_t>(ctx->nb, c0 + kAbCols);
        mx = std::max(mx, B.s_off_host[c1] - B.s_off_host[c0]);
      }
      const size_t lds = sizeof(double) * ctx->norb * ctx->norb + sizeof(uint32_t) * (size_t)mx;
      if (lds <= (size_t)ctx->lds_max) {
        ctx->ab_pass = 1;
        ctx->ab_lds = lds;
      }
    }
  }
  return ctx->ab_pass == 1;
}

// Row-wise form of the dense pass (k_sigma_ab_rows) for a vector of element
// size elem: columns per thread (0: the basis is too wide for it), whether
// two LDS buffers fit, and their bytes.  Builds the column-major record copy
// on first use.
static int sigma_ab_rows_cfg(sbd_ctx* ctx, size_t elem, int* C, bool* dbuf, size_t* lds,
                             hipStream_t st) {
  *C = 0;
  const uint64_t per = (ctx->nb + kAbRowBlock - 1) / kAbRowBlock;
  if (per > 40) return SBD_OK;  // wider rows take the column-block form
  *C = per <= 1    ? 1
       : per <= 2  ? 2
       : per <= 4  ? 4
       : per <= 8  ? 8
       : per <= 16 ? 16
       : per <= 24 ? 24
       : per <= 32 ? 32
                   : 40;
  const size_t row_bytes = (elem * ctx->nb + 15) & ~(size_t)15;
  const size_t buf = row_bytes + sizeof(double) * ctx->norb * ctx->norb;
  *dbuf = 2 * buf <= (size_t)ctx->lds_max;
  *lds = *dbuf ? 2 * buf : buf;
  if (*lds > (size_t)ctx->lds_max) {
    *C = 0;
    return SBD_OK;
  }
  if (!ctx->d_ab_sell) {
    const size_t n = (size_t)std::max<uint32_t>(ctx->tb.max_s, 1) * ctx->nb;
    SBD_HIP_TRY(ctx, hipMalloc(&ctx->d_ab_sell, sizeof(uint32_t) * n));
    hipLaunchKernelGGL(k_ab_sell, dim3((unsigned)((ctx->nb + 255) / 256)), dim3(256), 0, st,
                       (uint32_t)ctx->nb, ctx->tb.s_off, ctx->tb.s_rec, ctx->d_ab_sell);
    SBD_HIP_TRY(ctx, hipGetLastError());
  }
  return SBD_OK;
}

int launch_sigma_batch(BatchScratch& sc, const SigmaItem* items, int n, double* tw_out,
                       hipStream_t st) {
  if (n < 1) return SBD_OK;
  // jobs by form and rows per workgroup (1, 2, 4); the rest take their own
  // launch
  std::vector<int> direct, byrows[3], bysplit[3];
  for (int j = 0; j < n; j++) {
    sbd_ctx* c = items[j].ctx;
    if (!c->tables_built) return SBD_ERR_STATE;
    const bool lds = sizeof(double) * c->nb <= (size_t)c->lds_max;
    const int r = lds ? sigma_rows_for(c, sizeof(double)) : 1;
    const uint64_t groups = (c->na + r - 1) / r;
    const int ri = r == 4 ? 2 : r == 2 ? 1 : 0;
    if (sigma_use_elems(c, sizeof(double)) || sigma_ab_pass(c) || !sigma_cached(c) || !lds ||
        !c->tb.ell_packed || groups > kSigMaxGrid)
      direct.push_back(j);
    else if (sigma_use_split<double>(c))
      bysplit[ri].push_back(j);
    else
      byrows[ri].push_back(j);
  }
  // jobs outside the batched forms: their own launch (the same kernels, grids
  // and sums as alone)
  for (int j : direct) {
    int rc = launch_sigma<double>(items[j].ctx, items[j].hii, items[j].T, items[j].W, 0, st,
                                  &tw_out[j]);
    if (rc) return rc;
  }
  // the split form's streaming part of every split job, one launch per slab width
  for (int J : {1, 2, 4}) {
    std::vector<StreamJob<double>> sj;
    uint32_t blk = 0;
    for (int ri = 0; ri < 3; ri++)
      for (int j : bysplit[ri]) {
        const sbd_ctx* c = items[j].ctx;
        const int cj = c->stream_cols >= 1024 ? 4 : c->stream_cols >= 512 ? 2 : 1;
        if (cj != J) continue;
        StreamJob<double> q;
        q.p = sigma_args<double>(c, items[j].hii, items[j].T, items[j].W, 0);
        const uint64_t sw = (uint64_t)kStreamBlock * J;
        q.ntiles = ((c->nb + sw - 1) / sw) * c->na;
        q.ngrp = (uint32_t)std::min<uint64_t>(q.ntiles, (uint64_t)c->cu_count * 8);
        q.blk0 = blk;
        blk += q.ngrp;
        sj.push_back(q);
      }
    if (sj.empty()) continue;
    int rc = batch_reserve(sc, sizeof(StreamJob<double>) * sj.size(), 0, 0);
    if (rc) return rc;
    if (hipMemcpyAsync(sc.d_jobs, sj.data(), sizeof(StreamJob<double>) * sj.size(),
                       hipMemcpyHostToDevice, st) != hipSuccess)
      return SBD_ERR_DEVICE;
    const StreamJob<double>* dj = (const StreamJob<double>*)sc.d_jobs;
    switch (J) {
      case 4: hipLaunchKernelGGL((k_sigma_stream_batch<double, 4>), dim3(blk), dim3(kStreamBlock), 0, st, dj, (int)sj.size()); break;
      case 2: hipLaunchKernelGGL((k_sigma_stream_batch<double, 2>), dim3(blk), dim3(kStreamBlock), 0, st, dj, (int)sj.size()); break;
      default: hipLaunchKernelGGL((k_sigma_stream_batch<double, 1>), dim3(blk), dim3(kStreamBlock), 0, st, dj, (int)sj.size()); break;
    }
    if (hipGetLastError() != hipSuccess) return SBD_ERR_DEVICE;
    // the job table is read by the launch before the next copy overwrites it
    if (hipStreamSynchronize(st) != hipSuccess) return SBD_ERR_DEVICE;
  }
  for (int split = 0; split < 2; split++) {
    for (int pass = 0; pass < 3; pass++) {
      const std::vector<int>& sel = split ? bysplit[pass] : byrows[pass];
      if (sel.empty()) continue;
      const int R = 1 << pass;
      std::vector<SigmaJob<double>> jobs(sel.size());
      std::vector<SumJob> sums(sel.size());
      size_t parts = 0, shm = 0;
      uint32_t blk = 0;
      for (size_t q = 0; q < sel.size(); q++) {
        const SigmaItem& it = items[sel[q]];
        SigmaJob<double>& J = jobs[q];
        J.p = sigma_args<double>(it.ctx, it.hii, it.T, it.W, 0);
        J.ngrp = (uint32_t)((it.ctx->na + R - 1) / R);
        J.blk0 = blk;
        blk += J.ngrp;
        parts += (size_t)J.ngrp * (kSigBlock / 64);
        shm = std::max(shm, (size_t)R * sizeof(double) * it.ctx->nb);
      }
      int rc = batch_reserve(sc, std::max(sizeof(SigmaJob<double>), sizeof(SumJob)) * sel.size(),
                             parts, sel.size());
      if (rc) return rc;
      size_t off = 0;
      for (size_t q = 0; q < sel.size(); q++) {
        jobs[q].dot_part = sc.d_part + off;
        sums[q] = SumJob{sc.d_part + off, (uint64_t)jobs[q].ngrp * (kSigBlock / 64), sc.d_out + q};
        off += (size_t)jobs[q].ngrp * (kSigBlock / 64);
      }
      if (hipMemcpyAsync(sc.d_jobs, jobs.data(), sizeof(SigmaJob<double>) * jobs.size(),
                         hipMemcpyHostToDevice, st) != hipSuccess)
        return SBD_ERR_DEVICE;
      const SigmaJob<double>* dj = (const SigmaJob<double>*)sc.d_jobs;
      const int nj = (int)jobs.size();
      if (split) {
        switch (R) {
          case 4: hipLaunchKernelGGL((k_sigma_bd_batch<double, 4>), dim3(blk), dim3(kSigBlock), shm, st, dj, nj); break;
          case 2: hipLaunchKernelGGL((k_sigma_bd_batch<double, 2>), dim3(blk), dim3(kSigBlock), shm, st, dj, nj); break;
          default: hipLaunchKernelGGL((k_sigma_bd_batch<double, 1>), dim3(blk), dim3(kSigBlock), shm, st, dj, nj); break;
        }
      } else {
        switch (R) {
          case 4: hipLaunchKernelGGL((k_sigma_batch<double, 4>), dim3(blk), dim3(kSigBlock), shm, st, dj, nj); break;
          case 2: hipLaunchKernelGGL((k_sigma_batch<double, 2>), dim3(blk), dim3(kSigBlock), shm, st, dj, nj); break;
          default: hipLaunchKernelGGL((k_sigma_batch<double, 1>), dim3(blk), dim3(kSigBlock), shm, st, dj, nj); break;
        }
      }
      if (hipGetLastError() != hipSuccess) return SBD_ERR_DEVICE;
      if (hipMemcpyAsync(sc.d_fin, sums.data(), sizeof(SumJob) * sums.size(), hipMemcpyHostToDevice,
                         st) != hipSuccess)
        return SBD_ERR_DEVICE;
      hipLaunchKernelGGL(k_sum_fixed_batch, dim3((unsigned)sums.size()), dim3(kSumBlock), 0, st,
                         (const SumJob*)sc.d_fin);
      if (hipGetLastError() != hipSuccess) return SBD_ERR_DEVICE;
      if (hipMemcpyAsync(sc.h_out, sc.d_out, sizeof(double) * sel.size(), hipMemcpyDeviceToHost,
                         st) != hipSuccess)
        return SBD_ERR_DEVICE;
      if (hipStreamSynchronize(st) != hipSuccess) return SBD_ERR_DEVICE;
      for (size_t q = 0; q < sel.size(); q++) tw_out[sel[q]] = sc.h_out[q];
    }
  }
  return SBD_OK;
}

// dot-product scratch of a context: n partials + one result slot
static int sdot_reserve(sbd_ctx* ctx, size_t n) {
  if (ctx->sdot_cap >= n + 1) return SBD_OK;
  if (ctx->d_sdot) (void)hipFree(ctx->d_sdot);
  ctx->d_sdot = nullptr;
  ctx->sdot_cap = 0;
  SBD_HIP_TRY(ctx, hipMalloc(&ctx->d_sdot, sizeof(double) * (n + 1)));
  ctx->sdot_cap = n + 1;
  return SBD_OK;
}

// fixed-order sum of the np partials in part into *tw_dot (host), through
// part[np]; synchronises the stream
static int sdot_finish(sbd_ctx* ctx, double* part, uint64_t np, double* tw_dot, hipStream_t st) {
  hipLaunchKernelGGL(k_sum_fixed, dim3(1), dim3(kSumBlock), 0, st, part, np, part + np);
  SBD_HIP_TRY(ctx, hipGetLastError());
  SBD_HIP_TRY(ctx, hipMemcpyAsync(tw_dot, part + np, sizeof(double), hipMemcpyDeviceToHost, st));
  SBD_HIP_TRY(ctx, hipStreamSynchronize(st));
  return SBD_OK;
}

// Per-element kernel (sparse bases).  part: scratch of at least
// sigma_elems_parts(ctx) + 1 doubles when tw_dot is set.
static uint64_t elems_tiles(const sbd_ctx* ctx) {
  return ctx->na * ((ctx->nb + kElemTile - 1) / kElemTile);
}
static unsigned elems_grid(const sbd_ctx* ctx) {
  return (unsigned)std::min<uint64_t>(elems_tiles(ctx), (uint64_t)ctx->cu_count * 32);
}
template <typename VT>
static int launch_elems(sbd_ctx* ctx, const VT* hii, const VT* T, VT* W, int accumulate,
                        hipStream_t st, double* tw_dot, double* part) {
  const uint64_t ntiles = elems_tiles(ctx);
  const unsigned eg = elems_grid(ctx);
  const SigmaArgs<VT> args = sigma_args<VT>(ctx, hii, T, W, accumulate);
  double* epart = tw_dot ? part : nullptr;
  if (sigma_cached(ctx))
    hipLaunchKernelGGL((k_sigma_elems<VT, true>), dim3(eg), dim3(kElemBlock), 0, st, args,
                       ctx->tb.d_rec, ntiles, epart);
  else
    hipLaunchKernelGGL((k_sigma_elems<VT, false>), dim3(eg), dim3(kElemBlock), 0, st, args,
                       ctx->tb.d_rec, ntiles, epart);
  SBD_HIP_TRY(ctx, hipGetLastError());
  if (tw_dot) return sdot_finish(ctx, part, (uint64_t)eg * (kElemBlock / 64), tw_dot, st);
  return SBD_OK;
}

// Row kernel.  part: scratch of at least rows_parts(ctx) + 1 doubles when
// tw_dot is set.
template <typename VT>
static unsigned rows_grid(const sbd_ctx* ctx, int* rows_out) {
  const size_t row_bytes = sizeof(VT) * ctx->nb;
  const bool lds = row_bytes <= (size_t)ctx->lds_max;
  const int rows = lds ? sigma_rows_for(ctx, sizeof(VT)) : 1;
  const uint64_t groups = (ctx->na + rows - 1) / rows;
  if (rows_out) *rows_out = rows;
  return (unsigned)(groups < (uint64_t)kSigMaxGrid ? groups : kSigMaxGrid);
}
// The split sigma's two launches (see k_sigma_stream): slab width
// kStreamBlock * J columns from ctx->stream_cols (256, 512 or 1024).
template <typename VT, bool kCache>
static void launch_stream(const sbd_ctx* ctx, const SigmaArgs<VT>& args, hipStream_t st) {
  const int J = ctx->stream_cols >= 1024 ? 4 : ctx->stream_cols >= 512 ? 2 : 1;
  const uint64_t sw = (uint64_t)kStreamBlock * J;
  const uint64_t ntiles = ((ctx->nb + sw - 1) / sw) * ctx->na;
  const unsigned grid = (unsigned)std::min<uint64_t>(ntiles, 1u << 23);
  if (J == 4) hipLaunchKernelGGL((k_sigma_stream<VT, kCache, 4>), dim3(grid), dim3(kStreamBlock), 0, st, args, ntiles);
  else if (J == 2) hipLaunchKernelGGL((k_sigma_stream<VT, kCache, 2>), dim3(grid), dim3(kStreamBlock), 0, st, args, ntiles);
  else hipLaunchKernelGGL((k_sigma_stream<VT, kCache, 1>), dim3(grid), dim3(kStreamBlock), 0, st, args, ntiles);
}
template <typename VT, int kRows>
static void launch_bd(const sbd_ctx* ctx, const SigmaArgs<VT>& args, bool cache, unsigned grid,
                      size_t shm, double* dpart, hipStream_t st) {
  const bool pk = ctx->tb.ell_packed;
  if (pk && cache) hipLaunchKernelGGL((k_sigma_bd<VT, kRows, true, true>), dim3(grid), dim3(kSigBlock), shm, st, args, dpart);
  else if (pk) hipLaunchKernelGGL((k_sigma_bd<VT, kRows, true, false>), dim3(grid), dim3(kSigBlock), shm, st, args, dpart);
  else if (cache) hipLaunchKernelGGL((k_sigma_bd<VT, kRows, false, true>), dim3(grid), dim3(kSigBlock), shm, st, args, dpart);
  else hipLaunchKernelGGL((k_sigma_bd<VT, kRows, false, false>), dim3(grid), dim3(kSigBlock), shm, st, args, dpart);
}

template <typename VT>
static int launch_rows(sbd_ctx* ctx, const VT* hii, const VT* T, VT* W, int accumulate,
                       hipStream_t st, double* tw_dot, double* part) {
  const size_t row_bytes = sizeof(VT) * ctx->nb;
  const bool lds = row_bytes <= (size_t)ctx->lds_max;
  int rows = 1;
  const unsigned grid = rows_grid<VT>(ctx, &rows);
  const int skip_ab = sigma_ab_pass(ctx) ? 1 : 0;
  // with the dense pass the row kernel's W is not final: <T, W> comes after it
  double* dpart = (tw_dot && !skip_ab) ? part : nullptr;
  const size_t shm = lds ? rows * row_bytes : 0;
  const SpinTables& A = ctx->ta;
  const SpinTables& B = ctx->tb;
  const bool cache = sigma_cached(ctx);
  if (sigma_use_split<VT>(ctx)) {
    SigmaArgs<VT> args = sigma_args<VT>(ctx, hii, T, W, accumulate);
    args.skip_ab = skip_ab;
    if (cache) launch_stream<VT, true>(ctx, args, st);
    else launch_stream<VT, false>(ctx, args, st);
    SBD_HIP_TRY(ctx, hipGetLastError());
    switch (rows) {
      case 8:
        if constexpr (sizeof(VT) == 4) launch_bd<VT, 8>(ctx, args, cache, grid, shm, dpart, st);
        break;
      case 4: launch_bd<VT, 4>(ctx, args, cache, grid, shm, dpart, st); break;
      case 2: launch_bd<VT, 2>(ctx, args, cache, grid, shm, dpart, st); break;
      default: launch_bd<VT, 1>(ctx, args, cache, grid, shm, dpart, st); break;
    }
    SBD_HIP_TRY(ctx, hipGetLastError());
  } else {
#define SBD_SIGMA_ARGS                                                                          \
  ctx->d_a, ctx->d_b, ctx->na, ctx->nb, ctx->norb, hii, T, W, accumulate, A.s_off, A.s_rec,     \
      A.d_off, A.d_rec, B.s_off, B.s_rec, B.d_off, B.ell64, B.ell32, B.ell_ket, B.ell_off,      \
      B.ell_pos, B.col_rec, (const double*)ctx->d_Ea, (const double*)ctx->d_Eb, (uint32_t)B.n_s, \
      (const double*)ctx->d_G, (const double*)ctx->d_eri, (const double*)ctx->d_Vab, skip_ab, dpart
#define SBD_LAUNCH(LDS, CACHE, ROWS)                                                             \
  do {                                                                                           \
    if (B.ell_packed)                                                                            \
      hipLaunchKernelGGL((k_sigma_rows<VT, LDS, CACHE, ROWS, true>), dim3(grid), dim3(kSigBlock), \
                         shm, st, SBD_SIGMA_ARGS);                                               \
    else                                                                                         \
      hipLaunchKernelGGL((k_sigma_rows<VT, LDS, CACHE, ROWS, false>), dim3(grid),                \
                         dim3(kSigBlock), shm, st, SBD_SIGMA_ARGS);                              \
  } while (0)
  // precomputed single-excitation elements, when both caches exist
  if (lds && rows == 8) {
    if constexpr (sizeof(VT) == 4) {
      if (cache) SBD_LAUNCH(true, true, 8);
      else SBD_LAUNCH(true, false, 8);
    }
  } else if (lds && rows == 4) {
    if (cache) SBD_LAUNCH(true, true, 4);
    else SBD_LAUNCH(true, false, 4);
  } else if (lds && rows == 2) {
    if (cache) SBD_LAUNCH(true, true, 2);
    else SBD_LAUNCH(true, false, 2);
  } else if (lds) {
    if (cache) SBD_LAUNCH(true, true, 1);
    else SBD_LAUNCH(true, false, 1);
  } else {
    if (cache) SBD_LAUNCH(false, true, 1);
    else SBD_LAUNCH(false, false, 1);
  }
#undef SBD_LAUNCH
#undef SBD_SIGMA_ARGS
  SBD_HIP_TRY(ctx, hipGetLastError());
  }
  if (skip_ab) {
    const SigmaArgs<VT> args = sigma_args<VT>(ctx, hii, T, W, 1);
    int C = 0;
    bool dbuf = false;
    size_t lds_rows = 0;
    int rc = sigma_ab_rows_cfg(ctx, sizeof(VT), &C, &dbuf, &lds_rows, st);
    if (rc) return rc;
    if (C) {
      // one workgroup per CU over the rows
      const unsigned g = (unsigned)std::min<uint64_t>(ctx->na, (uint64_t)ctx->cu_count);
#define SBD_AB_ROWS(CC)                                                                           \
  do {                                                                                            \
    if (dbuf)                                                                                     \
      hipLaunchKernelGGL((k_sigma_ab_rows<VT, CC, true>), dim3(g), dim3(kAbRowBlock), lds_rows, st, \
                         args, (const uint32_t*)ctx->d_ab_sell);                                  \
    else                                                                                          \
      hipLaunchKernelGGL((k_sigma_ab_rows<VT, CC, false>), dim3(g), dim3(kAbRowBlock), lds_rows,  \
                         st, args, (const uint32_t*)ctx->d_ab_sell);                              \
  } while (0)
      switch (C) {
        case 1: SBD_AB_ROWS(1); break;
        case 2: SBD_AB_ROWS(2); break;
        case 4: SBD_AB_ROWS(4); break;
        case 8: SBD_AB_ROWS(8); break;
        case 16: SBD_AB_ROWS(16); break;
        case 24: SBD_AB_ROWS(24); break;
        case 32: SBD_AB_ROWS(32); break;
        default: SBD_AB_ROWS(40); break;
      }
#undef SBD_AB_ROWS
    } else {
      const uint32_t ncb = (uint32_t)((ctx->nb + kAbCols - 1) / kAbCols);
      // about eight workgroups per CU, each over a share of the rows
      const uint64_t want = ((uint64_t)ctx->cu_count * 8 + ncb - 1) / ncb;
      const uint32_t nrg = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(want, ctx->na));
      hipLaunchKernelGGL(k_sigma_ab<VT>, dim3(ncb * nrg), dim3(kAbCols), ctx->ab_lds, st, args,
                         ncb, nrg);
    }
    SBD_HIP_TRY(ctx, hipGetLastError());
    if (tw_dot) {
      const VT* X[1] = {T};
      return vec_multidot<VT>(ctx, X, 1, W, ctx->N, tw_dot, st);
    }
    return SBD_OK;
  }
  if (tw_dot) return sdot_finish(ctx, part, (uint64_t)grid * (kSigBlock / 64), tw_dot, st);
  return SBD_OK;
}

// doubles of dot-product scratch a sigma of this context needs (partials and
// the result slot)
template <typename VT>
static size_t sigma_scratch(const sbd_ctx* ctx) {
  if (sigma_use_elems(ctx, sizeof(VT))) return (size_t)elems_grid(ctx) * (kElemBlock / 64) + 1;
  return (size_t)rows_grid<VT>(ctx, nullptr) * (kSigBlock / 64) + 1;
}

template <typename VT>
static int launch_sigma_with(sbd_ctx* ctx, const VT* hii, const VT* T, VT* W, int accumulate,
                             hipStream_t st, double* tw_dot, double* part) {
  if (!ctx->tables_built) {
    ctx->last_error = "sigma called before sbd_make_helpers";
    return SBD_ERR_STATE;
  }
  if (ctx->na == 0 || ctx->nb == 0) return SBD_OK;
  // the output may not overlap the input vector or hii (the kernels' pointers
  // are restrict-qualified, and rows of T are read after W rows are written)
  const uintptr_t w0 = (uintptr_t)W, w1 = w0 + sizeof(VT) * ctx->N;
  for (uintptr_t x : {(uintptr_t)T, (uintptr_t)hii}) {
    if (x < w1 && w0 < x + sizeof(VT) * ctx->N) {
      ctx->last_error = "sigma: Wb overlaps Wk or hii";
      return SBD_ERR_ARG;
    }
  }
  if (sigma_use_elems(ctx, sizeof(VT)))
    return launch_elems<VT>(ctx, hii, T, W, accumulate, st, tw_dot, part);
  return launch_rows<VT>(ctx, hii, T, W, accumulate, st, tw_dot, part);
}

template <typename VT>
int launch_sigma(sbd_ctx* ctx, const VT* hii, const VT* T, VT* W, int accumulate, hipStream_t st,
                 double* tw_dot) {
  if (tw_dot) {
    int rc = sdot_reserve(ctx, sigma_scratch<VT>(ctx));
    if (rc) return rc;
  }
  return launch_sigma_with<VT>(ctx, hii, T, W, accumulate, st, tw_dot, ctx->d_sdot);
}

template int launch_sigma<double>(sbd_ctx*, const double*, const double*, double*, int,
                                  hipStream_t, double*);
template int launch_sigma<float>(sbd_ctx*, const float*, const float*, float*, int, hipStream_t,
                                 double*);

int sigma_set_lds_limit(sbd_ctx* ctx) {
  // allow the row-staged variants to use the device's full LDS
  int lds = 65536;
  SBD_HIP_TRY(ctx, hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock,
                                         ctx->device));
  const void* fns[] = {
      (const void*)k_sigma_rows<double, true, true, 1, true>,
      (const void*)k_sigma_rows<double, true, false, 1, true>,
      (const void*)k_sigma_rows<double, true, true, 2, true>,
      (const void*)k_sigma_rows<double, true, false, 2, true>,
      (const void*)k_sigma_rows<double, true, true, 4, true>,
      (const void*)k_sigma_rows<double, true, false, 4, true>,
      (const void*)k_sigma_rows<float, true, true, 1, true>,
      (const void*)k_sigma_rows<float, true, false, 1, true>,
      (const void*)k_sigma_rows<float, true, true, 2, true>,
      (const void*)k_sigma_rows<float, true, false, 2, true>,
      (const void*)k_sigma_rows<float, true, true, 4, true>,
      (const void*)k_sigma_rows<float, true, false, 4, true>,
      (const void*)k_sigma_rows<float, true, true, 8, true>,
      (const void*)k_sigma_rows<float, true, false, 8, true>,
      (const void*)k_sigma_rows<double, true, true, 1, false>,
      (const void*)k_sigma_rows<double, true, false, 1, false>,
      (const void*)k_sigma_rows<double, true, true, 2, false>,
      (const void*)k_sigma_rows<double, true, false, 2, false>,
      (const void*)k_sigma_rows<double, true, true, 4, false>,
      (const void*)k_sigma_rows<double, true, false, 4, false>,
      (const void*)k_sigma_rows<float, true, true, 1, false>,
      (const void*)k_sigma_rows<float, true, false, 1, false>,
      (const void*)k_sigma_rows<float, true, true, 2, false>,
      (const void*)k_sigma_rows<float, true, false, 2, false>,
      (const void*)k_sigma_rows<float, true, true, 4, false>,
      (const void*)k_sigma_rows<float, true, false, 4, false>,
      (const void*)k_sigma_rows<float, true, true, 8, false>,
      (const void*)k_sigma_rows<float, true, false, 8, false>,
#define SBD_BD_FNS(VT, R)                                                                         \
  (const void*)k_sigma_bd<VT, R, true, true>, (const void*)k_sigma_bd<VT, R, true, false>,        \
      (const void*)k_sigma_bd<VT, R, false, true>, (const void*)k_sigma_bd<VT, R, false, false>
      SBD_BD_FNS(double, 1), SBD_BD_FNS(double, 2), SBD_BD_FNS(double, 4),
      SBD_BD_FNS(float, 1), SBD_BD_FNS(float, 2), SBD_BD_FNS(float, 4), SBD_BD_FNS(float, 8),
#undef SBD_BD_FNS
      (const void*)k_sigma_bd_batch<double, 1>,
      (const void*)k_sigma_bd_batch<double, 2>,
      (const void*)k_sigma_bd_batch<double, 4>,
      (const void*)k_sigma_batch<double, 1>,
      (const void*)k_sigma_batch<double, 2>,
      (const void*)k_sigma_batch<double, 4>,
      (const void*)k_sigma_ab<double>,
      (const void*)k_sigma_ab<float>,
#define SBD_AB_FNS(VT)                                                                            \
  (const void*)k_sigma_ab_rows<VT, 1, true>, (const void*)k_sigma_ab_rows<VT, 2, true>,           \
      (const void*)k_sigma_ab_rows<VT, 4, true>, (const void*)k_sigma_ab_rows<VT, 8, true>,       \
      (const void*)k_sigma_ab_rows<VT, 1, false>, (const void*)k_sigma_ab_rows<VT, 2, false>,     \
      (const void*)k_sigma_ab_rows<VT, 4, false>, (const void*)k_sigma_ab_rows<VT, 8, false>, \
      (const void*)k_sigma_ab_rows<VT, 16, true>, (const void*)k_sigma_ab_rows<VT, 16, false>, \
      (const void*)k_sigma_ab_rows<VT, 24, true>, (const void*)k_sigma_ab_rows<VT, 24, false>, \
      (const void*)k_sigma_ab_rows<VT, 32, true>, (const void*)k_sigma_ab_rows<VT, 32, false>, \
      (const void*)k_sigma_ab_rows<VT, 40, true>, (const void*)k_sigma_ab_rows<VT, 40, false>
      SBD_AB_FNS(double), SBD_AB_FNS(float)
#undef SBD_AB_FNS
  };
  for (const void* f : fns)
    (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
  (void)hipGetLastError();
  ctx->lds_max = lds;
  return SBD_OK;
}

}  // namespace sbdx
