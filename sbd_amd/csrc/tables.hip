// Excitation tables on the GPU: the one-rank equivalent of MakeHelpers /
// GenerateExcitation (reference chemistry/tpb/helper.h:185-342, 729-873).
//
// For every bra string (one thread each) all ket strings are scanned in
// ascending order through 256-string LDS tiles; XOR + popcount gives the
// Hamming distance (difference(), determinants.h:380-404): 2 -> single,
// 4 -> double.  Pass 1 counts, a device scan gives CSR offsets, pass 2 writes
// records in ket-ascending order (the order GenerateExcitation pushes them).
// Each record carries the orbital difference (OrbitalDifference,
// determinants.h:406-449) and the sign/value parts that do not depend on the
// other spin's string, so the sigma kernel only has to add the other spin's
// parity.
#include <hipcub/hipcub.hpp>

#include <algorithm>

#include "sbd_internal.hpp"

namespace sbdx {

constexpr int kTabBlock = 256;

__global__ __launch_bounds__(kTabBlock) void k_tab_count(const uint64_t* __restrict__ s, uint32_t n,
                                                         uint32_t* __restrict__ sc,
                                                         uint32_t* __restrict__ dc) {
  __shared__ uint64_t tile[kTabBlock];
  const uint32_t i = blockIdx.x * kTabBlock + threadIdx.x;
  const uint64_t me = i < n ? s[i] : 0;
  uint32_t ns = 0, nd = 0;
  for (uint32_t t0 = 0; t0 < n; t0 += kTabBlock) {
    __syncthreads();
    if (t0 + threadIdx.x < n) tile[threadIdx.x] = s[t0 + threadIdx.x];
    __syncthreads();
    const uint32_t lim = min((uint32_t)kTabBlock, n - t0);
    for (uint32_t j = 0; j < lim; j++) {
      const int d = __popcll(me ^ tile[j]);
      ns += (d == 2);
      nd += (d == 4);
    }
  }
  if (i < n) {
    sc[i] = ns;
    dc[i] = nd;
  }
}

// Record of the pair (bra, ket) of spin s (0 alpha, 1 beta).
__device__ inline PairRec make_rec(uint64_t bra, uint64_t ket, uint32_t ket_idx, int s, int L2,
                                   const double* __restrict__ I1, const double* __restrict__ I2,
                                   bool dbl) {
  uint64_t xb = bra & ~ket, yb = ket & ~bra;
  const int x0 = __ffsll((unsigned long long)xb) - 1;
  const int y0 = __ffsll((unsigned long long)yb) - 1;
  PairRec r;
  r.ket = ket_idx;
  if (!dbl) {
    // same-spin bits strictly between x0 and y0 (parity(), determinants.h:242-294)
    const int lo = x0 < y0 ? x0 : y0, hi = x0 < y0 ? y0 : x0;
    const int par = __popcll(bra & bits_range(lo + 1, hi)) & 1;
    r.code = (uint32_t)x0 | ((uint32_t)y0 << 12) | ((uint32_t)par << 24);
    const int i = 2 * x0 + s, a = 2 * y0 + s;
    r.val = I1[a * L2 + i];  // OneExcite: energy = I1.Value(a, i)
  } else {
    xb &= xb - 1;
    yb &= yb - 1;
    const int x1 = __ffsll((unsigned long long)xb) - 1;
    const int y1 = __ffsll((unsigned long long)yb) - 1;
    const uint32_t lo1 = x0 < y0 ? x0 : y0, hi1 = x0 < y0 ? y0 : x0;
    const uint32_t lo2 = x1 < y1 ? x1 : y1, hi2 = x1 < y1 ? y1 : x1;
    r.code = lo1 | (hi1 << 6) | (lo2 << 12) | (hi2 << 18);
    // TwoExcite (determinants.h:500-518) with i=2x0+s, j=2x1+s, a=2y0+s, b=2y1+s:
    // I=i, J=j, A=a, B=b since x0<x1, y0<y1.
    const int I = 2 * x0 + s, J = 2 * x1 + s, A = 2 * y0 + s, B = 2 * y1 + s;
    int par = __popcll(bra & bits_range((x0 < y0 ? x0 : y0) + 1, x0 < y0 ? y0 : x0));
    par += __popcll(bra & bits_range((x1 < y1 ? x1 : y1) + 1, x1 < y1 ? y1 : x1));
    double sgn = (par & 1) ? -1.0 : 1.0;
    if (A > J || B < I) sgn *= -1.0;
    r.val = sgn * (eri_value(I2, A, I, B, J) - eri_value(I2, A, J, B, I));
  }
  return r;
}

__global__ __launch_bounds__(kTabBlock) void k_tab_fill(const uint64_t* __restrict__ s, uint32_t n,
                                                        int s_spin, int L,
                                                        const uint32_t* __restrict__ s_off,
                                                        const uint32_t* __restrict__ d_off,
                                                        PairRec* __restrict__ s_rec,
                                                        PairRec* __restrict__ d_rec,
                                                        const double* __restrict__ I1,
                                                        const double* __restrict__ I2) {
  __shared__ uint64_t tile[kTabBlock];
  const uint32_t i = blockIdx.x * kTabBlock + threadIdx.x;
  const uint64_t me = i < n ? s[i] : 0;
  uint32_t ps = i < n ? s_off[i] : 0, pd = i < n ? d_off[i] : 0;
  for (uint32_t t0 = 0; t0 < n; t0 += kTabBlock) {
    __syncthreads();
    if (t0 + threadIdx.x < n) tile[threadIdx.x] = s[t0 + threadIdx.x];
    __syncthreads();
    if (i >= n) continue;
    const uint32_t lim = min((uint32_t)kTabBlock, n - t0);
    for (uint32_t j = 0; j < lim; j++) {
      const uint64_t o = tile[j];
      const int d = __popcll(me ^ o);
      if (d == 2) s_rec[ps++] = make_rec(me, o, t0 + j, s_spin, 2 * L, I1, I2, false);
      else if (d == 4) d_rec[pd++] = make_rec(me, o, t0 + j, s_spin, 2 * L, I1, I2, true);
    }
  }
}

// Sliced-ELL copy of a beta doubles CSR table (one thread per block-sorted
// position), with the alpha-string sign mask of each record expanded from its
// code: a beta excitation between spin orbitals 2u+1 and 2w+1 picks up the
// parity of the alpha orbitals [u+1, w+1).
__global__ void k_tab_ell(uint32_t n, const uint32_t* __restrict__ off,
                          const PairRec* __restrict__ rec, const uint32_t* __restrict__ ell_off,
                          const EllPos* __restrict__ ell_pos, int packed,
                          EllSlot64* __restrict__ ell64, EllSlot32* __restrict__ ell32,
                          uint32_t* __restrict__ ell_ket) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t s = i / kEllSlice, l = i % kEllSlice;
  const uint32_t col = ell_pos[i].col;
  const uint32_t o0 = off[col], cnt = off[col + 1] - o0;
  const uint32_t w = (ell_off[s + 1] - ell_off[s]) / kEllSlice;
  for (uint32_t u = 0; u < w; u++) {
    uint32_t ket = col;
    double val = 0.0;
    uint64_t mask = 0;
    if (u < cnt) {
      const PairRec r = rec[o0 + u];
      const uint32_t c = r.code;
      ket = r.ket;
      val = r.val;
      mask = bits_range(code_f(c, 0) + 1, code_f(c, 1) + 1) ^
             bits_range(code_f(c, 2) + 1, code_f(c, 3) + 1);
    }
    const uint64_t mk = packed ? (mask | ((uint64_t)ket << kEllKetShift)) : mask;
    const size_t slot = ell_off[s] + (size_t)u * kEllSlice + l;
    ell64[slot] = EllSlot64{val, mk};
    ell32[slot] = EllSlot32{(float)val, (uint32_t)mk, (uint32_t)(mk >> 32)};
    if (!packed) ell_ket[slot] = ket;
  }
}

// G[s][x][y][j] = Value(a,i,j,j) - Value(a,j,j,i), a = 2y+s, i = 2x+s: the
// per-occupied-orbital term of OneExcite (determinants.h:486-495).
__global__ void k_gtable(const double* __restrict__ I2, int L, double* __restrict__ G) {
  const int L2 = 2 * L;
  const int64_t total = (int64_t)2 * L * L * L2;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int j = (int)(t % L2);
    int64_t r = t / L2;
    const int y = (int)(r % L);
    r /= L;
    const int x = (int)(r % L);
    const int s = (int)(r / L);
    const int a = 2 * y + s, i = 2 * x + s;
    G[t] = (eri_value(I2, a, i, j, j) - eri_value(I2, a, j, j, i));
  }
}

// Vab[((xa L + ya) L + xb) L + yb]: the integral part of TwoExcite(D, i, j, a, b)
// (determinants.h:500-518) for i = 2xa, a = 2ya (alpha) and j = 2xb+1, b = 2yb+1
// (beta), with the sign that depends only on the order of the four spin
// orbitals (-1 if A > J or B < I) applied: ((AI|BJ) - (AJ|BI)) negated or not,
// by the same operations.  The sigma multiplies it by the parity sign only.
__global__ void k_abtable(const double* __restrict__ I2, int L, double* __restrict__ V) {
  const int64_t total = (int64_t)L * L * L * L;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int yb = (int)(t % L), xb = (int)(t / L % L), ya = (int)(t / L / L % L),
              xa = (int)(t / L / L / L);
    const int i = 2 * xa, j = 2 * xb + 1, a = 2 * ya, b = 2 * yb + 1;
    const int I = i < j ? i : j, J = i < j ? j : i;
    const int A = a < b ? a : b, B = a < b ? b : a;
    double sgn = 1.0;
    if (A > J || B < I) sgn *= -1.0;
    V[t] = sgn * (eri_value(I2, A, I, B, J) - eri_value(I2, A, J, B, I));
  }
}

int build_abtable(sbd_ctx* ctx, hipStream_t st) {
  if (ctx->d_Vab || ctx->ta.n_s == 0 || ctx->tb.n_s == 0) return SBD_OK;
  const int L = ctx->norb;
  const size_t total = (size_t)L * L * L * L;
  SBD_HIP_TRY(ctx, hipMalloc(&ctx->d_Vab, sizeof(double) * total));
  const unsigned grid = (unsigned)std::min<size_t>((total + 255) / 256, 65535);
  hipLaunchKernelGGL(k_abtable, dim3(grid), dim3(256), 0, st, ctx->d_eri, L, ctx->d_Vab);
  SBD_HIP_TRY(ctx, hipGetLastError());
  return SBD_OK;
}

static int build_spin(sbd_ctx* ctx, const uint64_t* d_s, uint32_t n, int spin, SpinTables& T,
                      hipStream_t st) {
  uint32_t *sc = nullptr, *dc = nullptr;
  SBD_HIP_TRY(ctx, hipMalloc(&sc, sizeof(uint32_t) * (n + 1)));
  SBD_HIP_TRY(ctx, hipMalloc(&dc, sizeof(uint32_t) * (n + 1)));
  SBD_HIP_TRY(ctx, hipMemsetAsync(sc + n, 0, sizeof(uint32_t), st));
  SBD_HIP_TRY(ctx, hipMemsetAsync(dc + n, 0, sizeof(uint32_t), st));
  const int grid = (n + kTabBlock - 1) / kTabBlock;
  hipLaunchKernelGGL(k_tab_count, dim3(grid), dim3(kTabBlock), 0, st, d_s, n, sc, dc);
  SBD_HIP_TRY(ctx, hipGetLastError());
  // max records per bra and totals (host side, small)
  std::vector<uint32_t> hs(n), hd(n);
  SBD_HIP_TRY(ctx, hipMemcpyAsync(hs.data(), sc, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, st));
  SBD_HIP_TRY(ctx, hipMemcpyAsync(hd.data(), dc, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, st));
  SBD_HIP_TRY(ctx, hipStreamSynchronize(st));
  uint64_t tot_s = 0, tot_d = 0;
  uint32_t mx_s = 0, mx_d = 0;
  for (uint32_t i = 0; i < n; i++) {
    tot_s += hs[i];
    tot_d += hd[i];
    mx_s = hs[i] > mx_s ? hs[i] : mx_s;
    mx_d = hd[i] > mx_d ? hd[i] : mx_d;
  }
  if (tot_s >= 0xFFFFFFF0ULL || tot_d >= 0xFFFFFFF0ULL) {
    ctx->last_error = "excitation table exceeds 2^32 records";
    (void)hipFree(sc);
    (void)hipFree(dc);
    return SBD_ERR_ARG;
  }
  SBD_HIP_TRY(ctx, hipMalloc(&T.s_off, sizeof(uint32_t) * (n + 1)));
  SBD_HIP_TRY(ctx, hipMalloc(&T.d_off, sizeof(uint32_t) * (n + 1)));
  size_t tmp_bytes = 0, tmp2 = 0;
  (void)hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, sc, T.s_off, n + 1, st);
  (void)hipcub::DeviceScan::ExclusiveSum(nullptr, tmp2, dc, T.d_off, n + 1, st);
  if (tmp2 > tmp_bytes) tmp_bytes = tmp2;
  void* tmp = nullptr;
  SBD_HIP_TRY(ctx, hipMalloc(&tmp, tmp_bytes + 16));
  SBD_HIP_TRY(ctx, hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, sc, T.s_off, n + 1, st));
  SBD_HIP_TRY(ctx, hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, dc, T.d_off, n + 1, st));
  SBD_HIP_TRY(ctx, hipMalloc(&T.s_rec, sizeof(PairRec) * (tot_s + 1)));
  SBD_HIP_TRY(ctx, hipMalloc(&T.d_rec, sizeof(PairRec) * (tot_d + 1)));
  hipLaunchKernelGGL(k_tab_fill, dim3(grid), dim3(kTabBlock), 0, st, d_s, n, spin, ctx->norb,
                     T.s_off, T.d_off, T.s_rec, T.d_rec, ctx->d_I1, ctx->d_eri);
  SBD_HIP_TRY(ctx, hipGetLastError());
  SBD_HIP_TRY(ctx, hipStreamSynchronize(st));
  (void)hipFree(tmp);
  (void)hipFree(sc);
  (void)hipFree(dc);
  T.n_s = tot_s;
  T.n_d = tot_d;
  T.s_off_host.assign(n + 1, 0);
  for (uint32_t i = 0; i < n; i++) T.s_off_host[i + 1] = T.s_off_host[i] + hs[i];
  T.max_s = mx_s;
  T.max_d = mx_d;
  // strings with at least one single excitation
  std::vector<uint32_t> ws;
  for (uint32_t i = 0; i < n; i++)
    if (hs[i]) ws.push_back(i);
  T.n_with_s = (uint32_t)ws.size();
  SBD_HIP_TRY(ctx, hipMalloc(&T.with_s, sizeof(uint32_t) * (ws.size() + 1)));
  if (!ws.empty())
    SBD_HIP_TRY(ctx, hipMemcpy(T.with_s, ws.data(), sizeof(uint32_t) * ws.size(),
                               hipMemcpyHostToDevice));
  // sliced-ELL copy of the beta doubles over block-sorted positions
  if (spin == 1) {
    std::vector<uint32_t> col(n);
    for (uint32_t g0 = 0; g0 < n; g0 += kColBlock) {
      const uint32_t g1 = std::min<uint64_t>(n, (uint64_t)g0 + kColBlock);
      for (uint32_t i = g0; i < g1; i++) col[i] = i;
      // strings with single excitations first (the sigma waits for their
      // records only in the wavefronts that hold them), each group in
      // descending order of its doubles count
      std::stable_sort(col.begin() + g0, col.begin() + g1, [&](uint32_t x, uint32_t y) {
        const bool sx = hs[x] != 0, sy = hs[y] != 0;
        return sx != sy ? sx : hd[x] > hd[y];
      });
    }
    std::vector<EllPos> pos(n);
    std::vector<uint32_t> soff(n + 1, 0);
    for (uint32_t i = 0; i < n; i++) soff[i + 1] = soff[i] + hs[i];
    for (uint32_t i = 0; i < n; i++)
      pos[i] = EllPos{col[i], hd[col[i]], soff[col[i]], soff[col[i] + 1]};
    SBD_HIP_TRY(ctx, hipMalloc(&T.ell_pos, sizeof(EllPos) * n));
    SBD_HIP_TRY(ctx,
                hipMemcpy(T.ell_pos, pos.data(), sizeof(EllPos) * n, hipMemcpyHostToDevice));

    T.n_slices = (n + kEllSlice - 1) / kEllSlice;
    std::vector<uint32_t> eoff(T.n_slices + 1, 0);
    for (uint32_t s = 0; s < T.n_slices; s++) {
      uint32_t w = 0;
      for (uint32_t i = s * kEllSlice; i < std::min(n, (s + 1) * kEllSlice); i++)
        w = std::max(w, hd[col[i]]);
      eoff[s + 1] = eoff[s] + w * kEllSlice;
    }
    SBD_HIP_TRY(ctx, hipMalloc(&T.ell_off, sizeof(uint32_t) * (T.n_slices + 1)));
    SBD_HIP_TRY(ctx, hipMemcpy(T.ell_off, eoff.data(), sizeof(uint32_t) * (T.n_slices + 1),
                               hipMemcpyHostToDevice));
    T.n_slots = (uint64_t)eoff[T.n_slices];
    // packed slots: alpha strings below bit kEllKetShift, kets above it
    T.ell_packed = ctx->norb <= kEllKetShift && n <= (1u << (64 - kEllKetShift));
    // one zero slot (ket 0, value 0, mask 0) after the last slice: the kernel
    // reads it in place of slots past a slice's width, adding an exact zero
    const size_t ns = T.n_slots + 1;
    SBD_HIP_TRY(ctx, hipMalloc(&T.ell64, sizeof(EllSlot64) * ns));
    SBD_HIP_TRY(ctx, hipMalloc(&T.ell32, sizeof(EllSlot32) * ns));
    if (!T.ell_packed) SBD_HIP_TRY(ctx, hipMalloc(&T.ell_ket, sizeof(uint32_t) * ns));
    // all slots zero first: the lanes of the last slice past the last string
    // are never filled and must read as zero slots (value 0, mask 0, ket 0)
    SBD_HIP_TRY(ctx, hipMemsetAsync(T.ell64, 0, sizeof(EllSlot64) * ns, st));
    SBD_HIP_TRY(ctx, hipMemsetAsync(T.ell32, 0, sizeof(EllSlot32) * ns, st));
    if (T.ell_ket) SBD_HIP_TRY(ctx, hipMemsetAsync(T.ell_ket, 0, sizeof(uint32_t) * ns, st));
    hipLaunchKernelGGL(k_tab_ell, dim3((n + 255) / 256), dim3(256), 0, st, n, T.d_off, T.d_rec,
                       T.ell_off, T.ell_pos, T.ell_packed ? 1 : 0, T.ell64, T.ell32, T.ell_ket);
    SBD_HIP_TRY(ctx, hipGetLastError());
    // per-string records: string and single-excitation range
    std::vector<ColRec> cr(n);
    {
      std::vector<uint64_t> bs(n);
      SBD_HIP_TRY(ctx, hipMemcpy(bs.data(), d_s, sizeof(uint64_t) * n, hipMemcpyDeviceToHost));
      uint32_t o = 0;
      for (uint32_t i = 0; i < n; i++) {
        cr[i] = ColRec{bs[i], o, o + hs[i]};
        o += hs[i];
      }
    }
    SBD_HIP_TRY(ctx, hipMalloc(&T.col_rec, sizeof(ColRec) * n));
    SBD_HIP_TRY(ctx, hipMemcpy(T.col_rec, cr.data(), sizeof(ColRec) * n, hipMemcpyHostToDevice));
  }
  // bra string of every single record
  std::vector<uint32_t> sb(tot_s + 1, 0);
  for (uint32_t i = 0, q = 0; i < n; i++)
    for (uint32_t k = 0; k < hs[i]; k++) sb[q++] = i;
  SBD_HIP_TRY(ctx, hipMalloc(&T.s_bra, sizeof(uint32_t) * sb.size()));
  SBD_HIP_TRY(ctx, hipMemcpy(T.s_bra, sb.data(), sizeof(uint32_t) * sb.size(),
                             hipMemcpyHostToDevice));
  SBD_HIP_TRY(ctx, hipStreamSynchronize(st));
  return SBD_OK;
}

int build_gtable(sbd_ctx* ctx, hipStream_t st) {
  const int L = ctx->norb;
  const size_t total = (size_t)2 * L * L * 2 * L;
  SBD_HIP_TRY(ctx, hipMalloc(&ctx->d_G, sizeof(double) * total));
  hipLaunchKernelGGL(k_gtable, dim3((total + 255) / 256), dim3(256), 0, st, ctx->d_eri, L,
                     ctx->d_G);
  SBD_HIP_TRY(ctx, hipGetLastError());
  return SBD_OK;
}

int build_tables(sbd_ctx* ctx, hipStream_t st) {
  int rc = build_spin(ctx, ctx->d_a, (uint32_t)ctx->na, 0, ctx->ta, st);
  if (rc) return rc;
  rc = build_spin(ctx, ctx->d_b, (uint32_t)ctx->nb, 1, ctx->tb, st);
  if (rc) return rc;
  return SBD_OK;
}

}  // namespace sbdx
