// Measurements on a solved vector W: string weights, occupation density and
// carry-over strings.
//
// Reference: DiagonalAdetReducedDensityMatrix / DiagonalBdetReducedDensityMatrix
// (chemistry/tpb/rdmat.h:11-58, 136-182), OccupationDensity
// (chemistry/tpb/occupation.h:10-95), CarryOverAdet / CarryOverBdet
// (chemistry/tpb/rdmat.h:60-134 and their beta twins), as sbd::tpb::diag
// calls them (chemistry/tpb/sbdiag.h:666-697, 745-777).
//
// The one device pass is the string weights: Da[ia] = sum_ib |W[ia, ib]|^2 (a
// row sum) and Db[ib] = sum_ia |W[ia, ib]|^2 (a column sum), both with
// per-thread compensated sums combined in a fixed order, so the results are
// the same on every run.  Occupations and carry-over sets follow on the host
// from these na + nb numbers: the occupation of orbital o is the sum of the
// weights of the strings that hold o (the reference sums the same terms per
// determinant).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <numeric>
#include <vector>

#include "sbd_internal.hpp"

namespace sbdx {

namespace {

constexpr int kWBlock = 256;
constexpr int kColRows = 128;  // rows per column-partial chunk

struct Kahan2 {
  double s = 0.0, c = 0.0;
  __device__ inline void add(double x) {
    const double y = x - c;
    const double t = s + y;
    c = (t - s) - y;
    s = t;
  }
};

// Row sums: one workgroup per row (grid-stride); lanes stride the row.
__global__ __launch_bounds__(kWBlock) void k_row_weights(const double* __restrict__ W,
                                                         uint64_t na, uint64_t nb,
                                                         double* __restrict__ Da) {
  __shared__ double red[kWBlock];
  for (uint64_t ia = blockIdx.x; ia < na; ia += gridDim.x) {
    const double* __restrict__ row = W + ia * nb;
    Kahan2 k;
    for (uint64_t ib = threadIdx.x; ib < nb; ib += kWBlock) k.add(row[ib] * row[ib]);
    __syncthreads();
    red[threadIdx.x] = k.s;
    __syncthreads();
    for (int w = kWBlock / 2; w > 0; w >>= 1) {
      if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) Da[ia] = red[0];
  }
}

// Column partial sums over chunks of kColRows rows: thread (chunk, column)
// adds its column's squares (coalesced row by row).
__global__ __launch_bounds__(kWBlock) void k_col_partials(const double* __restrict__ W,
                                                          uint64_t na, uint64_t nb,
                                                          double* __restrict__ part) {
  const uint64_t ib = blockIdx.x * (uint64_t)kWBlock + threadIdx.x;
  const uint64_t chunk = blockIdx.y;
  if (ib >= nb) return;
  const uint64_t r0 = chunk * kColRows;
  const uint64_t r1 = std::min<uint64_t>(na, r0 + kColRows);
  Kahan2 k;
  for (uint64_t ia = r0; ia < r1; ia++) {
    const double x = W[ia * nb + ib];
    k.add(x * x);
  }
  part[chunk * nb + ib] = k.s;
}

// Column sums: each column's chunk partials in chunk order.
__global__ __launch_bounds__(kWBlock) void k_col_sum(const double* __restrict__ part,
                                                     uint64_t nchunk, uint64_t nb,
                                                     double* __restrict__ Db) {
  const uint64_t ib = blockIdx.x * (uint64_t)kWBlock + threadIdx.x;
  if (ib >= nb) return;
  Kahan2 k;
  for (uint64_t c = 0; c < nchunk; c++) k.add(part[c * nb + ib]);
  Db[ib] = k.s;
}

// The carry-over rule of CarryOverAdet (rdmat.h:63-99, threshold form) and
// (rdmat.h:102-134, count form), on weights D: strings by descending weight
// (ties by index); the threshold form keeps the strings up to and including
// the first whose cumulative weight exceeds 1 - threshold (std::upper_bound
// + 1; all strings when none does, where the reference would read past the
// end); the count form keeps `kept` strings.  Reports 1 - the kept weight.
void carry_select(const std::vector<double>& D, double threshold, bool by_count, uint64_t kept,
                  std::vector<uint32_t>& idx, double* truncated) {
  const uint64_t n = D.size();
  std::vector<uint32_t> order(n);
  std::iota(order.begin(), order.end(), 0u);
  std::stable_sort(order.begin(), order.end(),
                   [&](uint32_t i, uint32_t j) { return D[i] > D[j]; });
  uint64_t m = 0;
  double sum = 0.0;
  if (!by_count) {
    const double target = 1.0 - threshold;
    uint64_t first_above = n;  // upper_bound: first cumulative sum > target
    for (uint64_t i = 0; i < n; i++) {
      sum += D[order[i]];
      if (sum > target) {
        first_above = i;
        break;
      }
    }
    m = std::min<uint64_t>(n, first_above + 1);
    sum = 0.0;
    for (uint64_t i = 0; i < m; i++) sum += D[order[i]];
  } else {
    m = std::min<uint64_t>(n, kept);
    for (uint64_t i = 0; i < m; i++) sum += D[order[i]];
  }
  idx.assign(order.begin(), order.begin() + m);
  if (truncated) *truncated = 1.0 - sum;
}

}  // namespace

int string_weights(sbd_ctx* ctx, const double* W, double* Da_host, double* Db_host,
                   hipStream_t st) {
  const uint64_t na = ctx->na, nb = ctx->nb;
  if (!na || !nb) return SBD_OK;
  double *dA = nullptr, *dB = nullptr, *part = nullptr;
  const uint64_t nchunk = (na + kColRows - 1) / kColRows;
  auto cleanup = [&]() {
    if (dA) (void)hipFree(dA);
    if (dB) (void)hipFree(dB);
    if (part) (void)hipFree(part);
  };
  if (hipMalloc(&dA, 8 * na) != hipSuccess || hipMalloc(&dB, 8 * nb) != hipSuccess ||
      hipMalloc(&part, 8 * nchunk * nb) != hipSuccess) {
    cleanup();
    return set_err(ctx, SBD_ERR_MEMORY, "device allocation for string weights failed");
  }
  const unsigned rgrid = (unsigned)std::min<uint64_t>(na, 65535);
  hipLaunchKernelGGL(k_row_weights, dim3(rgrid), dim3(kWBlock), 0, st, W, na, nb, dA);
  const unsigned cblk = (unsigned)((nb + kWBlock - 1) / kWBlock);
  hipLaunchKernelGGL(k_col_partials, dim3(cblk, (unsigned)nchunk), dim3(kWBlock), 0, st, W, na,
                     nb, part);
  hipLaunchKernelGGL(k_col_sum, dim3(cblk), dim3(kWBlock), 0, st, part, nchunk, nb, dB);
  int rc = SBD_OK;
  if (hipGetLastError() != hipSuccess ||
      hipMemcpyAsync(Da_host, dA, 8 * na, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(Db_host, dB, 8 * nb, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    rc = set_err(ctx, SBD_ERR_DEVICE, "string weights: device error");
  cleanup();
  return rc;
}

void occupation_from_weights(const sbd_ctx* ctx, const double* Da, const double* Db,
                             double* density) {
  const int L = ctx->norb;
  std::vector<Kahan2> acc(2 * L);
  for (int o = 0; o < 2 * L; o++) acc[o] = Kahan2{};
  // host side of Kahan2::add is the same arithmetic
  auto add = [](Kahan2& k, double x) {
    const double y = x - k.c;
    const double t = k.s + y;
    k.c = (t - k.s) - y;
    k.s = t;
  };
  for (uint64_t ia = 0; ia < ctx->na; ia++) {
    const uint64_t A = ctx->a_host[ia];
    for (int o = 0; o < L; o++)
      if ((A >> o) & 1ULL) add(acc[2 * o], Da[ia]);
  }
  for (uint64_t ib = 0; ib < ctx->nb; ib++) {
    const uint64_t B = ctx->b_host[ib];
    for (int o = 0; o < L; o++)
      if ((B >> o) & 1ULL) add(acc[2 * o + 1], Db[ib]);
  }
  for (int o = 0; o < 2 * L; o++) density[o] = acc[o].s;
}

namespace {

// Strings of the determinants whose weight exceeds a cutoff
// (SinglesExtendHalfdets, chemistry/tpb/extend.h:14-58): rowflag[ia] /
// colflag[ib] set when any determinant of that string has W^2 > cutoff; the
// weight above the cutoff per row in rowsum (summed on the host in row order).
__global__ __launch_bounds__(kWBlock) void k_cut_flags(const double* __restrict__ W, uint64_t na,
                                                       uint64_t nb, double cutoff,
                                                       uint32_t* __restrict__ rowflag,
                                                       uint32_t* __restrict__ colflag,
                                                       double* __restrict__ rowsum) {
  __shared__ double red[kWBlock];
  __shared__ uint32_t any;
  for (uint64_t ia = blockIdx.x; ia < na; ia += gridDim.x) {
    __syncthreads();
    if (threadIdx.x == 0) any = 0;
    __syncthreads();
    const double* __restrict__ row = W + ia * nb;
    Kahan2 k;
    uint32_t f = 0;
    for (uint64_t ib = threadIdx.x; ib < nb; ib += kWBlock) {
      const double w2 = row[ib] * row[ib];
      if (w2 > cutoff) {
        k.add(w2);
        f = 1;
        colflag[ib] = 1;  // only ever set to 1: the result does not depend on order
      }
    }
    if (f) any = 1;
    red[threadIdx.x] = k.s;
    __syncthreads();
    for (int w = kWBlock / 2; w > 0; w >>= 1) {
      if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      rowflag[ia] = any;
      rowsum[ia] = red[0];
    }
  }
}

// Each string and all of its single excitations (single_from_hdet over the
// occupied x unoccupied orbitals), then sorted ascending without duplicates
// (sort_bitarray, framework/bit_manipulation.h:421-449).
std::vector<uint64_t> extend_singles(const std::vector<uint64_t>& in, int norb) {
  std::vector<uint64_t> out;
  out.reserve(in.size() * (1 + (size_t)norb * norb / 4));
  const uint64_t full = norb >= 64 ? ~0ULL : ((1ULL << norb) - 1ULL);
  for (uint64_t s : in) {
    out.push_back(s);
    for (uint64_t occ = s; occ; occ &= occ - 1) {
      const uint64_t i = occ & (~occ + 1);
      for (uint64_t vir = ~s & full; vir; vir &= vir - 1) {
        const uint64_t a = vir & (~vir + 1);
        out.push_back((s ^ i) | a);
      }
    }
  }
  std::sort(out.begin(), out.end());
  out.erase(std::unique(out.begin(), out.end()), out.end());
  return out;
}

int copy_out(const std::vector<uint64_t>& v, uint64_t* out, size_t cap, size_t* n_out) {
  *n_out = v.size();
  if (v.size() > cap) return SBD_ERR_ARG;
  std::copy(v.begin(), v.end(), out);
  return SBD_OK;
}

}  // namespace

}  // namespace sbdx

using namespace sbdx;

extern "C" {

int sbd_singles_extend_strings(int norb, const uint64_t* in, size_t n, uint64_t* out, size_t cap,
                               size_t* n_out) {
  if (norb < 1 || norb > 64 || (n && !in) || !out || !n_out) return SBD_ERR_ARG;
  return copy_out(extend_singles(std::vector<uint64_t>(in, in + n), norb), out, cap, n_out);
}

int sbd_singles_extend(sbd_ctx* ctx, const double* W, double cutoff, uint64_t* a_out,
                       size_t cap_a, size_t* na_out, uint64_t* b_out, size_t cap_b,
                       size_t* nb_out, double* total_weight, void* stream) {
  if (!ctx || !W || !a_out || !na_out || !b_out || !nb_out) return SBD_ERR_ARG;
  const uint64_t na = ctx->na, nb = ctx->nb;
  hipStream_t st = (hipStream_t)stream;
  uint32_t *rf = nullptr, *cf = nullptr;
  double* rs = nullptr;
  auto cleanup = [&]() {
    if (rf) (void)hipFree(rf);
    if (cf) (void)hipFree(cf);
    if (rs) (void)hipFree(rs);
  };
  if (hipMalloc(&rf, 4 * na) != hipSuccess || hipMalloc(&cf, 4 * nb) != hipSuccess ||
      hipMalloc(&rs, 8 * na) != hipSuccess) {
    cleanup();
    return set_err(ctx, SBD_ERR_MEMORY, "device allocation for singles extension failed");
  }
  std::vector<uint32_t> hrf(na), hcf(nb);
  std::vector<double> hrs(na);
  int rc = SBD_OK;
  if (hipMemsetAsync(cf, 0, 4 * nb, st) != hipSuccess) rc = SBD_ERR_DEVICE;
  if (!rc) {
    hipLaunchKernelGGL(k_cut_flags, dim3((unsigned)std::min<uint64_t>(na, 65535)), dim3(kWBlock),
                       0, st, W, na, nb, cutoff, rf, cf, rs);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(hrf.data(), rf, 4 * na, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(hcf.data(), cf, 4 * nb, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(hrs.data(), rs, 8 * na, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      rc = SBD_ERR_DEVICE;
  }
  cleanup();
  if (rc) return set_err(ctx, rc, "singles extension: device error");
  std::vector<uint64_t> ka, kb;
  double tw = 0.0;
  for (uint64_t i = 0; i < na; i++) {
    tw += hrs[i];
    if (hrf[i]) ka.push_back(ctx->a_host[i]);
  }
  for (uint64_t i = 0; i < nb; i++)
    if (hcf[i]) kb.push_back(ctx->b_host[i]);
  if (total_weight) *total_weight = tw;
  rc = copy_out(extend_singles(ka, ctx->norb), a_out, cap_a, na_out);
  const int rc2 = copy_out(extend_singles(kb, ctx->norb), b_out, cap_b, nb_out);
  if (rc || rc2) return set_err(ctx, SBD_ERR_ARG, "singles extension: output capacity too small");
  return SBD_OK;
}

int sbd_sort_strings(uint64_t* s, size_t n, size_t* n_out) {
  if ((n && !s) || !n_out) return SBD_ERR_ARG;
  std::sort(s, s + n);
  *n_out = (size_t)(std::unique(s, s + n) - s);
  return SBD_OK;
}

int sbd_decode_strings_bin(const char* path, int norb, uint64_t* out, size_t cap,
                           size_t* n_out) {
  if (!path || norb < 1 || norb > 64 || !n_out) return SBD_ERR_ARG;
  FILE* f = std::fopen(path, "rb");
  if (!f) return SBD_ERR_IO;
  const size_t nbytes = (norb + 7) / 8, pad = 8 * nbytes - norb;
  std::vector<uint8_t> rec(nbytes);
  size_t n = 0;
  int rc = SBD_OK;
  while (std::fread(rec.data(), 1, nbytes, f) == nbytes) {
    // bits most significant first; the first `pad` bits are skipped; the last
    // bit read is orbital 0 (DecodeAlphaDetsBits, makedeterminants.h:61-103)
    uint64_t v = 0;
    size_t bit = 0;
    for (size_t k = 0; k < nbytes; k++)
      for (int i = 7; i >= 0; i--, bit++)
        if (bit >= pad) v = (v << 1) | ((rec[k] >> i) & 1u);
    if (n < cap && out) out[n] = v;
    else rc = SBD_ERR_ARG;
    n++;
  }
  std::fclose(f);
  *n_out = n;
  return rc;
}

int sbd_encode_strings_bin(const char* path, int norb, const uint64_t* s, size_t n) {
  if (!path || norb < 1 || norb > 64 || (n && !s)) return SBD_ERR_ARG;
  FILE* f = std::fopen(path, "wb");
  if (!f) return SBD_ERR_IO;
  const size_t nbytes = (norb + 7) / 8;
  std::vector<uint8_t> rec(nbytes);
  bool ok = true;
  for (size_t q = 0; q < n; q++) {
    // orbital norb - 1 - j at bit j of the record, most significant bit of
    // byte 0 first, padding at the end (the carryover.bin writer of
    // apps/chemistry_tpb_selected_basis_diagonalization/main.cc:245-260)
    std::fill(rec.begin(), rec.end(), 0);
    for (int j = 0; j < norb; j++)
      if ((s[q] >> (norb - 1 - j)) & 1ULL) rec[j / 8] |= (uint8_t)(1u << (7 - j % 8));
    ok = ok && std::fwrite(rec.data(), 1, nbytes, f) == nbytes;
  }
  ok = (std::fclose(f) == 0) && ok;
  return ok ? SBD_OK : SBD_ERR_IO;
}


int sbd_string_weights(sbd_ctx* ctx, const double* W, double* wa, double* wb, void* stream) {
  if (!ctx || !W || !wa || !wb) return SBD_ERR_ARG;
  return string_weights(ctx, W, wa, wb, (hipStream_t)stream);
}

int sbd_occupation(sbd_ctx* ctx, const double* W, double* density, void* stream) {
  if (!ctx || !W || !density) return SBD_ERR_ARG;
  std::vector<double> Da(ctx->na), Db(ctx->nb);
  int rc = string_weights(ctx, W, Da.data(), Db.data(), (hipStream_t)stream);
  if (rc) return rc;
  occupation_from_weights(ctx, Da.data(), Db.data(), density);
  return SBD_OK;
}

int sbd_carryover(sbd_ctx* ctx, const double* W, double threshold, double ratio,
                  uint64_t* a_out, size_t* na_out, uint64_t* b_out, size_t* nb_out,
                  double* truncated, void* stream) {
  if (!ctx || !W || !a_out || !na_out || !b_out || !nb_out) return SBD_ERR_ARG;
  if (!(ratio >= 0.0 && ratio <= 1.0))
    return set_err(ctx, SBD_ERR_ARG, "carry-over ratio must be in [0, 1]");
  std::vector<double> Da(ctx->na), Db(ctx->nb);
  int rc = string_weights(ctx, W, Da.data(), Db.data(), (hipStream_t)stream);
  if (rc) return rc;
  // ratio == 0: threshold form; else keep floor(ratio * count) strings
  // (sbdiag.h:745-772)
  const bool by_count = ratio != 0.0;
  const uint64_t ka = (uint64_t)(ratio * (double)ctx->na);
  const uint64_t kb = (uint64_t)(ratio * (double)ctx->nb);
  std::vector<uint32_t> ia, ib;
  double ta = 0.0, tb = 0.0;
  carry_select(Da, threshold, by_count, ka, ia, &ta);
  carry_select(Db, threshold, by_count, kb, ib, &tb);
  for (size_t k = 0; k < ia.size(); k++) a_out[k] = ctx->a_host[ia[k]];
  for (size_t k = 0; k < ib.size(); k++) b_out[k] = ctx->b_host[ib[k]];
  *na_out = ia.size();
  *nb_out = ib.size();
  if (truncated) {
    truncated[0] = ta;
    truncated[1] = tb;
  }
  return SBD_OK;
}

}  // extern "C"
