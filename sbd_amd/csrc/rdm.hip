// Full one- and two-particle reduced density matrices (do_rdm != 0).
//
// Reference: Correlation, chemistry/tpb/correlation.h:379-643 (and its device
// form correlation_thrust.h), with ZeroDiffCorrelation, OneDiffCorrelation and
// TwoDiffCorrelation of chemistry/basic/correlation.h:16-50, 103-148, 190-247;
// called from sbd::tpb::diag, chemistry/tpb/sbdiag.h:704-728.  Outputs as the
// reference lays them out:
//   onebody[s * L^2 + o1 + L*o2]                      (s = 0 alpha, 1 beta)
//   twobody[(s1 + 2*s2) * L^4 + o1 + L*o2 + L^2*o3 + L^3*o4]
// Every pair of determinants (bra I, ket J) that the reference visits adds
// W_I W_J times the parity of the interleaved bra determinant to entries that
// depend only on the orbitals that differ (and, for single excitations, on
// each occupied spectator orbital).  The device therefore reduces W products
// per string pair, and the host spreads the reduced sums over the entries.
//
// Device passes (all sums in a fixed order, so results do not vary between
// runs):
//  1. k_rdm_pairs: for each pair of strings (r1, r2) of one spin (a single or
//     double excitation, or r1 == r2 for the diagonal), over every string c of
//     the other spin (the spectator):
//        T0   = sum_c M[r1,c] M[r2,c] (-1)^popcount(spec_c & m)
//        T_q  = the same sum restricted to spectators with orbital q occupied
//     where m holds the spectator orbitals that lie between the excited spin
//     orbitals in the interleaved determinant.  M = W for alpha pairs and the
//     transpose of W for beta pairs, so rows are read contiguously.
//  2. k_rdm_ab: alpha single x beta single terms (task type 0).  Alpha single
//     records are grouped by their orbitals (x, y) into chunks; beta singles are
//     grouped by theirs (c, d).  Thread k of a chunk's workgroup owns key
//     (c, d) and sums W[ia, ib] W[ja, jb] sgn over the chunk's alpha records
//     and the key's beta records.
// The host adds the chunks of equal (x, y) in order and writes the entries.
#include <algorithm>
#include <cstring>
#include <vector>

#include "sbd_internal.hpp"

using namespace sbdx;

namespace {

constexpr int kRdmBlock = 256;  // 4 waves, one pair record per wave
constexpr int kAbChunk = 32;    // alpha records per type-0 chunk
constexpr int kAbBlock = 256;

struct alignas(16) RdmRec {
  uint32_t r1, r2;
  uint64_t m;  // spectator orbitals between the excited spin orbitals
};

struct alignas(16) AbRec {  // one alpha single (bra ia, ket ja) and the bra string
  uint32_t ia, ja;
  uint64_t A;
};

struct alignas(16) AbEnt {  // one beta single (bra ib, ket jb) and the bra string
  uint32_t ib, jb;
  uint64_t B;
};

struct alignas(16) AbChunk {
  uint32_t start, count, x, y;
};

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// One wave per record.  Records [0, nq_rec) also produce T_q for q < L and
// write L + 1 values at out[r * (L + 1)]; the others write T0 alone at
// out[nq_rec * (L + 1) + (r - nq_rec)].
template <int NQ>
__global__ __launch_bounds__(kRdmBlock) void k_rdm_pairs(const double* __restrict__ M,
                                                         uint64_t n2,
                                                         const uint64_t* __restrict__ spec,
                                                         const RdmRec* __restrict__ rec,
                                                         uint64_t nrec, uint64_t nq_rec, int L,
                                                         double* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const uint64_t nwaves = (uint64_t)gridDim.x * (kRdmBlock / 64);
  for (uint64_t r = (uint64_t)blockIdx.x * (kRdmBlock / 64) + (threadIdx.x >> 6); r < nrec;
       r += nwaves) {
    const RdmRec q = rec[r];
    const double* x = M + (uint64_t)q.r1 * n2;
    const double* y = M + (uint64_t)q.r2 * n2;
    double* o = r < nq_rec ? out + r * (uint64_t)(L + 1)
                           : out + nq_rec * (uint64_t)(L + 1) + (r - nq_rec);
    double t0 = 0.0;
    if (r < nq_rec) {
      double acc[NQ];
#pragma unroll
      for (int k = 0; k < NQ; k++) acc[k] = 0.0;
      for (uint64_t c = lane; c < n2; c += 64) {
        const uint64_t s = spec[c];
        double v = x[c] * y[c];
        if (__popcll(s & q.m) & 1) v = -v;
        t0 += v;
#pragma unroll
        for (int k = 0; k < NQ; k++) acc[k] += ((s >> k) & 1ULL) ? v : 0.0;
      }
      t0 = wave_sum(t0);
#pragma unroll
      for (int k = 0; k < NQ; k++) {
        if (k < L) {
          const double a = wave_sum(acc[k]);
          if (lane == 0) o[1 + k] = a;
        }
      }
    } else {
      for (uint64_t c = lane; c < n2; c += 64) {
        double v = x[c] * y[c];
        if (__popcll(spec[c] & q.m) & 1) v = -v;
        t0 += v;
      }
      t0 = wave_sum(t0);
    }
    if (lane == 0) o[0] = t0;
  }
}

// Spin orbitals strictly between u < w in the interleaved determinant:
// alpha orbitals p with u < 2p < w, beta orbitals q with u < 2q+1 < w.
__host__ __device__ inline uint64_t alpha_between(int u, int w) {
  return bits_range(u / 2 + 1, (w + 1) / 2);
}
__host__ __device__ inline uint64_t beta_between(int u, int w) {
  return bits_range((u + 1) / 2, w / 2);
}

__global__ __launch_bounds__(kAbBlock) void k_rdm_ab(const double* __restrict__ W, uint64_t nb,
                                                     const AbRec* __restrict__ arec,
                                                     const AbChunk* __restrict__ chunks,
                                                     uint32_t nchunks,
                                                     const uint32_t* __restrict__ key_off,
                                                     const AbEnt* __restrict__ ent, int L,
                                                     double* __restrict__ out) {
  __shared__ AbRec sr[kAbChunk];
  const int LL = L * L;
  for (uint32_t ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
    const AbChunk C = chunks[ch];
    __syncthreads();
    if (threadIdx.x < C.count) sr[threadIdx.x] = arec[C.start + threadIdx.x];
    __syncthreads();
    for (int k = threadIdx.x; k < LL; k += kAbBlock) {
      const uint32_t e0 = key_off[k], e1 = key_off[k + 1];
      double acc = 0.0;
      if (e0 < e1) {
        // TwoDiffCorrelation(i = alpha bra-only, j = beta bra-only,
        //                    a = alpha ket-only, b = beta ket-only)
        const int c = k / L, d = k % L;
        const int i = 2 * (int)C.x, j = 2 * c + 1, a = 2 * (int)C.y, b = 2 * d + 1;
        const int I = min(i, j), J = max(i, j), A = min(a, b), B = max(a, b);
        const int u1 = min(I, A), w1 = max(I, A), u2 = min(J, B), w2 = max(J, B);
        const uint64_t am = alpha_between(u1, w1) ^ alpha_between(u2, w2);
        const uint64_t bm = beta_between(u1, w1) ^ beta_between(u2, w2);
        for (uint32_t r = 0; r < C.count; r++) {
          const AbRec ar = sr[r];
          const uint32_t ap = __popcll(ar.A & am) & 1u;
          const double* wi = W + (uint64_t)ar.ia * nb;
          const double* wj = W + (uint64_t)ar.ja * nb;
          for (uint32_t e = e0; e < e1; e++) {
            const AbEnt en = ent[e];
            double v = wi[en.ib] * wj[en.jb];
            if ((__popcll(en.B & bm) ^ ap) & 1u) v = -v;
            acc += v;
          }
        }
        if (A > J || B < I) acc = -acc;
      }
      out[(uint64_t)ch * LL + k] = acc;
    }
  }
}

__global__ void k_transpose(const double* __restrict__ W, uint64_t na, uint64_t nb,
                            double* __restrict__ WT) {
  __shared__ double tile[64][65];
  const uint64_t r0 = (uint64_t)blockIdx.y * 64, c0 = (uint64_t)blockIdx.x * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;  // 256 threads: 4 rows per pass
  for (int k = ty; k < 64; k += 4) {
    const uint64_t r = r0 + k, c = c0 + tx;
    if (r < na && c < nb) tile[k][tx] = W[r * nb + c];
  }
  __syncthreads();
  for (int k = ty; k < 64; k += 4) {
    const uint64_t c = c0 + k, r = r0 + tx;
    if (r < na && c < nb) WT[c * na + r] = tile[tx][k];
  }
}

struct DevBuf {
  void* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  template <typename T>
  T* as() const {
    return (T*)p;
  }
};

// Host-side copy of one spin's CSR excitation tables.
struct HostTab {
  std::vector<uint32_t> s_off, d_off;
  std::vector<PairRec> s_rec, d_rec;
};

int fetch_tab(sbd_ctx* ctx, const SpinTables& t, uint64_t n, HostTab& h, hipStream_t st) {
  h.s_off.resize(n + 1);
  h.d_off.resize(n + 1);
  h.s_rec.resize(t.n_s);
  h.d_rec.resize(t.n_d);
  SBD_HIP_TRY(ctx, hipMemcpyAsync(h.s_off.data(), t.s_off, 4 * (n + 1), hipMemcpyDeviceToHost, st));
  SBD_HIP_TRY(ctx, hipMemcpyAsync(h.d_off.data(), t.d_off, 4 * (n + 1), hipMemcpyDeviceToHost, st));
  if (t.n_s)
    SBD_HIP_TRY(ctx, hipMemcpyAsync(h.s_rec.data(), t.s_rec, sizeof(PairRec) * t.n_s,
                                    hipMemcpyDeviceToHost, st));
  if (t.n_d)
    SBD_HIP_TRY(ctx, hipMemcpyAsync(h.d_rec.data(), t.d_rec, sizeof(PairRec) * t.n_d,
                                    hipMemcpyDeviceToHost, st));
  SBD_HIP_TRY(ctx, hipStreamSynchronize(st));
  return SBD_OK;
}

inline int lowbit(uint64_t v) { return __builtin_ctzll(v); }
inline int highbit(uint64_t v) { return 63 - __builtin_clzll(v); }

// Spectator mask of an excitation of spin s between same-spin orbital ranges
// (lo1, hi1) and optionally (lo2, hi2): the other spin's orbitals strictly
// between the spin orbitals 2*lo+s and 2*hi+s.
inline uint64_t spectator_mask(int s, int lo, int hi) {
  return s == 0 ? beta_between(2 * lo, 2 * hi) : alpha_between(2 * lo + 1, 2 * hi + 1);
}
inline uint64_t open_range(int lo, int hi) { return bits_range(lo + 1, hi); }

struct Entry4 {
  int L;
  size_t L2, L3, L4;
  double* two;
  void add(int blk, int p0, int p1, int p2, int p3, double v) const {
    two[(size_t)blk * L4 + p0 + (size_t)L * p1 + L2 * p2 + L3 * p3] += v;
  }
};

}  // namespace

extern "C" int sbd_correlation(sbd_ctx* ctx, const double* W, double* onebody, double* twobody,
                               void* stream) {
  if (!ctx || !W || !onebody || !twobody) return SBD_ERR_ARG;
  if (!ctx->tables_built) return set_err(ctx, SBD_ERR_STATE, "correlation before sbd_make_helpers");
  if (ctx->na >= (1ULL << 32) || ctx->nb >= (1ULL << 32) || (ctx->na + 63) / 64 > 65535)
    return set_err(ctx, SBD_ERR_ARG, "correlation: at most 65535*64 alpha strings");
  hipStream_t st = (hipStream_t)stream;
  const int L = ctx->norb;
  const uint64_t na = ctx->na, nb = ctx->nb;
  const size_t L2 = (size_t)L * L, L3 = L2 * L, L4 = L3 * L;
  std::memset(onebody, 0, sizeof(double) * 2 * L2);
  std::memset(twobody, 0, sizeof(double) * 4 * L4);
  HostTab ha, hb;
  int rc = fetch_tab(ctx, ctx->ta, na, ha, st);
  if (!rc) rc = fetch_tab(ctx, ctx->tb, nb, hb, st);
  if (rc) return rc;
  const auto& As = ctx->a_host;
  const auto& Bs = ctx->b_host;

  // ---- pair records ----
  // A pair's sums are symmetric in (bra, ket): the products commute and the
  // spectator mask depends on the orbital ranges only.  So each unordered pair
  // gets one record (from its lower string); `*_rec[k]` maps table entry k to
  // its pair's record for either direction.
  // alpha side (M = W, spectator beta): diagonal and singles with T_q, then doubles
  // beta side (M = W^T, spectator alpha): singles with T_q, then diagonal and doubles
  std::vector<RdmRec> ra, rb;
  ra.reserve(na + (ha.s_rec.size() + ha.d_rec.size()) / 2);
  rb.reserve(nb + (hb.s_rec.size() + hb.d_rec.size()) / 2);
  for (uint64_t ia = 0; ia < na; ia++) ra.push_back({(uint32_t)ia, (uint32_t)ia, 0});
  // record of every table entry: entries with ket > bra own one; the reverse
  // entry (ket < bra) is found in the ket's row (kets ascending)
  auto pair_records = [&](std::vector<RdmRec>& out, const std::vector<uint32_t>& off,
                          const std::vector<PairRec>& rec, uint64_t n,
                          const std::vector<uint64_t>& S, int s, bool dbl) {
    std::vector<uint64_t> idx(rec.size());
    for (uint64_t r = 0; r < n; r++)
      for (uint32_t k = off[r]; k < off[r + 1]; k++) {
        const uint32_t ket = rec[k].ket;
        if (ket < r) continue;
        const uint64_t bo = S[r] & ~S[ket], ko = S[ket] & ~S[r];
        uint64_t m;
        if (dbl) {
          const int x0 = lowbit(bo), x1 = highbit(bo), y0 = lowbit(ko), y1 = highbit(ko);
          m = spectator_mask(s, std::min(x0, y0), std::max(x0, y0)) ^
              spectator_mask(s, std::min(x1, y1), std::max(x1, y1));
        } else {
          const int x = lowbit(bo), y = lowbit(ko);
          m = spectator_mask(s, std::min(x, y), std::max(x, y));
        }
        idx[k] = out.size();
        out.push_back({(uint32_t)r, ket, m});
      }
    for (uint64_t r = 0; r < n; r++)
      for (uint32_t k = off[r]; k < off[r + 1]; k++) {
        const uint32_t ket = rec[k].ket;
        if (ket > r) continue;
        const auto b = rec.begin() + off[ket], e = rec.begin() + off[ket + 1];
        const auto it = std::lower_bound(
            b, e, (uint32_t)r, [](const PairRec& q, uint32_t v) { return q.ket < v; });
        idx[k] = idx[it - rec.begin()];
      }
    return idx;
  };
  const std::vector<uint64_t> as_idx = pair_records(ra, ha.s_off, ha.s_rec, na, As, 0, false);
  const uint64_t nq_a = ra.size();
  const std::vector<uint64_t> ad_idx = pair_records(ra, ha.d_off, ha.d_rec, na, As, 0, true);
  const std::vector<uint64_t> bs_idx = pair_records(rb, hb.s_off, hb.s_rec, nb, Bs, 1, false);
  const uint64_t nq_b = rb.size();
  for (uint64_t ib = 0; ib < nb; ib++) rb.push_back({(uint32_t)ib, (uint32_t)ib, 0});
  const std::vector<uint64_t> bd_idx = pair_records(rb, hb.d_off, hb.d_rec, nb, Bs, 1, true);

  // ---- type 0: alpha single chunks by (x, y), beta singles by (c, d) ----
  std::vector<AbRec> arec;
  std::vector<AbChunk> chunks;
  std::vector<uint32_t> key_off(L2 + 1, 0);
  std::vector<AbEnt> ent;
  {
    std::vector<std::vector<AbRec>> by_key(L2);
    for (uint64_t ia = 0; ia < na; ia++)
      for (uint32_t k = ha.s_off[ia]; k < ha.s_off[ia + 1]; k++) {
        const uint32_t ja = ha.s_rec[k].ket;
        const int x = lowbit(As[ia] & ~As[ja]), y = lowbit(As[ja] & ~As[ia]);
        by_key[(size_t)x * L + y].push_back({(uint32_t)ia, ja, As[ia]});
      }
    for (size_t key = 0; key < L2; key++) {
      const auto& v = by_key[key];
      for (size_t s = 0; s < v.size(); s += kAbChunk) {
        const uint32_t cnt = (uint32_t)std::min<size_t>(kAbChunk, v.size() - s);
        chunks.push_back({(uint32_t)arec.size(), cnt, (uint32_t)(key / L), (uint32_t)(key % L)});
        arec.insert(arec.end(), v.begin() + s, v.begin() + s + cnt);
      }
    }
    std::vector<std::vector<AbEnt>> eb(L2);
    for (uint64_t ib = 0; ib < nb; ib++)
      for (uint32_t k = hb.s_off[ib]; k < hb.s_off[ib + 1]; k++) {
        const uint32_t jb = hb.s_rec[k].ket;
        const int c = lowbit(Bs[ib] & ~Bs[jb]), d = lowbit(Bs[jb] & ~Bs[ib]);
        eb[(size_t)c * L + d].push_back({(uint32_t)ib, jb, Bs[ib]});
      }
    for (size_t key = 0; key < L2; key++) {
      key_off[key] = (uint32_t)ent.size();
      ent.insert(ent.end(), eb[key].begin(), eb[key].end());
    }
    key_off[L2] = (uint32_t)ent.size();
  }

  // ---- device work ----
  auto out_size = [&](uint64_t nrec, uint64_t nq) { return nq * (uint64_t)(L + 1) + (nrec - nq); };
  DevBuf dWT, dra, drb, douta, doutb, darec, dch, dko, dent, dab;
  const uint64_t nc = chunks.size();
  auto alloc = [&](DevBuf& b, size_t bytes) {
    return hipMalloc(&b.p, std::max<size_t>(bytes, 16)) == hipSuccess;
  };
  if (!alloc(dWT, 8 * na * nb) || !alloc(dra, sizeof(RdmRec) * ra.size()) ||
      !alloc(drb, sizeof(RdmRec) * rb.size()) || !alloc(douta, 8 * out_size(ra.size(), nq_a)) ||
      !alloc(doutb, 8 * out_size(rb.size(), nq_b)) || !alloc(darec, sizeof(AbRec) * arec.size()) ||
      !alloc(dch, sizeof(AbChunk) * nc) || !alloc(dko, 4 * key_off.size()) ||
      !alloc(dent, sizeof(AbEnt) * ent.size()) || !alloc(dab, 8 * nc * L2))
    return set_err(ctx, SBD_ERR_MEMORY, "device allocation for the density matrices failed");
  auto up = [&](DevBuf& b, const void* src, size_t bytes) {
    return bytes == 0 || hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, st) == hipSuccess;
  };
  if (!up(dra, ra.data(), sizeof(RdmRec) * ra.size()) ||
      !up(drb, rb.data(), sizeof(RdmRec) * rb.size()) ||
      !up(darec, arec.data(), sizeof(AbRec) * arec.size()) ||
      !up(dch, chunks.data(), sizeof(AbChunk) * nc) ||
      !up(dko, key_off.data(), 4 * key_off.size()) ||
      !up(dent, ent.data(), sizeof(AbEnt) * ent.size()))
    return set_err(ctx, SBD_ERR_DEVICE, "correlation: upload failed");
  hipLaunchKernelGGL(k_transpose, dim3((unsigned)((nb + 63) / 64), (unsigned)((na + 63) / 64)),
                     dim3(256), 0, st, W, na, nb, dWT.as<double>());
  auto pairs = [&](const double* M, uint64_t n2, const uint64_t* spec, const DevBuf& r,
                   uint64_t nrec, uint64_t nq, DevBuf& out) {
    if (!nrec) return;
    const unsigned grid =
        (unsigned)std::min<uint64_t>((nrec + 3) / 4, (uint64_t)ctx->cu_count * 16);
    if (L <= 32)
      hipLaunchKernelGGL(k_rdm_pairs<32>, dim3(grid), dim3(kRdmBlock), 0, st, M, n2, spec,
                         r.as<RdmRec>(), nrec, nq, L, out.as<double>());
    else
      hipLaunchKernelGGL(k_rdm_pairs<64>, dim3(grid), dim3(kRdmBlock), 0, st, M, n2, spec,
                         r.as<RdmRec>(), nrec, nq, L, out.as<double>());
  };
  pairs(W, nb, ctx->d_b, dra, ra.size(), nq_a, douta);
  pairs(dWT.as<double>(), na, ctx->d_a, drb, rb.size(), nq_b, doutb);
  if (nc) {
    const unsigned grid = (unsigned)std::min<uint64_t>(nc, (uint64_t)ctx->cu_count * 8);
    hipLaunchKernelGGL(k_rdm_ab, dim3(grid), dim3(kAbBlock), 0, st, W, nb, darec.as<AbRec>(),
                       dch.as<AbChunk>(), (uint32_t)nc, dko.as<uint32_t>(), dent.as<AbEnt>(), L,
                       dab.as<double>());
  }
  std::vector<double> oa(out_size(ra.size(), nq_a)), ob(out_size(rb.size(), nq_b)), oab(nc * L2);
  if (hipGetLastError() != hipSuccess ||
      hipMemcpyAsync(oa.data(), douta.p, 8 * oa.size(), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(ob.data(), doutb.p, 8 * ob.size(), hipMemcpyDeviceToHost, st) != hipSuccess ||
      (nc && hipMemcpyAsync(oab.data(), dab.p, 8 * oab.size(), hipMemcpyDeviceToHost, st) !=
                 hipSuccess) ||
      hipStreamSynchronize(st) != hipSuccess)
    return set_err(ctx, SBD_ERR_DEVICE, "correlation: device error");

  // ---- spread the reduced sums over the entries ----
  const Entry4 E{L, L2, L3, L4, twobody};
  const size_t stride = (size_t)L + 1;
  // ZeroDiffCorrelation: per spin, Z[p][q] (p < q) = sum of W^2 over the
  // determinants with both p and q occupied; Zab[p][q] for alpha p, beta q
  std::vector<double> Zs(2 * L2, 0.0), Zab(L2, 0.0);
  auto zero_same = [&](int s, uint64_t S, double t0) {
    for (uint64_t u = S; u; u &= u - 1) {
      const int p = lowbit(u);
      onebody[s * L2 + p + (size_t)L * p] += t0;
      for (uint64_t v = u & (u - 1); v; v &= v - 1) Zs[s * L2 + (size_t)p * L + lowbit(v)] += t0;
    }
  };
  for (uint64_t ia = 0; ia < na; ia++) {
    const double* o = &oa[ia * stride];
    zero_same(0, As[ia], o[0]);
    for (uint64_t u = As[ia]; u; u &= u - 1) {
      double* z = &Zab[(size_t)lowbit(u) * L];
      for (int q = 0; q < L; q++) z[q] += o[1 + q];
    }
  }
  const size_t ob_diag = nq_b * stride;  // beta diagonal records follow the singles
  for (uint64_t ib = 0; ib < nb; ib++) zero_same(1, Bs[ib], ob[ob_diag + ib]);
  for (int s = 0; s < 2; s++)
    for (int p = 0; p < L; p++)
      for (int q = p + 1; q < L; q++) {
        const double z = Zs[s * L2 + (size_t)p * L + q];
        if (z == 0.0) continue;
        E.add(3 * s, p, q, p, q, z);
        E.add(3 * s, q, p, q, p, z);
        E.add(3 * s, p, q, q, p, -z);
        E.add(3 * s, q, p, p, q, -z);
      }
  for (int p = 0; p < L; p++)
    for (int q = 0; q < L; q++) {
      const double z = Zab[(size_t)p * L + q];
      E.add(2, p, q, p, q, z);
      E.add(1, q, p, q, p, z);
    }
  // OneDiffCorrelation per single excitation of spin s (bra string S, bra-only
  // x, ket-only y): the one-body entry, then the two-body entries of each
  // same-spin spectator oj in S (x included) with weight T0 and of each
  // other-spin spectator q with weight T_q.  The two-body weights are first
  // summed per (x, y, oj) and (x, y, q) (L^3 each), then written out.
  std::vector<double> A3(2 * L3, 0.0), B3(2 * L3, 0.0);
  auto one_diff = [&](int s, uint64_t S, int x, int y, const double* o) {
    const double sg = (__builtin_popcountll(S & open_range(std::min(x, y), std::max(x, y))) & 1)
                          ? -1.0 : 1.0;
    const double v0 = sg * o[0];
    onebody[s * L2 + x + (size_t)L * y] += v0;
    double* a3 = &A3[s * L3 + ((size_t)x * L + y) * L];
    for (uint64_t u = S; u; u &= u - 1) a3[lowbit(u)] += v0;
    double* b3 = &B3[s * L3 + ((size_t)x * L + y) * L];
    for (int q = 0; q < L; q++) b3[q] += sg * o[1 + q];
  };
  // TwoDiffCorrelation per same-spin double excitation (bra-only x0 < x1,
  // ket-only y0 < y1): the signed sum is kept per orbital quadruple, D[s]
  // indexed by the pairs (x0, x1) and (y0, y1)
  std::vector<uint32_t> pidx(L2, 0);
  uint32_t npair = 0;
  for (int a = 0; a < L; a++)
    for (int b = a + 1; b < L; b++) pidx[(size_t)a * L + b] = npair++;
  std::vector<double> D(2 * (size_t)npair * npair, 0.0);
  auto two_diff = [&](int s, uint64_t S, uint64_t K, double t0) {
    const uint64_t bo = S & ~K, ko = K & ~S;
    const int x0 = lowbit(bo), x1 = highbit(bo), y0 = lowbit(ko), y1 = highbit(ko);
    int par = __builtin_popcountll(S & open_range(std::min(x0, y0), std::max(x0, y0))) +
              __builtin_popcountll(S & open_range(std::min(x1, y1), std::max(x1, y1)));
    if (y0 > x1 || y1 < x0) par++;
    D[(size_t)s * npair * npair + (size_t)pidx[(size_t)x0 * L + x1] * npair +
      pidx[(size_t)y0 * L + y1]] += (par & 1) ? -t0 : t0;
  };
  // every table entry (both directions of each pair), bra order as the
  // reference visits them
  auto singles = [&](int s, const HostTab& h, uint64_t n, const std::vector<uint64_t>& S,
                     const std::vector<uint64_t>& idx, const std::vector<double>& out) {
    for (uint64_t r = 0; r < n; r++)
      for (uint32_t k = h.s_off[r]; k < h.s_off[r + 1]; k++) {
        const uint32_t ket = h.s_rec[k].ket;
        one_diff(s, S[r], lowbit(S[r] & ~S[ket]), lowbit(S[ket] & ~S[r]), &out[idx[k] * stride]);
      }
  };
  auto doubles = [&](int s, const HostTab& h, uint64_t n, const std::vector<uint64_t>& S,
                     const std::vector<uint64_t>& idx, const std::vector<double>& out,
                     uint64_t nq) {
    for (uint64_t r = 0; r < n; r++)
      for (uint32_t k = h.d_off[r]; k < h.d_off[r + 1]; k++)
        two_diff(s, S[r], S[h.d_rec[k].ket], out[nq * stride + (idx[k] - nq)]);
  };
  singles(0, ha, na, As, as_idx, oa);
  doubles(0, ha, na, As, ad_idx, oa, nq_a);
  singles(1, hb, nb, Bs, bs_idx, ob);
  doubles(1, hb, nb, Bs, bd_idx, ob, nq_b);
  for (int s = 0; s < 2; s++) {
    const int so = 1 - s;
    for (int x = 0; x < L; x++)
      for (int y = 0; y < L; y++) {
        const double* a3 = &A3[s * L3 + ((size_t)x * L + y) * L];
        const double* b3 = &B3[s * L3 + ((size_t)x * L + y) * L];
        for (int j = 0; j < L; j++) {
          if (a3[j] != 0.0) {
            E.add(3 * s, y, j, x, j, a3[j]);
            E.add(3 * s, j, y, j, x, a3[j]);
            E.add(3 * s, y, j, j, x, -a3[j]);
            E.add(3 * s, j, y, x, j, -a3[j]);
          }
          if (b3[j] != 0.0) {
            E.add(s + 2 * so, y, j, x, j, b3[j]);
            E.add(so + 2 * s, j, y, j, x, b3[j]);
          }
        }
      }
    // same-spin block: entry (p0, p1, p2, p3) takes the quadruple with
    // (x0, x1) = the pair {p2, p3} and (y0, y1) = {p0, p1}: +D for
    // (y0, y1, x0, x1) and (y1, y0, x1, x0), -D for (y0, y1, x1, x0) and
    // (y1, y0, x0, x1); written in index order
    const double* Ds = &D[(size_t)s * npair * npair];
    double* blk = twobody + (size_t)(3 * s) * L4;
    for (int p3 = 0; p3 < L; p3++)
      for (int p2 = 0; p2 < L; p2++) {
        if (p2 == p3) continue;
        const size_t px = pidx[(size_t)std::min(p2, p3) * L + std::max(p2, p3)];
        const double* row = Ds + px * npair;
        double* dst = blk + (size_t)L2 * p2 + L3 * p3;
        for (int p1 = 0; p1 < L; p1++)
          for (int p0 = 0; p0 < L; p0++) {
            if (p0 == p1) continue;
            const double v = row[pidx[(size_t)std::min(p0, p1) * L + std::max(p0, p1)]];
            dst[p0 + (size_t)L * p1] += ((p0 < p1) == (p2 < p3)) ? v : -v;
          }
      }
  }
  // type 0: chunks of equal (x, y) are consecutive; add them in order, then
  // write the reference's entries (both of its branches: spin(I) == spin(A)
  // and spin(I) == spin(B), the latter into block spin(I) + 2 spin(J) as the
  // reference does)
  std::vector<double> X(L2);
  for (uint64_t c0 = 0; c0 < nc;) {
    uint64_t c1 = c0;
    std::fill(X.begin(), X.end(), 0.0);
    while (c1 < nc && chunks[c1].x == chunks[c0].x && chunks[c1].y == chunks[c0].y) {
      for (size_t k = 0; k < L2; k++) X[k] += oab[c1 * L2 + k];
      c1++;
    }
    const int i = 2 * (int)chunks[c0].x, a = 2 * (int)chunks[c0].y;
    for (int c = 0; c < L; c++)
      for (int d = 0; d < L; d++) {
        const double v = X[(size_t)c * L + d];
        if (v == 0.0) continue;
        const int j = 2 * c + 1, b = 2 * d + 1;
        const int I = std::min(i, j), J = std::max(i, j), A = std::min(a, b), B = std::max(a, b);
        const int oi = I / 2, si = I % 2, oa_ = A / 2, sa = A % 2;
        const int oj = J / 2, sj = J % 2, ob_ = B / 2, sb = B % 2;
        if (si == sa) {
          E.add(si + 2 * sj, oa_, ob_, oi, oj, v);
          E.add(sj + 2 * si, ob_, oa_, oj, oi, v);
        }
        if (si == sb) {
          E.add(si + 2 * sj, oa_, ob_, oj, oi, -v);
          E.add(sj + 2 * si, ob_, oa_, oi, oj, -v);
        }
      }
    c0 = c1;
  }
  return SBD_OK;
}
