// Internal definitions shared by the HIP translation units of libsbd_tpb.
//
// Bit conventions (reference: chemistry/basic/determinants.h:104-150):
//   half string   : one uint64 per alpha or beta string, orbital p -> bit p (norb <= 64)
//   determinant   : interleaved, alpha p -> spin orbital 2p, beta p -> 2p+1,
//                   held as two uint64 words (spin orbitals 0..63, 64..127)
//   vector layout : W[ia * nb + ib], beta index fastest (tpb/davidson.h:171-173)
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/sbd_tpb.h"

namespace sbdx {

// Excitation record: one (bra, ket) pair of strings of one spin that differ by
// a single (2 bits) or double (4 bits) excitation.  16 bytes.
//   ket  : ket string index
//   code : single: bra-only orbital x (bits 0-5), ket-only orbital y (12-17),
//          bit 24 = same-spin parity (bra bits of this spin strictly between
//          x and y, mod 2);
//          double: the two orbital ranges that the sign depends on,
//          lo1 = min(x0,y0) (0-5), hi1 = max(x0,y0) (6-11), lo2 = min(x1,y1)
//          (12-17), hi2 = max(x1,y1) (18-23), with x0 < x1 bra-only and
//          y0 < y1 ket-only orbitals
//   val  : single: one-body integral h[y][x] (= I1.Value(a,i));
//          double: same-spin sign x extra sign x ((AI|BJ) - (AJ|BI))
struct alignas(16) PairRec {
  uint32_t ket;
  uint32_t code;
  double val;
};

__host__ __device__ inline int code_x0(uint32_t c) { return c & 63; }
__host__ __device__ inline int code_y0(uint32_t c) { return (c >> 12) & 63; }
__host__ __device__ inline int code_par(uint32_t c) { return (c >> 24) & 1; }
__host__ __device__ inline int code_f(uint32_t c, int k) { return (c >> (6 * k)) & 63; }

// bits [lo, hi) of a 64-bit word, 0 <= lo <= hi <= 64
__host__ __device__ inline uint64_t bits_range(int lo, int hi) {
  uint64_t h = hi >= 64 ? ~0ULL : ((1ULL << hi) - 1ULL);
  uint64_t l = lo >= 64 ? ~0ULL : ((1ULL << lo) - 1ULL);
  return h & ~l;
}

// The sign of an excitation between spin orbitals 2u+s and 2w+s (u < w)
// picks up the parity of the OTHER spin's bits strictly between them in the
// interleaved determinant: s = 0 -> beta orbitals [u, w); s = 1 -> alpha
// orbitals [u+1, w+1).
//
// Alpha excitations (other = beta): mask over beta orbitals.
__host__ __device__ inline uint64_t beta_mask_single(uint32_t c) {
  const int x = code_x0(c), y = code_y0(c);
  return bits_range(x < y ? x : y, x < y ? y : x);
}
__host__ __device__ inline uint64_t beta_mask_double(uint32_t c) {
  return bits_range(code_f(c, 0), code_f(c, 1)) ^ bits_range(code_f(c, 2), code_f(c, 3));
}
// Beta excitations (other = alpha): with PA = inclusive prefix parity of the
// alpha string (bit j = parity of alpha bits 0..j), the parity of alpha
// orbitals [u+1, w+1) is bit u ^ bit w of PA.
__host__ __device__ inline uint64_t prefix_parity(uint64_t x) {
  x ^= x << 1;
  x ^= x << 2;
  x ^= x << 4;
  x ^= x << 8;
  x ^= x << 16;
  x ^= x << 32;
  return x;
}
__host__ __device__ inline uint32_t alpha_par_single(uint64_t PA, uint32_t c) {
  return (uint32_t)((PA >> code_x0(c)) ^ (PA >> code_y0(c))) & 1u;
}
__host__ __device__ inline uint32_t alpha_par_double(uint64_t PA, uint32_t c) {
  return (uint32_t)((PA >> code_f(c, 0)) ^ (PA >> code_f(c, 1)) ^ (PA >> code_f(c, 2)) ^
                    (PA >> code_f(c, 3))) & 1u;
}

// Seeded random vector elements (Randomize, framework/dm_vector.h:101-191):
// element i of a vector drawn with seed s is uniform_pm1(elem_hash(s, i)), so
// any element is computed on its own and the vector is the same for any split
// of the work.  elem_hash mixes (seed, index) linearly and applies the
// splitmix64 output finaliser; uniform_pm1 maps the top 53 bits to [-1, 1).
__host__ __device__ inline uint64_t elem_hash(uint64_t seed, uint64_t idx) {
  uint64_t x = seed * 6364136223846793005ULL + idx * 2862933555777941757ULL;
  x ^= x >> 30;
  x *= 0xbf58476d1ce4e5b9ULL;
  x ^= x >> 27;
  x *= 0x94d049bb133111ebULL;
  return x ^ (x >> 31);
}
__host__ __device__ inline double uniform_pm1(uint64_t h) {
  return (double)(h >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

__host__ __device__ inline uint32_t parity64(uint64_t x) {
  return (uint32_t)__builtin_popcountll(x) & 1u;
}

__host__ __device__ inline uint64_t spread32(uint64_t v) {
  v &= 0xFFFFFFFFULL;
  v = (v | (v << 16)) & 0x0000FFFF0000FFFFULL;
  v = (v | (v << 8)) & 0x00FF00FF00FF00FFULL;
  v = (v | (v << 4)) & 0x0F0F0F0F0F0F0F0FULL;
  v = (v | (v << 2)) & 0x3333333333333333ULL;
  v = (v | (v << 1)) & 0x5555555555555555ULL;
  return v;
}

// packed 8-fold ERI index (chemistry/basic/integrals.h:94-102), spatial orbitals
__host__ __device__ inline int64_t eri_index(int I, int J, int K, int L) {
  int mij = I > J ? I : J, nij = I > J ? J : I;
  int mkl = K > L ? K : L, nkl = K > L ? L : K;
  int64_t ij = (int64_t)mij * (mij + 1) / 2 + nij;
  int64_t kl = (int64_t)mkl * (mkl + 1) / 2 + nkl;
  int64_t a = ij > kl ? ij : kl, b = ij > kl ? kl : ij;
  return a * (a + 1) / 2 + b;
}

// twoInt::Value over spin orbitals: zero unless spin(i)==spin(j) and spin(k)==spin(l)
__host__ __device__ inline double eri_value(const double* I2, int i, int j, int k, int l) {
  if (!((i % 2 == j % 2) && (k % 2 == l % 2))) return 0.0;
  return I2[eri_index(i / 2, j / 2, k / 2, l / 2)];
}

// One beta-double slot of the sliced ELL: the value and the alpha-string mask
// of the element's sign (the element is -val when popcount(alpha & mask) is
// odd).  When every alpha string fits below bit kEllKetShift and the ket
// index fits the bits above it ("packed" tables), the ket string index rides
// in those bits, so a slot is one 16-byte load (FP64) or one 12-byte load
// (FP32, value rounded to FP32: the FP32 kernel rounds every element to FP32
// before its product, so both give the same sum); otherwise the kets are a
// separate array.  Padding slots hold value 0 and mask 0 (their product adds
// an exact zero) with the position's own string as ket.
constexpr int kEllKetShift = 40;
struct alignas(16) EllSlot64 {
  double val;
  uint64_t mk;
};
struct alignas(4) EllSlot32 {
  float val;
  uint32_t mlo, mhi;
};

// Per-column record of the beta strings read by the sigma kernel in one load:
// the string and the range [s0, s1) of its single-excitation records.
struct alignas(16) ColRec {
  uint64_t B;
  uint32_t s0, s1;
};

// Column (string) at one block-sorted ELL position: its doubles count and the
// range [s0, s1) of its single-excitation records.
struct alignas(16) EllPos {
  uint32_t col;
  uint32_t cnt;
  uint32_t s0, s1;
};

// One spin's excitation tables on the device (CSR by bra string).
struct SpinTables {
  uint32_t* s_off = nullptr;  // n+1
  uint32_t* d_off = nullptr;  // n+1
  PairRec* s_rec = nullptr;
  PairRec* d_rec = nullptr;
  uint64_t n_s = 0, n_d = 0;
  uint32_t max_s = 0, max_d = 0;  // max records per bra
  uint32_t* with_s = nullptr;     // ascending list of strings that have single excitations
  uint32_t n_with_s = 0;
  // Doubles in sliced-ELL order for per-column access by a wavefront (built
  // for the beta tables only: alpha doubles are read by row from the CSR).  The
  // strings are taken in blocks of kColBlock; within a block, positions list
  // the block's strings in descending order of their doubles count (stable),
  // ell_pos[position] = {string, its doubles count}.  Slice s holds positions
  // [64s, 64s+64) and is as wide as its largest count; record u of position
  // 64s+l is at slot ell_off[s] + 64u + l (zero padding past its count).
  EllSlot64* ell64 = nullptr;
  EllSlot32* ell32 = nullptr;   // the same slots for the FP32 sigma
  uint32_t* ell_ket = nullptr;  // ket of each slot, when the tables are not packed
  bool ell_packed = false;
  uint64_t n_slots = 0;
  uint32_t* ell_off = nullptr;  // n_slices + 1
  EllPos* ell_pos = nullptr;   // n: string and count at each block-sorted position
  ColRec* col_rec = nullptr;   // n: string and single-excitation range of each string
  uint32_t n_slices = 0;
  uint32_t* s_bra = nullptr;    // bra string of each single record
  std::vector<uint32_t> s_off_host;  // singles offsets (host copy, n+1)
};

constexpr int kEllSlice = 64;
constexpr int kColBlock = 1024;  // columns per sigma workgroup step (= its thread count)

}  // namespace sbdx

// The opaque context behind the C ABI.
struct sbd_ctx {
  int norb = 0;
  int device = 0;
  uint64_t na = 0, nb = 0, N = 0;
  double I0 = 0.0;
  std::vector<double> h1;    // L*L one-body (spatial)
  std::vector<double> eri;   // packed
  std::vector<uint64_t> a_host, b_host;
  // device copies
  uint64_t* d_a = nullptr;
  uint64_t* d_b = nullptr;
  double* d_I1 = nullptr;   // (2L)x(2L) spin-orbital one-body, index i*2L+j
  double* d_eri = nullptr;  // packed ERI
  double* d_J = nullptr;    // L x L  (pp|qq)
  double* d_K = nullptr;    // L x L  (pq|qp)
  double* d_G = nullptr;    // 2 x L x L x 2L: single-excitation integral sums per spin
  double* d_Vab = nullptr;  // L^4: alpha-single x beta-single integral parts (when both exist)
  sbdx::SpinTables ta, tb;
  bool tables_built = false;
  // Single-excitation matrix elements, precomputed once per context (same
  // operations as the on-the-fly evaluation, so bit-identical):
  //   Ea[q * nb + ib] : alpha single record q (bra row s_bra[q]) at beta column ib
  //   Eb[ia * Sb + q] : beta single record q (bra column s_bra[q]) at alpha row ia
  // Left null (on-the-fly evaluation) when they would exceed the cache budget.
  double* d_Ea = nullptr;
  double* d_Eb = nullptr;
  uint64_t single_cache_budget = 32ULL << 30;
  // scratch for reductions
  double* d_sdot = nullptr;  // sigma <T, W> partials (+ result slot)
  size_t sdot_cap = 0;
  double* d_partials = nullptr;
  double* d_results = nullptr;
  double* h_results = nullptr;  // pinned
  size_t partial_cap = 0;
  int lds_max = 65536;
  int cu_count = 256;
  // stream of the solver passes (Davidson, Lanczos, block Davidson, diag):
  // the null stream unless a caller such as the serving engine sets its own
  hipStream_t stream = nullptr;
  int sigma_kernel = 0;  // 0 = by sparsity, 1 = row kernel, 2 = per-element, 3 = rows, alpha x beta inline, 4 = split
  int stream_cols = 512;  // slab width of the split sigma's streaming kernel (256, 512, 1024;
                          // set from nb at creation, sbd_set_sigma_slab)
  int ab_pass = -1;      // dense alpha x beta pass: -1 not yet decided, 0 no, 1 yes
  size_t ab_lds = 0;     // its LDS bytes per workgroup
  uint32_t* d_ab_sell = nullptr;  // its column-major packed beta single records
  std::string last_error;
};

#define SBD_HIP_TRY(ctx, expr)                                                        \
  do {                                                                                \
    hipError_t _e = (expr);                                                           \
    if (_e != hipSuccess) {                                                           \
      if (ctx) (ctx)->last_error = std::string(#expr) + ": " + hipGetErrorString(_e); \
      return SBD_ERR_DEVICE;                                                          \
    }                                                                                 \
  } while (0)

// kernel launchers (defined in the .hip files)
namespace sbdx {
int build_tables(sbd_ctx* ctx, hipStream_t st);
int build_gtable(sbd_ctx* ctx, hipStream_t st);
int build_abtable(sbd_ctx* ctx, hipStream_t st);
int build_single_cache(sbd_ctx* ctx, hipStream_t st);
int launch_hdiag(sbd_ctx* ctx, double* hii, hipStream_t st);
int launch_hdiag_f32(sbd_ctx* ctx, float* hii, hipStream_t st);
// W (+)= H T.  With tw_dot set (and accumulate 0), also returns <T, W> summed in
// a fixed order (the Davidson diagonal element), synchronising the stream.
template <typename VT>
int launch_sigma(sbd_ctx* ctx, const VT* hii, const VT* T, VT* W, int accumulate, hipStream_t st,
                 double* tw_dot = nullptr);

// Davidson vector kernels; all reductions are deterministic (fixed grid,
// per-thread Kahan sums, fixed-order final sum) and accumulate in FP64.
constexpr int kMaxVecs = 40;
template <typename VT>
int vec_multidot(sbd_ctx* ctx, const VT* const* X, int m, const VT* y, uint64_t n, double* out,
                 hipStream_t st);
// out: [0] |r|^2, [1] |w|^2, [2] |t|^2, [3, 3+k) <C_j, t>, [3+k, 3+2k) <HC_j, t>
template <typename VT>
// wout: the last step of a restart cycle; t receives the restart vector
// sum u_j C_j and out[2] its |.|^2 (as vec_lincomb forms them), no overlaps
int vec_ritz_precond(sbd_ctx* ctx, const VT* const* C, const VT* const* HC, int k, const double* u,
                     double E0, const VT* dii, VT* t, uint64_t n, double* out, hipStream_t st,
                     bool wout = false);
template <typename VT>
int vec_gs_update(sbd_ctx* ctx, const VT* const* C, int k, const double* coef, VT* t, double scale,
                  uint64_t n, double* out_norm2, hipStream_t st);
template <typename VT>
int vec_lincomb(sbd_ctx* ctx, const VT* const* C, int k, const double* u, double scale, VT* out,
                uint64_t n, double* out_norm2, hipStream_t st);
template <typename VT>
int vec_scale(sbd_ctx* ctx, VT* x, double s, uint64_t n, hipStream_t st);
template <typename VS, typename VD>
int vec_convert(sbd_ctx* ctx, const VS* x, VD* y, uint64_t n, hipStream_t st);
// x = the normalised seeded vector of BasisInitVector's init = 1 (Randomize:
// elements uniform_pm1(elem_hash(seed, i)), compensated squared norm, scaled
// by 1 / norm in FP64, then stored as VT)
template <typename VT>
int vec_random(sbd_ctx* ctx, VT* x, uint64_t n, uint64_t seed, hipStream_t st);
template <typename VT>
int vec_fill(sbd_ctx* ctx, VT* x, uint64_t n, double v, hipStream_t st);

// W[0, N) = the start vector of sbd_diag in W's element type, on ctx->stream
// (synchronised): the caller's W_init (host FP64) if given, else the unit
// vector e_unit when unit >= 0 (BasisInitVectorFromDeterminants), else
// BasisInitVector's init 0 (e_0) or init 1 (Randomize with opt->seed).
template <typename VT>
int start_vector(sbd_ctx* ctx, const sbd_options* opt, const double* W_init, VT* W,
                 int64_t unit = -1);

// Seconds on a steady clock; records msg as the context's last error and
// returns code.
double now_s();
int set_err(sbd_ctx* ctx, int code, const std::string& msg);
// W[0, N) = the normalised seeded random vector of BasisInitVector's init = 1
// (Randomize: elements from elem_hash, squared norm summed with Kahan
// compensation in index order, then scaled by 1 / norm); seed 0 draws a seed
// from the clock, as the reference does.
void randomize(double* W, uint64_t N, uint64_t seed);

// Restarted Lanczos (chemistry/tpb/lanczos.h:162-327) on device vectors; W is
// the start vector on entry and sum_ib U[ib, 0] C[ib] of the last pass on exit.
// hist (optional) receives (E0, beta, seconds since t_start, 64) per step.
int lanczos_core(sbd_ctx* ctx, const double* hii, double* W, int max_it, int nb, double eps,
                 int verbose, double* energy, std::vector<double>* hist, sbd_report* rep,
                 double t_start);

// String weights of a device vector W (na x nb): Da[ia] = sum_ib W^2 and
// Db[ib] = sum_ia W^2 into host arrays (fixed-order compensated sums).
int string_weights(sbd_ctx* ctx, const double* W, double* Da_host, double* Db_host,
                   hipStream_t st);
// density[2o] / density[2o+1] = alpha / beta occupation of orbital o.
void occupation_from_weights(const sbd_ctx* ctx, const double* Da, const double* Db,
                             double* density);

// Scratch of a batching engine (device jobs, partials and results; pinned
// host results), grown on demand.
struct BatchScratch {
  void* d_jobs = nullptr;
  size_t jobs_cap = 0;
  double* d_part = nullptr;
  size_t part_cap = 0;
  double* d_out = nullptr;
  size_t out_cap = 0;
  double* h_out = nullptr;
  size_t hout_cap = 0;
  void* d_fin = nullptr;
  size_t fin_cap = 0;
};
int batch_reserve(BatchScratch& sc, size_t job_bytes, size_t parts, size_t outs);
void batch_free(BatchScratch& sc);

// One job of a batched Ritz pass (arguments of vec_ritz_precond).
template <typename VT>
struct RitzArgs {
  const VT* const* C;
  const VT* const* HC;
  int k;
  const double* u;
  double E0;
  const VT* dii;
  VT* t;
  uint64_t n;
  bool wout = false;  // restart-vector form (see vec_ritz_precond)
};
// out[j]: 3 + 2 k_j sums in the layout of vec_ritz_precond
template <typename VT>
int vec_ritz_batch(BatchScratch& sc, const RitzArgs<VT>* a, int njobs, double* const* out,
                   hipStream_t st);
// One job of a batched Gram-Schmidt pass (arguments of vec_gs_update).
template <typename VT>
struct GsArgs {
  const VT* const* C;
  int k;
  const double* o;
  VT* t;
  double scale;
  uint64_t n;
};
template <typename VT>
int vec_gs_batch(BatchScratch& sc, const GsArgs<VT>* a, int njobs, double* out_norm2,
                 hipStream_t st);
// One job of a batched sigma: W = H T (accumulate 0) with <T, W> returned.
struct SigmaItem {
  sbd_ctx* ctx;
  const double* hii;
  const double* T;
  double* W;
};
int launch_sigma_batch(BatchScratch& sc, const SigmaItem* items, int n, double* tw_out,
                       hipStream_t st);
}  // namespace sbdx
