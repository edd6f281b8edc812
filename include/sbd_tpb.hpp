/*
 * sbd_tpb.hpp -- C++ host layer over the C ABI of libsbd_tpb (sbd_tpb.h),
 * shaped like the reference's host API for the tensor-product-basis path so
 * that a caller of the reference finds the same names, arguments and outputs
 * at one rank (no communicators).
 *
 *   reference (include/sbd/...)                      here (namespace sbd_amd)
 *   FCIDump, LoadFCIDump   framework/fcidump.h:20-95  FCIDump, LoadFCIDump
 *   from_string            framework/bit_manipulation.h:219-235  from_string
 *   from_string_checked    framework/determinant_initialization.h:18-34
 *                                                     from_string_checked
 *   sort_bitarray          framework/bit_manipulation.h:421-449  sort_bitarray
 *   LoadAlphaDets          chemistry/basic/makedeterminants.h:167-191
 *                                                     LoadAlphaDets
 *   det_vector<size_t, det_kind::half>                det_vector (rows of words)
 *   tpb::SBD               chemistry/tpb/sbdiag.h:24-54  tpb::SBD
 *   generate_sbd_data      chemistry/tpb/sbdiag.h:56-134  tpb::generate_sbd_data
 *   tpb::diag (two forms)  chemistry/tpb/sbdiag.h:151-163, 846-857  tpb::diag
 *                          (with loadname / savename, the explicit initial
 *                          determinant and dump_matrix_form_wf)
 *   MultTPBThrust (makeQChamDiagTerms, run, correlation)
 *                          chemistry/tpb/mult_thrust.h:61-96  tpb::MultTPB
 *   Davidson(hii, W, mult, max_it, nb, eps, max_time)
 *                          chemistry/basic/davidson_thrust.h:97-104  tpb::Davidson
 *
 * Errors: every failing call throws std::runtime_error with the library's
 * message (the reference throws for bad input and exits on device errors).
 * Device vectors are HIP device pointers; the host-vector entry points copy
 * through HIP.  Header only: link with -lsbd_tpb and the HIP runtime
 * (libamdhip64); compile with -D__HIP_PLATFORM_AMD__ -I/opt/rocm/include when
 * the compiler is not hipcc.
 */
#pragma once

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <iterator>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "sbd_tpb.h"

namespace sbd_amd {

// The reference's half det_vector: one row of ceil(L / bit_length) words per
// string, bit_length bits per word, orbital p in word p / bit_length.
using det_vector = std::vector<std::vector<size_t>>;

inline void check(int status, const sbd_ctx* ctx = nullptr) {
  if (status == SBD_OK) return;
  std::string msg = sbd_status_string(status);
  if (ctx && sbd_last_error(ctx)[0]) msg += std::string(": ") + sbd_last_error(ctx);
  throw std::runtime_error("sbd_amd: " + msg);
}

inline void check_hip(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string("sbd_amd: ") + what + ": " +
                                                hipGetErrorString(e));
}

struct FCIDump {
  std::map<std::string, std::string> header;  // NORB and NELEC
  std::vector<std::tuple<double, int, int, int, int>> integrals;  // file order
};

inline FCIDump LoadFCIDump(const std::string& filename) {
  int norb = 0, nelec = 0;
  size_t n = 0;
  check(sbd_fcidump_read(filename.c_str(), &norb, &nelec, &n));
  std::vector<double> v(n);
  std::vector<int> idx(4 * n);
  check(sbd_fcidump_entries(filename.c_str(), v.data(), idx.data(), n));
  FCIDump f;
  f.header["NORB"] = std::to_string(norb);
  f.header["NELEC"] = std::to_string(nelec);
  f.integrals.reserve(n);
  for (size_t k = 0; k < n; k++)
    f.integrals.emplace_back(v[k], idx[4 * k], idx[4 * k + 1], idx[4 * k + 2], idx[4 * k + 3]);
  return f;
}

inline size_t norb_of(const FCIDump& f) {
  const auto it = f.header.find("NORB");
  if (it == f.header.end()) throw std::runtime_error("sbd_amd: FCIDUMP without NORB");
  return (size_t)std::stoul(it->second);
}

// Text bit string -> words: character s[L-1-i] is orbital i.
inline std::vector<size_t> from_string(const std::string& s, size_t bit_length, size_t L) {
  if (s.size() < L) throw std::invalid_argument("sbd_amd: bit string shorter than norb");
  std::vector<size_t> r((L + bit_length - 1) / bit_length, 0);
  for (size_t i = 0; i < L; i++)
    if (s[L - 1 - i] == '1') r[i / bit_length] |= (size_t)1 << (i % bit_length);
  return r;
}

// from_string_checked (framework/determinant_initialization.h:18-34): a bit
// string of exactly total_bits characters, each '0' or '1'.
inline std::vector<size_t> from_string_checked(const std::string& bitstring, size_t bit_length,
                                               size_t total_bits) {
  if (bit_length == 0 || bit_length > 8 * sizeof(size_t))
    throw std::invalid_argument("invalid determinant bit length");
  if (bitstring.size() != total_bits)
    throw std::invalid_argument("determinant bitstring length does not match the basis");
  for (const char v : bitstring)
    if (v != '0' && v != '1')
      throw std::invalid_argument("determinant bitstring must contain only 0 and 1");
  return from_string(bitstring, bit_length, total_bits);
}

// Ascending by the last word first (the bit string's value), duplicates removed.
inline void sort_bitarray(det_vector& d) {
  auto less_from_back = [](const std::vector<size_t>& a, const std::vector<size_t>& b) {
    for (size_t k = a.size(); k > 0; k--) {
      if (a[k - 1] != b[k - 1]) return a[k - 1] < b[k - 1];
    }
    return false;
  };
  std::sort(d.begin(), d.end(), less_from_back);
  d.erase(std::unique(d.begin(), d.end()), d.end());
}

// det_vector rows <-> one uint64 per string (norb <= 64)
inline std::vector<uint64_t> pack(const det_vector& d, size_t bit_length, size_t L) {
  const size_t w = (L + bit_length - 1) / bit_length;
  std::vector<size_t> flat;
  flat.reserve(d.size() * w);
  for (const auto& row : d) {
    if (row.size() != w) throw std::invalid_argument("sbd_amd: string row of the wrong width");
    flat.insert(flat.end(), row.begin(), row.end());
  }
  std::vector<uint64_t> out(d.size());
  check(sbd_pack_strings(flat.data(), d.size(), bit_length, (int)L, out.data()));
  return out;
}

inline det_vector unpack(const uint64_t* s, size_t n, size_t bit_length, size_t L) {
  const size_t w = (L + bit_length - 1) / bit_length;
  det_vector d(n, std::vector<size_t>(w, 0));
  for (size_t k = 0; k < n; k++)
    for (size_t p = 0; p < L; p++)
      if ((s[k] >> p) & 1ULL) d[k][p / bit_length] |= (size_t)1 << (p % bit_length);
  return d;
}

// Text (one bit string per line) or binary (.bin, DecodeAlphaDets) string
// files; the result is sorted with duplicates removed.
inline void LoadAlphaDets(const std::string& adetfile, det_vector& adet, size_t bit_length,
                          size_t total_bit_length) {
  const size_t L = total_bit_length;
  const auto dot = adetfile.rfind('.');
  const std::string ext = dot == std::string::npos ? "" : adetfile.substr(dot + 1);
  adet.clear();
  if (ext == "bin") {
    size_t n = 0;
    uint64_t probe = 0;  // capacity 0: the call reports the count
    const int st = sbd_decode_strings_bin(adetfile.c_str(), (int)L, &probe, 0, &n);
    if (st != SBD_OK && st != SBD_ERR_ARG) check(st);
    std::vector<uint64_t> s(n);
    check(sbd_decode_strings_bin(adetfile.c_str(), (int)L, s.data(), n, &n));
    adet = unpack(s.data(), n, bit_length, L);
  } else {
    FILE* f = std::fopen(adetfile.c_str(), "r");
    if (!f) throw std::runtime_error("Failed to open alpha det file.");
    std::string line;
    int c;
    auto flush = [&]() {
      while (!line.empty() && (line.back() == '\r' || line.back() == ' ')) line.pop_back();
      if (!line.empty()) adet.push_back(from_string(line, bit_length, L));
      line.clear();
    };
    while ((c = std::fgetc(f)) != EOF) {
      if (c == '\n') flush();
      else line.push_back((char)c);
    }
    flush();
    std::fclose(f);
  }
  sort_bitarray(adet);
}

namespace tpb {

// The reference's options (sbdiag.h:24-54) with its defaults, plus the two
// additions of this library (precision, fp32_eps).  Communicator sizes are
// kept for source compatibility and must be 1.
struct SBD {
  int task_comm_size = 1;
  int adet_comm_size = 1;
  int bdet_comm_size = 1;
  int h_comm_size = 1;
  int method = 0;
  int max_it = 1;
  int max_nb = 10;
  double eps = 1.0e-4;
  double max_time = 86400.0;
  int init = 0;
  std::string initial_adeterminant_bitstring;
  std::string initial_bdeterminant_bitstring;
  size_t seed = 1729;
  int do_shuffle = 0;
  int do_rdm = 0;
  int carryover_type = 0;
  double ratio = 0.0;
  double threshold = 0.01;
  size_t bit_length = 20;
  std::string dump_matrix_form_wf;
  int precision = 0;       // added: 1 = FP32 vectors and sigma, then an FP64 sweep
  double fp32_eps = 1e-2;  // added: residual at which FP32 hands over to FP64
};

// generate_sbd_data (sbdiag.h:56-134): the example programs' command-line
// flags into an SBD, scanned as the reference scans them (each flag takes the
// next argument, read with atoi / atof; unknown arguments are skipped).  The
// reference's Thrust-only flags (--use_precalculated_dets,
// --max_memory_gb_for_determinants, --thrust_collapse_loops) are accepted and
// have no effect here.  A flag without its value throws std::invalid_argument
// (the reference would read past argv).
inline SBD generate_sbd_data(int argc, char* argv[]) {
  SBD d;
  for (int i = 0; i < argc; i++) {
    auto val = [&]() -> const char* {
      if (i + 1 >= argc) throw std::invalid_argument(std::string(argv[i]) + " needs a value");
      return argv[++i];
    };
    auto is = [&](const char* a) { return std::string(argv[i]) == a; };
    if (is("--adet_comm_size")) d.adet_comm_size = std::atoi(val());
    if (is("--bdet_comm_size")) d.bdet_comm_size = std::atoi(val());
    if (is("--task_comm_size")) d.task_comm_size = std::atoi(val());
    if (is("--method")) d.method = std::atoi(val());
    if (is("--init")) d.init = std::atoi(val());
    if (is("--initial_adeterminant_bitstring") || is("--initial-adeterminant-bitstring"))
      d.initial_adeterminant_bitstring = val();
    if (is("--initial_bdeterminant_bitstring") || is("--initial-bdeterminant-bitstring"))
      d.initial_bdeterminant_bitstring = val();
    if (is("--seed")) d.seed = (size_t)std::atoi(val());
    if (is("--iteration")) d.max_it = std::atoi(val());
    if (is("--block")) d.max_nb = std::atoi(val());
    if (is("--tolerance")) d.eps = std::atof(val());
    if (is("--carryover_type")) d.carryover_type = std::atoi(val());
    if (is("--carryover_ratio")) d.ratio = std::atof(val());
    if (is("--carryover_threshold")) d.threshold = std::atof(val());
    if (is("--max_time")) d.max_time = std::atof(val());
    if (is("--shuffle")) d.do_shuffle = std::atoi(val());
    if (is("--rdm")) d.do_rdm = std::atoi(val());
    if (is("--bit_length")) d.bit_length = (size_t)std::atoi(val());
    if (is("--dump_matrix_form_wf")) d.dump_matrix_form_wf = val();
    if (is("--use_precalculated_dets") || is("--max_memory_gb_for_determinants") ||
        is("--thrust_collapse_loops"))
      (void)val();
  }
  return d;
}

inline sbd_options options_of(const SBD& s) {
  if (s.task_comm_size != 1 || s.adet_comm_size != 1 || s.bdet_comm_size != 1 ||
      s.h_comm_size != 1)
    throw std::invalid_argument("sbd_amd: one rank per context (communicator sizes must be 1)");
  sbd_options o;
  sbd_options_default(&o);
  o.method = s.method;
  o.max_it = s.max_it;
  o.max_nb = s.max_nb;
  o.eps = s.eps;
  o.max_time = s.max_time;
  o.init = s.init;
  o.seed = s.seed;
  o.precision = s.precision;
  o.fp32_eps = s.fp32_eps;
  return o;
}

namespace detail {
struct CtxDel {
  void operator()(sbd_ctx* c) const { sbd_ctx_destroy(c); }
};
using Ctx = std::unique_ptr<sbd_ctx, CtxDel>;

inline Ctx make_ctx(const FCIDump& fd, const std::vector<uint64_t>& a,
                    const std::vector<uint64_t>& b) {
  const size_t n = fd.integrals.size();
  std::vector<double> v(n);
  std::vector<int> idx(4 * n);
  for (size_t k = 0; k < n; k++) {
    v[k] = std::get<0>(fd.integrals[k]);
    idx[4 * k] = std::get<1>(fd.integrals[k]);
    idx[4 * k + 1] = std::get<2>(fd.integrals[k]);
    idx[4 * k + 2] = std::get<3>(fd.integrals[k]);
    idx[4 * k + 3] = std::get<4>(fd.integrals[k]);
  }
  int st = SBD_OK;
  Ctx c(sbd_ctx_create_fcidump((int)norb_of(fd), n, v.data(), idx.data(), a.data(), a.size(),
                               b.data(), b.size(), &st));
  check(st);
  if (!c) check(SBD_ERR_ARG);
  check(sbd_make_helpers(c.get()), c.get());
  return c;
}

struct DevVec {
  double* p = nullptr;
  explicit DevVec(size_t n) { check_hip(hipMalloc(&p, 8 * std::max<size_t>(n, 1)), "hipMalloc"); }
  ~DevVec() {
    if (p) (void)hipFree(p);
  }
  DevVec(const DevVec&) = delete;
  DevVec& operator=(const DevVec&) = delete;
};
}  // namespace detail

// The operator plug-in (MultTPBThrust at one rank): device vectors of size().
class MultTPB {
 public:
  MultTPB(const det_vector& adet, const det_vector& bdet, size_t bit_length, size_t norbs,
          const FCIDump& fcidump)
      : ctx_(detail::make_ctx(fcidump, pack(adet, bit_length, norbs),
                              pack(bdet, bit_length, norbs))) {
    if (norb_of(fcidump) != norbs) throw std::invalid_argument("sbd_amd: norbs != NORB");
  }
  size_t size() const {
    uint64_t d[4];
    check(sbd_ctx_dims(ctx_.get(), d), ctx_.get());
    return (size_t)d[2];
  }
  // hii[size()] on the device
  void makeQChamDiagTerms(double* hii) { check(sbd_hdiag(ctx_.get(), hii, nullptr), ctx_.get()); }
  // Wb += H Wk (MultBase::run)
  void run(const double* hii, const double* Wk, double* Wb) {
    check(sbd_sigma(ctx_.get(), hii, Wk, Wb, 1, nullptr), ctx_.get());
  }
  // one- and two-particle RDMs of device vector W (do_rdm != 0)
  void correlation(const double* W, std::vector<std::vector<double>>& onebody,
                   std::vector<std::vector<double>>& twobody) {
    const size_t L = norbs(), L2 = L * L, L4 = L2 * L2;
    std::vector<double> one(2 * L2), two(4 * L4);
    check(sbd_correlation(ctx_.get(), W, one.data(), two.data(), nullptr), ctx_.get());
    onebody.assign(2, {});
    twobody.assign(4, {});
    for (size_t s = 0; s < 2; s++) onebody[s].assign(one.begin() + s * L2, one.begin() + (s + 1) * L2);
    for (size_t s = 0; s < 4; s++) twobody[s].assign(two.begin() + s * L4, two.begin() + (s + 1) * L4);
  }
  size_t norbs() const {
    uint64_t d[4];
    check(sbd_ctx_dims(ctx_.get(), d), ctx_.get());
    return (size_t)d[3];
  }
  sbd_ctx* handle() const { return ctx_.get(); }

 private:
  detail::Ctx ctx_;
};

// Davidson(hii, W, mult, max_it, nb, eps, max_time) on the plug-in: hii on the
// device (from makeQChamDiagTerms), W on the host (initial vector in,
// normalised Ritz vector out).  Returns the Ritz value of the last step.
inline double Davidson(const double* hii, std::vector<double>& W, MultTPB& mult, int max_it,
                       int nb, double eps, double max_time) {
  if (W.size() != mult.size()) throw std::invalid_argument("sbd_amd: W has the wrong size");
  sbd_options o;
  sbd_options_default(&o);
  o.max_it = max_it;
  o.max_nb = nb;
  o.eps = eps;
  o.max_time = max_time;
  detail::DevVec dW(W.size());
  check_hip(hipMemcpy(dW.p, W.data(), 8 * W.size(), hipMemcpyHostToDevice), "hipMemcpy");
  double e = 0.0;
  sbd_report rep;
  check(sbd_davidson(mult.handle(), hii, dW.p, &o, &e, nullptr, &rep), mult.handle());
  check_hip(hipMemcpy(W.data(), dW.p, 8 * W.size(), hipMemcpyDeviceToHost), "hipMemcpy");
  return e;
}

// sbd::tpb::diag at one rank: solve, then the density (do_rdm == 0) or the
// full RDMs with the density off their diagonal (do_rdm != 0), and the
// carry-over strings (carryover_type 1: dominant strings; 2: those extended
// by singles; 3: singles extension of the strings of the determinants with
// W^2 > threshold).  loadname: warm start from a saved wavefunction;
// savename: save the result.  Strings must be sorted without duplicates
// (LoadAlphaDets leaves them so).
inline void diag(const SBD& sbd_data, const FCIDump& fcidump, const det_vector& adet,
                 const det_vector& bdet, const std::string& loadname, const std::string& savename,
                 double& energy, std::vector<double>& density, det_vector& co_adet,
                 det_vector& co_bdet, std::vector<std::vector<double>>& one_p_rdm,
                 std::vector<std::vector<double>>& two_p_rdm) {
  const sbd_options o = options_of(sbd_data);
  const size_t L = norb_of(fcidump), bl = sbd_data.bit_length;
  // explicit initial determinant (sbdiag.h:303-335): checked before the solve
  // as BasisInitVectorFromDeterminants checks it (davidson.h:97-110)
  const bool has_a = !sbd_data.initial_adeterminant_bitstring.empty();
  const bool has_b = !sbd_data.initial_bdeterminant_bitstring.empty();
  uint64_t init_a = 0, init_b = 0;
  if (loadname.empty() && has_a) {
    const auto ra = from_string_checked(sbd_data.initial_adeterminant_bitstring, bl, L);
    const auto rb = from_string_checked(
        has_b ? sbd_data.initial_bdeterminant_bitstring : sbd_data.initial_adeterminant_bitstring,
        bl, L);
    auto unique_in = [](const det_vector& basis, const std::vector<size_t>& want,
                        const char* name) {
      const auto it = std::find(basis.begin(), basis.end(), want);
      if (it == basis.end())
        throw std::invalid_argument(std::string("initial ") + name +
                                    " is not present in the basis");
      if (std::find(std::next(it), basis.end(), want) != basis.end())
        throw std::invalid_argument(std::string("initial ") + name +
                                    " occurs more than once in the basis");
    };
    unique_in(adet, ra, "alpha determinant");
    unique_in(bdet, rb, "beta determinant");
    init_a = pack({ra}, bl, L)[0];
    init_b = pack({rb}, bl, L)[0];
  } else if (loadname.empty() && has_b) {
    throw std::invalid_argument("an initial beta determinant requires an initial alpha determinant");
  }
  MultTPB m(adet, bdet, bl, L, fcidump);
  sbd_ctx* c = m.handle();
  const size_t N = m.size();
  std::vector<double> W(N);
  sbd_report rep;
  if (loadname.empty()) {
    if (has_a)
      check(sbd_diag_from_det(c, &o, init_a, init_b, &energy, W.data(), nullptr, 0, &rep), c);
    else
      check(sbd_diag(c, &o, &energy, W.data(), nullptr, 0, &rep), c);
  } else {
    std::vector<double> W0(N);
    uint64_t placed = 0;
    check(sbd_load_wavefunction(c, loadname.c_str(), bl, W0.data(), &placed), c);
    if (placed == 0) throw std::runtime_error("sbd_amd: " + loadname + ": no coefficient placed");
    check(sbd_diag_from(c, &o, W0.data(), &energy, W.data(), nullptr, 0, &rep), c);
  }
  if (!savename.empty()) check(sbd_save_wavefunction(c, savename.c_str(), bl, W.data()), c);
  if (!sbd_data.dump_matrix_form_wf.empty()) {
    // SaveMatrixFormWF (sbdiag.h:808-813)
    const std::vector<uint64_t> pa = pack(adet, bl, L), pb = pack(bdet, bl, L);
    const int st = sbd_save_matrix_form_wf(sbd_data.dump_matrix_form_wf.c_str(), (int)L,
                                           pa.data(), pa.size(), pb.data(), pb.size(), W.data());
    if (st != SBD_OK)
      throw std::runtime_error("sbd_amd: cannot write " + sbd_data.dump_matrix_form_wf);
  }
  detail::DevVec dW(N);
  check_hip(hipMemcpy(dW.p, W.data(), 8 * N, hipMemcpyHostToDevice), "hipMemcpy");
  density.assign(2 * L, 0.0);
  one_p_rdm.clear();
  two_p_rdm.clear();
  if (sbd_data.do_rdm == 0) {
    check(sbd_occupation(c, dW.p, density.data(), nullptr), c);
  } else {
    m.correlation(dW.p, one_p_rdm, two_p_rdm);
    for (size_t io = 0; io < L; io++) {
      density[2 * io] = one_p_rdm[0][io + L * io];
      density[2 * io + 1] = one_p_rdm[1][io + L * io];
    }
  }
  co_adet.clear();
  co_bdet.clear();
  const int ct = sbd_data.carryover_type;
  if (ct == 1 || ct == 2) {
    std::vector<uint64_t> ca(adet.size()), cb(bdet.size());
    size_t na = 0, nb = 0;
    check(sbd_carryover(c, dW.p, sbd_data.threshold, sbd_data.ratio, ca.data(), &na, cb.data(),
                        &nb, nullptr, nullptr),
          c);
    ca.resize(na);
    cb.resize(nb);
    if (ct == 2) {
      auto extend = [&](std::vector<uint64_t>& s) {
        size_t n = 0;
        uint64_t probe = 0;  // capacity 0: the call reports the count
        const int st = sbd_singles_extend_strings((int)L, s.data(), s.size(), &probe, 0, &n);
        if (st != SBD_OK && st != SBD_ERR_ARG) check(st);
        std::vector<uint64_t> out(n);
        check(sbd_singles_extend_strings((int)L, s.data(), s.size(), out.data(), n, &n));
        out.resize(n);
        s.swap(out);
      };
      extend(ca);
      extend(cb);
    }
    co_adet = unpack(ca.data(), ca.size(), bl, L);
    co_bdet = unpack(cb.data(), cb.size(), bl, L);
  } else if (ct == 3) {
    size_t na = 0, nb = 0;
    uint64_t pa = 0, pb = 0;  // capacity 0: the call reports the counts
    const int st = sbd_singles_extend(c, dW.p, sbd_data.threshold, &pa, 0, &na, &pb, 0, &nb,
                                      nullptr, nullptr);
    if (st != SBD_OK && st != SBD_ERR_ARG) check(st, c);
    std::vector<uint64_t> ca(na), cb(nb);
    check(sbd_singles_extend(c, dW.p, sbd_data.threshold, ca.data(), na, &na, cb.data(), nb, &nb,
                             nullptr, nullptr),
          c);
    co_adet = unpack(ca.data(), na, bl, L);
    co_bdet = unpack(cb.data(), nb, bl, L);
  }
}

// The file form (sbdiag.h:846-857): FCIDUMP and one string file used for both
// spins (do_shuffle only reorders the basis across ranks and is ignored).
inline void diag(const SBD& sbd_data, const std::string& fcidumpfile, const std::string& adetfile,
                 const std::string& loadname, const std::string& savename, double& energy,
                 std::vector<double>& density, det_vector& co_adet, det_vector& co_bdet,
                 std::vector<std::vector<double>>& one_p_rdm,
                 std::vector<std::vector<double>>& two_p_rdm) {
  const FCIDump fcidump = LoadFCIDump(fcidumpfile);
  det_vector adet;
  LoadAlphaDets(adetfile, adet, sbd_data.bit_length, norb_of(fcidump));
  diag(sbd_data, fcidump, adet, adet, loadname, savename, energy, density, co_adet, co_bdet,
       one_p_rdm, two_p_rdm);
}

}  // namespace tpb
}  // namespace sbd_amd
