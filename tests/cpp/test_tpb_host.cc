// Tests of the C++ host layer (include/sbd_tpb.hpp), written like the
// reference's own functionality tests (tests/functionality/*.cc: plain
// functions with TEST_ASSERT, a main that runs them).
//
//   test_tpb_host                      host-only tests (no device calls)
//   test_tpb_host --gpu FCIDUMP ALPHA  also the solves on H2O 1e-3 (needs a GPU)
#include <cmath>
#include <cstdio>
#include <fstream>
#include <cstdlib>
#include <string>
#include <vector>

#include "sbd_tpb.hpp"

static int g_fail = 0, g_pass = 0;
#define TEST_ASSERT(cond)                                                   \
  do {                                                                      \
    if (cond) {                                                             \
      g_pass++;                                                             \
    } else {                                                                \
      g_fail++;                                                             \
      std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond);           \
    }                                                                       \
  } while (0)

using namespace sbd_amd;

// test_initial_determinant.cc:16-24: "10000001", bit_length 4 -> {1, 8}
void test_from_string_kat() {
  const auto w = from_string("10000001", 4, 8);
  TEST_ASSERT(w.size() == 2 && w[0] == 1 && w[1] == 8);
  const auto v = from_string("0110", 20, 4);
  TEST_ASSERT(v.size() == 1 && v[0] == 6);
}

void test_sort_bitarray() {
  det_vector d = {{3, 1}, {0, 2}, {3, 1}, {5, 0}, {1, 1}};
  sort_bitarray(d);
  TEST_ASSERT(d.size() == 4);
  TEST_ASSERT((d[0] == std::vector<size_t>{5, 0}));
  TEST_ASSERT((d[1] == std::vector<size_t>{1, 1}));
  TEST_ASSERT((d[2] == std::vector<size_t>{3, 1}));
  TEST_ASSERT((d[3] == std::vector<size_t>{0, 2}));
}

void test_pack_roundtrip() {
  det_vector d = {{0b1011, 0b1}, {0b0111, 0b0}};
  const auto p = pack(d, 4, 6);
  TEST_ASSERT(p.size() == 2 && p[0] == 0b11011 && p[1] == 0b0111);
  TEST_ASSERT(unpack(p.data(), p.size(), 4, 6) == d);
}

// sbdiag.h:24-54 defaults
void test_options_defaults() {
  tpb::SBD s;
  const sbd_options o = tpb::options_of(s);
  TEST_ASSERT(o.method == 0 && o.max_it == 1 && o.max_nb == 10 && o.eps == 1.0e-4);
  TEST_ASSERT(o.max_time == 86400.0 && o.init == 0 && o.seed == 1729);
  TEST_ASSERT(s.bit_length == 20 && s.threshold == 0.01 && s.ratio == 0.0 && s.do_rdm == 0);
  s.adet_comm_size = 2;
  bool threw = false;
  try {
    tpb::options_of(s);
  } catch (const std::invalid_argument&) {
    threw = true;
  }
  TEST_ASSERT(threw);
}

void test_load_inputs(const std::string& fcidump, const std::string& alpha) {
  const FCIDump fd = LoadFCIDump(fcidump);
  TEST_ASSERT(norb_of(fd) == 24 && fd.header.at("NELEC") == "10");
  TEST_ASSERT(fd.integrals.size() > 1000);
  det_vector a;
  LoadAlphaDets(alpha, a, 20, 24);
  TEST_ASSERT(a.size() > 10 && a[0].size() == 2);
  // sorted ascending, first = Hartree-Fock (5 lowest orbitals)
  TEST_ASSERT(a[0][0] == 0b11111 && a[0][1] == 0);
  bool threw = false;
  try {
    LoadAlphaDets(alpha + ".missing", a, 20, 24);
  } catch (const std::runtime_error&) {
    threw = true;
  }
  TEST_ASSERT(threw);
}

// H2O 1e-3 (data/h2o/README.md:6): -76.23594663 Ha
void test_diag_h2o(const std::string& fcidump, const std::string& alpha) {
  tpb::SBD s;
  s.max_it = 20;
  s.max_nb = 10;
  s.eps = 1e-8;
  s.carryover_type = 1;
  double E = 0.0;
  std::vector<double> density;
  det_vector co_a, co_b;
  std::vector<std::vector<double>> one, two;
  tpb::diag(s, fcidump, alpha, "", "", E, density, co_a, co_b, one, two);
  std::printf("  diag (file form): E = %.10f\n", E);
  TEST_ASSERT(std::fabs(E - (-76.23594663)) < 6e-9);
  double na = 0.0, nb = 0.0;
  for (size_t o = 0; o < density.size() / 2; o++) {
    na += density[2 * o];
    nb += density[2 * o + 1];
  }
  TEST_ASSERT(std::fabs(na - 5.0) < 1e-10 && std::fabs(nb - 5.0) < 1e-10);
  TEST_ASSERT(!co_a.empty() && co_a[0][0] == 0b11111 && one.empty() && two.empty());

  // object form with RDMs, saved, then a warm start from the saved file
  const FCIDump fd = LoadFCIDump(fcidump);
  det_vector a;
  LoadAlphaDets(alpha, a, s.bit_length, 24);
  s.carryover_type = 0;
  s.do_rdm = 1;
  double E1 = 0.0;
  tpb::diag(s, fd, a, a, "", "h2o_wf", E1, density, co_a, co_b, one, two);
  TEST_ASSERT(E1 == E);
  TEST_ASSERT(one.size() == 2 && one[0].size() == 24 * 24 && two.size() == 4);
  double tr = 0.0;
  for (size_t o = 0; o < 24; o++) tr += one[0][o + 24 * o];
  TEST_ASSERT(std::fabs(tr - 5.0) < 1e-10);
  double E2 = 0.0;
  s.do_rdm = 0;
  tpb::diag(s, fd, a, a, "h2o_wf", "", E2, density, co_a, co_b, one, two);
  std::printf("  warm start: E = %.10f\n", E2);
  TEST_ASSERT(std::fabs(E2 - E) < 1e-9);
  std::remove("h2o_wf000000");

  // the operator plug-in and Davidson on it
  tpb::MultTPB m(a, a, s.bit_length, 24, fd);
  const size_t N = m.size();
  TEST_ASSERT(N == a.size() * a.size());
  tpb::detail::DevVec hii(N);
  m.makeQChamDiagTerms(hii.p);
  std::vector<double> W(N, 0.0);
  W[0] = 1.0;
  const double Ed = tpb::Davidson(hii.p, W, m, 20, 10, 1e-8, 86400.0);
  std::printf("  Davidson on MultTPB: E0 = %.10f\n", Ed);
  TEST_ASSERT(std::fabs(Ed - E) < 1e-8);
  double nrm = 0.0;
  for (double x : W) nrm += x * x;
  TEST_ASSERT(std::fabs(nrm - 1.0) < 1e-10);
  // run accumulates: Wb = H W twice over gives 2 H W
  tpb::detail::DevVec dW(N), dWb(N);
  check_hip(hipMemcpy(dW.p, W.data(), 8 * N, hipMemcpyHostToDevice), "hipMemcpy");
  check_hip(hipMemset(dWb.p, 0, 8 * N), "hipMemset");
  m.run(hii.p, dW.p, dWb.p);
  std::vector<double> HW(N), HW2(N);
  check_hip(hipMemcpy(HW.data(), dWb.p, 8 * N, hipMemcpyDeviceToHost), "hipMemcpy");
  m.run(hii.p, dW.p, dWb.p);
  check_hip(hipMemcpy(HW2.data(), dWb.p, 8 * N, hipMemcpyDeviceToHost), "hipMemcpy");
  double e = 0.0, dmax = 0.0;
  for (size_t i = 0; i < N; i++) {
    e += W[i] * HW[i];
    dmax = std::fmax(dmax, std::fabs(HW2[i] - 2.0 * HW[i]));
  }
  TEST_ASSERT(std::fabs(e - E) < 1e-8 && dmax < 1e-12);
}

// from_string_checked: the reference's known answers
// (tests/functionality/test_initial_determinant.cc:16-27) and the bit-length
// checks of determinant_initialization.h:21-23
void test_from_string_checked() {
  const auto r = from_string_checked("10000001", 4, 8);
  TEST_ASSERT(r.size() == 2 && r[0] == 1 && r[1] == 8);
  const auto v = from_string_checked("00101", 20, 5);
  TEST_ASSERT(v.size() == 1 && v[0] == 0b101);
  auto throws = [](const std::string& s, size_t bl, size_t n) {
    try {
      from_string_checked(s, bl, n);
    } catch (const std::invalid_argument&) {
      return true;
    }
    return false;
  };
  TEST_ASSERT(throws("1000000", 4, 8));   // one character short
  TEST_ASSERT(throws("1000000x", 4, 8));  // not a bit
  TEST_ASSERT(throws("0101", 20, 5));    // too short
  TEST_ASSERT(throws("001011", 20, 5));  // too long
  TEST_ASSERT(throws("00a01", 20, 5));   // not a bit
  TEST_ASSERT(throws("00101", 0, 5));    // bit length 0
  TEST_ASSERT(throws("00101", 65, 5));   // bit length above the word
}

// SaveMatrixFormWF (restart.h:663-780) layouts, host only
void test_save_matrix_form_wf() {
  const uint64_t a[2] = {0b011, 0b101}, b[2] = {0b011, 0b110};
  const double W[4] = {0.5, -1.25e-7, 1.0 / 3.0, 0.0};
  TEST_ASSERT(sbd_save_matrix_form_wf("mf.bin", 3, a, 2, b, 2, W) == SBD_OK);
  FILE* f = std::fopen("mf.bin", "rb");
  double got[5] = {0, 0, 0, 0, 0};
  const size_t n = f ? std::fread(got, sizeof(double), 5, f) : 0;
  if (f) std::fclose(f);
  TEST_ASSERT(n == 4 && got[0] == W[0] && got[1] == W[1] && got[2] == W[2] && got[3] == W[3]);
  TEST_ASSERT(sbd_save_matrix_form_wf("mf.TXT", 3, a, 2, b, 2, W) == SBD_OK);
  std::ifstream t("mf.TXT");
  std::vector<std::string> lines;
  for (std::string l; std::getline(t, l);) lines.push_back(l);
  TEST_ASSERT(lines.size() == 4);
  TEST_ASSERT(lines.size() == 4 && lines[0] == "0.5 # 0: 011 0: 011" &&
              lines[1] == "-1.25e-07 # 0: 011 1: 110" && lines[2] == "0.333333 # 1: 101 0: 011" &&
              lines[3] == "0 # 1: 101 1: 110");
  // any other extension writes nothing
  std::remove("mf.wf");
  TEST_ASSERT(sbd_save_matrix_form_wf("mf.wf", 3, a, 2, b, 2, W) == SBD_OK);
  TEST_ASSERT(std::fopen("mf.wf", "rb") == nullptr);
  std::remove("mf.bin");
  std::remove("mf.TXT");
}

// explicit initial determinant and matrix-form dump through tpb::diag
void test_initial_determinant_h2o(const std::string& fcidump, const std::string& alpha) {
  const FCIDump fd = LoadFCIDump(fcidump);
  det_vector a;
  tpb::SBD s;
  s.max_it = 20;
  s.max_nb = 10;
  s.eps = 1e-8;
  LoadAlphaDets(alpha, a, s.bit_length, 24);
  double E0 = 0.0, E1 = 0.0;
  std::vector<double> density;
  det_vector co_a, co_b;
  std::vector<std::vector<double>> one, two;
  tpb::diag(s, fd, a, a, "", "", E0, density, co_a, co_b, one, two);
  // the Hartree-Fock string is the first one: the same start as init 0
  s.initial_adeterminant_bitstring = "000000000000000000011111";
  s.dump_matrix_form_wf = "h2o_mf.bin";
  tpb::diag(s, fd, a, a, "", "", E1, density, co_a, co_b, one, two);
  std::printf("  explicit initial determinant: E = %.10f (init 0: %.10f)\n", E1, E0);
  TEST_ASSERT(E1 == E0);
  FILE* f = std::fopen("h2o_mf.bin", "rb");
  TEST_ASSERT(f != nullptr);
  if (f) {
    std::fseek(f, 0, SEEK_END);
    const long bytes = std::ftell(f);
    std::fclose(f);
    TEST_ASSERT((size_t)bytes == 8 * a.size() * a.size());
  }
  std::remove("h2o_mf.bin");
  s.dump_matrix_form_wf.clear();
  auto throws_msg = [&](const std::string& sa, const std::string& sb, const std::string& want) {
    tpb::SBD t = s;
    t.initial_adeterminant_bitstring = sa;
    t.initial_bdeterminant_bitstring = sb;
    double E = 0.0;
    try {
      tpb::diag(t, fd, a, a, "", "", E, density, co_a, co_b, one, two);
    } catch (const std::invalid_argument& ex) {
      return std::string(ex.what()) == want;
    }
    return false;
  };
  TEST_ASSERT(throws_msg("", "000000000000000000011111",
                         "an initial beta determinant requires an initial alpha determinant"));
  TEST_ASSERT(throws_msg("111110000000000000000000", "",
                         "initial alpha determinant is not present in the basis"));
  TEST_ASSERT(throws_msg("000000000000000000011111", "111110000000000000000000",
                         "initial beta determinant is not present in the basis"));
  det_vector dup = a;
  dup.push_back(a[0]);
  tpb::SBD t = s;
  t.initial_adeterminant_bitstring = "000000000000000000011111";
  bool threw = false;
  try {
    double E = 0.0;
    tpb::diag(t, fd, dup, a, "", "", E, density, co_a, co_b, one, two);
  } catch (const std::invalid_argument& ex) {
    threw = std::string(ex.what()) == "initial alpha determinant occurs more than once in the basis";
  }
  TEST_ASSERT(threw);
}

// generate_sbd_data (sbdiag.h:56-134)
void test_generate_sbd_data() {
  const char* av[] = {"prog", "--iteration", "20", "--block", "8", "--tolerance", "1e-8",
                      "--rdm", "1", "--initial-adeterminant-bitstring", "0011",
                      "--dump_matrix_form_wf", "w.txt", "--thrust_collapse_loops", "1",
                      "--bdet_comm_size", "2", "other"};
  const int ac = (int)(sizeof(av) / sizeof(av[0]));
  const tpb::SBD d = tpb::generate_sbd_data(ac, const_cast<char**>(av));
  TEST_ASSERT(d.max_it == 20 && d.max_nb == 8 && d.eps == 1e-8 && d.do_rdm == 1);
  TEST_ASSERT(d.initial_adeterminant_bitstring == "0011" && d.dump_matrix_form_wf == "w.txt");
  TEST_ASSERT(d.bdet_comm_size == 2 && d.method == 0 && d.seed == 1729);
  bool threw = false;
  try {
    tpb::options_of(d);  // communicator sizes other than 1 are refused at the solve
  } catch (const std::invalid_argument&) {
    threw = true;
  }
  TEST_ASSERT(threw);
}

int main(int argc, char** argv) {
  test_from_string_kat();
  test_sort_bitarray();
  test_pack_roundtrip();
  test_options_defaults();
  test_from_string_checked();
  test_save_matrix_form_wf();
  test_generate_sbd_data();
  if (argc >= 4) test_load_inputs(argv[2], argv[3]);
  if (argc >= 4 && std::string(argv[1]) == "--gpu") {
    try {
      test_diag_h2o(argv[2], argv[3]);
      test_initial_determinant_h2o(argv[2], argv[3]);
    } catch (const std::exception& ex) {
      std::printf("FAIL exception: %s\n", ex.what());
      g_fail++;
    }
  }
  std::printf("%d passed, %d failed\n", g_pass, g_fail);
  return g_fail ? 1 : 0;
}
