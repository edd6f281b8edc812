// Wavefunction files: SaveWavefunction / LoadWavefunction
// (chemistry/tpb/restart.h:20-65, 414-657) and SaveMatrixFormWF
// (restart.h:663-780) at one rank.
//
// File layout (one file per rank; at one rank the name is the prefix followed
// by "000000", to_padded_string(rank, 6), restart.h:14-18, 52):
//   size_t na, size_t nb, size_t words_per_string,
//   na alpha strings, then nb beta strings, each as words_per_string size_t
//   words of bit_length bits (the det_vector layout, orbital p in word
//   p / bit_length, bit p % bit_length),
//   na * nb coefficients (double), W[ia * nb + ib].
// Loading places every saved coefficient whose alpha and beta strings are both
// in the current basis at their current position (others are dropped), then
// normalises, as LoadWavefunction does.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "sbd_internal.hpp"

using namespace sbdx;

namespace {

std::string rank_file(const char* prefix) { return std::string(prefix) + "000000"; }

void unpack_words(uint64_t s, size_t bit_length, size_t wps, size_t* out) {
  for (size_t w = 0; w < wps; w++) out[w] = 0;
  for (int p = 0; p < 64; p++)
    if ((s >> p) & 1ULL) out[p / bit_length] |= (size_t)1 << (p % bit_length);
}

bool pack_words(const size_t* words, size_t bit_length, size_t wps, int norb, uint64_t* out) {
  uint64_t v = 0;
  for (size_t w = 0; w < wps; w++) {
    for (size_t b = 0; b < bit_length; b++) {
      if (!((words[w] >> b) & 1)) continue;
      const size_t p = w * bit_length + b;
      if (p >= (size_t)norb) return false;
      v |= 1ULL << p;
    }
  }
  *out = v;
  return true;
}

// get_extension (framework/bit_manipulation.h:1145-1166): the text after the
// last '.', lower-cased, with NUL and space characters dropped past its first
// character; empty when the name has no '.'.
std::string file_extension(const std::string& path) {
  const size_t dot = path.rfind('.');
  if (dot == std::string::npos) return "";
  std::string ext = path.substr(dot + 1);
  for (char& ch : ext) ch = (char)std::tolower((unsigned char)ch);
  std::string out;
  for (size_t i = 0; i < ext.size(); i++)
    if (i == 0 || (ext[i] != 0 && ext[i] != ' ')) out += ext[i];
  return out;
}

// makestring (framework/bit_manipulation.h:196-211) of a one-word string:
// orbitals norb - 1 down to 0 as '1' / '0'.
std::string bit_text(uint64_t s, int norb) {
  std::string t((size_t)norb, '0');
  for (int p = 0; p < norb; p++)
    if ((s >> p) & 1ULL) t[(size_t)(norb - 1 - p)] = '1';
  return t;
}

}  // namespace

extern "C" {

int sbd_save_matrix_form_wf(const char* filename, int norb, const uint64_t* a, size_t na,
                            const uint64_t* b, size_t nb, const double* W) {
  if (!filename || norb < 1 || norb > 64 || (na && !a) || (nb && !b) || (na * nb && !W))
    return SBD_ERR_ARG;
  const std::string ext = file_extension(filename);
  if (ext == "bin") {
    // raw coefficients, W[ia * nb + ib] in index order
    FILE* f = std::fopen(filename, "wb");
    if (!f) return SBD_ERR_IO;
    bool ok = std::fwrite(W, sizeof(double), na * nb, f) == na * nb;
    ok = (std::fclose(f) == 0) && ok;
    return ok ? SBD_OK : SBD_ERR_IO;
  }
  if (ext != "dat" && ext != "txt") return SBD_OK;  // other extensions write nothing
  // one line per determinant: "<value> # <ia>: <alpha string> <ib>: <beta
  // string>", the value as an ostream with its default format writes a
  // double (six significant digits, %g)
  FILE* f = std::fopen(filename, "w");
  if (!f) return SBD_ERR_IO;
  bool ok = true;
  std::vector<std::string> bt(nb);
  for (size_t ib = 0; ib < nb; ib++) bt[ib] = bit_text(b[ib], norb);
  for (size_t ia = 0; ia < na && ok; ia++) {
    const std::string at = bit_text(a[ia], norb);
    for (size_t ib = 0; ib < nb && ok; ib++)
      ok = std::fprintf(f, "%g # %zu: %s %zu: %s\n", W[ia * nb + ib], ia, at.c_str(), ib,
                        bt[ib].c_str()) > 0;
  }
  ok = (std::fclose(f) == 0) && ok;
  return ok ? SBD_OK : SBD_ERR_IO;
}

int sbd_save_wavefunction(sbd_ctx* ctx, const char* prefix, size_t bit_length, const double* W) {
  if (!ctx || !prefix || !W || bit_length == 0 || bit_length > 64) return SBD_ERR_ARG;
  const size_t wps = (ctx->norb + bit_length - 1) / bit_length;
  const std::string fn = rank_file(prefix);
  FILE* f = std::fopen(fn.c_str(), "wb");
  if (!f) return set_err(ctx, SBD_ERR_IO, "cannot write " + fn);
  const size_t hdr[3] = {(size_t)ctx->na, (size_t)ctx->nb, wps};
  bool ok = std::fwrite(hdr, sizeof(size_t), 3, f) == 3;
  std::vector<size_t> words(wps);
  for (const auto* S : {&ctx->a_host, &ctx->b_host})
    for (uint64_t s : *S) {
      unpack_words(s, bit_length, wps, words.data());
      ok = ok && std::fwrite(words.data(), sizeof(size_t), wps, f) == wps;
    }
  ok = ok && std::fwrite(W, sizeof(double), ctx->N, f) == ctx->N;
  ok = (std::fclose(f) == 0) && ok;
  return ok ? SBD_OK : set_err(ctx, SBD_ERR_IO, "short write to " + fn);
}

int sbd_load_wavefunction(sbd_ctx* ctx, const char* prefix, size_t bit_length, double* W,
                          uint64_t* placed) {
  if (!ctx || !prefix || !W || bit_length == 0 || bit_length > 64) return SBD_ERR_ARG;
  const std::string fn = rank_file(prefix);
  FILE* f = std::fopen(fn.c_str(), "rb");
  if (!f) return set_err(ctx, SBD_ERR_IO, "failed to open input file " + fn);
  size_t hdr[3] = {0, 0, 0};
  auto fail = [&](const std::string& m) {
    std::fclose(f);
    return set_err(ctx, SBD_ERR_IO, fn + ": " + m);
  };
  if (std::fread(hdr, sizeof(size_t), 3, f) != 3) return fail("truncated header");
  const size_t la = hdr[0], lb = hdr[1], wps = hdr[2];
  if (wps != (ctx->norb + bit_length - 1) / bit_length)
    return fail("words per string do not match norb and bit_length");
  // current positions of the saved strings (-1 = not in this basis)
  std::unordered_map<uint64_t, int64_t> apos, bpos;
  for (uint64_t i = 0; i < ctx->na; i++) apos.emplace(ctx->a_host[i], (int64_t)i);
  for (uint64_t i = 0; i < ctx->nb; i++) bpos.emplace(ctx->b_host[i], (int64_t)i);
  std::vector<int64_t> ja(la, -1), jb(lb, -1);
  std::vector<size_t> words(wps);
  for (size_t k = 0; k < la + lb; k++) {
    if (std::fread(words.data(), sizeof(size_t), wps, f) != wps) return fail("truncated strings");
    uint64_t s = 0;
    if (!pack_words(words.data(), bit_length, wps, ctx->norb, &s)) continue;
    auto& m = k < la ? apos : bpos;
    auto it = m.find(s);
    if (it != m.end()) (k < la ? ja[k] : jb[k - la]) = it->second;
  }
  std::fill(W, W + ctx->N, 0.0);
  std::vector<double> row(lb);
  uint64_t n_placed = 0;
  for (size_t ia = 0; ia < la; ia++) {
    if (std::fread(row.data(), sizeof(double), lb, f) != lb) return fail("truncated coefficients");
    if (ja[ia] < 0) continue;
    for (size_t ib = 0; ib < lb; ib++) {
      if (jb[ib] < 0) continue;
      W[(uint64_t)ja[ia] * ctx->nb + (uint64_t)jb[ib]] += row[ib];
      n_placed++;
    }
  }
  std::fclose(f);
  // Normalize (framework/dm_vector.h:54-87): compensated sum of squares
  double s = 0.0, c = 0.0;
  for (uint64_t i = 0; i < ctx->N; i++) {
    const double y = W[i] * W[i] - c;
    const double t = s + y;
    c = (t - s) - y;
    s = t;
  }
  if (s > 0.0) {
    const double fct = 1.0 / std::sqrt(s);
    for (uint64_t i = 0; i < ctx->N; i++) W[i] *= fct;
  }
  if (placed) *placed = n_placed;
  return SBD_OK;
}

}  // extern "C"
