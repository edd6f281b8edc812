/*
 * sbd_tpb.h -- C ABI of the MI355X tensor-product-basis (TPB) selected-basis
 * diagonalisation library (libsbd_tpb.so).
 *
 * It replaces the reference's TPB hot path behind its own host API and
 * operator plug-in interface (paths relative to the reference repository,
 * include/sbd/...):
 *
 *   reference                                            here
 *   -------------------------------------------------   ---------------------------
 *   sbd::SetupIntegrals  chemistry/basic/makeintegrals.h:124-174
 *                                                        sbd_ctx_create_fcidump
 *   sbd::MakeHelpers     chemistry/tpb/helper.h:729-873 (+GenerateExcitation :185-342)
 *                                                        sbd_make_helpers, sbd_table_*
 *   sbd::makeQChamDiagTerms chemistry/tpb/qcham.h:290-340
 *                                                        sbd_hdiag
 *   MultBase<ElemT>::run chemistry/basic/mult.h:94-102 / MultTPBThrust::run
 *                        chemistry/tpb/mult_thrust.h:88-90, 1170-1580;
 *                        CPU sbd::mult chemistry/tpb/mult.h:625-881
 *                                                        sbd_sigma (accumulate = 1)
 *   sbd::Davidson        chemistry/tpb/davidson.h:707-970,
 *                        chemistry/basic/davidson_thrust.h:97-104
 *                                                        sbd_davidson
 *   sbd::Lanczos         chemistry/tpb/lanczos.h:162-327 sbd_lanczos
 *   sbd::tpb::diag       chemistry/tpb/sbdiag.h:151-526 (methods 0-3)
 *                                                        sbd_diag
 *   sbd::InnerProduct    framework/dm_vector.h:23-51     sbd_inner
 *   sbd::DiagonalAdet/BdetReducedDensityMatrix chemistry/tpb/rdmat.h:11-58,
 *                        136-182                         sbd_string_weights
 *   sbd::OccupationDensity chemistry/tpb/occupation.h:10-95
 *                                                        sbd_occupation
 *   sbd::CarryOverAdet / CarryOverBdet chemistry/tpb/rdmat.h:60-134 (+ beta)
 *                                                        sbd_carryover
 *   sbd::Correlation     chemistry/tpb/correlation.h:379-643 (do_rdm != 0;
 *                        device form correlation_thrust.h) sbd_correlation
 *   sbd::SaveWavefunction / LoadWavefunction chemistry/tpb/restart.h:20-65,
 *                        414-657                         sbd_save_wavefunction,
 *                                                        sbd_load_wavefunction
 *   diag with loadname   chemistry/tpb/sbdiag.h:339-347  sbd_diag_from
 *   BasisInitVectorFromDeterminants chemistry/tpb/davidson.h:89-127 (diag with
 *                        initial_*determinant_bitstring, sbdiag.h:303-335)
 *                                                        sbd_diag_from_det
 *   sbd::SaveMatrixFormWF chemistry/tpb/restart.h:663-780 (dump_matrix_form_wf,
 *                        sbdiag.h:808-813)               sbd_save_matrix_form_wf
 *   sbd::SinglesExtendHalfdets chemistry/tpb/extend.h:14-185 (amplitude form)
 *                                                        sbd_singles_extend
 *                        chemistry/tpb/extend.h:186-324 (string form)
 *                                                        sbd_singles_extend_strings
 *   sbd::sort_bitarray   framework/bit_manipulation.h:421-449 sbd_sort_strings
 *   sbd::DecodeAlphaDets chemistry/basic/makedeterminants.h:61-114
 *                                                        sbd_decode_strings_bin
 *   carryover.bin writer apps/chemistry_tpb_selected_basis_diagonalization/
 *                        main.cc:245-260                 sbd_encode_strings_bin
 *
 * Conventions
 *   - Strings are packed one uint64 per alpha/beta string (orbital p -> bit p);
 *     norb <= 64.  sbd_pack_strings converts the reference's det_vector words
 *     (bit_length bits per size_t, little-endian) to this form.  Lists must be
 *     sorted ascending and unique, as LoadAlphaDets/sort_bitarray leave them.
 *   - Vectors are indexed W[ia * nb + ib] (beta fastest), as the reference's
 *     one-rank layout (tpb/davidson.h:171-173).
 *   - Functions taking vectors take DEVICE pointers and a HIP stream
 *     (hipStream_t passed as void*, NULL = the null stream).
 *   - Every function returns an sbd_status; sbd_last_error(ctx) gives text.
 *     The library never calls exit() and never throws across the ABI.
 *   - A context is used by one host thread at a time; distinct contexts may be
 *     used concurrently.
 */
#ifndef SBD_TPB_H
#define SBD_TPB_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  SBD_OK = 0,
  SBD_ERR_ARG = 1,      /* invalid argument (reference: std::invalid_argument) */
  SBD_ERR_IO = 2,       /* unreadable file (reference: std::runtime_error) */
  SBD_ERR_DEVICE = 3,   /* HIP error (reference: exit(EXIT_FAILURE)) */
  SBD_ERR_MEMORY = 4,   /* device allocation failed */
  SBD_ERR_STATE = 5,    /* call out of order (e.g. sigma before helpers) */
  SBD_ERR_NOT_CONVERGED = 6
} sbd_status;

typedef struct sbd_ctx sbd_ctx;

/* Options of sbd::tpb::SBD (tpb/sbdiag.h:24-54) that apply to one rank, plus
 * the fields this library adds (precision, verbose).  sbd_options_default
 * fills the reference's defaults. */
typedef struct {
  int method;        /* 0 = Davidson, 2 = Lanczos; 1 and 3 (the reference's
                        stored-H forms of 0 and 2) run as 0 and 2             */
  int max_it;        /* restarts; default 1                                    */
  int max_nb;        /* subspace size; default 10                              */
  double eps;        /* residual-norm tolerance; default 1e-4                  */
  double max_time;   /* seconds; default 86400                                 */
  int init;          /* 0 = Hartree-Fock determinant (W[0] = 1); 1 = random    */
  uint64_t seed;     /* default 1729 (used by init = 1)                        */
  int precision;     /* 0 = FP64; 1 = FP32 vectors/sigma + FP64 reductions,
                        Ritz and Gram-Schmidt, then an FP64 correction sweep   */
  double fp32_eps;   /* precision = 1: residual at which FP32 hands over to FP64
                        (default 1e-2)                                         */
  int verbose;       /* 1 = print per-step lines like the reference            */
} sbd_options;

/* Per-solve report. */
typedef struct {
  int steps;              /* Davidson steps taken (sigma builds inside Davidson) */
  int converged;          /* 1 if the residual fell below eps                    */
  double residual;        /* last residual norm                                  */
  double energy;          /* Ritz value of the last step                         */
  double t_tables;        /* seconds: excitation tables                          */
  double t_hdiag;         /* seconds: diagonal                                   */
  double t_davidson;      /* seconds: Davidson                                   */
  double t_sigma;         /* seconds inside sigma kernels during Davidson        */
  int steps_fp32;         /* precision = 1: steps taken in FP32                  */
  int nb_used;            /* subspace size of the FP64 phase (max_nb, or fewer
                             when 2 nb + 2 vectors would not fit free memory)    */
  int nb_used_fp32;       /* the same for the FP32 phase (precision = 1)         */
} sbd_report;

const char* sbd_status_string(int status);
const char* sbd_last_error(const sbd_ctx* ctx);
void sbd_options_default(sbd_options* opt);

/* Packs n strings given as the reference's det_vector rows (words_per_string
 * = ceil(norb / bit_length) size_t words, bit_length bits each; replaces
 * from_string / det_vector storage, framework/bit_manipulation.h:219-235)
 * into one uint64 each. */
int sbd_pack_strings(const size_t* words, size_t n, size_t bit_length, int norb, uint64_t* out);

/* Parses a text FCIDUMP exactly as sbd::LoadFCIDump (framework/fcidump.h:29-95)
 * and returns NORB, NELEC and the entry count.  Entries are read with
 * sbd_fcidump_entries into caller arrays (values[n], idx[4n]). */
int sbd_fcidump_read(const char* path, int* norb, int* nelec, size_t* n_entries);
int sbd_fcidump_entries(const char* path, double* values, int* idx, size_t n_entries);

/* Creates a context on the current HIP device from FCIDUMP entries (applied in
 * order, as SetupIntegrals does) and packed alpha/beta strings. */
sbd_ctx* sbd_ctx_create_fcidump(int norb, size_t n_entries, const double* values, const int* idx,
                                const uint64_t* a_strings, size_t na, const uint64_t* b_strings,
                                size_t nb, int* status);
/* Same, from I0, the L x L one-body matrix h (row-major) and the packed ERI
 * (size B(B+1)/2, B = L(L+1)/2, chemistry/basic/integrals.h:67-110). */
sbd_ctx* sbd_ctx_create(int norb, double I0, const double* h1, const double* eri_packed,
                        const uint64_t* a_strings, size_t na, const uint64_t* b_strings,
                        size_t nb, int* status);
void sbd_ctx_destroy(sbd_ctx* ctx);

/* Sigma kernel choice (no counterpart in the reference: a tuning knob of this
 * library).  0 (default): bases with fewer than 2 off-diagonal pairs per
 * alpha string plus per beta string, or whose bra rows do not fit LDS, use the
 * per-element kernel; others the split form (a streaming kernel for the
 * diagonal and alpha terms over column slabs, then a row kernel for the beta
 * and alpha-single x beta-single terms) when two or more rows fit LDS
 * together, else the row kernel; in bases with at least 4 alpha-single x
 * beta-single terms per determinant those terms go to a pass of their own.
 * 1: row kernel (with that pass where it applies); 2: per-element kernel;
 * 3: row kernel with every term inline; 4: split form (where one row fits
 * LDS).  All give the same bits, and the serving engine's batches follow the
 * same rule. */
int sbd_set_sigma_kernel(sbd_ctx* ctx, int kernel);
/* Slab width (columns: 256, 512 or 1024) of the split form's streaming kernel
 * (no counterpart in the reference: a tuning knob; results do not depend on it). */
int sbd_set_sigma_slab(sbd_ctx* ctx, int cols);

/* Dimensions: out[0] = na, out[1] = nb, out[2] = N = na*nb, out[3] = norb. */
int sbd_ctx_dims(const sbd_ctx* ctx, uint64_t* out);

/* Builds the alpha and beta single/double excitation tables on the GPU
 * (MakeHelpers at one rank: all bras x all kets, kets ascending). */
int sbd_make_helpers(sbd_ctx* ctx);

/* Table sizes: out[0..3] = alpha singles, alpha doubles, beta singles, beta
 * doubles (total pairs). */
int sbd_table_counts(const sbd_ctx* ctx, uint64_t* out);

/* Copies one table to host memory in the reference's TaskHelpers layout:
 * which = 0 SinglesFromAlpha/SinglesAlphaCrAn, 1 DoublesFromAlpha/DoublesAlphaCrAn,
 * 2 SinglesFromBeta/SinglesBetaCrAn, 3 DoublesFromBeta/DoublesBetaCrAn.
 * off[n+1] (CSR by bra), ket[count], cran[2*count] (singles) or [4*count]
 * (doubles), spin-orbital indices as helper.h:239-335 writes them. */
int sbd_table_get(const sbd_ctx* ctx, int which, int64_t* off, int64_t* ket, int32_t* cran);

/* hii[N] = <D|H|D> for every determinant (qcham.h:290-340, ZeroExcite order). */
int sbd_hdiag(sbd_ctx* ctx, double* hii, void* stream);
int sbd_hdiag_f32(sbd_ctx* ctx, float* hii, void* stream);

/* Sigma build, MultBase::run semantics: Wb += H * Wk with the diagonal taken
 * from hii (accumulate = 1).  accumulate = 0 computes Wb = H * Wk without
 * reading Wb (bit-identical to accumulating onto zeros).  Deterministic: one
 * writer per element, fixed summation order (diagonal, alpha singles, alpha
 * doubles, beta singles, beta doubles, alpha-beta) as tpb/mult.h:688-828. */
int sbd_sigma(sbd_ctx* ctx, const double* hii, const double* Wk, double* Wb, int accumulate,
              void* stream);
/* FP32 storage and FP32 per-determinant accumulation (matrix elements are
 * formed in FP64 and rounded once). */
int sbd_sigma_f32(sbd_ctx* ctx, const float* hii, const float* Wk, float* Wb, int accumulate,
                  void* stream);

/* FP64 inner product on device vectors (Kahan per thread, fixed-order sum). */
int sbd_inner(sbd_ctx* ctx, const double* x, const double* y, uint64_t n, double* out,
              void* stream);

/* sbd_inner without any device allocation: no counterpart in the reference;
 * for callers that cannot allocate device memory at call time (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same kernels and grid as sbd_inner, so the same bits.  Its reduction
 * scratch is a fixed device array of 1,088 doubles (8.5 KiB) per device,
 * held until the process ends; any n fits (at most 1,024 block partials and
 * the result).  Every call on a device shares that array: calls that overlap
 * on one device race.  Synchronises the stream; *out is host memory. */
int sbd_inner_noalloc(sbd_ctx* ctx, const double* x, const double* y, uint64_t n, double* out,
                      void* stream);

/* Single-root Davidson on device vectors (tpb/davidson.h:707-970): W is the
 * initial vector on entry and the normalised Ritz vector on exit; *energy is
 * the Ritz value of the last step.  hist (optional, host, 4*max_it*max_nb
 * doubles) receives (E0, |R|, seconds, 64) per step. */
int sbd_davidson(sbd_ctx* ctx, const double* hii, double* W, const sbd_options* opt,
                 double* energy, double* hist, sbd_report* rep);

/* Restarted Lanczos on device vectors (tpb/lanczos.h:162-327): each of
 * max_it passes runs the three-term recurrence from C[0] = W for up to max_nb
 * steps with no reorthogonalisation, stops when E0 changes by less than eps
 * between steps, and sets W = sum U[ib, 0] C[ib] (not renormalised, as the
 * reference).  *energy is the last E0; hist (optional, host, 4*max_it*max_nb
 * doubles) receives (E0, beta, seconds, 64) per step (beta 0 when the step
 * ended before computing it).  rep->converged = 1 when the E0 test stopped it,
 * rep->residual = the last |E0 change|. */
int sbd_lanczos(sbd_ctx* ctx, const double* hii, double* W, const sbd_options* opt,
                double* energy, double* hist, sbd_report* rep);

/* sbd::tpb::diag, one rank (tpb/sbdiag.h:151-526): tables, hii, initial
 * vector (opt->init), Davidson (method 0/1) or Lanczos (method 2/3), then
 * E = <W, H W> with a fresh sigma.
 * W_out (optional, host, N doubles) receives the final vector.  hist
 * (optional, host, 4*hist_cap doubles) receives per Davidson step
 * (E0, |R|, seconds since the solver started, 32 or 64 = vector precision). */
int sbd_diag(sbd_ctx* ctx, const sbd_options* opt, double* energy, double* W_out, double* hist,
             int hist_cap, sbd_report* rep);

/* sbd_diag from a given start vector W_init (host, N doubles; opt->init is
 * ignored): the warm start of sbd::tpb::diag with a loadname
 * (tpb/sbdiag.h:339-347), e.g. from sbd_load_wavefunction. */
int sbd_diag_from(sbd_ctx* ctx, const sbd_options* opt, const double* W_init, double* energy,
                  double* W_out, double* hist, int hist_cap, sbd_report* rep);

/* sbd_diag from the unit vector on the determinant (a_string, b_string) (one
 * uint64 per string, orbital p -> bit p; opt->init is ignored): the explicit
 * initial determinant of sbd::tpb::diag (initial_adeterminant_bitstring /
 * initial_bdeterminant_bitstring, tpb/sbdiag.h:303-335) through
 * BasisInitVectorFromDeterminants (tpb/davidson.h:89-127).  SBD_ERR_ARG with
 * "initial alpha (beta) determinant is not present in the basis" when a
 * string is not one of the context's. */
int sbd_diag_from_det(sbd_ctx* ctx, const sbd_options* opt, uint64_t a_string, uint64_t b_string,
                      double* energy, double* W_out, double* hist, int hist_cap,
                      sbd_report* rep);

/* Wavefunction files (tpb/restart.h:20-65, 414-657) at one rank: the file is
 * prefix + "000000" and holds size_t na, nb, words per string, the alpha then
 * beta strings as det_vector words of bit_length bits, then W (host, N
 * doubles, W[ia * nb + ib]).  Loading places each saved coefficient whose
 * alpha and beta strings are in this context's basis at their position here
 * (others are dropped) and normalises W; *placed (optional) = coefficients
 * placed (0: W is all zeros). */
int sbd_save_wavefunction(sbd_ctx* ctx, const char* prefix, size_t bit_length, const double* W);
int sbd_load_wavefunction(sbd_ctx* ctx, const char* prefix, size_t bit_length, double* W,
                          uint64_t* placed);

/* SaveMatrixFormWF (tpb/restart.h:663-780, called for dump_matrix_form_wf at
 * tpb/sbdiag.h:808-813) at one rank, from host arrays: by the file's extension
 * (get_extension, framework/bit_manipulation.h:1145-1166), "bin" writes the
 * na * nb doubles of W as they lie; "dat" or "txt" writes one line per
 * determinant, "<W> # <ia>: <alpha bits> <ib>: <beta bits>" (W as a default
 * ostream writes a double, strings as makestring: orbital norb - 1 first);
 * any other extension writes nothing.  Strings are one uint64 each. */
int sbd_save_matrix_form_wf(const char* filename, int norb, const uint64_t* a, size_t na,
                            const uint64_t* b, size_t nb, const double* W);

/* The nroots lowest states by block Davidson (Davidson-Liu, restarted GD+k
 * style from the current and previous Ritz vectors): sbd::tpb::diag
 * (chemistry/tpb/sbdiag.h:151-526) finds the lowest root only
 * (chemistry/tpb/davidson.h:707-970); this is its multi-root form.  nroots is
 * 1..13; max_nb is the largest subspace (raised to 3 nroots, at most 40), max_it the
 * number of block iterations, eps the residual every root must reach; init 0
 * starts from the determinants with the lowest diagonal, init 1 from seeded
 * random vectors (seed, seed + 1, ...).  Where H splits into blocks that do
 * not couple, the init 0 start finds the lowest states of the blocks its
 * determinants lie in only (on the norb-48 open-shell test basis it misses the
 * ground state); init 1 reaches every block.  precision = 1 runs the FP32 block
 * phase to fp32_eps first.  energies[nroots] = <W_r, H W_r> (fresh FP64
 * sigma), residuals[nroots] (optional), W_out (optional, host, nroots * N,
 * root-major).  Returns SBD_OK also when max_it ends the iterations first
 * (rep->converged tells). */
int sbd_diag_roots(sbd_ctx* ctx, const sbd_options* opt, int nroots, double* energies,
                   double* residuals, double* W_out, double* hist, int hist_cap,
                   sbd_report* rep);

/* Measurements on a device vector W (N doubles, the layout above), as
 * sbd::tpb::diag makes them after the solve (tpb/sbdiag.h:666-777).
 *
 * String weights (DiagonalAdet/BdetReducedDensityMatrix, tpb/rdmat.h:11-58,
 * 136-182): wa[ia] = sum_ib W^2, wb[ib] = sum_ia W^2 (host arrays of na and nb
 * doubles; fixed-order compensated sums). */
int sbd_string_weights(sbd_ctx* ctx, const double* W, double* wa, double* wb, void* stream);
/* Occupation density (OccupationDensity, tpb/occupation.h:10-95, for every
 * orbital): density[2o] = alpha and density[2o+1] = beta occupation of
 * orbital o (host array of 2 norb doubles), the reference's `density` output. */
int sbd_occupation(sbd_ctx* ctx, const double* W, double* density, void* stream);
/* One- and two-particle reduced density matrices (Correlation,
 * tpb/correlation.h:379-643, with ZeroDiff/OneDiff/TwoDiffCorrelation of
 * basic/correlation.h:16-247; tpb/sbdiag.h:704-728 for do_rdm != 0).  Host
 * outputs in the reference's layout:
 *   onebody[s*L^2 + o1 + L*o2]          (2 L^2 doubles; s = 0 alpha, 1 beta)
 *   twobody[(s1 + 2*s2)*L^4 + o1 + L*o2 + L^2*o3 + L^3*o4]   (4 L^4 doubles)
 * with L = norb; same-spin blocks equal <a+_{o1} a+_{o2} a_{o4} a_{o3}>.  The
 * alpha x beta double-excitation terms are written where the reference writes
 * them, including the terms whose lowest bra-only spin orbital has the spin
 * of the highest ket-only one, which the reference adds with the block index
 * s1 + 2*s2 of the other spin assignment.  Sums are taken in a fixed order;
 * the reference's order (atomics) differs, so results agree with it to
 * rounding.  Requires sbd_make_helpers; uses 8 N bytes of device scratch. */
int sbd_correlation(sbd_ctx* ctx, const double* W, double* onebody, double* twobody,
                    void* stream);
/* Carry-over strings (CarryOverAdet / CarryOverBdet, tpb/rdmat.h:60-134; the
 * selection of tpb/sbdiag.h:745-772): strings in descending order of weight
 * (ties by index).  ratio == 0: keep up to and including the first string
 * whose cumulative weight exceeds 1 - threshold (all strings if none does);
 * ratio > 0: keep (size_t)(ratio * count) strings.  a_out / b_out (host,
 * capacity na / nb) receive the kept strings in that order, *na_out / *nb_out
 * their numbers, truncated[2] (optional) 1 - the kept weight per spin. */
int sbd_carryover(sbd_ctx* ctx, const double* W, double threshold, double ratio,
                  uint64_t* a_out, size_t* na_out, uint64_t* b_out, size_t* nb_out,
                  double* truncated, void* stream);

/* Singles extension (carry-over types 2 and 3 of tpb/sbdiag.h:773-791).
 * Amplitude form (SinglesExtendHalfdets, tpb/extend.h:14-185): the strings
 * of the determinants with W^2 > cutoff (device W), each with all of its
 * single excitations, sorted ascending without duplicates; *total_weight =
 * the weight of those determinants.  String form (tpb/extend.h:186-324): the
 * same extension of given strings.  Outputs are host arrays of capacity cap_*;
 * a too small capacity is SBD_ERR_ARG with the needed size in *n*_out. */
int sbd_singles_extend(sbd_ctx* ctx, const double* W, double cutoff, uint64_t* a_out,
                       size_t cap_a, size_t* na_out, uint64_t* b_out, size_t cap_b,
                       size_t* nb_out, double* total_weight, void* stream);
int sbd_singles_extend_strings(int norb, const uint64_t* in, size_t n, uint64_t* out, size_t cap,
                               size_t* n_out);
/* sort_bitarray (framework/bit_manipulation.h:421-449): ascending, duplicates
 * removed, in place; *n_out = strings kept. */
int sbd_sort_strings(uint64_t* s, size_t n, size_t* n_out);
/* Binary string files of ceil(norb/8) bytes per string.  Decoding follows
 * DecodeAlphaDets (chemistry/basic/makedeterminants.h:61-114): the record is a
 * big-endian number whose low norb bits are the string (orbital p = bit p);
 * strings come out in file order, unsorted.  Encoding follows the
 * carryover.bin writer (main.cc:245-260): orbital norb-1 is the first bit of
 * the record and the padding bits are the last, so for norb % 8 != 0 the
 * reference's writer and reader disagree; both are kept as the reference has
 * them.  Decoding into too small a capacity is SBD_ERR_ARG with *n_out = the
 * number of strings in the file. */
int sbd_decode_strings_bin(const char* path, int norb, uint64_t* out, size_t cap, size_t* n_out);
int sbd_encode_strings_bin(const char* path, int norb, const uint64_t* s, size_t n);

/* Device bytes sbd_diag would allocate for these options (vectors + tables +
 * integrals), for admission control. */
int sbd_diag_footprint(const sbd_ctx* ctx, const sbd_options* opt, uint64_t* bytes);

/* ------------------------------------------------------------------ serving
 * A server runs many independent solves on the current GPU with continuous
 * batching: requests are admitted as device memory (mem_budget bytes) and
 * slots (max_active) allow, and requests that arrive while others run join
 * them as memory and slots free up.  Single-root FP64 solves (the
 * reference's sbd::tpb::diag, chemistry/tpb/sbdiag.h:151-526, method 0) are
 * advanced one Davidson step per engine round, each kind of device pass for
 * all of them in one launch; each computes exactly as alone (same kernels,
 * grids and fixed-order sums), so its energy is bit-identical alone, in any
 * mix and to sbd_diag with the same options.  Requests that the batched form
 * does not take -- precision 1 (FP32 Davidson + FP64 sweep) or nroots > 1
 * (block Davidson, sbd_diag_roots) -- run whole on worker threads of their
 * own, each on its own stream, concurrently with the batches; their results
 * equal sbd_diag / sbd_diag_roots with the same options.  sbd_server_create
 * reserves one device arena of three quarters of mem_budget (less if the
 * device has less free memory) from which the batched solves' Davidson vectors are
 * handed out and to which they return, with no hipMalloc / hipFree per solve;
 * tables and the whole-solve workers allocate per solve.  The reference has no
 * serving mode.  Server calls are thread-safe. */
typedef struct sbd_server sbd_server;

#define SBD_MAX_ROOTS 13

typedef struct {
  int norb;
  double I0;
  const double* h1;         /* norb x norb one-body integrals, row-major */
  const double* eri_packed; /* B(B+1)/2, B = norb(norb+1)/2 */
  const uint64_t* a_strings;
  size_t na;
  const uint64_t* b_strings;
  size_t nb;
  sbd_options opt; /* method 0; precision 0 or 1; max_it, max_nb, eps, init, seed,
                      fp32_eps */
  int nroots;      /* 0 or 1: lowest root; 2..SBD_MAX_ROOTS: block Davidson */
} sbd_request;

typedef struct {
  double energy;   /* <W, H W> of the final vector, as sbd_diag (lowest root) */
  double residual; /* |H W - E0 W| of the last step (largest over the roots) */
  int converged;
  int steps;
  int status;      /* sbd_status of the solve */
  double t_submit, t_admit, t_done; /* seconds on the server clock */
  int nroots;      /* roots solved for */
  double energies[SBD_MAX_ROOTS]; /* <W_r, H W_r> per root, ascending */
  int steps_fp32;  /* precision 1: steps taken in FP32 */
} sbd_result;

sbd_server* sbd_server_create(uint64_t mem_budget, int max_active, int* status);
/* Copies the request's inputs and queues it; returns at once. */
int sbd_server_submit(sbd_server* sv, const sbd_request* rq, uint64_t* id);
/* Blocks until the solve id has finished. */
int sbd_server_wait(sbd_server* sv, uint64_t id, sbd_result* out);
int sbd_server_poll(sbd_server* sv, uint64_t id, sbd_result* out, int* done);
double sbd_server_now(const sbd_server* sv);
/* Admission order (no counterpart in the reference: a serving addition).
 * policy 0 (default): arrival order.  policy 1: the queued request with the
 * smallest device footprint first, except that a request that has waited
 * age_limit_s seconds or more goes first.  Results do not depend on it. */
int sbd_server_set_policy(sbd_server* sv, int policy, double age_limit_s);
/* out[0] = engine rounds so far, out[1] = most solves active in one round. */
int sbd_server_stats(const sbd_server* sv, uint64_t* out);
/* Finishes every queued solve, then frees the server. */
void sbd_server_destroy(sbd_server* sv);

#ifdef __cplusplus
}
#endif

#endif /* SBD_TPB_H */
