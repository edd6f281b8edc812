/*
 * sbd_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the reference's tensor-product-basis (TPB) path: excitation
 * tables, diagonal, sigma (H*v), single-root Davidson and Lanczos, and the
 * measurements on a solved vector (string weights, occupations, carry-over,
 * the full one- and two-particle RDMs).  It is the parity
 * checker for the HIP library and the timed CPU baseline; it is never part of
 * the product path.  Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may load it.
 *
 * Each function follows the cited reference lines (paths relative to the
 * reference repository root, include/sbd/...).  Strings are packed as one
 * uint64 word per half string (norb <= 64), i.e. the reference's half
 * det_vector with bit_length = 64; the full (interleaved) determinant is two
 * uint64 words (bit_length 64, D_size = 2).  Every formula below is
 * independent of bit_length, so results equal the reference's for any
 * bit_length.
 *
 * Build: see oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off).
 * Parity pins: tests/test_oracle.py and tests/test_host_abi.py (golden vectors
 * from the reference's own tests, reference-held energies, brute-force second
 * quantisation: the Hamiltonian matrix and the RDMs, closed forms).
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* integrals: chemistry/basic/integrals.h:13-25 (oneInt), :67-110 (twoInt) */

static inline double eri_value(const double *I2, int i, int j, int k, int l) {
  /* twoInt::Value, integrals.h:94-102 */
  if (!((i % 2 == j % 2) && (k % 2 == l % 2))) return 0.0;
  int I = i / 2, J = j / 2, K = k / 2, L = l / 2;
  int mij = I > J ? I : J, nij = I > J ? J : I;
  int mkl = K > L ? K : L, nkl = K > L ? L : K;
  long ij = (long)mij * (mij + 1) / 2 + nij;
  long kl = (long)mkl * (mkl + 1) / 2 + nkl;
  long a = ij > kl ? ij : kl, b = ij > kl ? kl : ij;
  return I2[a * (a + 1) / 2 + b];
}

static inline long eri_index(int I, int J, int K, int L) {
  int mij = I > J ? I : J, nij = I > J ? J : I;
  int mkl = K > L ? K : L, nkl = K > L ? L : K;
  long ij = (long)mij * (mij + 1) / 2 + nij;
  long kl = (long)mkl * (mkl + 1) / 2 + nkl;
  long a = ij > kl ? ij : kl, b = ij > kl ? kl : ij;
  return a * (a + 1) / 2 + b;
}

/*
 * SetupIntegrals (RHF), chemistry/basic/makeintegrals.h:124-174.
 * Entries are applied in file order; later entries overwrite earlier ones.
 *   I1: (2L)x(2L), index i*2L+j.  I2: packed, B(B+1)/2 with B = L(L+1)/2.
 *   Dm/Xm: Direct/Exchange, (2L)x(2L), index i + 2L*j.
 */
void orc_setup_integrals(int norb, long n, const double *val, const int *ii, const int *jj,
                         const int *kk, const int *ll, double *I0, double *I1, double *I2,
                         double *Dm, double *Xm) {
  int L = norb, L2 = 2 * norb;
  long B = (long)L * (L + 1) / 2;
  memset(I1, 0, sizeof(double) * L2 * L2);
  memset(I2, 0, sizeof(double) * (B * (B + 1) / 2));
  for (long e = 0; e < n; e++) {
    int i = ii[e], j = jj[e], k = kk[e], l = ll[e];
    double v = val[e];
    if (i == 0 && k == 0 && j == 0 && l == 0) {
      *I0 = v;
    } else if (k == l && k == 0) {
      I1[(2 * (i - 1)) * L2 + 2 * (j - 1)] = v;
      I1[(2 * (i - 1) + 1) * L2 + 2 * (j - 1) + 1] = v;
      I1[(2 * (j - 1)) * L2 + 2 * (i - 1)] = v;
      I1[(2 * (j - 1) + 1) * L2 + 2 * (i - 1) + 1] = v;
    } else {
      I2[eri_index(i - 1, j - 1, k - 1, l - 1)] = v;
    }
  }
  for (int i = 0; i < L; i++) {
    for (int j = 0; j < L; j++) {
      Dm[(2 * i) + L2 * (2 * j)] = eri_value(I2, 2 * i, 2 * i, 2 * j, 2 * j);
      Dm[(2 * i) + L2 * (2 * j + 1)] = eri_value(I2, 2 * i, 2 * i, 2 * j + 1, 2 * j + 1);
      Dm[(2 * i + 1) + L2 * (2 * j)] = eri_value(I2, 2 * i + 1, 2 * i + 1, 2 * j, 2 * j);
      Dm[(2 * i + 1) + L2 * (2 * j + 1)] = eri_value(I2, 2 * i + 1, 2 * i + 1, 2 * j + 1, 2 * j + 1);
      Xm[(2 * i) + L2 * (2 * j)] = eri_value(I2, 2 * i, 2 * j, 2 * j, 2 * i);
      Xm[(2 * i) + L2 * (2 * j + 1)] = eri_value(I2, 2 * i, 2 * j + 1, 2 * j + 1, 2 * i);
      Xm[(2 * i + 1) + L2 * (2 * j)] = eri_value(I2, 2 * i + 1, 2 * j, 2 * j, 2 * i + 1);
      Xm[(2 * i + 1) + L2 * (2 * j + 1)] = eri_value(I2, 2 * i + 1, 2 * j + 1, 2 * j + 1, 2 * i + 1);
    }
  }
}

/* ------------------------------------------------------------------ */
/* determinants: chemistry/basic/determinants.h */

/* DetFromAlphaBeta (non-TRADMODE), determinants.h:104-150, bit_length 64. */
static inline void det_from_ab(uint64_t a, uint64_t b, int L, uint64_t D[2]) {
  D[0] = D[1] = 0;
  for (int i = 0; i < L; i++) {
    if (a & (1ULL << i)) D[(2 * i) / 64] |= 1ULL << ((2 * i) % 64);
    if (b & (1ULL << i)) D[(2 * i + 1) / 64] |= 1ULL << ((2 * i + 1) % 64);
  }
}

void orc_det_from_ab(const uint64_t *a, const uint64_t *b, long n, int L, uint64_t *D) {
  for (long k = 0; k < n; k++) det_from_ab(a[k], b[k], L, D + 2 * k);
}

static inline int getocc(const uint64_t *D, int x) { return (int)((D[x / 64] >> (x % 64)) & 1ULL); }

/* parity(dets, bit_length, start, end, sgn), determinants.h:242-294, bit_length 64. */
static inline void parity(const uint64_t *D, int start, int end, double *sgn) {
  int bs = start / 64, bi = start % 64, be = end / 64, bj = end % 64;
  int nz = 0;
  if (bs == be) {
    uint64_t mask = ((1ULL << bj) - 1) ^ ((1ULL << bi) - 1);
    nz += __builtin_popcountll(D[bs] & mask);
  } else {
    if (bi != 0) {
      uint64_t mask = ~((1ULL << bi) - 1);
      nz += __builtin_popcountll(D[bs] & mask);
      bs++;
    }
    for (int w = bs; w < be; w++) nz += __builtin_popcountll(D[w]);
    if (bj != 0) {
      uint64_t mask = (1ULL << bj) - 1;
      nz += __builtin_popcountll(D[be] & mask);
    }
  }
  *sgn *= (-2. * (nz % 2) + 1);
  if ((D[start / 64] >> (start % 64)) & 1ULL) *sgn *= -1.;
}

/* ZeroExcite, determinants.h:451-474 (Direct/Exchange index i + 2L*j). */
static inline double zero_excite(const uint64_t *D, int L, double I0, const double *I1,
                                 const double *Dm, const double *Xm) {
  double energy = 0.0;
  int L2 = 2 * L;
  for (int i = 0; i < L2; i++) {
    if (getocc(D, i)) {
      energy += I1[i * L2 + i];
      for (int j = i + 1; j < L2; j++) {
        if (getocc(D, j)) {
          energy += Dm[i + L2 * j];
          if ((i % 2) == (j % 2)) energy -= Xm[i + L2 * j];
        }
      }
    }
  }
  return energy + I0;
}

/* OneExcite, determinants.h:476-498.  det.size() = D_size words of 64 bits. */
static inline double one_excite(const uint64_t *D, int dsize, int i, int a, int L2,
                                const double *I1, const double *I2) {
  double sgn = 1.0;
  parity(D, i < a ? i : a, i < a ? a : i, &sgn);
  double energy = I1[a * L2 + i];
  for (int x = 0; x < dsize; x++) {
    uint64_t bits = D[x];
    for (int pos = 0; pos < 64; pos++) {
      if ((bits & 1ULL) == 1ULL) {
        int j = x * 64 + pos;
        energy += (eri_value(I2, a, i, j, j) - eri_value(I2, a, j, j, i));
      }
      bits >>= 1;
    }
  }
  energy *= sgn;
  return energy;
}

/* TwoExcite, determinants.h:500-518. */
static inline double two_excite(const uint64_t *D, int i, int j, int a, int b, const double *I2) {
  double sgn = 1.0;
  int I = i < j ? i : j, J = i < j ? j : i;
  int A = a < b ? a : b, B = a < b ? b : a;
  parity(D, I < A ? I : A, I < A ? A : I, &sgn);
  parity(D, J < B ? J : B, J < B ? B : J, &sgn);
  if (A > J || B < I) sgn *= -1.0;
  return sgn * (eri_value(I2, A, I, B, J) - eri_value(I2, A, J, B, I));
}

double orc_zero_excite(const uint64_t *D, int L, double I0, const double *I1, const double *Dm,
                       const double *Xm) {
  return zero_excite(D, L, I0, I1, Dm, Xm);
}
double orc_one_excite(const uint64_t *D, int L, int i, int a, const double *I1, const double *I2) {
  return one_excite(D, (2 * L + 63) / 64, i, a, 2 * L, I1, I2);
}
double orc_two_excite(const uint64_t *D, int i, int j, int a, int b, const double *I2) {
  return two_excite(D, i, j, a, b, I2);
}
double orc_parity(const uint64_t *D, int start, int end) {
  double s = 1.0;
  parity(D, start, end, &s);
  return s;
}

/* ------------------------------------------------------------------ */
/* excitation tables: chemistry/tpb/helper.h:185-342 (GenerateExcitation),
 * difference / OrbitalDifference determinants.h:380-449.
 * spin = 0 (alpha, CrAn 2*orb) or 1 (beta, CrAn 2*orb+1).
 * Pass 1 (kets == NULL): fills s_cnt/d_cnt per bra.  Pass 2: fills CSR rows at
 * the given offsets, kets ascending. */

static inline void orbital_difference(uint64_t a, uint64_t b, int L, int *cr, int *an) {
  int nc = 0, na = 0;
  uint64_t da = a & ~b, db = b & ~a;
  for (int p = 0; p < L; p++) {
    if (da & (1ULL << p)) cr[nc++] = p;
    if (db & (1ULL << p)) an[na++] = p;
  }
}

void orc_tables(const uint64_t *s, long n, int L, int spin, long *s_cnt, long *d_cnt,
                const long *s_off, const long *d_off, long *s_ket, int *s_cran, long *d_ket,
                int *d_cran) {
  uint64_t mask = (L >= 64) ? ~0ULL : ((1ULL << L) - 1);
#pragma omp parallel for schedule(dynamic, 16)
  for (long ia = 0; ia < n; ia++) {
    long ns = 0, nd = 0;
    for (long ja = 0; ja < n; ja++) {
      int d = __builtin_popcountll((s[ia] ^ s[ja]) & mask);
      if (d == 2) {
        if (s_ket) {
          int cr[2], an[2];
          orbital_difference(s[ia], s[ja], L, cr, an);
          long o = s_off[ia] + ns;
          s_ket[o] = ja;
          s_cran[2 * o + 0] = 2 * cr[0] + spin;
          s_cran[2 * o + 1] = 2 * an[0] + spin;
        }
        ns++;
      } else if (d == 4) {
        if (d_ket) {
          int cr[2], an[2];
          orbital_difference(s[ia], s[ja], L, cr, an);
          long o = d_off[ia] + nd;
          d_ket[o] = ja;
          d_cran[4 * o + 0] = 2 * cr[0] + spin;
          d_cran[4 * o + 1] = 2 * cr[1] + spin;
          d_cran[4 * o + 2] = 2 * an[0] + spin;
          d_cran[4 * o + 3] = 2 * an[1] + spin;
        }
        nd++;
      }
    }
    if (s_cnt) s_cnt[ia] = ns;
    if (d_cnt) d_cnt[ia] = nd;
  }
}

/* ------------------------------------------------------------------ */
/* system context */

typedef struct {
  int L;
  long na, nb;
  uint64_t *a, *b;
  double I0;
  double *I1, *I2, *Dm, *Xm;
  long *as_off, *ad_off, *bs_off, *bd_off; /* n+1 */
  long *as_ket, *ad_ket, *bs_ket, *bd_ket;
  int *as_cr, *ad_cr, *bs_cr, *bd_cr;
} orc_sys;

static void build_tables(const uint64_t *s, long n, int L, int spin, long **s_off, long **d_off,
                         long **s_ket, int **s_cr, long **d_ket, int **d_cr) {
  long *sc = (long *)calloc(n + 1, sizeof(long)), *dc = (long *)calloc(n + 1, sizeof(long));
  orc_tables(s, n, L, spin, sc, dc, NULL, NULL, NULL, NULL, NULL, NULL);
  long *so = (long *)malloc((n + 1) * sizeof(long)), *dof = (long *)malloc((n + 1) * sizeof(long));
  so[0] = dof[0] = 0;
  for (long i = 0; i < n; i++) {
    so[i + 1] = so[i] + sc[i];
    dof[i + 1] = dof[i] + dc[i];
  }
  *s_ket = (long *)malloc((so[n] + 1) * sizeof(long));
  *s_cr = (int *)malloc((2 * so[n] + 2) * sizeof(int));
  *d_ket = (long *)malloc((dof[n] + 1) * sizeof(long));
  *d_cr = (int *)malloc((4 * dof[n] + 4) * sizeof(int));
  orc_tables(s, n, L, spin, NULL, NULL, so, dof, *s_ket, *s_cr, *d_ket, *d_cr);
  *s_off = so;
  *d_off = dof;
  free(sc);
  free(dc);
}

orc_sys *orc_sys_create(int L, const uint64_t *a, long na, const uint64_t *b, long nb, double I0,
                        const double *I1, const double *I2, const double *Dm, const double *Xm) {
  orc_sys *S = (orc_sys *)calloc(1, sizeof(orc_sys));
  long B = (long)L * (L + 1) / 2, nI2 = B * (B + 1) / 2, L2 = 2 * L;
  S->L = L;
  S->na = na;
  S->nb = nb;
  S->a = (uint64_t *)malloc(na * 8);
  memcpy(S->a, a, na * 8);
  S->b = (uint64_t *)malloc(nb * 8);
  memcpy(S->b, b, nb * 8);
  S->I0 = I0;
  S->I1 = (double *)malloc(L2 * L2 * 8);
  memcpy(S->I1, I1, L2 * L2 * 8);
  S->I2 = (double *)malloc(nI2 * 8);
  memcpy(S->I2, I2, nI2 * 8);
  S->Dm = (double *)malloc(L2 * L2 * 8);
  memcpy(S->Dm, Dm, L2 * L2 * 8);
  S->Xm = (double *)malloc(L2 * L2 * 8);
  memcpy(S->Xm, Xm, L2 * L2 * 8);
  build_tables(S->a, na, L, 0, &S->as_off, &S->ad_off, &S->as_ket, &S->as_cr, &S->ad_ket, &S->ad_cr);
  build_tables(S->b, nb, L, 1, &S->bs_off, &S->bd_off, &S->bs_ket, &S->bs_cr, &S->bd_ket, &S->bd_cr);
  return S;
}

void orc_sys_free(orc_sys *S) {
  if (!S) return;
  void *p[] = {S->a, S->b, S->I1, S->I2, S->Dm, S->Xm, S->as_off, S->ad_off, S->bs_off, S->bd_off,
               S->as_ket, S->ad_ket, S->bs_ket, S->bd_ket, S->as_cr, S->ad_cr, S->bs_cr, S->bd_cr};
  for (unsigned k = 0; k < sizeof(p) / sizeof(p[0]); k++) free(p[k]);
  free(S);
}

/* pair counts: out[0..3] = alpha singles, alpha doubles, beta singles, beta doubles */
void orc_sys_counts(const orc_sys *S, long *out) {
  out[0] = S->as_off[S->na];
  out[1] = S->ad_off[S->na];
  out[2] = S->bs_off[S->nb];
  out[3] = S->bd_off[S->nb];
}

/* copy one table out: which = 0 as, 1 ad, 2 bs, 3 bd */
void orc_sys_table(const orc_sys *S, int which, long *off, long *ket, int *cran) {
  long n = (which < 2) ? S->na : S->nb;
  const long *o = which == 0 ? S->as_off : which == 1 ? S->ad_off : which == 2 ? S->bs_off : S->bd_off;
  const long *k = which == 0 ? S->as_ket : which == 1 ? S->ad_ket : which == 2 ? S->bs_ket : S->bd_ket;
  const int *c = which == 0 ? S->as_cr : which == 1 ? S->ad_cr : which == 2 ? S->bs_cr : S->bd_cr;
  int w = (which % 2 == 0) ? 2 : 4;
  memcpy(off, o, (n + 1) * sizeof(long));
  memcpy(ket, k, o[n] * sizeof(long));
  memcpy(cran, c, w * o[n] * sizeof(int));
}

/* makeQChamDiagTerms, chemistry/tpb/qcham.h:290-340 */
void orc_hdiag_rows(const orc_sys *S, double *hii, long ia0, long ia1);
void orc_sigma_rows(const orc_sys *S, const double *hii, const double *T, double *Wb, long ia0,
                    long ia1);
void orc_hdiag(const orc_sys *S, double *hii) { orc_hdiag_rows(S, hii, 0, S->na); }

/*
 * mult (default, non-TRADMODE), chemistry/tpb/mult.h:625-881, one rank:
 * Wb += hii o Wk, then task type 2 (alpha singles, alpha doubles), type 1
 * (beta singles, beta doubles), type 0 (alpha single x beta single), each bra
 * row owned by one thread (round-robin ia).  Accumulates into Wb.
 */
void orc_sigma(const orc_sys *S, const double *hii, const double *T, double *Wb) {
  orc_sigma_rows(S, hii, T, Wb, 0, S->na);
}

/* ------------------------------------------------------------------ */
/* vector ops: framework/dm_vector.h:23-87 (per-thread Kahan, thread-order sum) */

double orc_inner(const double *X, const double *Y, long n) {
  int nth = omp_get_max_threads();
  double *arr = (double *)calloc(nth, sizeof(double));
  double sum = 0.0;
#pragma omp parallel
  {
    int tid = omp_get_thread_num();
    double mysum = 0.0, eps = 0.0, val, tmp;
#pragma omp for schedule(static)
    for (long is = 0; is < n; is++) {
      val = X[is] * Y[is] - eps;
      tmp = mysum + val;
      eps = (tmp - mysum) - val;
      mysum = tmp;
    }
    arr[tid] = mysum;
  }
  for (int i = 0; i < nth; i++) sum += arr[i];
  free(arr);
  return sum;
}

double orc_normalize(double *X, long n) {
  double res = sqrt(orc_inner(X, X, n));
  double f = 1.0 / res;
#pragma omp parallel for schedule(static)
  for (long is = 0; is < n; is++) X[is] *= f;
  return res;
}

/* ------------------------------------------------------------------ */
/* small symmetric eigenproblem.  The reference calls LAPACK dsyev
 * (framework/hp_numeric.h:503-519): eigenvalues ascending, eigenvectors in
 * the columns of A (column-major, leading dimension lda).  Restated here by
 * cyclic Jacobi rotations followed by an ascending sort. */
void orc_syev(int n, double *A, int lda, double *w) {
  double *V = (double *)calloc(n * n, sizeof(double));
  for (int i = 0; i < n; i++) V[i + i * n] = 1.0;
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0.0, tot = 0.0;
    for (int j = 0; j < n; j++)
      for (int i = 0; i < n; i++) {
        double x = A[i + j * lda] * A[i + j * lda];
        tot += x;
        if (i != j) off += x;
      }
    if (off <= 1e-30 * (tot > 0 ? tot : 1.0)) break;
    for (int p = 0; p < n - 1; p++)
      for (int q = p + 1; q < n; q++) {
        double apq = A[p + q * lda];
        if (fabs(apq) < 1e-300) continue;
        double app = A[p + p * lda], aqq = A[q + q * lda];
        double theta = (aqq - app) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; k++) {
          double akp = A[k + p * lda], akq = A[k + q * lda];
          A[k + p * lda] = c * akp - s * akq;
          A[k + q * lda] = s * akp + c * akq;
        }
        for (int k = 0; k < n; k++) {
          double apk = A[p + k * lda], aqk = A[q + k * lda];
          A[p + k * lda] = c * apk - s * aqk;
          A[q + k * lda] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; k++) {
          double vkp = V[k + p * n], vkq = V[k + q * n];
          V[k + p * n] = c * vkp - s * vkq;
          V[k + q * n] = s * vkp + c * vkq;
        }
      }
  }
  int *idx = (int *)malloc(n * sizeof(int));
  for (int i = 0; i < n; i++) idx[i] = i;
  for (int i = 1; i < n; i++) { /* insertion sort by eigenvalue */
    int k = idx[i];
    int j = i - 1;
    while (j >= 0 && A[idx[j] + idx[j] * lda] > A[k + k * lda]) {
      idx[j + 1] = idx[j];
      j--;
    }
    idx[j + 1] = k;
  }
  for (int i = 0; i < n; i++) w[i] = A[idx[i] + idx[i] * lda];
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) A[i + j * lda] = V[i + idx[j] * n];
  free(idx);
  free(V);
}

/*
 * Davidson, chemistry/tpb/davidson.h:707-970, one rank: single root, modified
 * Gram-Schmidt, restart with C0 = W.  W is in/out.  hist (optional, size
 * max_it*nb*2) receives (E0, |R|) per step.  Returns the number of steps.
 */
int orc_davidson(const orc_sys *S, const double *hii, double *W, int max_it, int nb, double eps,
                 double *hist) {
  long N = S->na * S->nb;
  const double eps_reg = 1.0e-12;
  double **C = (double **)malloc(nb * sizeof(double *));
  double **HC = (double **)malloc(nb * sizeof(double *));
  for (int k = 0; k < nb; k++) {
    C[k] = (double *)malloc(N * 8);
    HC[k] = (double *)malloc(N * 8);
  }
  double *R = (double *)malloc(N * 8);
  double *H = (double *)calloc(nb * nb, 8), *U = (double *)calloc(nb * nb, 8);
  double *E = (double *)malloc(nb * 8);
  int steps = 0, stop = 0;
  for (int it = 0; it < max_it && !stop; it++) {
    memcpy(C[0], W, N * 8);
    for (int ib = 0; ib < nb; ib++) {
      memset(HC[ib], 0, N * 8);
      orc_sigma(S, hii, C[ib], HC[ib]);
      for (int jb = 0; jb <= ib; jb++) {
        H[jb + nb * ib] = orc_inner(C[jb], HC[ib], N);
        H[ib + nb * jb] = H[jb + nb * ib];
      }
      for (int jb = 0; jb <= ib; jb++)
        for (int kb = 0; kb <= ib; kb++) U[jb + nb * kb] = H[jb + nb * kb];
      orc_syev(ib + 1, U, nb, E);
      double x = U[0];
#pragma omp parallel for
      for (long is = 0; is < N; is++) W[is] = C[0][is] * x;
      x = -1.0 * U[0];
#pragma omp parallel for
      for (long is = 0; is < N; is++) R[is] = HC[0][is] * x;
      for (int kb = 1; kb <= ib; kb++) {
        x = U[kb];
#pragma omp parallel for
        for (long is = 0; is < N; is++) W[is] += C[kb][is] * x;
        x = -1.0 * U[kb];
#pragma omp parallel for
        for (long is = 0; is < N; is++) R[is] += HC[kb][is] * x;
      }
#pragma omp parallel for
      for (long is = 0; is < N; is++) R[is] += E[0] * W[is];
      orc_normalize(W, N);
      double norm_R = orc_normalize(R, N);
      if (hist) {
        hist[2 * steps] = E[0];
        hist[2 * steps + 1] = norm_R;
      }
      steps++;
      if (norm_R < eps) {
        stop = 1;
        break;
      }
      if (ib < nb - 1) {
#pragma omp parallel for
        for (long is = 0; is < N; is++) {
          if (fabs(E[0] - hii[is]) > eps_reg)
            C[ib + 1][is] = R[is] / (E[0] - hii[is]);
          else
            C[ib + 1][is] = R[is] / (E[0] - hii[is] - eps_reg);
        }
        for (int kb = 0; kb < ib + 1; kb++) {
          double olap = -1.0 * orc_inner(C[kb], C[ib + 1], N);
#pragma omp parallel for
          for (long is = 0; is < N; is++) C[ib + 1][is] += C[kb][is] * olap;
        }
        orc_normalize(C[ib + 1], N);
      }
    }
  }
  for (int k = 0; k < nb; k++) {
    free(C[k]);
    free(HC[k]);
  }
  free(C);
  free(HC);
  free(R);
  free(H);
  free(U);
  free(E);
  return steps;
}

/*
 * Lanczos, chemistry/tpb/lanczos.h:162-327 (the operator form that diag calls
 * for method 2, sbdiag.h:461-468), one rank.  Each iteration restarts the
 * three-term recurrence from C[0] = W with no reorthogonalisation:
 *   HC = 0; for ib: HC += H C[ib]; A[ib,ib] = <C[ib], HC>; E = eig(A[0:n,0:n]);
 *   stop if |E[0] - E_old| < eps; HC -= A[ib,ib] C[ib]; C[ib+1] = HC;
 *   A[ib,ib+1] = A[ib+1,ib] = |C[ib+1]| (then normalised); break if that is
 *   below eps; HC = -A[ib,ib+1] C[ib].
 * After the steps, W = sum_ib U[ib,0] C[ib] (not renormalised).  hist
 * (optional, size max_it*nb*3) receives (E0, alpha, beta) per step, beta = 0
 * when the step ended before computing it.  Returns the number of steps.
 */
int orc_lanczos(const orc_sys *S, const double *hii, double *W, int max_it, int nb, double eps,
                double *hist) {
  long N = S->na * S->nb;
  double **C = (double **)malloc(nb * sizeof(double *));
  for (int k = 0; k < nb; k++) C[k] = (double *)malloc(N * 8);
  double *HC = (double *)malloc(N * 8);
  double *A = (double *)calloc(nb * nb, 8), *U = (double *)calloc(nb * nb, 8);
  double *E = (double *)malloc(nb * 8);
  int steps = 0;
  for (int it = 0; it < max_it; it++) {
    int n = 0, stop = 0;
    double E_old = 1.0e+8;
    memcpy(C[0], W, N * 8);
    memset(HC, 0, N * 8);
    for (int ib = 0; ib < nb; ib++) {
      n++;
      int ii = ib + nb * ib, ij = ib + nb * (ib + 1), ji = ib + 1 + nb * ib;
      orc_sigma(S, hii, C[ib], HC);
      double Aii = orc_inner(C[ib], HC, N);
      A[ii] = Aii;
      for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) U[i + nb * j] = A[i + nb * j];
      orc_syev(n, U, nb, E);
      if (hist) {
        hist[3 * steps] = E[0];
        hist[3 * steps + 1] = Aii;
        hist[3 * steps + 2] = 0.0;
      }
      steps++;
      if (fabs(E[0] - E_old) < eps) {
        stop = 1;
        break;
      }
      E_old = E[0];
      if (ib + 1 == nb) break;
#pragma omp parallel for
      for (long is = 0; is < N; is++) HC[is] -= Aii * C[ib][is];
      memcpy(C[ib + 1], HC, N * 8);
      A[ij] = orc_normalize(C[ib + 1], N);
      A[ji] = A[ij];
      if (hist) hist[3 * (steps - 1) + 2] = A[ij];
      if (fabs(A[ij]) < eps) break;
#pragma omp parallel for
      for (long is = 0; is < N; is++) HC[is] = -A[ij] * C[ib][is];
    }
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) U[i + nb * j] = A[i + nb * j];
    orc_syev(n, U, nb, E);
    for (int ib = 0; ib < n; ib++) {
      if (ib == 0) {
#pragma omp parallel for
        for (long is = 0; is < N; is++) W[is] = U[0] * C[ib][is];
      } else {
#pragma omp parallel for
        for (long is = 0; is < N; is++) W[is] += U[ib] * C[ib][is];
      }
    }
    if (stop) break;
  }
  for (int k = 0; k < nb; k++) free(C[k]);
  free(C);
  free(HC);
  free(A);
  free(U);
  free(E);
  return steps;
}

/*
 * DiagonalAdetReducedDensityMatrix / DiagonalBdetReducedDensityMatrix,
 * chemistry/tpb/rdmat.h:11-58, 136-182 (one rank): Da[ia] = sum_ib W^2 in ib
 * order, Db[ib] = sum_ia W^2 in ia order.
 */
void orc_weights(const orc_sys *S, const double *W, double *Da, double *Db) {
  for (long ia = 0; ia < S->na; ia++) {
    double d = 0.0;
    for (long ib = 0; ib < S->nb; ib++) d += W[ia * S->nb + ib] * W[ia * S->nb + ib];
    Da[ia] = d;
  }
  for (long ib = 0; ib < S->nb; ib++) {
    double d = 0.0;
    for (long ia = 0; ia < S->na; ia++) d += W[ia * S->nb + ib] * W[ia * S->nb + ib];
    Db[ib] = d;
  }
}

/*
 * OccupationDensity, chemistry/tpb/occupation.h:10-95 (one rank, every
 * orbital): res[2o] = sum over determinants of W^2 if orbital o is occupied in
 * the alpha string, res[2o+1] the same for beta (sbdiag.h:666-693 stores it as
 * density[2o], density[2o+1]).
 */
void orc_occupation(const orc_sys *S, const double *W, double *res) {
  const int L = S->L;
  for (int o = 0; o < 2 * L; o++) res[o] = 0.0;
  for (long ia = 0; ia < S->na; ia++)
    for (long ib = 0; ib < S->nb; ib++) {
      const double w2 = W[ia * S->nb + ib] * W[ia * S->nb + ib];
      for (int o = 0; o < L; o++) {
        if ((S->a[ia] >> o) & 1ULL) res[2 * o] += w2;
        if ((S->b[ib] >> o) & 1ULL) res[2 * o + 1] += w2;
      }
    }
}

/*
 * CarryOverAdet / CarryOverBdet, chemistry/tpb/rdmat.h:60-134 (spin 0 alpha, 1
 * beta): strings sorted by descending weight (ties by index: the reference's
 * std::sort leaves their order unspecified).  ratio == 0: keep up to and
 * including the first string whose cumulative weight exceeds 1 - threshold
 * (std::upper_bound + 1, clamped to all strings); else keep
 * (size_t)(ratio * count) strings (sbdiag.h:745-772).  out receives the kept
 * string indices; returns their number; *truncated = 1 - their weight.
 */
static const double *g_sort_w;
static int cmp_desc(const void *x, const void *y) {
  const long i = *(const long *)x, j = *(const long *)y;
  if (g_sort_w[i] > g_sort_w[j]) return -1;
  if (g_sort_w[i] < g_sort_w[j]) return 1;
  return i < j ? -1 : (i > j ? 1 : 0);
}
long orc_carryover(const orc_sys *S, const double *W, int spin, double threshold, double ratio,
                   long *out, double *truncated) {
  double *Da = (double *)malloc(S->na * 8), *Db = (double *)malloc(S->nb * 8);
  orc_weights(S, W, Da, Db);
  const double *D = spin ? Db : Da;
  const long n = spin ? S->nb : S->na;
  long *order = (long *)malloc(n * sizeof(long));
  for (long i = 0; i < n; i++) order[i] = i;
  g_sort_w = D;
  qsort(order, n, sizeof(long), cmp_desc);
  long m;
  if (ratio == 0.0) {
    double sum = 0.0, target = 1.0 - threshold;
    long first = n;
    for (long i = 0; i < n; i++) {
      sum += D[order[i]];
      if (sum > target) {
        first = i;
        break;
      }
    }
    m = first + 1 < n ? first + 1 : n;
  } else {
    m = (long)(ratio * (double)n);
    if (m > n) m = n;
  }
  double kept = 0.0;
  for (long i = 0; i < m; i++) {
    out[i] = order[i];
    kept += D[order[i]];
  }
  if (truncated) *truncated = 1.0 - kept;
  free(order);
  free(Da);
  free(Db);
  return m;
}

/* Rows [ia0, ia1) of hdiag and sigma (same arithmetic as orc_hdiag /
 * orc_sigma); used for a bounded CPU-baseline sample of a large workload.
 * hii, T and Wb are full-size vectors (index ia*nb+ib). */
void orc_hdiag_rows(const orc_sys *S, double *hii, long ia0, long ia1) {
  long nb = S->nb;
#pragma omp parallel for schedule(static)
  for (long ia = ia0; ia < ia1; ia++) {
    uint64_t D[2];
    for (long ib = 0; ib < nb; ib++) {
      det_from_ab(S->a[ia], S->b[ib], S->L, D);
      hii[ia * nb + ib] = zero_excite(D, S->L, S->I0, S->I1, S->Dm, S->Xm);
    }
  }
}

void orc_sigma_rows(const orc_sys *S, const double *hii, const double *T, double *Wb, long ia0,
                    long ia1) {
  long nb = S->nb;
  int L = S->L, L2 = 2 * L, dsize = (2 * L + 63) / 64;
#pragma omp parallel for schedule(static)
  for (long i = ia0 * nb; i < ia1 * nb; i++) Wb[i] += hii[i] * T[i];
  for (int task_type = 2; task_type >= 0; task_type--) {
#pragma omp parallel
    {
      int nth = omp_get_num_threads(), tid = omp_get_thread_num();
      uint64_t D[2];
      for (long ia = ia0 + tid; ia < ia1; ia += nth) {
        for (long ib = 0; ib < nb; ib++) {
          long bra = ia * nb + ib;
          det_from_ab(S->a[ia], S->b[ib], L, D);
          if (task_type == 2) {
            for (long o = S->as_off[ia]; o < S->as_off[ia + 1]; o++) {
              long ket = S->as_ket[o] * nb + ib;
              double e = one_excite(D, dsize, S->as_cr[2 * o], S->as_cr[2 * o + 1], L2, S->I1, S->I2);
              Wb[bra] += e * T[ket];
            }
            for (long o = S->ad_off[ia]; o < S->ad_off[ia + 1]; o++) {
              long ket = S->ad_ket[o] * nb + ib;
              const int *c = S->ad_cr + 4 * o;
              Wb[bra] += two_excite(D, c[0], c[1], c[2], c[3], S->I2) * T[ket];
            }
          } else if (task_type == 1) {
            for (long o = S->bs_off[ib]; o < S->bs_off[ib + 1]; o++) {
              long ket = ia * nb + S->bs_ket[o];
              double e = one_excite(D, dsize, S->bs_cr[2 * o], S->bs_cr[2 * o + 1], L2, S->I1, S->I2);
              Wb[bra] += e * T[ket];
            }
            for (long o = S->bd_off[ib]; o < S->bd_off[ib + 1]; o++) {
              long ket = ia * nb + S->bd_ket[o];
              const int *c = S->bd_cr + 4 * o;
              Wb[bra] += two_excite(D, c[0], c[1], c[2], c[3], S->I2) * T[ket];
            }
          } else {
            for (long o = S->as_off[ia]; o < S->as_off[ia + 1]; o++) {
              long ja = S->as_ket[o];
              for (long q = S->bs_off[ib]; q < S->bs_off[ib + 1]; q++) {
                long ket = ja * nb + S->bs_ket[q];
                Wb[bra] += two_excite(D, S->as_cr[2 * o], S->bs_cr[2 * q], S->as_cr[2 * o + 1],
                                      S->bs_cr[2 * q + 1], S->I2) * T[ket];
              }
            }
          }
        }
      }
    }
  }
}

/* ------------------------------------------------------------------ */
/*
 * Correlation (do_rdm != 0), chemistry/tpb/correlation.h:379-643 at one rank,
 * with ZeroDiffCorrelation, OneDiffCorrelation and TwoDiffCorrelation of
 * chemistry/basic/correlation.h:16-50, 103-148, 190-247.  one[s*L*L + o1 + L*o2]
 * (s = 0 alpha, 1 beta) and two[(s1 + 2*s2)*L^4 + o1 + L*o2 + L^2*o3 + L^3*o4]
 * are written as the reference writes onebody[s] and twobody[s1 + 2*s2]:
 *   diagonal: every (ia, ib);
 *   task type 2: alpha singles and doubles (same beta string);
 *   task type 1: beta singles and doubles (same alpha string);
 *   task type 0: an alpha single times a beta single.
 * Each pair (bra I, ket J) adds W_I * W_J times the parity of the interleaved
 * bra determinant, in the loop order of the reference's one-thread run.
 */
static void zero_diff(const uint64_t *D, double w, int L, double *one, double *two) {
  const long L2 = (long)L * L, L3 = L2 * L, L4 = L3 * L;
  const double w2 = w * w;
  for (int i = 0; i < 2 * L; i++) {
    if (!getocc(D, i)) continue;
    const int oi = i / 2, si = i % 2;
    one[si * L2 + oi + (long)L * oi] += w2;
    for (int j = i + 1; j < 2 * L; j++) {
      if (!getocc(D, j)) continue;
      const int oj = j / 2, sj = j % 2;
      two[(si + 2 * sj) * L4 + oi + L * oj + L2 * oi + L3 * oj] += w2;
      two[(sj + 2 * si) * L4 + oj + L * oi + L2 * oj + L3 * oi] += w2;
      if (si == sj) {
        two[(si + 2 * sj) * L4 + oi + L * oj + L2 * oj + L3 * oi] += -w2;
        two[(sj + 2 * si) * L4 + oj + L * oi + L2 * oi + L3 * oj] += -w2;
      }
    }
  }
}

static void one_diff(const uint64_t *D, double wi, double wj, int L, int i, int a, double *one,
                     double *two) {
  const long L2 = (long)L * L, L3 = L2 * L, L4 = L3 * L;
  double sgn = 1.0;
  parity(D, i < a ? i : a, i < a ? a : i, &sgn);
  const int oi = i / 2, si = i % 2, oa = a / 2;
  const double v = wi * wj * sgn;
  one[si * L2 + oi + (long)L * oa] += v;
  for (int x = 0; x < 2; x++) {
    uint64_t bits = D[x];
    while (bits) {
      const int pos = __builtin_ctzll(bits);
      const int soj = x * 64 + pos, oj = soj / 2, sj = soj % 2;
      two[(si + 2 * sj) * L4 + oa + oj * L + oi * L2 + oj * L3] += v;
      two[(sj + 2 * si) * L4 + oj + oa * L + oj * L2 + oi * L3] += v;
      if (si == sj) {
        two[(si + 2 * sj) * L4 + oa + oj * L + oj * L2 + oi * L3] += -v;
        two[(sj + 2 * si) * L4 + oj + oa * L + oi * L2 + oj * L3] += -v;
      }
      bits &= bits - 1;
    }
  }
}

static void two_diff(const uint64_t *D, double wi, double wj, int L, int i, int j, int a, int b,
                     double *two) {
  const long L2 = (long)L * L, L4 = L2 * L2;
  double sgn = 1.0;
  const int I = i < j ? i : j, J = i < j ? j : i;
  const int A = a < b ? a : b, B = a < b ? b : a;
  parity(D, I < A ? I : A, I < A ? A : I, &sgn);
  parity(D, J < B ? J : B, J < B ? B : J, &sgn);
  if (A > J || B < I) sgn *= -1.0;
  const int oi = I / 2, si = I % 2, oa = A / 2, sa = A % 2;
  const int oj = J / 2, sj = J % 2, ob = B / 2, sb = B % 2;
  const double v = sgn * wi * wj;
  if (si == sa) {
    two[(si + 2 * sj) * L4 + oa + L * ob + L2 * (oi + L * oj)] += v;
    two[(sj + 2 * si) * L4 + ob + L * oa + L2 * (oj + L * oi)] += v;
  }
  if (si == sb) {
    two[(si + 2 * sj) * L4 + oa + L * ob + L2 * (oj + L * oi)] += -v;
    two[(sj + 2 * si) * L4 + ob + L * oa + L2 * (oi + L * oj)] += -v;
  }
}

void orc_correlation(const orc_sys *S, const double *W, double *one, double *two) {
  const int L = S->L;
  const long na = S->na, nb = S->nb, L2 = (long)L * L;
  memset(one, 0, 2 * L2 * sizeof(double));
  memset(two, 0, 4 * L2 * L2 * sizeof(double));
  uint64_t D[2];
  for (long ia = 0; ia < na; ia++)
    for (long ib = 0; ib < nb; ib++) {
      det_from_ab(S->a[ia], S->b[ib], L, D);
      zero_diff(D, W[ia * nb + ib], L, one, two);
    }
  /* task type 2: alpha singles, then alpha doubles, per bra (ia, ib) */
  for (long ia = 0; ia < na; ia++)
    for (long ib = 0; ib < nb; ib++) {
      det_from_ab(S->a[ia], S->b[ib], L, D);
      const double wi = W[ia * nb + ib];
      for (long k = S->as_off[ia]; k < S->as_off[ia + 1]; k++)
        one_diff(D, wi, W[S->as_ket[k] * nb + ib], L, S->as_cr[2 * k], S->as_cr[2 * k + 1], one,
                 two);
      for (long k = S->ad_off[ia]; k < S->ad_off[ia + 1]; k++)
        two_diff(D, wi, W[S->ad_ket[k] * nb + ib], L, S->ad_cr[4 * k], S->ad_cr[4 * k + 1],
                 S->ad_cr[4 * k + 2], S->ad_cr[4 * k + 3], two);
    }
  /* task type 1: beta singles, then beta doubles */
  for (long ia = 0; ia < na; ia++)
    for (long ib = 0; ib < nb; ib++) {
      det_from_ab(S->a[ia], S->b[ib], L, D);
      const double wi = W[ia * nb + ib];
      for (long k = S->bs_off[ib]; k < S->bs_off[ib + 1]; k++)
        one_diff(D, wi, W[ia * nb + S->bs_ket[k]], L, S->bs_cr[2 * k], S->bs_cr[2 * k + 1], one,
                 two);
      for (long k = S->bd_off[ib]; k < S->bd_off[ib + 1]; k++)
        two_diff(D, wi, W[ia * nb + S->bd_ket[k]], L, S->bd_cr[4 * k], S->bd_cr[4 * k + 1],
                 S->bd_cr[4 * k + 2], S->bd_cr[4 * k + 3], two);
    }
  /* task type 0: alpha single x beta single, TwoDiffCorrelation(i = alpha cr,
   * j = beta cr, a = alpha an, b = beta an) */
  for (long ia = 0; ia < na; ia++)
    for (long ib = 0; ib < nb; ib++) {
      det_from_ab(S->a[ia], S->b[ib], L, D);
      const double wi = W[ia * nb + ib];
      for (long k = S->as_off[ia]; k < S->as_off[ia + 1]; k++)
        for (long m = S->bs_off[ib]; m < S->bs_off[ib + 1]; m++)
          two_diff(D, wi, W[S->as_ket[k] * nb + S->bs_ket[m]], L, S->as_cr[2 * k],
                   S->bs_cr[2 * m], S->as_cr[2 * k + 1], S->bs_cr[2 * m + 1], two);
    }
}
