/* refcpu.h -- CPU oracle for the R-scape pairwise-covariation hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: the tests, the
 * smoke check and the CPU-baseline leg of the benchmark may load it; the
 * product path (r-scape_amd/) never links, loads or calls it.
 *
 * It restates, in plain serial FP64 C, the loops of the reference R-scape
 * v2.6.11 sources that the GPU path replaces.  Every function cites the
 * reference file:line it follows.  No reference source is copied; the
 * loop order is kept so FP64 sums are reproducible.
 *
 * Parts that live in the un-vendored Easel library (histogram binning,
 * tail-by-mass, binned exponential/Gamma fits, PB/GSC weights) are
 * restated from their published methods; see the "parity" notes on each.
 *
 * Conventions
 *   ax     N x L uint8, row-major per sequence, 0-based columns.  Codes follow
 *          the Easel RNA digital alphabet: 0..3 = A,C,G,U (canonical),
 *          4 = gap, 5..14 degenerate, 15 = N, 16 = '*' nonresidue, 17 = '~' missing.
 *   pp     L x L x 16 FP64, pp[(i*L+j)*16 + x*4 + y]   (IDX(x,y,4), ribosum_matrix.h:15)
 *   pm     L x 4, ps L x 5, nseff / ngap / COV  L x L.
 */
#ifndef RSCAPE_REFCPU_H
#define RSCAPE_REFCPU_H
#include <stdint.h>

#define REF_OK    0
#define REF_FAIL  1

/* COVTYPE (correlators.h:44-89) -- base statistics only; +1 = "p" (APC), +2 = "a" (ASC) */
enum { REF_CHI = 0, REF_GT = 3, REF_MI = 6, REF_MIr = 9, REF_MIg = 12,
       REF_OMES = 15, REF_RAF = 18, REF_RAFS = 21, REF_CCF = 24 };
/* COVCLASS (correlators.h:37-42) */
enum { REF_C16 = 0, REF_C2 = 1, REF_CWC = 2, REF_CSELECT = 3 };
/* ACTYPE (correlators.h:91-94) */
enum { REF_APC = 0, REF_ASC = 1 };

/* ---- counts, probabilities (correlators.c) ---- */
int ref_naive_pp(const uint8_t *ax, const double *wgt, int64_t N, int64_t L,
                 double *pp, double *nseff, double *ngap);
int ref_naive_ps(const uint8_t *ax, const double *wgt, int64_t N, int64_t L, double *ps);
int ref_marginals(int64_t L, const double *pp, const double *nseff, double tol, double *pm, char *errbuf);
int ref_validate_probs(int64_t L, const double *pp, const double *pm, const double *ps, double tol, char *errbuf);
int ref_probs(const uint8_t *ax, const double *wgt, int64_t N, int64_t L, double tol,
              double *pp, double *nseff, double *ngap, double *pm, double *ps, char *errbuf);
/* threads of the pair-count loops (default 1, the reference's serial run); results are
   bit-identical for any count */
void ref_set_threads(int n);
int  ref_get_threads(void);
/* integer 4x4 pair counts (unit weights), cnt[(i*L+j)*16+x*4+y] for i<j */
int ref_pair_counts(const uint8_t *ax, int64_t N, int64_t L, int32_t *cnt);

/* ---- statistics (correlators.c:49-1061) ---- */
int ref_cov_stat(int covtype, int covclass, int64_t nseq, int64_t L, int nseqthresh, int alenthresh,
                 const double *allowpair, const double *pp, const double *pm,
                 const double *nseff, const double *ngap, double *COV, double *minmax, char *errbuf);
int ref_raf(const uint8_t *ax, int64_t N, int64_t L, const double *allowpair, double *COV, double *minmax);
int ref_rafs(const uint8_t *ax, int64_t N, int64_t L, const double *allowpair, double *COV, double *minmax);
/* RAF from unweighted pair counts (closed form of the literal loop); RAFS stencil over a RAF matrix */
int ref_raf_counts(const int32_t *cnt, int64_t N, int64_t L, const double *allowpair, double *COV, double *minmax);
int ref_rafs_from(const double *B, int64_t L, double *COV, double *minmax);
int ref_cov_corrected(int actype, int64_t L, double *COV, double *minmax, int shiftnonneg, char *errbuf);

/* ---- histogram with Easel semantics (esl_histogram, restated) ---- */
typedef struct {
  double    bmin, bmax, w;
  int       nb;
  uint64_t *obs;
  int       imin, imax, cmin;
  double    xmin, xmax, phi;
  uint64_t  n, Nc, No;
  int       tailfit;        /* 0 = complete data, 1 = virtually censored at phi */
} ref_hist;

ref_hist *ref_hist_create(double bmin, double bmax, double w);
/* the .null text format (cov_WriteNullHistogram / cov_ReadNullHistogram) */
int       ref_write_null(const char *path, const ref_hist *h);
ref_hist *ref_read_null(const char *path);
void      ref_hist_destroy(ref_hist *h);
int       ref_hist_score2bin(const ref_hist *h, double x, int *ret_b);
double    ref_hist_lbound(const ref_hist *h, int b);
double    ref_hist_ubound(const ref_hist *h, int b);
int       ref_hist_add(ref_hist *h, double x);
int       ref_hist_set_tail_by_mass(ref_hist *h, double pmass, double *ret_newmass);
int       ref_exp_fit_binned(const ref_hist *h, double *ret_mu, double *ret_lambda);
int       ref_gam_fit_binned(const ref_hist *h, double *ret_mu, double *ret_lambda, double *ret_tau);
double    ref_exp_surv(double x, double mu, double lambda);
double    ref_gam_surv(double x, double mu, double lambda, double tau);
double   *ref_survfit(const ref_hist *h, double pmass, int gamma, const double *ep);
double    ref_hist_pmass(const ref_hist *h, double target_pmass, double target_fracfit);
double    ref_cov2evalue(double cov, int64_t Nc, const ref_hist *h, const double *survfit);
double    ref_evalue2cov(double eval_thresh, int64_t Nc, const ref_hist *h, const double *survfit);

/* ranking pieces (covariation.c:325-530, R-scape.c:1280-1371,1564-1612) */
ref_hist *ref_rank_hist(int64_t L, const double *COV, double maxCOV, double bmin, double w, double tol);
int       ref_null_add(ref_hist **cum, const ref_hist *h);
double    ref_width(double maxCOV, double minCOV, double bmin, double w_old, int hpts, double tol);
int       ref_null_fit(ref_hist *null, double pmass_target, double fracfit, int doexpfit,
                       double **ret_survfit, double *ret_fit /* mu, lambda, tau, newmass */);
int       ref_evalues(int64_t L, const double *COV, int64_t Nc, const ref_hist *null,
                      const double *survfit, double *Eval);

/* ---- MSA-level helpers (msamanip.c, R-scape.c) ---- */
int  ref_gapcol_filter(const uint8_t *ax, const double *wgt, int64_t N, int64_t L, double gapthresh, int *useme);
void ref_canonicalize(uint8_t *ax, int64_t n);
int  ref_weights_pb(const uint8_t *ax, int64_t N, int64_t L, double *wgt);
int  ref_weights_gsc(const uint8_t *ax, int64_t N, int64_t L, double *wgt);
int  ref_nshuffle_default(int64_t nseq, int64_t alen);

/* ---- counter-based RNG + rate-matrix evolution (cov_simulate.c) ---- */
void     ref_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
uint32_t ref_rng_u32(uint64_t seed, uint32_t stream, uint32_t node, uint32_t pos);
int ref_evolve_gtr(int64_t nleaf, int64_t L, const int *left, const int *right, const double *ld,
                   const double *rd, const double *model, uint64_t seed, uint32_t stream, uint32_t gapthr,
                   const uint8_t *root, uint8_t *leaves);
int ref_evolve_pairs(int64_t nleaf, int64_t L, const int *left, const int *right, const double *ld,
                     const double *rd, const double *model, const double *m16, const int *ct, uint64_t seed,
                     uint32_t stream, uint32_t gapthr, const uint8_t *root, uint8_t *leaves);
/* e1 HMM transitions (AF family, e2_GLOBAL, RenormStateE); rates {rI, rM, rD, muAM, muAD, muAI, ldI, muI} */
int ref_e1_transitions(const double *rates, double time, int64_t L, double *t16);
int ref_add_indels(int64_t N, int64_t L0, const int *left, const int *right, const double *ld, const double *rd,
                   const double *rates, const double *insf, uint64_t seed, uint32_t stream, const uint8_t *leaves_in,
                   const int *ct_in, int64_t cap, uint8_t *leaves_out, int *ct_out, int64_t *alen);
int ref_fitch_null(const uint8_t *ax, int64_t N, int64_t L, const int *left, const int *right,
                   uint64_t seed, uint8_t *out);
int ref_tree_substitutions(const uint8_t *ax, int64_t N, int64_t L, const int *left, const int *right, uint64_t seed,
                           int includegaps, int32_t *nsubs, int32_t *ndouble, int32_t *njoin);
int ref_fitch_null_rows(const uint8_t *ax, int64_t N, int64_t L, const int *left, const int *right,
                        uint64_t seed, uint8_t *all_out, uint8_t *sh_out);

/* ---- surrogate MLP reference (FP64 forward / backward) ---- */
int ref_mlp_forward(int nl, const int *dims, const double *const *W, const double *const *b,
                    const double *x, int64_t B, double *y, double *acts);
int ref_mlp_backward(int nl, const int *dims, const double *const *W, const double *const *b,
                     const double *x, int64_t B, const double *acts, const double *dy,
                     double *const *dW, double *const *db);
#endif
