/* rscape_hip.h -- C-ABI of the MI355X pairwise-covariation library (librscape_hip.so).
 *
 * Plain C types only: pointers, sizes, ints, doubles.  Every entry point returns
 * RSG_OK (0, Easel's eslOK) or RSG_FAIL (1, eslFAIL) and writes a message into a
 * caller-provided errbuf of RSG_ERRBUFSIZE bytes (Easel's eslERRBUFSIZE); no C++
 * exception crosses this boundary.  HIP errors and device-side failures (NaN in a
 * correction, impossible histogram bins) are turned into RSG_FAIL plus text.
 *
 * Threading: one context = one HIP stream; a context is not re-entrant.  Calls are
 * synchronous at this boundary (work is queued on the context stream, and the call
 * returns after it finished), except where marked "async".  rsg_cov_Calculate returns as
 * soon as its last kernel has written the result block to pinned host memory (every
 * device output is complete by then); later calls on the context follow on its stream.
 *
 * Each entry point names the reference interface it replaces (R-scape v2.6.11).
 * Symbol conventions follow the reference: digital RNA codes 0..3 = A,C,G,U,
 * 4 = gap, 15 = N (src/msamanip.c:284-330); matrices are L x L row-major with
 * COV[i][i] = -inf, Eval[i][i] = +inf (src/correlators.c:1254-1268).
 */
#ifndef RSCAPE_HIP_H
#define RSCAPE_HIP_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define RSG_OK          0
#define RSG_FAIL        1
#define RSG_ERRBUFSIZE  128

/* COVTYPE values (src/correlators.h:44-89). */
enum { RSG_CHI = 0, RSG_CHIp, RSG_CHIa, RSG_GT, RSG_GTp, RSG_GTa, RSG_MI, RSG_MIp, RSG_MIa,
       RSG_MIr, RSG_MIrp, RSG_MIra, RSG_MIg, RSG_MIgp, RSG_MIga, RSG_OMES, RSG_OMESp, RSG_OMESa,
       RSG_RAF, RSG_RAFp, RSG_RAFa, RSG_RAFS, RSG_RAFSp, RSG_RAFSa, RSG_CCF, RSG_CCFp, RSG_CCFa };
/* COVCLASS (src/correlators.h:37-42) and ACTYPE (:91-94). */
enum { RSG_C16 = 0, RSG_C2 = 1, RSG_CWC = 2, RSG_CSELECT = 3 };
enum { RSG_APC = 0, RSG_ASC = 1 };

/* How pair counts are formed. */
enum {
  RSG_COUNTS_INT8  = 0,   /* unit weights: int8 one-hot Gram, int32 accumulation, bit-exact;
                             uint16 counts kept, so at most 65535 sequences */
  RSG_COUNTS_BF16  = 1,   /* sequence weights: split-BF16 (hi+lo) Gram, FP32 accumulation  */
  RSG_COUNTS_BF16X3 = 2,  /* sequence weights: 3-term BF16 split, FP32 accumulation        */
  RSG_COUNTS_W8    = 3    /* sequence weights rounded to 21-bit fixed point (relative to the
                             largest), three 7-bit limbs through the int8 Gram with exact int32
                             accumulation, combined in FP64: counts carry only the weight
                             rounding, |dw| <= 2^-22 max w; at most 2^21 sequences */
};
/* Arithmetic of the per-pair statistics. */
enum { RSG_STAT_FP32 = 0, RSG_STAT_FP64 = 1 };

typedef struct rsg_ctx rsg_ctx;   /* device, stream, scratch, timers */
typedef struct rsg_msa rsg_msa;   /* one alignment resident in HBM + its pair tables */

/* ---- context ---- */
rsg_ctx *rsg_ctx_create(int device, char *errbuf);
void     rsg_ctx_destroy(rsg_ctx *ctx);
/* Per-kernel timing of the next calls (HIP events on the context stream): on = 1 every
 * stage, 2 the Gram launches only, 0 off.  The durations of the most recent launches, in
 * ms, are returned by rsg_ctx_timings: names[k] / ms[k] for k < *n. */
int      rsg_ctx_set_timing(rsg_ctx *ctx, int on);
int      rsg_ctx_timings(rsg_ctx *ctx, const char **names, double *ms, int *n, int cap);
/* Gram tile shape: 0 = autotuned / default, else an index into the shape table
 * (rsg_gram_shape_name gives its text).  Used by the autotuner. */
int         rsg_gram_nshapes(void);
const char *rsg_gram_shape_name(int shape);
int         rsg_ctx_set_gram_shape(rsg_ctx *ctx, int shape);
/* The autotuner's table (no reference counterpart): "logN logL shape" lines giving the fastest
 * shape for alignments with ceil(log2 nseq) = logN and ceil(log2 alen) = logL.  The library
 * reads gram_tiles.tsv next to itself on first use; this replaces that table for the process.
 * Sizes without an entry (and shape 0 contexts only) use the built-in choice. */
int         rsg_gram_tiles_load(const char *path, char *errbuf);
/* 1 (default): when the tile count leaves a last partial wave of work groups, those tiles
 * are split along the sequence axis over the idle CUs and their partial counts summed
 * (integers, so the result is identical); 0: one work group per tile. */
int         rsg_ctx_set_gram_split(rsg_ctx *ctx, int on);
/* Batched null (rsg_null_batch, rsg_null_batch_fitch): 1 (default) runs the batched Gram's
 * 32 x 32-column tiles (alignments under 256 sequences or columns) in a persistent kernel fed
 * by per-XCD tile queues (longest K first); 0 launches one work group per tile.  Results are
 * identical (no reference counterpart). */
int         rsg_ctx_set_batch_persistent(rsg_ctx *ctx, int on);
/* How long rsg_cov_Calculate spins on its completion stamp in pinned host memory before it
 * waits in a stream synchronisation instead (default 2000 microseconds; 0 = always
 * synchronise).  Results are the same either way (no reference counterpart). */
int         rsg_ctx_set_stamp_wait(rsg_ctx *ctx, int microseconds);
/* Counter calibration (no reference counterpart): runs the plain copy kernel
 * k_copy_probe over `bytes` (multiple of 16) three times, reading and writing `bytes` per
 * launch, so HBM counters of other kernels can be converted to bytes. */
int         rsg_ctx_copy_probe(rsg_ctx *ctx, size_t bytes, char *errbuf);

/* ---- alignment upload (replaces the ESL_MSA the corr_* calls read: ax[s][1..L], wgt[s]) ----
 * ax: nseq x alen digital codes, row-major, 0-based columns (msa->ax[s][i+1] in Easel).
 * wgt: nseq FP64 sequence weights, or NULL for unit weights. */
int  rsg_msa_create(rsg_ctx *ctx, const uint8_t *ax, const double *wgt, int64_t nseq, int64_t alen,
                    rsg_msa **ret_msa, char *errbuf);
/* Replace the symbols / weights of an existing alignment of the same shape (async). */
int  rsg_msa_update(rsg_ctx *ctx, rsg_msa *msa, const uint8_t *ax, const double *wgt, char *errbuf);
void rsg_msa_destroy(rsg_ctx *ctx, rsg_msa *msa);

/* ---- corr_Probs (src/correlators.c:1423-1496, correlators.h:478) ----
 * Pair counts (int8 or split-BF16 Gram), nseff/ngap, pm (corr_Marginals, FP64) and ps
 * (corr_NaivePS, FP64) on the device.  Validation as corr_ValidateProbs with tol. */
int rsg_corr_Probs(rsg_ctx *ctx, rsg_msa *msa, int counts_mode, double tol, char *errbuf);
/* Materialise the probability tables in the reference layout (any pointer may be NULL):
 * pp L*L*16 (pp[i][j][IDX(x,y,4)], both triangles), pm L*4, ps L*5, nseff L*L, ngap L*L
 * (ngap written for i<j only, as the reference). */
int rsg_corr_GetProbs(rsg_ctx *ctx, rsg_msa *msa, double *pp, double *pm, double *ps,
                      double *nseff, double *ngap, char *errbuf);
/* Raw pair counts for i<j: cnt[(i*L+j)*16 + x*4+y] (unit weights: integers). */
int rsg_corr_GetCounts(rsg_ctx *ctx, rsg_msa *msa, double *cnt, char *errbuf);

/* ---- corr_Calculate{CHI,OMES,GT,MI,MIr,MIg,CCF} (src/correlators.c:49-1061) ----
 * covtype is a base type (RSG_CHI, RSG_GT, ...); covclass C16/C2/CWC/CSELECT with the
 * CSELECT rule of src/correlators.c:335-341.  allowpair: 4x4 (NULL = WC + GU). */
int rsg_corr_CalculateCOV(rsg_ctx *ctx, rsg_msa *msa, int covtype, int covclass,
                          const double *allowpair, int nseqthresh, int alenthresh,
                          int stat_precision, char *errbuf);
/* corr_CalculateRAF / corr_CalculateRAFS (src/correlators.c:876-982): unweighted, from counts. */
int rsg_corr_CalculateRAF(rsg_ctx *ctx, rsg_msa *msa, int rafs, const double *allowpair,
                          int stat_precision, char *errbuf);
/* corr_CalculateCOVCorrected (src/correlators.c:1063-1157). */
int rsg_corr_CalculateCOVCorrected(rsg_ctx *ctx, rsg_msa *msa, int actype, int shiftnonneg, char *errbuf);
/* mi->COV and mi->minCOV / mi->maxCOV. COV may be NULL. */
int rsg_corr_GetCOV(rsg_ctx *ctx, rsg_msa *msa, double *COV, double *minCOV, double *maxCOV, char *errbuf);

/* ---- cov_Calculate (src/covariation.c:63-306) without the ranking:
 * corr_Probs + statistic + optional correction in one call (the hot path). */
int rsg_cov_Calculate(rsg_ctx *ctx, rsg_msa *msa, int covtype, int covclass, int counts_mode,
                      int stat_precision, const double *allowpair, int nseqthresh, int alenthresh,
                      double tol, double *minCOV, double *maxCOV, char *errbuf);

/* ---- cov_Calculate without any allocation (no reference counterpart) ----
 * For callers that cannot allocate device memory: code under HIP stream capture, tight serving
 * loops, harnesses that forbid hipMalloc.  The same chain as rsg_cov_Calculate (unit weights,
 * int8 counts; covtype CHI/GT/MI/MIr/MIg/OMES with or without APC/ASC) on an alignment already
 * in device memory: d_ax = nseq x alen codes (row-major, device), d_COV = alen x alen FP64
 * output (device; diagonal -inf).  Every intermediate buffer comes from one fixed device array
 * per GPU inside the library (about 42 MB, held until the process ends), so the call makes no
 * hipMalloc / hipFree: at most RSG_NA_MAXN sequences and RSG_NA_MAXL columns (RSG_FAIL with a
 * message beyond).  Calls on one GPU share that array: calls that overlap on one GPU (from two
 * host threads or contexts) race.  Results equal rsg_cov_Calculate's. */
#define RSG_NA_MAXN 4096
#define RSG_NA_MAXL 1024
int rsg_cov_Calculate_noalloc(rsg_ctx *ctx, const uint8_t *d_ax, int64_t nseq, int64_t alen, int covtype,
                              int covclass, int stat_precision, const double *allowpair, int nseqthresh,
                              int alenthresh, double tol, double *d_COV, double *minCOV, double *maxCOV,
                              char *errbuf);

/* ---- ranking, null histogram, tail fit, E-values (src/covariation.c:325-530, 1914-1975,
 *      2369-2506; src/R-scape.c:1280-1371, 1564-1612) ----
 * Histograms follow Easel's esl_histogram binning: bin b covers (bmin + b w, bmin + (b+1) w]. */
typedef struct {
  double    bmin, bmax, w;
  int32_t   nb, imin, imax, cmin;
  double    xmin, xmax, phi;
  uint64_t  n, Nc, No;
  int32_t   tailfit;
  uint64_t *obs;             /* nb counts, owned by the library */
} rsg_hist;

/* calculate_width_histo (R-scape.c:1353-1360): w from the current COV extremes. */
double rsg_width(double maxCOV, double minCOV, double bmin, double w_old, int hpts, double tol);
/* Histogram of max(COV_ij, bmin + w) over i<j, bmax = maxCOV + 5w (covariation.c:414-432). */
int  rsg_cov_RankHist(rsg_ctx *ctx, rsg_msa *msa, double bmin, double w, double tol,
                      rsg_hist **ret_hist, char *errbuf);
/* null_add2cumranklist (R-scape.c:1564-1612): merge h into *cum (created if NULL). */
int  rsg_hist_NullAdd(rsg_hist **cum, const rsg_hist *h, char *errbuf);
/* cov_histogram_pmass + SetTailByMass + cov_NullFit{Exponential,Gamma} + SetSurvFitTail.
 * Runs in FP64 on the device.  fit = {mu, lambda, tau, newmass, pmass}. survfit: 2*nb doubles
 * owned by the caller (allocated by the library with rsg_free). NULL if the fit is infinite. */
int  rsg_hist_NullFit(rsg_ctx *ctx, rsg_hist *null, double pmass, double fracfit, int doexpfit,
                      double **ret_survfit, double *fit, char *errbuf);
/* cov2evalue for every pair: Eval[i][j] = p * Nc (covariation.c:846-856), FP64 on the device.
 * The matrix stays on the device for rsg_cov_Hits; Eval (may be NULL) receives a copy. */
int  rsg_cov_Evalues(rsg_ctx *ctx, rsg_msa *msa, const rsg_hist *null, const double *survfit,
                     int64_t Nc, double *Eval, char *errbuf);
/* The significant pairs of the last E-value call (cov_CreateHitList's selection,
 * src/covariation.c:778-1006): pairs i < j with E <= thresh, sorted by (i, j), up to cap of
 * them; *nhits receives the total (RSG_FAIL with the first cap returned if it exceeds cap). */
int  rsg_cov_Hits(rsg_ctx *ctx, rsg_msa *msa, double thresh, int cap, int *nhits, int *i, int *j,
                  double *score, double *evalue, char *errbuf);
/* evalue2cov (covariation.c:2404-2436). */
double rsg_evalue2cov(double eval_thresh, int64_t Nc, const rsg_hist *h, const double *survfit);
double rsg_cov2evalue(double cov, int64_t Nc, const rsg_hist *h, const double *survfit);
rsg_hist *rsg_hist_create(double bmin, double bmax, double w);
void rsg_hist_destroy(rsg_hist *h);
void rsg_free(void *p);
/* cov_WriteNullHistogram / cov_ReadNullHistogram (covariation.c:1717-1868): the .null
 * text format, one "%f %llu" line per bin. */
int  rsg_hist_WriteNull(const char *path, const rsg_hist *h, char *errbuf);
int  rsg_hist_ReadNull(const char *path, rsg_hist **ret, char *errbuf);

/* ---- substitutions on the tree and the power of base pairs to covary (the .cov and .power
 *      columns; src/msatree.c:1423-1555, src/power.c:65-217, 532-650) ---- */
enum { RSG_SINGLE_SUBS = 0, RSG_JOIN_SUBS = 1, RSG_DOUBLE_SUBS = 2 };   /* POWERTYPE */
/* BPTYPE values the hit list uses (lib/R-view/src/rview_contacts.h:22-62): WWc, the first
 * non-base-pair contact type, and none */
enum { RSG_BP_WWC = 0, RSG_BP_STACKED = 32, RSG_BP_CONTACT = 33, RSG_BP_NONE = 34 };
/* SAMPLESIZE (src/correlators.h:26-31) */
enum { RSG_SAMPLE_CONTACTS = 0, RSG_SAMPLE_BP = 1, RSG_SAMPLE_WC = 2, RSG_SAMPLE_ALL = 3 };
/* Tree_Substitutions: Fitch ancestral states on the tree (random tie-breaks keyed by seed, as
 * rsg_null_fitch), then nsubs[c] (alen) = branches substituting at column c; ndouble / njoin
 * (alen x alen, [i * alen + j] for i < j, 0 elsewhere) = branches substituting at both / at
 * either of columns i and j.  Without includegaps a branch counts only where parent and child
 * are canonical (--powergaps off).  Any output may be NULL. */
int rsg_tree_Substitutions(rsg_ctx *ctx, const uint8_t *ax, int64_t nseq, int64_t alen, const int *left, const int *right,
                           uint64_t seed, int includegaps, int32_t *nsubs, int32_t *ndouble, int32_t *njoin, char *errbuf);
/* power_Read: a table of (substitutions, power) rows, two whitespace-separated columns; the
 * power is clamped to [0, 1] and stays 1 after the first row that exceeds 1. */
typedef struct { int ns; double *subs; double *prob; } rsg_power;
int    rsg_power_Read(const char *path, rsg_power **ret, char *errbuf);
void   rsg_power_Destroy(rsg_power *p);
/* the power of a pair with nsubs substitutions: the power of the last row whose count it exceeds */
double rsg_power_Prob(const rsg_power *p, int64_t nsubs);
/* SPAIR (src/correlators.h / power.h): one entry per pair i < j in the order i, then j */
typedef struct {
  int64_t i, j, iabs, jabs;               /* analysed columns (0-based); original positions (1-based) */
  double  nseff, ngap, wcf_f, wobble_f, noncan_f;
  int32_t powertype;
  int64_t nsubs, nsubs_join, nsubs_double;
  double  power, power_join, power_double;
  int32_t bptype_given;                   /* RSG_BP_WWC for a pair of the given structure */
  int32_t covary;                         /* set by rsg_cov_CreateHitList */
  double  sc, Pval, Eval;
} rsg_spair;
/* power_SPAIR_Create + power_SPAIR_AddMI after corr_Probs: alen (alen - 1) / 2 entries into spair;
 * msamap[c] = original position of column c - 1 (NULL = identity); ct[c] = partner column or -1
 * (the given structure, may be NULL); any of nsubs / njoin / ndouble (not all NULL) sets the
 * power type as the reference does (double, else join, else single). */
int rsg_power_SPAIR_Create(rsg_ctx *ctx, rsg_msa *msa, const int *msamap, const rsg_power *power, int powertype,
                           const int *ct, const int32_t *nsubs, const int32_t *njoin, const int32_t *ndouble,
                           rsg_spair *spair, char *errbuf);

/* ---- hit list and output files (src/covariation.c:778-1006, 1139-1406, 1637-1673, 1870-1890;
 *      src/power.c:261-378) ---- */
typedef struct {
  int64_t i, j;                           /* analysed columns, i < j */
  double  sc, Eval, pval;
  int64_t nsubs;
  double  power;
  int32_t bptype, is_compatible;
} rsg_hit;
/* cov_CreateHitList's selection over every pair i < j of the last scoring: pval =
 * cov2evalue(COV, 1) against the null, E = pval * Nb for pairs counted as base pairs under
 * samplesize, pval * expBP while fewer than expBP hits were found, else pval * Nt; a pair is a
 * hit when E < thresh (every pair when thresh > 1000).  bptype (alen x alen, may be NULL) gives
 * each pair's BPTYPE, else ct gives WWc for its pairs; is_compatible as the reference computes
 * it (column i unpaired in ct).  spair (may be NULL) receives covary, sc, Pval, Eval.  *ret_hits
 * is allocated by the library (rsg_free). */
int rsg_cov_CreateHitList(rsg_ctx *ctx, rsg_msa *msa, const rsg_hist *null, const double *survfit, int samplesize,
                          const int *ct, const uint8_t *bptype, int64_t Nt, int64_t Nb, double expBP, double thresh,
                          rsg_spair *spair, rsg_hit **ret_hits, int *ret_nhit, char *errbuf);
/* .cov (ranked = 0: cov_WriteHitList with its header) or .sorted.cov (ranked = 1:
 * cov_WriteRankedHitList, by E-value, or by score when by_score), with the summary line
 * "# covtype thresh [cov_min,cov_max] [FP | TP True Found | Sen PPV F]"; tnum = pairs of the
 * structure counted under samplesize (True). */
int rsg_cov_WriteHitList(const char *path, int ranked, int by_score, const char *covtype, double thresh,
                         double cov_min, double cov_max, int tnum, int64_t Nb, int64_t Nt, const rsg_hit *hits,
                         int nhit, const rsg_spair *spair, int64_t alen, const int *msamap, int firstpos,
                         int samplesize, char *errbuf);
/* .power: power_SPAIR_Write for the given structure's pairs (in_given = 1). */
int rsg_power_SPAIR_Write(const char *path, const rsg_spair *spair, int64_t dim, int in_given, char *errbuf);
/* .surv: survival curves of the base-pair and other-pair score histograms and of the null with
 * its fit (cov_histogram_PlotSurvival); any histogram may be NULL. */
int rsg_cov_WriteSurv(const char *path, const rsg_hist *hb, const rsg_hist *ht, const rsg_hist *null,
                      const double *survfit, char *errbuf);

/* ---- sequence weights and preprocessing (src/R-scape.c:1544-1562; src/msamanip.c:284-541) ---- */
/* esl_msaweight_PB: position-based weights, FP64 on the device. */
int rsg_weights_PB(rsg_ctx *ctx, const uint8_t *ax, int64_t nseq, int64_t alen, double *wgt, char *errbuf);
/* esl_msaweight_GSC: identities on the device (a sequence x sequence one-hot Gram),
 * UPGMA + GSC on the host in FP64. */
int rsg_weights_GSC(rsg_ctx *ctx, const uint8_t *ax, int64_t nseq, int64_t alen, double *wgt, char *errbuf);
/* msamanip_RemoveGapColumns column mask (gapthresh, weights as given). Returns the kept count. */
int rsg_gapcol_mask(const uint8_t *ax, const double *wgt, int64_t nseq, int64_t alen, double gapthresh,
                    int *useme);
/* msamanip_RemoveGapColumns (src/msamanip.c:461-541) on a resident alignment: with
 * gapthresh < 1 a column is kept iff its residues carry weight r > 0 and r / (residues + gaps)
 * >= 1 - gapthresh (the msa's weights; R-scape filters before weighting, so with unit
 * weights), computed on the device; with gapthresh >= 1 every column is kept.  *ret receives
 * a new resident alignment of the kept columns, built on the device (same weights);
 * useme (src alen ints) and map (new column -> src column, 0-based; up to src alen ints) may
 * be NULL.  No column kept: RSG_FAIL "no positions left after gap trimming". */
int rsg_msa_RemoveGapColumns(rsg_ctx *ctx, const rsg_msa *src, double gapthresh, int *useme, int *map,
                             rsg_msa **ret, char *errbuf);

/* ---- phylogenetic null simulation (cov_GenerateAlignment, src/cov_simulate.c:60-859;
 *      trees as R-scape-sim's create_tree, src/R-scape-sim.c:608-680) ----
 * Trees use the Easel ESL_TREE layout: internal nodes 0..N-2 (root 0, parents first),
 * children > 0 internal, <= 0 leaf -child.  Substitutions follow a 4-state GTR P(t) per
 * branch; every draw is Philox4x32-10 keyed by (seed, stream, node row, column). */
/* GTR model from stationary pi[4] and exchangeabilities exch[6] (AC AG AU CG CU GU), scaled
 * to one substitution per unit time; model receives 40 doubles (eigendecomposition). */
int rsg_gtr_model(const double *pi, const double *exch, double *model);
/* P(t) inverse-CDF tables on the device (FP64): cdf[k*12 + a*3 + b] = floor(2^32 sum_{c<=b} P(t_k)[a][c]). */
int rsg_gtr_cdf(rsg_ctx *ctx, const double *model, int64_t nt, const double *t, uint32_t *cdf, char *errbuf);
/* Yule tree with n leaves, branch lengths rescaled to mean mean_bl (esl_tree_Simulate +
 * esl_tree_er_RescaleAverageBL). */
int rsg_tree_yule(int64_t n, double mean_bl, uint64_t seed, uint32_t stream,
                  int *left, int *right, double *ld, double *rd);
typedef struct {
  int64_t  nleaf;            /* N sequences (rsg_null_batch: at most 65535) */
  int64_t  alen;             /* L columns */
  const int *left, *right;   /* N-1 internal nodes */
  const double *ld, *rd;     /* branch lengths */
  const double *model;       /* rsg_gtr_model output; the root is drawn from its pi */
  uint64_t seed;
  uint32_t stream;
  double   gapfrac;          /* i.i.d. gap probability per leaf symbol (0 = none) */
  int32_t  group;            /* rsg_null_batch: replicates of one null share a group id */
} rsg_null_job;
/* R-scape's null alignment of a given alignment (null_rscape, src/R-scape.c:1615-1724):
 * Fitch ancestral states on the tree (src/msatree.c:173-229,1700-1931; gap is a state, other
 * non-canonical codes enter as one random state), one shuffle of all columns, then per branch
 * the original (old -> new) substitution counts replayed at random positions of the shuffled
 * parent holding old (src/msamanip.c:1449-1530,1597-1780).  ax: nseq x alen codes; leaves
 * receive the nseq x alen shuffled alignment (codes 0..4).  Draws: Philox keyed by seed. */
int rsg_null_fitch(rsg_ctx *ctx, const uint8_t *ax, int64_t nseq, int64_t alen, const int *left, const int *right,
                   uint64_t seed, uint8_t *leaves, char *errbuf);
/* Simulate one alignment; leaves = nleaf x alen codes (the leaf rows, leaf order). */
int rsg_null_simulate(rsg_ctx *ctx, const rsg_null_job *job, uint8_t *leaves, char *errbuf);
/* 16-state pair model for structured simulation: Q16 = the ribosum pair rate matrix bprsQ
 * (16 x 16 row-major over pair states a*4 + b, 5' residue a, 3' residue b; rows sum to 0) and
 * M16 = its stationary pair frequencies bprsM (src/cov_simulate.c:393-394).  Q16 must be
 * reversible with respect to M16.  model16 receives 544 doubles (eigendecomposition for
 * P(t) = exp(t Q16), ratematrix_ConditionalsFromRate src/ratematrix.c:185-233).  RSG_FAIL for
 * an invalid or non-reversible model.  Host only. */
int rsg_pair_model(const double *Q16, const double *M16, double *model16);
/* cov_GenerateAlignmentUngapped with a structure (src/cov_simulate.c:88-159, 288-451, 724-798;
 * noss off, no indels): as rsg_null_simulate, but every base-paired column pair evolves as one
 * 16-state character under model16 (root pair from M16, cov_addpair).  ct[c] = partner column
 * of c (0-based) or -1 (the reference's 1-based ct with 0 = unpaired, shifted). */
int rsg_null_simulate_pairs(rsg_ctx *ctx, const rsg_null_job *job, const int *ct, const double *model16,
                            uint8_t *leaves, char *errbuf);
/* Rates of the e1 indel model (E1_RATE fields read by e1_model_transitions_AF, src/e1_model.c:784-868):
 * for R-scape-sim's default AIF model (src/e1_rate.c:431-519) rM = rD = 0 and muAD = muAI = muAM, from
 * the parameter file's rI, muAM, ldI, muI (data/training/Pfam.seed.S1000.trainGD.AIF.param). */
typedef struct { double rI, rM, rD, muAM, muAD, muAI, ldI, muI; } rsg_e1_rates;
/* e1_model_Create's transitions for a branch of length time in e2_GLOBAL mode, renormalised with
 * no end state (e1_model_RenormStateE): t16 in the E1_MODEL order, rows B, S, D, I of {S, D, I, E}.
 * L (the parent's length) matters only when the insertion term beta reaches 1. */
int rsg_e1_transitions(const rsg_e1_rates *rates, double time, int64_t L, double *t16);
/* cov_GenerateAlignment (src/cov_simulate.c:60-86): the job's tree evolves root (L = job->alen
 * codes 0..3; NULL: drawn from the model's pi) without gaps -- with ct and model16 (rsg_pair_model)
 * base-paired columns as 16-state characters, as rsg_null_simulate_pairs -- then, unless rates is
 * NULL (noindels), the e1 indel model of cov_GenerateAlignmentAddGaps (:161-190, 326-361, 495-563,
 * 634-722, 803-859): branch by branch in tree order an HMM walk over the current alignment deletes
 * cells of the child and inserts columns (gap in every other row, the child's residue drawn from
 * insf[4], NULL: the model's pi).  *ret_leaves receives N x *ret_alen codes and *ret_ct (may be
 * NULL) the partner of every column or -1; both are freed with rsg_free.  Draws are Philox keyed as
 * the oracle's (the reference's Easel generator is not restated: parity is in distribution). */
int rsg_cov_GenerateAlignment(rsg_ctx *ctx, const rsg_null_job *job, const uint8_t *root, const int *ct,
                              const double *model16, const rsg_e1_rates *rates, const double *insf,
                              uint8_t **ret_leaves, int64_t *ret_alen, int **ret_ct, char *errbuf);
/* Simulate and score a batch of alignments of any sizes (GT/MI/... + correction, unit
 * weights, CSELECT) entirely on the device, then build one null per group of jobs: jobs that
 * share a group id (ids contiguous, ascending from 0) are the replicates of one null, as
 * R-scape's nshuffle loop makes them: the histogram width comes from the group's first job
 * (calculate_width_histo), every replicate's pairs accumulate into one histogram
 * (null_add2cumranklist) and one FP64 tail fit is made per group.  Outputs per group g:
 * fits[g*8 + {0..7}] = {mu, lambda, tau, newmass, pmass, w, minCOV, maxCOV}; hists (may be
 * NULL) receives each group's histogram; feats (may be NULL) receives RSG_NULL_NFEAT surrogate features
 * per group (rsg_null_features of its first replicate).  use_graph = k >= 1 records the whole batch in one
 * hipGraph and launches it k times (each launch recomputes everything); 0 runs eagerly.  The
 * instantiated graph stays with the context: a later call whose launch geometry is the same
 * (job sizes, groups, tree shapes by depth, covtype and fit arguments) uploads its inputs
 * (trees, branch lengths, models, seeds) and replays it without a new capture. */
int rsg_null_batch(rsg_ctx *ctx, const rsg_null_job *jobs, int njobs, int covtype,
                   double bmin, double pmass, double fracfit, int doexpfit,
                   rsg_hist **hists, double *fits, float *feats, int use_graph, char *errbuf);
/* Calls of rsg_null_batch / rsg_null_batch_fitch on this context that replayed the cached
 * graph instead of capturing one (no counterpart in the reference: a diagnostic). */
int rsg_ctx_graph_reuses(const rsg_ctx *ctx);
/* R-scape's null of one alignment in one batch: null_rscape's nshuffle loop
 * (src/R-scape.c:1615-1724) with replicate r = rsg_null_fitch(ax, tree, seeds[r]) (Fitch
 * ancestors, column shuffle, substitution replay), every replicate scored (covtype, unit
 * weights, CSELECT), the histogram width from replicate 0 (calculate_width_histo), all pairs
 * merged into one histogram (null_add2cumranklist) and one FP64 tail fit: fit[0..7] as in
 * rsg_null_batch, *hist (may be NULL) the merged histogram.  Each Fitch kernel covers every
 * replicate in one launch per tree depth; use_graph as in rsg_null_batch.  nseq <= 65535. */
int rsg_null_batch_fitch(rsg_ctx *ctx, const uint8_t *ax, int64_t nseq, int64_t alen, const int *left, const int *right,
                         const uint64_t *seeds, int nrep, int covtype, double bmin, double pmass, double fracfit,
                         int doexpfit, rsg_hist **hist, double *fit, int use_graph, char *errbuf);

/* ---- learned null-tail surrogate (no reference counterpart) ---- */
/* The surrogate's RSG_NULL_NFEAT inputs of an alignment and its tree (column statistics on the
 * device, tree statistics on the host): ln N, ln L, ln total branch length, 16-bin
 * column-entropy and 16-bin gap histograms, A/C/G composition, 10 tree-shape moments, and the
 * split spectrum (branch lengths by the minor clade share of their split, 8 bins, two sums). */
#define RSG_NULL_NFEAT 64
int  rsg_null_features(rsg_ctx *ctx, rsg_msa *msa, const int *left, const int *right,
                       const double *ld, const double *rd, float *feat, char *errbuf);
/* E-values from a parametric tail P(score > x) = p0 exp(-lambda (x - m)) (the surrogate's
 * output): Eval[i][j] = Nc * min(1, p0 exp(-lambda (min(COV[i][j], xcap) - m))) for i != j,
 * +inf on the diagonal (cov2evalue's role, src/covariation.c:2369-2436, for a fitted tail).
 * xcap is where cov2evalue stops following the fit: scores in or past the last bin of the
 * 2 nb survival array (src/covariation.c:2380) all get the E-value of its upper bound,
 * bmin + 2 nb w; +inf for no cap.  Kept on the device for rsg_cov_Hits; Eval may be NULL. */
int  rsg_cov_EvaluesTail(rsg_ctx *ctx, rsg_msa *msa, double m, double lambda, double p0, double xcap,
                         int64_t Nc, double *Eval, char *errbuf);
typedef struct rsg_mlp rsg_mlp;
int  rsg_mlp_create(rsg_ctx *ctx, int nlayers, const int *dims, uint64_t seed, rsg_mlp **ret, char *errbuf);
void rsg_mlp_destroy(rsg_ctx *ctx, rsg_mlp *m);
/* bf16 = 1 (default): BF16 weights and activations into v_mfma_f32_32x32x16_bf16, FP32
 * accumulation; bf16 = 0: FP32 operands (v_mfma_f32_32x32x2f32), for gradient checks. */
int  rsg_mlp_set_precision(rsg_mlp *m, int bf16);
/* Input / output scalers kept with the network: the kernels use x' = (x - xmean) / xstd and
 * y = y' * ystd + ymean (targets are scaled the same way for training).  NULL = identity. */
int  rsg_mlp_set_scaling(rsg_ctx *ctx, rsg_mlp *m, const float *xmean, const float *xstd,
                         const float *ymean, const float *ystd, char *errbuf);
/* Training range of the inputs; rsg_mlp_in_range is 1 when every input lies inside it (the
 * caller falls back to the full simulated null otherwise).  NULL clears the range. */
int  rsg_mlp_set_range(rsg_mlp *m, const float *xmin, const float *xmax);
int  rsg_mlp_in_range(const rsg_mlp *m, const float *x);
int  rsg_mlp_set_params(rsg_ctx *ctx, rsg_mlp *m, const float *const *W, const float *const *b, char *errbuf);
int  rsg_mlp_get_params(rsg_ctx *ctx, rsg_mlp *m, float *const *W, float *const *b, char *errbuf);
int  rsg_mlp_forward(rsg_ctx *ctx, rsg_mlp *m, const float *x, int64_t B, float *y, char *errbuf);
/* One forward+backward on a batch, returning parameter gradients (FP32) and the MSE loss. */
int  rsg_mlp_grad(rsg_ctx *ctx, rsg_mlp *m, const float *x, const float *t, int64_t B,
                  float *const *dW, float *const *db, double *loss, char *errbuf);
/* Adam training over a dataset resident on the device. */
int  rsg_mlp_train(rsg_ctx *ctx, rsg_mlp *m, const float *x, const float *t, int64_t n,
                   int epochs, int batch, double lr, uint64_t seed, double *final_loss, char *errbuf);

#ifdef __cplusplus
}
#endif
#endif
