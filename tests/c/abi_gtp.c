/* abi_gtp.c -- a plain C99 caller of librscape_hip.so through include/rscape_hip.h only.
 *
 * Reads a binary alignment (int64 nseq, int64 alen, then nseq*alen uint8 codes), runs the
 * cov_Calculate chain for GTp (rsg_cov_Calculate: int8 counts, FP64 statistics, CSELECT, WC+GU
 * allowed pairs), fetches mi->COV and the extremes (rsg_corr_GetCOV) and writes them as raw
 * doubles: minCOV, maxCOV, then the alen*alen matrix.  Exit status 0 on success.  Usage:
 *   abi_gtp in.bin out.bin */
#include <stdio.h>
#include <stdlib.h>
#include "rscape_hip.h"

int main(int argc, char **argv)
{
  char errbuf[RSG_ERRBUFSIZE];
  FILE *f;
  int64_t hdr[2];
  uint8_t *ax;
  double *COV, mn = 0., mx = 0., mn2 = 0., mx2 = 0.;
  rsg_ctx *ctx;
  rsg_msa *msa = NULL;
  size_t n;

  if (argc != 3) { fprintf(stderr, "usage: abi_gtp in.bin out.bin\n"); return 2; }
  if (!(f = fopen(argv[1], "rb")) || fread(hdr, sizeof(int64_t), 2, f) != 2) { fprintf(stderr, "bad input\n"); return 2; }
  n = (size_t)hdr[0] * (size_t)hdr[1];
  ax = (uint8_t *)malloc(n);
  COV = (double *)malloc(sizeof(double) * (size_t)hdr[1] * (size_t)hdr[1]);
  if (!ax || !COV || fread(ax, 1, n, f) != n) { fprintf(stderr, "bad input\n"); return 2; }
  fclose(f);

  if (!(ctx = rsg_ctx_create(0, errbuf))) { fprintf(stderr, "rsg_ctx_create: %s\n", errbuf); return 1; }
  if (rsg_msa_create(ctx, ax, NULL, hdr[0], hdr[1], &msa, errbuf) != RSG_OK ||
      rsg_cov_Calculate(ctx, msa, RSG_GTp, RSG_CSELECT, RSG_COUNTS_INT8, RSG_STAT_FP64, NULL, 8, 50, 1e-6,
                        &mn, &mx, errbuf) != RSG_OK ||
      rsg_corr_GetCOV(ctx, msa, COV, &mn2, &mx2, errbuf) != RSG_OK) {
    fprintf(stderr, "scoring failed: %s\n", errbuf);
    return 1;
  }
  if (mn != mn2 || mx != mx2) { fprintf(stderr, "extremes differ between calls\n"); return 1; }
  if (!(f = fopen(argv[2], "wb"))) return 2;
  fwrite(&mn, sizeof(double), 1, f);
  fwrite(&mx, sizeof(double), 1, f);
  fwrite(COV, sizeof(double), (size_t)hdr[1] * (size_t)hdr[1], f);
  fclose(f);
  rsg_msa_destroy(ctx, msa);
  rsg_ctx_destroy(ctx);
  free(ax); free(COV);
  return 0;
}
