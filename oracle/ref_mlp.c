/* ref_mlp.c -- oracle: FP64 forward / backward of the null-tail surrogate MLP.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).
 * The surrogate has no counterpart in the reference; this FP64 restatement is
 * the gradient check for the HIP training kernels.
 *
 * Layer l (0..nl-1): z = h W_l^T + b_l with W_l of shape dims[l+1] x dims[l];
 * h_{l+1} = GELU(z) (erf form) for hidden layers, identity for the last.
 * acts holds, per sample, [z_0 | h_1 | z_1 | h_2 | ... | z_{nl-1}].
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "refcpu.h"

static double gelu(double z)  { return 0.5 * z * (1. + erf(z / sqrt(2.))); }
static double dgelu(double z) { return 0.5 * (1. + erf(z / sqrt(2.))) + z * exp(-0.5 * z * z) / sqrt(2. * M_PI); }

static int64_t acts_width(int nl, const int *dims)
{
  int64_t w = 0;
  for (int l = 0; l < nl; l++) w += dims[l + 1] + ((l < nl - 1) ? dims[l + 1] : 0);
  return w;
}

int ref_mlp_forward(int nl, const int *dims, const double *const *W, const double *const *b,
                    const double *x, int64_t B, double *y, double *acts)
{
  int64_t aw = acts_width(nl, dims);
  for (int64_t s = 0; s < B; s++) {
    const double *h = x + s * dims[0];
    double *a = acts + s * aw;
    for (int l = 0; l < nl; l++) {
      int din = dims[l], dout = dims[l + 1];
      double *z = a;
      for (int o = 0; o < dout; o++) {
        double acc = b[l][o];
        for (int i = 0; i < din; i++) acc += W[l][o * din + i] * h[i];
        z[o] = acc;
      }
      a += dout;
      if (l < nl - 1) {
        for (int o = 0; o < dout; o++) a[o] = gelu(z[o]);
        h = a; a += dout;
      } else {
        memcpy(y + s * dout, z, sizeof(double) * dout);
      }
    }
  }
  return REF_OK;
}

int ref_mlp_backward(int nl, const int *dims, const double *const *W, const double *const *b,
                     const double *x, int64_t B, const double *acts, const double *dy,
                     double *const *dW, double *const *db)
{
  (void)b;
  int64_t aw = acts_width(nl, dims);
  int maxd = 0;
  for (int l = 0; l <= nl; l++) if (dims[l] > maxd) maxd = dims[l];
  double *g = malloc(sizeof(double) * maxd), *gp = malloc(sizeof(double) * maxd);
  int64_t *zoff = malloc(sizeof(int64_t) * nl), *hoff = malloc(sizeof(int64_t) * nl);
  int64_t off = 0;
  for (int l = 0; l < nl; l++) {
    zoff[l] = off; off += dims[l + 1];
    if (l < nl - 1) { hoff[l + 1 < nl ? l + 1 : 0] = off; off += dims[l + 1]; }
  }
  for (int l = 0; l < nl; l++) {
    memset(dW[l], 0, sizeof(double) * dims[l] * dims[l + 1]);
    memset(db[l], 0, sizeof(double) * dims[l + 1]);
  }
  for (int64_t s = 0; s < B; s++) {
    const double *a = acts + s * aw;
    memcpy(g, dy + s * dims[nl], sizeof(double) * dims[nl]);        /* dL/dz of the last layer */
    for (int l = nl - 1; l >= 0; l--) {
      int din = dims[l], dout = dims[l + 1];
      const double *hin = (l == 0) ? x + s * dims[0] : a + hoff[l];
      for (int o = 0; o < dout; o++) {
        db[l][o] += g[o];
        for (int i = 0; i < din; i++) dW[l][o * din + i] += g[o] * hin[i];
      }
      if (l > 0) {
        const double *zin = a + zoff[l - 1];
        for (int i = 0; i < din; i++) {
          double acc = 0.;
          for (int o = 0; o < dout; o++) acc += W[l][o * din + i] * g[o];
          gp[i] = acc * dgelu(zin[i]);
        }
        memcpy(g, gp, sizeof(double) * din);
      }
    }
  }
  free(g); free(gp); free(zoff); free(hoff);
  return REF_OK;
}
