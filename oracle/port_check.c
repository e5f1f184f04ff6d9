/*
 * port_check.c -- drives the C restatement (spektri_port.c) for the
 * sanitizer build (TEST INFRASTRUCTURE ONLY): every entry point on a small
 * synthetic s16 stream, real and complex spectra, outputs collected, under
 * -fsanitize=address,undefined.  Exit status 0 when every output is finite
 * and the two channel-stage paths (spkport_run, spkport_post on the
 * spectra of spkport_spectra) agree to 1e-5 of full scale.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

double spkport_run(const int16_t* x, int64_t nbuf_max, int N, int F, int A, int nch, const int64_t* firsts,
                   const int32_t* Ls, const float* weights, double seconds, int threads, int* done,
                   float* chan_out_all, float* welch_out, int* nrec);
void spkport_spectra(const int16_t* x, int nblk, int N, float* out_re, float* out_im);
void* spkport_post_create(int nch, const int32_t* Ls, int64_t out_blocks);
void spkport_post_destroy(void* h);
void spkport_post(void* h, const float* X, int nblk, int N, int complex_in, int A, float* acc, int* accn,
                  const int64_t* firsts, const int32_t* Ls, const float* weights, int threads, float* chan_out,
                  int64_t blk0, float* welch_out, int* nrec);

int main(void) {
  const int N = 1024, F = 4, A = 3, H = N - N / 4, K = N / 2 + 1, nbuf = 3, nb = nbuf * F;
  const int nch = 4;
  const int64_t firsts[4] = {-20, 100, K - 30, 3 * N + 8};
  const int32_t Ls[4] = {64, 256, 32, 16};
  int wsum = 0, osum = 0;
  for (int c = 0; c < nch; ++c) {
    wsum += Ls[c];
    osum += 6 * (Ls[c] / 8);
  }
  float* w = malloc(sizeof(float) * wsum);
  for (int i = 0; i < wsum; ++i) w[i] = 0.25f + 0.5f * (float)((i * 37) % 17) / 17.f;
  const size_t ns = (size_t)nbuf * F * H;
  int16_t* x = malloc(sizeof(int16_t) * ns);
  for (size_t i = 0; i < ns; ++i) x[i] = (int16_t)(8000.0 * sin(0.0123 * i) + (double)((i * 7919) % 2001) - 1000.0);
  float* ch_a = calloc((size_t)2 * nb * osum, sizeof(float));
  float* ch_b = calloc((size_t)2 * nb * osum, sizeof(float));
  float* rec_a = calloc((size_t)(nb / A + 1) * K, sizeof(float));
  float* rec_b = calloc((size_t)(nb / A + 1) * K, sizeof(float));
  int done = 0, nra = 0, nrb = 0;
  spkport_run(x, nbuf, N, F, A, nch, firsts, Ls, w, 1e9, 2, &done, ch_a, rec_a, &nra);
  if (done != nbuf || nra != nb / A) return 10;
  /* the same blocks through spkport_spectra + spkport_post (in pairs: the
     first block's history is zeros, as in spkport_run) */
  float* re = malloc(sizeof(float) * (size_t)nb * K);
  float* im = malloc(sizeof(float) * (size_t)nb * K);
  spkport_spectra(x, nb, N, re, im);
  float* X = malloc(sizeof(float) * 2 * (size_t)nb * K);
  for (size_t i = 0; i < (size_t)nb * K; ++i) {
    X[2 * i] = re[i];
    X[2 * i + 1] = im[i];
  }
  void* h = spkport_post_create(nch, Ls, nb);
  float* acc = calloc(K, sizeof(float));
  int accn = 0;
  for (int b = 0; b < nbuf; ++b)
    spkport_post(h, X + 2 * (size_t)b * F * K, F, N, 0, A, acc, &accn, firsts, Ls, w, 2, ch_b, (int64_t)b * F, rec_b,
                 &nrb);
  spkport_post_destroy(h);
  if (nrb != nra) return 11;
  double e = 0;
  for (size_t i = 0; i < (size_t)2 * nb * osum; ++i) {
    if (!isfinite(ch_a[i]) || !isfinite(ch_b[i])) return 12;
    e = fmax(e, fabs((double)ch_a[i] - ch_b[i]));
  }
  double er = 0, rm = 0;
  for (size_t i = 0; i < (size_t)nra * K; ++i) {
    if (!isfinite(rec_a[i])) return 13;
    er = fmax(er, fabs((double)rec_a[i] - rec_b[i]));
    rm = fmax(rm, fabs((double)rec_a[i]));
  }
  /* complex spectra through spkport_post (wrap-around get_bin) */
  void* hc = spkport_post_create(nch, Ls, nb);
  float* accc = calloc(N, sizeof(float));
  float* Xc = calloc((size_t)2 * F * N, sizeof(float));
  for (size_t i = 0; i < (size_t)2 * F * N; ++i) Xc[i] = (float)sin(0.37 * i);
  int accnc = 0, nrc = 0;
  spkport_post(hc, Xc, F, N, 1, A, accc, &accnc, firsts, Ls, w, 2, NULL, 0, NULL, &nrc);
  spkport_post_destroy(hc);
  printf("port_check: channel max diff %.3g, record max diff %.3g of %.3g\n", e, er, rm);
  free(x); free(w); free(ch_a); free(ch_b); free(rec_a); free(rec_b); free(re); free(im); free(X); free(acc);
  free(accc); free(Xc);
  return (e < 1e-5 && er <= 1e-5 * rm) ? 0 : 14;
}
