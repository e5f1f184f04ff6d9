/*
 * spektri_port.c -- C restatement of spektri's CPU FCFB + Welch path.
 *
 * TEST INFRASTRUCTURE ONLY: the CPU baseline of bench.py and a second,
 * f32 checker of the oracle.  Nothing in the product links it.
 *
 * Follows the reference loop structure:
 *   main.rs:154-193      mainloop_real: overlap carry (copy_within), s16 -> f32
 *   dsp.rs:138-160       process_real: blocks at i * 3N/4 of the buffer
 *   multifft.rs:51-96    two real blocks packed into one complex N-point FFT,
 *                        window multiply (rectangular, scaling / N), untangle;
 *                        block pairs in parallel (rayon par_chunks(2))
 *   spectrum.rs:89-116   Welch 3-tap accumulation, serial on the calling thread
 *   fcfb.rs:66-100,154-188  channels in parallel: gather (get_bin), weights,
 *                        fftshift, unnormalised inverse FFT, keep [L/8, 7L/8)
 * The FFT is this file's own radix-4/radix-2 Stockham transform over split
 * real/imaginary arrays (vectorisable inner loops), twiddles from double.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef struct {
  int n;
  int sign;
  int nst;         /* number of stages */
  int rad[32];     /* radix of stage */
  int len[32];     /* sub-sequence length entering the stage */
  float* tw[32];   /* per stage: (rad-1) * (len/rad) complex twiddles, split */
} fft_plan;

static void plan_init(fft_plan* p, int n, int sign) {
  memset(p, 0, sizeof(*p));
  p->n = n;
  p->sign = sign;
  int len = n;
  while (len > 1) {
    int r = (len % 4 == 0) ? 4 : 2;
    int m = len / r;
    float* t = (float*)malloc(sizeof(float) * 2 * (size_t)(r - 1) * (size_t)(m > 0 ? m : 1));
    for (int j = 1; j < r; ++j)
      for (int q = 0; q < m; ++q) {
        double a = 2.0 * M_PI * (double)((long)j * q) / (double)len;
        t[2 * ((j - 1) * m + q)] = (float)cos(a);
        t[2 * ((j - 1) * m + q) + 1] = (float)(sign * sin(a));
      }
    p->rad[p->nst] = r;
    p->len[p->nst] = len;
    p->tw[p->nst] = t;
    p->nst++;
    len = m;
  }
}

static void plan_free(fft_plan* p) {
  for (int i = 0; i < p->nst; ++i) free(p->tw[i]);
}

/* Decimation in frequency Stockham: stage with sub-length len, stride s
 * (s * len = n).  Input element (q, k) at x[q + s k]; output (q, r k + j) at
 * y[q + s (r k + j)] holds sub-transform j of the next stage, twiddled. */
static void fft_exec(const fft_plan* p, float* xr, float* xi, float* yr, float* yi) {
  const int n = p->n;
  const float sg = (float)p->sign;
  int s = 1;
  float *ar = xr, *ai = xi, *br = yr, *bi = yi;
  for (int st = 0; st < p->nst; ++st) {
    const int len = p->len[st], r = p->rad[st], m = len / r;
    const float* t = p->tw[st];
    if (r == 4) {
      for (int k = 0; k < m; ++k) {
        const float w1r = t[2 * k], w1i = t[2 * k + 1];
        const float w2r = t[2 * (m + k)], w2i = t[2 * (m + k) + 1];
        const float w3r = t[2 * (2 * m + k)], w3i = t[2 * (2 * m + k) + 1];
        const float *a0r = ar + s * k, *a0i = ai + s * k;
        const float *a1r = ar + s * (k + m), *a1i = ai + s * (k + m);
        const float *a2r = ar + s * (k + 2 * m), *a2i = ai + s * (k + 2 * m);
        const float *a3r = ar + s * (k + 3 * m), *a3i = ai + s * (k + 3 * m);
        float *o0r = br + s * (4 * k), *o0i = bi + s * (4 * k);
        float *o1r = o0r + s, *o1i = o0i + s, *o2r = o0r + 2 * s, *o2i = o0i + 2 * s;
        float *o3r = o0r + 3 * s, *o3i = o0i + 3 * s;
        for (int q = 0; q < s; ++q) {
          const float spr = a0r[q] + a2r[q], spi = a0i[q] + a2i[q];
          const float smr = a0r[q] - a2r[q], smi = a0i[q] - a2i[q];
          const float tpr = a1r[q] + a3r[q], tpi = a1i[q] + a3i[q];
          /* sign * i * (a1 - a3) */
          const float tmr = -sg * (a1i[q] - a3i[q]), tmi = sg * (a1r[q] - a3r[q]);
          const float y0r = spr + tpr, y0i = spi + tpi;
          const float y1r = smr + tmr, y1i = smi + tmi;
          const float y2r = spr - tpr, y2i = spi - tpi;
          const float y3r = smr - tmr, y3i = smi - tmi;
          o0r[q] = y0r;
          o0i[q] = y0i;
          o1r[q] = y1r * w1r - y1i * w1i;
          o1i[q] = y1r * w1i + y1i * w1r;
          o2r[q] = y2r * w2r - y2i * w2i;
          o2i[q] = y2r * w2i + y2i * w2r;
          o3r[q] = y3r * w3r - y3i * w3i;
          o3i[q] = y3r * w3i + y3i * w3r;
        }
      }
    } else {
      for (int k = 0; k < m; ++k) {
        const float w1r = t[2 * k], w1i = t[2 * k + 1];
        const float *a0r = ar + s * k, *a0i = ai + s * k;
        const float *a1r = ar + s * (k + m), *a1i = ai + s * (k + m);
        float *o0r = br + s * (2 * k), *o0i = bi + s * (2 * k), *o1r = o0r + s, *o1i = o0i + s;
        for (int q = 0; q < s; ++q) {
          const float dr = a0r[q] - a1r[q], di = a0i[q] - a1i[q];
          o0r[q] = a0r[q] + a1r[q];
          o0i[q] = a0i[q] + a1i[q];
          o1r[q] = dr * w1r - di * w1i;
          o1i[q] = dr * w1i + di * w1r;
        }
      }
    }
    float* tr = ar; ar = br; br = tr;
    float* ti = ai; ai = bi; bi = ti;
    s *= r;
  }
  /* natural order result is in (ar, ai); move to x if needed */
  if (ar != xr) {
    memcpy(xr, ar, sizeof(float) * (size_t)n);
    memcpy(xi, ai, sizeof(float) * (size_t)n);
  }
}

/* get_bin (fftutil.rs:8-22) on split arrays of K bins of an N-point real FFT. */
static inline void get_bin(const float* Xr, const float* Xi, int N, int K, long i, float* re, float* im) {
  long m = ((i % N) + N) % N;
  if (m < K) {
    *re = Xr[m];
    *im = Xi[m];
  } else {
    *re = Xr[N - m];
    *im = -Xi[N - m];
  }
}

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/*
 * Process whole buffers of F blocks from the s16 stream x until `seconds`
 * have elapsed or nbuf_max buffers are done.  Returns elapsed seconds;
 * *done receives the buffer count.  The reference serialises its outputs to
 * bytes; that cost is omitted.  When chan_out is not NULL, every channel's
 * output of every processed block is stored there (channel c at float
 * offset 2 * nbuf_max * F * sum_{c' < c} 6 (L_c' / 8), interleaved re, im);
 * when welch_out is not NULL, every finished record acc / A is stored there
 * (K floats each) and *nrec receives the count.  Otherwise the channel
 * outputs go to a per-buffer scratch and only one bin of each record is
 * kept.
 */
double spkport_run(const int16_t* x, int64_t nbuf_max, int N, int F, int A, int nch, const int64_t* firsts,
                   const int32_t* Ls, const float* weights, double seconds, int threads, int* done,
                   float* chan_out_all, float* welch_out, int* nrec) {
  const int O = N / 4, H = N - O, K = N / 2 + 1;
  const float scaling = 1.0f / 32767.0f;
  const float wv = scaling / (float)N; /* dsp.rs:70-73 */
  if (nch < 0 || nch > 1 << 20 || N <= 0 || F <= 0 || A <= 0) {
    *done = 0;
    return 0.0;
  }
  if (threads > 0) omp_set_num_threads(threads);
  fft_plan fwd;
  plan_init(&fwd, N, -1);
  int maxL = 0;
  for (int c = 0; c < nch; ++c) maxL = Ls[c] > maxL ? Ls[c] : maxL;
  fft_plan* inv = (fft_plan*)calloc((size_t)nch, sizeof(fft_plan));
  size_t* woff = (size_t*)calloc((size_t)nch, sizeof(size_t));
  size_t wsum = 0;
  for (int c = 0; c < nch; ++c) {
    plan_init(&inv[c], Ls[c], +1);
    woff[c] = wsum;
    wsum += (size_t)Ls[c];
  }
  float* buf = (float*)calloc((size_t)O + (size_t)F * H, sizeof(float));
  float* specr = (float*)malloc(sizeof(float) * (size_t)F * K);
  float* speci = (float*)malloc(sizeof(float) * (size_t)F * K);
  float* acc = (float*)calloc((size_t)K, sizeof(float));
  unsigned accn = 0;
  float* chan_out = (float*)malloc(sizeof(float) * 2 * (size_t)nch * F * (3 * (size_t)maxL / 4 + 1));
  size_t* coff = (size_t*)calloc((size_t)nch, sizeof(size_t)); /* float offsets in chan_out_all */
  {
    size_t o = 0;
    for (int c = 0; c < nch; ++c) {
      coff[c] = o;
      o += 2 * (size_t)nbuf_max * F * (6 * (size_t)(Ls[c] / 8));
    }
  }
  int recs = 0;
  volatile float sink = 0.f;

  const double t0 = now_s();
  int64_t b = 0;
  for (; b < nbuf_max; ++b) {
    /* main.rs:173-181: keep the overlap, convert the new samples */
    memmove(buf, buf + (size_t)F * H, sizeof(float) * (size_t)O);
    const int16_t* src = x + (size_t)b * F * H;
    for (int64_t i = 0; i < (int64_t)F * H; ++i) buf[O + i] = (float)src[i];

    /* multifft.rs:51-96: pairs of blocks, one complex N-FFT per pair */
#pragma omp parallel for schedule(static)
    for (int pr = 0; pr < F / 2; ++pr) {
      float* zr = (float*)malloc(sizeof(float) * 4 * (size_t)N);
      float *zi = zr + N, *wr = zr + 2 * N, *wi = zr + 3 * N;
      const float* in0 = buf + (size_t)(2 * pr) * H;
      const float* in1 = buf + (size_t)(2 * pr + 1) * H;
      for (int n = 0; n < N; ++n) {
        zr[n] = in0[n] * wv;
        zi[n] = in1[n] * wv;
      }
      fft_exec(&fwd, zr, zi, wr, wi);
      float *o0r = specr + (size_t)(2 * pr) * K, *o0i = speci + (size_t)(2 * pr) * K;
      float *o1r = o0r + K, *o1i = o0i + K;
      for (int k = 0; k < K; ++k) {
        const int mk = k == 0 ? 0 : N - k;
        const float rr = zr[k], ri = zi[k], mr = zr[mk], mi = zi[mk];
        o0r[k] = 0.5f * (rr + mr);
        o0i[k] = 0.5f * (ri - mi);
        o1r[k] = 0.5f * (ri + mi);
        o1i[k] = 0.5f * (-rr + mr);
      }
      free(zr);
    }

    /* spectrum.rs:89-116: serial Welch accumulation */
    for (int g = 0; g < F; ++g) {
      const float* Xr = specr + (size_t)g * K;
      const float* Xi = speci + (size_t)g * K;
      for (int e = 0; e < 2; ++e) {
        const long i = e == 0 ? 0 : K - 1;
        float r0, i0, rm, im_, rp, ip;
        get_bin(Xr, Xi, N, K, i, &r0, &i0);
        get_bin(Xr, Xi, N, K, i - 1, &rm, &im_);
        get_bin(Xr, Xi, N, K, i + 1, &rp, &ip);
        const float cr = r0 - (rm + rp) * 0.5f, ci = i0 - (im_ + ip) * 0.5f;
        acc[i] += cr * cr + ci * ci;
      }
      for (int k = 1; k < K - 1; ++k) {
        const float cr = Xr[k] - (Xr[k - 1] + Xr[k + 1]) * 0.5f;
        const float ci = Xi[k] - (Xi[k - 1] + Xi[k + 1]) * 0.5f;
        acc[k] += cr * cr + ci * ci;
      }
      if (++accn >= (unsigned)A) {
        if (welch_out) {
          float* r = welch_out + (size_t)recs * K;
          for (int k = 0; k < K; ++k) r[k] = acc[k] / (float)accn;
        } else {
          sink += acc[K / 2] / (float)accn;
        }
        ++recs;
        memset(acc, 0, sizeof(float) * (size_t)K);
        accn = 0;
      }
    }

    /* fcfb.rs:74,154-188: channels in parallel */
#pragma omp parallel for schedule(dynamic)
    for (int c = 0; c < nch; ++c) {
      const int L = Ls[c];
      float* tr = (float*)malloc(sizeof(float) * 4 * (size_t)L);
      float *ti = tr + L, *ur = tr + 2 * L, *ui = tr + 3 * L;
      const float* W = weights + woff[c];
      const size_t per_blk = 2 * (size_t)(6 * (L / 8));
      float* out0 = chan_out_all ? chan_out_all + coff[c] + (size_t)b * F * per_blk
                                 : chan_out + 2 * (size_t)c * F * (3 * (size_t)maxL / 4 + 1);
      for (int g = 0; g < F; ++g) {
        float* out = out0 + (chan_out_all ? (size_t)g * per_blk : 0);
        const float* Xr = specr + (size_t)g * K;
        const float* Xi = speci + (size_t)g * K;
        for (int i = 0; i < L; ++i) {
          const int j = (i + L / 2) % L;
          float re, im;
          get_bin(Xr, Xi, N, K, firsts[c] + j, &re, &im);
          tr[i] = W[j] * re;
          ti[i] = W[j] * im;
        }
        fft_exec(&inv[c], tr, ti, ur, ui);
        for (int i = L / 8, o = 0; i < (L / 8) * 7; ++i, ++o) {
          out[2 * o] = tr[i];
          out[2 * o + 1] = ti[i];
        }
      }
      free(tr);
    }
    if (now_s() - t0 >= seconds) {
      ++b;
      break;
    }
  }
  const double el = now_s() - t0;
  *done = (int)b;
  if (nrec) *nrec = recs;
  free(coff);
  plan_free(&fwd);
  for (int c = 0; c < nch; ++c) plan_free(&inv[c]);
  free(inv);
  free(woff);
  free(buf);
  free(specr);
  free(speci);
  free(acc);
  free(chan_out);
  (void)sink;
  return el;
}

/* Checker entry: forward spectra of nblk blocks of one real stream by the
 * pair-packed path (for the oracle cross-check), split re/im [nblk][K]. */
void spkport_spectra(const int16_t* x, int nblk, int N, float* out_re, float* out_im) {
  const int O = N / 4, H = N - O, K = N / 2 + 1;
  const float wv = (1.0f / 32767.0f) / (float)N;
  fft_plan fwd;
  plan_init(&fwd, N, -1);
  float* zr = (float*)malloc(sizeof(float) * 4 * (size_t)N);
  float *zi = zr + N, *wr = zr + 2 * N, *wi = zr + 3 * N;
  for (int g = 0; g < nblk; g += 2) {
    for (int n = 0; n < N; ++n) {
      const long s0 = (long)g * H - O + n, s1 = (long)(g + 1) * H - O + n;
      zr[n] = (s0 >= 0 ? (float)x[s0] : 0.f) * wv;
      zi[n] = (g + 1 < nblk && s1 >= 0) ? (float)x[s1] * wv : 0.f;
    }
    fft_exec(&fwd, zr, zi, wr, wi);
    for (int k = 0; k < K; ++k) {
      const int mk = k == 0 ? 0 : N - k;
      out_re[(size_t)g * K + k] = 0.5f * (zr[k] + zr[mk]);
      out_im[(size_t)g * K + k] = 0.5f * (zi[k] - zi[mk]);
      if (g + 1 < nblk) {
        out_re[(size_t)(g + 1) * K + k] = 0.5f * (zi[k] + zi[mk]);
        out_im[(size_t)(g + 1) * K + k] = 0.5f * (-zr[k] + zr[mk]);
      }
    }
  }
  free(zr);
  plan_free(&fwd);
}

/*
 * Welch and channel stage of spektri_port's loop on spectra computed
 * elsewhere (the pocketfft forward FFTs of oracle/cpu_port.py).
 * spkport_post_create plans every channel's inverse FFT once;
 * spkport_post takes nblk blocks of K complex bins (interleaved re, im;
 * K = N/2 + 1 for real input, N for complex).  Welch is accumulated block
 * after block on the calling thread (spectrum.rs:89-116) into acc / *accn
 * across calls; a finished record is stored in welch_out (when not NULL)
 * and counted in *nrec.  Channels run in parallel (fcfb.rs:74,154-188);
 * their outputs go to chan_out (when not NULL) at float offset
 * 2 * out_blocks * sum_{c' < c} 6 (L_c' / 8), block blk0 + g of this call
 * at 2 * (blk0 + g) * 6 (L_c / 8) after that.
 */
typedef struct {
  int nch;
  fft_plan* inv;
  size_t *woff, *coff;
  int64_t out_blocks;
} post_ctx;

void* spkport_post_create(int nch, const int32_t* Ls, int64_t out_blocks) {
  post_ctx* p = (post_ctx*)calloc(1, sizeof(post_ctx));
  p->nch = nch;
  p->out_blocks = out_blocks;
  p->inv = (fft_plan*)calloc((size_t)nch, sizeof(fft_plan));
  p->woff = (size_t*)calloc((size_t)nch, sizeof(size_t));
  p->coff = (size_t*)calloc((size_t)nch, sizeof(size_t));
  size_t wsum = 0;
  for (int c = 0; c < nch; ++c) {
    plan_init(&p->inv[c], Ls[c], +1);
    p->woff[c] = wsum;
    wsum += (size_t)Ls[c];
    p->coff[c] = c == 0 ? 0 : p->coff[c - 1] + 2 * (size_t)out_blocks * (6 * (size_t)(Ls[c - 1] / 8));
  }
  return p;
}

void spkport_post_destroy(void* h) {
  post_ctx* p = (post_ctx*)h;
  for (int c = 0; c < p->nch; ++c) plan_free(&p->inv[c]);
  free(p->inv);
  free(p->woff);
  free(p->coff);
  free(p);
}

void spkport_post(void* h, const float* X, int nblk, int N, int complex_in, int A, float* acc, int* accn,
                  const int64_t* firsts, const int32_t* Ls, const float* weights, int threads, float* chan_out,
                  int64_t blk0, float* welch_out, int* nrec) {
  const post_ctx* p = (const post_ctx*)h;
  const int K = complex_in ? N : N / 2 + 1;
  const int nch = p->nch;
  if (threads > 0) omp_set_num_threads(threads);
  for (int g = 0; g < nblk; ++g) {
    const float* B = X + 2 * (size_t)g * K;
    /* the two end bins take their outer neighbour through get_bin
       (complex: wrap around; real: conjugate mirror) */
    for (int e = 0; e < 2; ++e) {
      const int k = e == 0 ? 0 : K - 1;
      float lr, li, rr, ri;
      if (complex_in) {
        const int km = (k - 1 + N) % N, kp = (k + 1) % N;
        lr = B[2 * km]; li = B[2 * km + 1];
        rr = B[2 * kp]; ri = B[2 * kp + 1];
      } else if (k == 0) {
        lr = B[2]; li = -B[3];
        rr = B[2]; ri = B[3];
      } else {
        lr = B[2 * (k - 1)]; li = B[2 * (k - 1) + 1];
        rr = B[2 * (N - k - 1)]; ri = -B[2 * (N - k - 1) + 1];
      }
      const float cr = B[2 * k] - (lr + rr) * 0.5f, ci = B[2 * k + 1] - (li + ri) * 0.5f;
      acc[k] += cr * cr + ci * ci;
    }
    for (int k = 1; k < K - 1; ++k) {
      const float cr = B[2 * k] - (B[2 * k - 2] + B[2 * k + 2]) * 0.5f;
      const float ci = B[2 * k + 1] - (B[2 * k - 1] + B[2 * k + 3]) * 0.5f;
      acc[k] += cr * cr + ci * ci;
    }
    if (++*accn >= A) {
      if (welch_out) {
        float* r = welch_out + (size_t)*nrec * K;
        for (int k = 0; k < K; ++k) r[k] = acc[k] / (float)*accn;
      }
      ++*nrec;
      memset(acc, 0, sizeof(float) * (size_t)K);
      *accn = 0;
    }
  }
#pragma omp parallel for schedule(dynamic)
  for (int c = 0; c < nch; ++c) {
    const int L = Ls[c];
    float* tr = (float*)malloc(sizeof(float) * 4 * (size_t)L);
    float *ti = tr + L, *ur = tr + 2 * L, *ui = tr + 3 * L;
    const float* W = weights + p->woff[c];
    const size_t per_blk = 2 * (size_t)(6 * (L / 8));
    for (int g = 0; g < nblk; ++g) {
      const float* B = X + 2 * (size_t)g * K;
      for (int i = 0; i < L; ++i) {
        const int j = (i + L / 2) % L;
        long m = ((firsts[c] + j) % N + N) % N;
        float re, im;
        if (complex_in || m < K) {
          re = B[2 * m];
          im = B[2 * m + 1];
        } else {
          re = B[2 * (N - m)];
          im = -B[2 * (N - m) + 1];
        }
        tr[i] = W[j] * re;
        ti[i] = W[j] * im;
      }
      fft_exec(&p->inv[c], tr, ti, ur, ui);
      if (chan_out) {
        float* out = chan_out + p->coff[c] + (size_t)(blk0 + g) * per_blk;
        for (int i = L / 8, o = 0; i < (L / 8) * 7; ++i, ++o) {
          out[2 * o] = tr[i];
          out[2 * o + 1] = ti[i];
        }
      }
    }
    free(tr);
  }
}
