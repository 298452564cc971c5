/*
 * ORACLE -- test infrastructure, not part of the product.
 *
 * A plain-C restatement of the reference's CPU phasing loop used to check the
 * HIP engine.  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg load it, and only as the checker (or as the timed CPU baseline).
 *
 * Built twice from this one file:
 *   liboracle_f32.so  OR_REAL = float: the reference's FP32 arithmetic
 *                     (DOUBLE_PRECISION=OFF, reference CMakeLists.txt:45)
 *   liboracle_f64.so  OR_REAL = double: the FP64 oracle every result's error
 *                     is stated against
 *
 * Functions and the reference code they restate:
 *   or_fft              src/fft.c:79,95,163,185 (FFTW3 c2c over (z,y,x),
 *                       unnormalised, forward sign -1).  FFTW itself is not
 *                       available here, so the transform is this file's own
 *                       mixed-radix FFT evaluated in double precision and
 *                       rounded to OR_REAL.
 *   or_module_projection         src/phasing.c:1070-1100
 *   or_phased_projection         src/phasing.c:1103-1109
 *   or_apply_fourier_constraints src/phasing.c:1061-1068
 *   or_apply_constraints         src/phasing.c:1026-1059
 *   iterate_er/hio/raar/dm       src/phasing.c:1111-1274
 *   or_iterate (scheduler)       src/phasing.c:895-1024
 *   or_efourier / or_ereal / or_support_fraction  src/phasing.c:803-893
 *   or_gaussian_blur     src/image_filter.c:80-91 with dist_to_top_left,
 *                        src/image_util.c:1012-1032
 *   or_support_area_update       src/support_update.c:136-151,392-400,443-454
 *   or_support_threshold_update  src/support_update.c:167-179
 *   or_bezier            include/spimage/map.h:103-141 (solved here by
 *                        bisection on the curve, not by the closed form)
 *
 * How each is pinned (tests/test_oracle_*.py): the FFT against a brute-force
 * O(N^2) DFT and numpy's pocketfft; projections, updates and constraints
 * against closed forms, fixed points and special cases (beta = 0, beta = 1,
 * ER = HIO inside the support); the blur against its closed-form transfer
 * function, integral conservation and maximum position (the reference's
 * tests/ImageTests.c:194-243); the area threshold against a brute-force
 * sort; the Bezier schedule against its end points, midpoint and the
 * engine's Cardano solution; the scheduler against the reference's integer
 * recurrence enumerated by hand.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef OR_DOUBLE
typedef double R;
#define OR_SUFFIX f64
#else
typedef float R;
#define OR_SUFFIX f32
#endif

typedef struct {
  R re, im;
} oc;

typedef struct {
  double re, im;
} dc;

#define EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------------ */
/* FFT backends                                                              */
/*   0: this file's recursive mixed-radix FFT in double precision (default;  */
/*      the checker)                                                         */
/*   1: an FFTW3-API library loaded at run time (Intel MKL's libmkl_rt when  */
/*      present), single precision with a plan made per call like the        */
/*      reference's sp_image_fft_fast (src/fft.c:177-189); used only to time */
/*      the CPU baseline                                                      */
/* ------------------------------------------------------------------------ */
#include <dlfcn.h>

static int g_backend = 0;

typedef struct {
  long n;
  dc *w; /* w[k] = exp(-2 pi i k / n) */
} twtab;
static twtab g_tw[16];

static const dc *tw_for(long n) {
  for (int i = 0; i < 16; ++i)
    if (g_tw[i].n == n) return g_tw[i].w;
  int slot = 0;
  for (int i = 0; i < 16; ++i)
    if (g_tw[i].n == 0) {
      slot = i;
      break;
    }
  free(g_tw[slot].w);
  g_tw[slot].n = n;
  g_tw[slot].w = (dc *)malloc(sizeof(dc) * n);
  for (long k = 0; k < n; ++k) {
    const double a = -2.0 * M_PI * (double)k / (double)n;
    g_tw[slot].w[k].re = cos(a);
    g_tw[slot].w[k].im = sin(a);
  }
  return g_tw[slot].w;
}

/* X[k + j m] = sum_q W_n^{q (k + j m)} Y_q[k], Y_q the length-m transform of
 * x[q], x[q+p], ...; W taken from the table of the top-level length */
static void fft1d_rec(const dc *in, long istride, dc *out, long n, int sign, const dc *w, long tstride,
                      dc *scratch) {
  if (n == 1) {
    out[0] = in[0];
    return;
  }
  long p = 2;
  while (n % p) ++p;
  const long m = n / p;
  for (long q = 0; q < p; ++q) fft1d_rec(in + q * istride, istride * p, out + q * m, m, sign, w, tstride * p, scratch);
  for (long k = 0; k < m; ++k) {
    for (long q = 0; q < p; ++q) scratch[q] = out[q * m + k];
    for (long j = 0; j < p; ++j) {
      double sr = 0.0, si = 0.0;
      const long kk = k + j * m;
      for (long q = 0; q < p; ++q) {
        const dc t = w[((q * kk) % n) * tstride];
        const double ti = sign < 0 ? t.im : -t.im;
        sr += scratch[q].re * t.re - scratch[q].im * ti;
        si += scratch[q].re * ti + scratch[q].im * t.re;
      }
      out[j * m + k].re = sr;
      out[j * m + k].im = si;
    }
  }
}

/* FFTW3 single-precision entry points, resolved at run time */
typedef void *(*plan3d_fn)(int, int, int, void *, void *, int, unsigned);
typedef void (*plan_fn)(void *);
static plan3d_fn p_plan3d;
static plan_fn p_execute, p_destroy;

/* 0 own FFT, 1 FFTW3 API from `libpath`; returns 0 on success */
EXPORT int or_set_fft_backend(int backend, const char *libpath) {
  if (backend == 0) {
    g_backend = 0;
    return 0;
  }
  void *h = dlopen(libpath, RTLD_NOW | RTLD_GLOBAL);
  if (!h) return -1;
  p_plan3d = (plan3d_fn)dlsym(h, "fftwf_plan_dft_3d");
  p_execute = (plan_fn)dlsym(h, "fftwf_execute");
  p_destroy = (plan_fn)dlsym(h, "fftwf_destroy_plan");
  if (!p_plan3d || !p_execute || !p_destroy) return -2;
  g_backend = 1;
  return 0;
}

/* unnormalised c2c over (nz, ny, nx), x fastest; sign -1 forward, +1 inverse */
EXPORT void or_fft(const oc *in, oc *out, int nx, int ny, int nz, int sign) {
  const long N = (long)nx * ny * nz;
#ifndef OR_DOUBLE
  if (g_backend == 1) {
    /* FFTW_MEASURE = 0 in the FFTW3 API; FFTW_FORWARD = -1, FFTW_BACKWARD = +1 */
    void *plan = p_plan3d(nz, ny, nx, (void *)in, (void *)out, sign, 0u);
    p_execute(plan);
    p_destroy(plan);
    return;
  }
#endif
  dc *a = (dc *)malloc(sizeof(dc) * N);
  for (long i = 0; i < N; ++i) {
    a[i].re = in[i].re;
    a[i].im = in[i].im;
  }
  long maxn = nx > ny ? nx : ny;
  if (nz > maxn) maxn = nz;
  dc *line = (dc *)malloc(sizeof(dc) * maxn);
  dc *res = (dc *)malloc(sizeof(dc) * maxn);
  dc *scr = (dc *)malloc(sizeof(dc) * maxn);
  const int dims[3] = {nx, ny, nz};
  const long strides[3] = {1, nx, (long)nx * ny};
  for (int ax = 0; ax < 3; ++ax) {
    const long n = dims[ax], st = strides[ax];
    if (n == 1) continue;
    const dc *w = tw_for(n);
    const long outer = N / (n * st);
    for (long o = 0; o < outer; ++o)
      for (long r = 0; r < st; ++r) {
        const long base = o * n * st + r;
        for (long i = 0; i < n; ++i) line[i] = a[base + i * st];
        fft1d_rec(line, 1, res, n, sign, w, 1, scr);
        for (long i = 0; i < n; ++i) a[base + i * st] = res[i];
      }
  }
  for (long i = 0; i < N; ++i) {
    out[i].re = (R)a[i].re;
    out[i].im = (R)a[i].im;
  }
  free(a);
  free(line);
  free(res);
  free(scr);
}

/* ------------------------------------------------------------------------ */
/* helpers with the reference's FP evaluation (include/spimage/linear_alg.h) */
/* ------------------------------------------------------------------------ */
static inline R cabs_ref(oc a) { return (R)sqrt(a.im * a.im + a.re * a.re); }
static inline oc cscale(oc a, R b) {
  a.re *= b;
  a.im *= b;
  return a;
}

enum { SUPPORT = 1, MEASURED = 2 };
enum { HIO = 1, RAAR = 2, DIFFMAP = 3, ER = 4 };
enum { REAL_OBJ = 1, POS_REAL = 2, POS_COMPLEX = 4, FLIP = 8, CENTRO = 16, MARGIN = 32 };

/* ------------------------------------------------------------------------ */
/* cubic-Bezier schedule: x-control points (p0, m, m, p3), y-control points  */
/* (y0, y0, y3, y3); t found by bisection on the monotone x(t)               */
/* ------------------------------------------------------------------------ */
EXPORT R or_bezier(const R *keys, const R *vals, int n, R x) {
  int idx = 0;
  while (idx < n && !(x < keys[idx])) ++idx;
  if (idx == 0) return vals[0];
  if (idx == n) return vals[n - 1];
  const double p0 = keys[idx - 1], p3 = keys[idx], y0 = vals[idx - 1], y3 = vals[idx];
  const double m = 0.5 * (p0 + p3);
  double lo = 0.0, hi = 1.0;
  for (int it = 0; it < 200; ++it) {
    const double t = 0.5 * (lo + hi), w = 1.0 - t;
    const double xt = p0 * w * w * w + 3.0 * m * t * w * w + 3.0 * m * t * t * w + p3 * t * t * t;
    if (xt < x) lo = t;
    else hi = t;
  }
  const double t = 0.5 * (lo + hi), w = 1.0 - t;
  return (R)(y0 * w * w * w + 3.0 * y0 * t * w * w + 3.0 * y3 * t * t * w + y3 * t * t * t);
}

/* ------------------------------------------------------------------------ */
/* phaser state                                                              */
/* ------------------------------------------------------------------------ */
typedef struct {
  int type, constraints;
  int nbeta;
  R beta_keys[32], beta_vals[32];
  R gamma1, gamma2;
} OrAlg;

typedef struct {
  int type; /* 0 none, 1 threshold, 2 area, 5 static */
  int period;
  int nblur;
  R blur_keys[32], blur_vals[32];
  int nval;
  R val_keys[32], val_vals[32]; /* area or relative threshold */
} OrSup;

typedef struct {
  int nx, ny, nz;
  long N;
  int iteration;
  int objective;   /* 0 phases, 1 amplitudes */
  int have_margins;
  oc *g0, *g1, *gp;
  R *amp, *amp_min, *amp_max;
  oc *phased_amp;
  int *flags;
  R last_threshold; /* threshold of the last support update (for tests) */
  int nsup;         /* support updates so far, and the iterations they took */
  int sup_log[256];
} OrPhaser;

EXPORT OrPhaser *or_phaser_create(int nx, int ny, int nz) {
  OrPhaser *p = (OrPhaser *)calloc(1, sizeof(OrPhaser));
  p->nx = nx;
  p->ny = ny;
  p->nz = nz;
  p->N = (long)nx * ny * nz;
  p->g0 = (oc *)calloc(p->N, sizeof(oc));
  p->g1 = (oc *)calloc(p->N, sizeof(oc));
  p->gp = (oc *)calloc(p->N, sizeof(oc));
  p->amp = (R *)calloc(p->N, sizeof(R));
  p->amp_min = (R *)calloc(p->N, sizeof(R));
  p->amp_max = (R *)calloc(p->N, sizeof(R));
  p->phased_amp = (oc *)calloc(p->N, sizeof(oc));
  p->flags = (int *)calloc(p->N, sizeof(int));
  return p;
}

EXPORT void or_phaser_free(OrPhaser *p) {
  free(p->g0);
  free(p->g1);
  free(p->gp);
  free(p->amp);
  free(p->amp_min);
  free(p->amp_max);
  free(p->phased_amp);
  free(p->flags);
  free(p);
}

/* array accessors for the ctypes wrapper */
EXPORT oc *or_g0(OrPhaser *p) { return p->g0; }
EXPORT oc *or_g1(OrPhaser *p) { return p->g1; }
EXPORT oc *or_gp(OrPhaser *p) { return p->gp; }
EXPORT R *or_amp(OrPhaser *p) { return p->amp; }
EXPORT R *or_amp_min(OrPhaser *p) { return p->amp_min; }
EXPORT R *or_amp_max(OrPhaser *p) { return p->amp_max; }
EXPORT oc *or_phased_amp(OrPhaser *p) { return p->phased_amp; }
EXPORT int *or_flags(OrPhaser *p) { return p->flags; }
EXPORT int *or_iteration(OrPhaser *p) { return &p->iteration; }
EXPORT int *or_objective(OrPhaser *p) { return &p->objective; }
EXPORT int *or_have_margins(OrPhaser *p) { return &p->have_margins; }
EXPORT R or_last_threshold(OrPhaser *p) { return p->last_threshold; }
EXPORT int or_support_log(OrPhaser *p, int *out, int max) {
  const int n = p->nsup < 256 ? p->nsup : 256;
  for (int i = 0; i < n && i < max; ++i) out[i] = p->sup_log[i];
  return p->nsup;
}

/* ------------------------------------------------------------------------ */
/* projections and constraints                                              */
/* ------------------------------------------------------------------------ */
EXPORT void or_apply_fourier_constraints(oc *a, long N, int c) {
  for (long i = 0; i < N; ++i)
    if (c & CENTRO) a[i].im = 0;
}

EXPORT int or_module_projection(oc *a, const R *amp, const R *amp_min, const R *amp_max, const int *flags, long N,
                                int c, int have_margins) {
  if ((c & MARGIN) && !have_margins) return -1; /* the reference is fatal here */
  for (long i = 0; i < N; ++i) {
    R m = 1.;
    if (flags[i] & MEASURED) {
      if (!(c & MARGIN)) {
        m = amp[i] / cabs_ref(a[i]);
      } else {
        const R aa = cabs_ref(a[i]);
        if (aa < amp_min[i]) m = amp_min[i] / aa;
        else if (aa > amp_max[i]) m = amp_max[i] / aa;
      }
      if (isfinite(m)) {
        a[i] = cscale(a[i], m);
      } else {
        a[i].re = amp[i];
        a[i].im = 0;
      }
    }
  }
  return 0;
}

EXPORT void or_phased_projection(oc *a, const oc *phased, const int *flags, long N) {
  for (long i = 0; i < N; ++i)
    if (flags[i] & MEASURED) a[i] = phased[i];
}

EXPORT void or_apply_constraints(oc *g, const int *flags, long N, int c) {
  for (long i = 0; i < N; ++i) {
    if (!(flags[i] & SUPPORT)) continue;
    if (c & REAL_OBJ) {
      g[i].im = 0;
    } else if (c & POS_REAL) {
      if (g[i].re < 0) g[i].re = (c & FLIP) ? (R)fabs(g[i].re) : 0;
      g[i].im = 0;
    } else if (c & POS_COMPLEX) {
      if (g[i].re < 0) g[i].re = (c & FLIP) ? (R)fabs(g[i].re) : 0;
      if (g[i].im < 0) g[i].im = (c & FLIP) ? (R)fabs(g[i].im) : 0;
    }
  }
}

static void scale_inplace(oc *a, long N, R s) {
  /* sp_cmul with (s, 0), src/image_util.c:1967-1970 */
  for (long i = 0; i < N; ++i) {
    const R re = a[i].re * s - a[i].im * 0;
    const R im = a[i].re * 0 + s * a[i].im;
    a[i].re = re;
    a[i].im = im;
  }
}

static R beta_at(const OrAlg *alg, int it) { return or_bezier(alg->beta_keys, alg->beta_vals, alg->nbeta, (R)it); }

static int fourier_step(OrPhaser *p, oc *G, int c) {
  if (p->objective == 0)
    return or_module_projection(G, p->amp, p->amp_min, p->amp_max, p->flags, p->N, c, p->have_margins);
  or_phased_projection(G, p->phased_amp, p->flags, p->N);
  return 0;
}

/* ER / HIO / RAAR share everything but the outside-support update */
static int iterate_simple(OrPhaser *p, const OrAlg *alg, int n) {
  const long N = p->N;
  const R s = (R)(1.0 / (double)N);
  for (int it = 0; it < n; ++it) {
    const R beta = (alg->type == ER) ? 0 : beta_at(alg, p->iteration);
    oc *sw = p->g0;
    p->g0 = p->g1;
    p->g1 = sw;
    or_fft(p->g0, p->g1, p->nx, p->ny, p->nz, -1);
    or_apply_fourier_constraints(p->g1, N, alg->constraints);
    if (fourier_step(p, p->g1, alg->constraints)) return -1;
    or_fft(p->g1, p->gp, p->nx, p->ny, p->nz, +1);
    scale_inplace(p->gp, N, s);
    for (long i = 0; i < N; ++i) {
      if (p->flags[i] & SUPPORT) {
        p->g1[i] = p->gp[i];
      } else if (alg->type == HIO) {
        p->g1[i].re = p->g0[i].re - p->gp[i].re * beta;
        p->g1[i].im = p->g0[i].im - p->gp[i].im * beta;
      } else if (alg->type == RAAR) {
        const R c1 = 1 - 2 * beta;
        p->g1[i].re = p->gp[i].re * c1 + p->g0[i].re * beta;
        p->g1[i].im = p->gp[i].im * c1 + p->g0[i].im * beta;
      } else {
        p->g1[i].re = 0;
        p->g1[i].im = 0;
      }
    }
    or_apply_constraints(p->g1, p->flags, N, alg->constraints);
    p->iteration++;
  }
  return 0;
}

static int iterate_dm(OrPhaser *p, const OrAlg *alg, int n) {
  const long N = p->N;
  const R size = (R)N;
  oc *f1 = (oc *)malloc(sizeof(oc) * N);
  for (int it = 0; it < n; ++it) {
    const R beta = beta_at(alg, p->iteration);
    oc *sw = p->g0;
    p->g0 = p->g1;
    p->g1 = sw;
    or_fft(p->g0, p->g1, p->nx, p->ny, p->nz, -1);
    or_apply_fourier_constraints(p->g1, N, alg->constraints);
    for (long i = 0; i < N; ++i) {
      f1[i] = p->g0[i];
      if (!(p->flags[i] & SUPPORT)) {
        f1[i].re = -alg->gamma1 * f1[i].re;
        f1[i].im = -alg->gamma1 * f1[i].im;
      }
    }
    or_fft(f1, f1, p->nx, p->ny, p->nz, -1);
    if (or_module_projection(f1, p->amp, p->amp_min, p->amp_max, p->flags, N, alg->constraints, p->have_margins)) {
      free(f1);
      return -1;
    }
    or_fft(f1, f1, p->nx, p->ny, p->nz, +1);
    if (or_module_projection(p->g1, p->amp, p->amp_min, p->amp_max, p->flags, N, alg->constraints, p->have_margins)) {
      free(f1);
      return -1;
    }
    or_fft(p->g1, p->gp, p->nx, p->ny, p->nz, +1); /* unscaled, like the CPU engine */
    for (long i = 0; i < N; ++i) {
      const oc g0 = p->g0[i], rho = p->gp[i];
      oc g1;
      if (p->flags[i] & SUPPORT) {
        g1.re = g0.re + beta * ((1 + alg->gamma2) * rho.re / size - alg->gamma2 * g0.re);
        g1.im = g0.im + beta * ((1 + alg->gamma2) * rho.im / size - alg->gamma2 * g0.im);
      } else {
        g1 = g0;
      }
      g1.re -= beta * f1[i].re / size;
      g1.im -= beta * f1[i].im / size;
      p->g1[i] = g1;
    }
    or_apply_constraints(p->g1, p->flags, N, alg->constraints);
    p->iteration++;
  }
  free(f1);
  return 0;
}

static int run_alg(OrPhaser *p, const OrAlg *alg, int n) {
  if (n <= 0) return 0;
  if (alg->type == DIFFMAP) return iterate_dm(p, alg, n);
  return iterate_simple(p, alg, n);
}

/* ------------------------------------------------------------------------ */
/* Gaussian blur and support updates                                         */
/* ------------------------------------------------------------------------ */
EXPORT void or_gaussian_blur(const oc *in, oc *out, int nx, int ny, int nz, R radius) {
  const long N = (long)nx * ny * nz;
  oc *F = (oc *)malloc(sizeof(oc) * N);
  or_fft(in, F, nx, ny, nz, -1);
  for (long i = 0; i < N; ++i) {
    const long x = i % nx, y = (i / nx) % ny, z = i / ((long)nx * ny);
    const R dx = (nx - x < x) ? (R)(nx - x) : (R)x;
    const R dy = (ny - y < y) ? (R)(ny - y) : (R)y;
    const R dz = (nz - z < z) ? (R)(nz - z) : (R)z;
    const R dist = (R)sqrt((double)dx * dx + (double)dy * dy + (double)dz * dz);
    const R r = dist / (R)nx;
    const R h = (R)(exp(-2.0 * (M_PI * M_PI) * ((double)r * r) * ((double)radius * radius)) / (double)(R)N);
    const R re = F[i].re * h - F[i].im * 0;
    const R im = F[i].re * 0 + h * F[i].im;
    F[i].re = re;
    F[i].im = im;
  }
  or_fft(F, out, nx, ny, nz, +1);
  free(F);
}

static int cmp_desc(const void *a, const void *b) {
  const R x = *(const R *)a, y = *(const R *)b;
  return (x < y) - (x > y);
}

/* blur of |g1| and its magnitudes */
static R *blur_abs(OrPhaser *p, R radius) {
  const long N = p->N;
  oc *tmp = (oc *)calloc(N, sizeof(oc));
  for (long i = 0; i < N; ++i) {
    tmp[i].re = cabs_ref(p->g1[i]);
    tmp[i].im = 0;
  }
  oc *bl = (oc *)malloc(sizeof(oc) * N);
  or_gaussian_blur(tmp, bl, p->nx, p->ny, p->nz, radius);
  R *m = (R *)malloc(sizeof(R) * N);
  for (long i = 0; i < N; ++i) m[i] = cabs_ref(bl[i]);
  free(tmp);
  free(bl);
  return m;
}

static void mask_from_threshold(OrPhaser *p, const R *m, R t) {
  for (long i = 0; i < p->N; ++i) {
    if (m[i] > t) p->flags[i] |= SUPPORT;
    else p->flags[i] &= ~SUPPORT;
  }
  p->last_threshold = t;
}

EXPORT void or_support_area_update(OrPhaser *p, R radius, R area) {
  R *m = blur_abs(p, radius);
  R *sorted = (R *)malloc(sizeof(R) * p->N);
  memcpy(sorted, m, sizeof(R) * p->N);
  qsort(sorted, p->N, sizeof(R), cmp_desc);
  long k = (long)(int)((float)p->N * (float)area);
  if (k >= p->N) k = p->N - 1;
  mask_from_threshold(p, m, sorted[k]);
  free(sorted);
  free(m);
}

EXPORT void or_support_threshold_update(OrPhaser *p, R radius, R rel) {
  R *m = blur_abs(p, radius);
  R mx = m[0];
  for (long i = 1; i < p->N; ++i)
    if (m[i] > mx) mx = m[i];
  mask_from_threshold(p, m, mx * rel);
  free(m);
}

static void support_update(OrPhaser *p, const OrSup *s) {
  if (s->type == 5) return;
  const R radius = or_bezier(s->blur_keys, s->blur_vals, s->nblur, (R)p->iteration);
  const R v = or_bezier(s->val_keys, s->val_vals, s->nval, (R)p->iteration);
  if (s->type == 2) or_support_area_update(p, radius, v);
  else if (s->type == 1) or_support_threshold_update(p, radius, v);
}

/* the reference's support scheduler (src/phasing.c:1000-1022) */
EXPORT int or_iterate(OrPhaser *p, const OrAlg *alg, const OrSup *sup, int iterations) {
  int ret = 0;
  if (sup && sup->type != 0) {
    while (iterations) {
      int to_update = sup->period - 1 - p->iteration % sup->period;
      const int to_iterate = iterations < to_update ? iterations : to_update;
      ret = run_alg(p, alg, to_iterate);
      if (ret) return ret;
      iterations -= to_iterate;
      to_update -= to_iterate;
      if (to_update == 0 && iterations > 0) {
        if (p->nsup < 256) p->sup_log[p->nsup] = p->iteration;
        p->nsup++;
        support_update(p, sup);
        p->iteration++;
        iterations -= 1;
      }
    }
  } else if (iterations) {
    ret = run_alg(p, alg, iterations);
  }
  return ret;
}

/* ------------------------------------------------------------------------ */
/* metrics (src/phasing.c:803-893); accumulators are R like the reference    */
/* ------------------------------------------------------------------------ */
EXPORT R or_efourier(OrPhaser *p, R *fcfo) {
  oc *F = (oc *)malloc(sizeof(oc) * p->N);
  or_fft(p->g1, F, p->nx, p->ny, p->nz, -1);
  R den = 0, nom = 0, fden = 0, fnum = 0;
  for (long i = 0; i < p->N; ++i) {
    if (p->flags[i] & MEASURED) {
      const oc a = {p->amp[i], 0};
      fnum += cabs_ref(a);
      fden += cabs_ref(F[i]);
      nom += pow(cabs_ref(F[i]) - cabs_ref(a), 2);
      den += pow(cabs_ref(a), 2);
    } else {
      den += pow(cabs_ref(F[i]), 2);
    }
  }
  free(F);
  if (fcfo) *fcfo = fnum / fden;
  return (R)sqrt(nom / den);
}

EXPORT R or_ereal(OrPhaser *p) {
  R nom = 0, den = 0;
  for (long i = 0; i < p->N; ++i) {
    if (!(p->flags[i] & SUPPORT)) nom += pow(cabs_ref(p->gp[i]), 2);
    den += pow(cabs_ref(p->gp[i]), 2);
  }
  return (R)sqrt(nom / den);
}

EXPORT R or_support_fraction(OrPhaser *p) {
  long c = 0;
  for (long i = 0; i < p->N; ++i)
    if (p->flags[i] & SUPPORT) ++c;
  return (R)c / (R)p->N;
}

/* model change g1 - g0 (src/phasing.c:533-554) */
EXPORT void or_model_change(OrPhaser *p, oc *out) {
  for (long i = 0; i < p->N; ++i) {
    out[i].re = p->g1[i].re - p->g0[i].re;
    out[i].im = p->g1[i].im - p->g0[i].im;
  }
}

/* the difference map's automatic gammas (src/phasing.c:16-47) */
EXPORT void or_dm_gammas(R beta0, R *gamma1, R *gamma2) {
  const R sigma = (R)0.1;
  *gamma1 = -(4 + (2 + beta0) * sigma + beta0 * sigma * sigma) / (beta0 * (4 - sigma + sigma * sigma));
  *gamma2 = (3 - beta0) / (2 * beta0);
}
