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
 *   or_template_prepare, or_support_template_update
 *                        src/support_update.c:47-85,181-207
 *   or_support_close_update      src/support_update.c:213-280
 *   or_support_centre_image      src/support_update.c:282-389 with
 *                        src/image_util.c:3587-3611 and src/linear_alg.c:444-475
 *   or_iterate_multi     src/phasing.c:995-1024 with src/support_update.c:489-494
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
 * recurrence enumerated by hand; the closing against scipy.ndimage's binary
 * dilation and erosion with the cross structure; the centre-image update
 * against np.roll of the support by the shift that moves its float64
 * centre of mass to the origin, and against hand-built cases of the
 * half-image shift; the template threshold against a brute-force sort.
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

/* Threads of the FFTW3-API library's plans (the reference's sp_init_fft,
 * src/fft.c:34-49: fftwf_init_threads + fftwf_plan_with_nthreads); only for
 * the all-cores CPU baseline.  0 on success, -1 without a loaded library or
 * thread support. */
EXPORT int or_set_fft_threads(int nthreads) {
  if (g_backend != 1) return -1;
  typedef int (*init_fn)(void);
  typedef void (*nthr_fn)(int);
  void *h = dlopen(NULL, RTLD_NOW | RTLD_GLOBAL);
  init_fn init = (init_fn)dlsym(h, "fftwf_init_threads");
  nthr_fn with = (nthr_fn)dlsym(h, "fftwf_plan_with_nthreads");
  if (!with) return -1;
  if (init && !init()) return -1;
  with(nthreads);
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
  int type; /* 0 none, 1 threshold, 2 area, 3 centre image, 4 template, 5 static, 6 close */
  int period;
  int nblur;
  R blur_keys[32], blur_vals[32];
  int nval;
  R val_keys[32], val_vals[32]; /* area or relative threshold */
  int close_size;               /* close: dilations (= erosions) */
  const R *tmpl_abs;            /* template: |blurred template|, N values */
  const R *tmpl_sorted;         /* template: the same sorted descending */
  R tmpl_area;                  /* template: original area fraction */
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

/* |blur(template, radius)|, the same values sorted descending, and the
 * template's area fraction: the set-up of sp_support_template_alloc
 * (src/support_update.c:47-85).  The fraction counts pixels with a non-zero
 * real or imaginary part, accumulated in R like the reference's. */
EXPORT void or_template_prepare(const oc *tmpl, int nx, int ny, int nz, R radius, R *abs_out, R *sorted_out,
                                R *area_out) {
  const long N = (long)nx * ny * nz;
  oc *bl = (oc *)malloc(sizeof(oc) * N);
  or_gaussian_blur(tmpl, bl, nx, ny, nz, radius);
  R area = 0;
  for (long i = 0; i < N; ++i) {
    abs_out[i] = cabs_ref(bl[i]);
    if (tmpl[i].re != 0 || tmpl[i].im != 0) area += (R)1.0;
  }
  *area_out = area / (R)N;
  memcpy(sorted_out, abs_out, sizeof(R) * N);
  qsort(sorted_out, N, sizeof(R), cmp_desc);
  free(bl);
}

/* src/support_update.c:181-207: threshold at index (int)(N * area * fraction)
 * of the sorted template, computed in R (clamped to the array) */
EXPORT void or_support_template_update(OrPhaser *p, const R *tmpl_abs, const R *sorted, R area, R fraction) {
  long k = (long)(int)((R)p->N * area * fraction);
  if (k >= p->N) k = p->N - 1;
  if (k < 0) k = 0;
  mask_from_threshold(p, tmpl_abs, sorted[k]);
}

static int flag_at(const int *f, int nx, int ny, int x, int y, int z) {
  return f[((long)z * ny + y) * nx + x] & SUPPORT;
}

/* src/support_update.c:213-280: `size` dilations, then `size` erosions, each
 * from one buffer into the other */
EXPORT void or_support_close_update(OrPhaser *p, int size) {
  const int nx = p->nx, ny = p->ny, nz = p->nz;
  int *a = (int *)malloc(sizeof(int) * p->N), *b = (int *)malloc(sizeof(int) * p->N);
  memcpy(a, p->flags, sizeof(int) * p->N);
  memcpy(b, p->flags, sizeof(int) * p->N);
  for (int pass = 0; pass < 2 * size; ++pass) {
    const int dilate = pass < size;
    for (int x = 0; x < nx; ++x)
      for (int y = 0; y < ny; ++y)
        for (int z = 0; z < nz; ++z) {
          int n[6], k = 0;
          if (x != 0) n[k++] = flag_at(a, nx, ny, x - 1, y, z);
          if (x != nx - 1) n[k++] = flag_at(a, nx, ny, x + 1, y, z);
          if (y != 0) n[k++] = flag_at(a, nx, ny, x, y - 1, z);
          if (y != ny - 1) n[k++] = flag_at(a, nx, ny, x, y + 1, z);
          if (z != 0) n[k++] = flag_at(a, nx, ny, x, y, z - 1);
          if (z != nz - 1) n[k++] = flag_at(a, nx, ny, x, y, z + 1);
          int v = flag_at(a, nx, ny, x, y, z) != 0;
          for (int q = 0; q < k; ++q) v = dilate ? (v || n[q]) : (v && n[q]);
          const long i = ((long)z * ny + y) * nx + x;
          b[i] = v ? (b[i] | SUPPORT) : (b[i] & ~SUPPORT);
        }
    int *t = a;
    a = b;
    b = t;
  }
  memcpy(p->flags, a, sizeof(int) * p->N);
  free(a);
  free(b);
}

/* sp_pixel_flags_translate_mask (src/image_util.c:3587-3611): the support
 * bit moves cyclically by (dx, dy, dz); the other bits stay in place */
static void translate_support(OrPhaser *p, int dx, int dy, int dz) {
  const int nx = p->nx, ny = p->ny, nz = p->nz;
  while (dx < 0) dx += nx;
  while (dy < 0) dy += ny;
  while (dz < 0) dz += nz;
  int *t = (int *)malloc(sizeof(int) * p->N);
  for (int x = 0; x < nx; ++x)
    for (int y = 0; y < ny; ++y)
      for (int z = 0; z < nz; ++z) {
        const long src = ((long)z * ny + y) * nx + x;
        const long dst = ((long)((z + dz) % nz) * ny + (y + dy) % ny) * nx + (x + dx) % nx;
        t[dst] = (p->flags[src] & SUPPORT) ? (p->flags[dst] | SUPPORT) : (p->flags[dst] & ~SUPPORT);
      }
  memcpy(p->flags, t, sizeof(int) * p->N);
  free(t);
}

/* src/support_update.c:282-389 on the support bits (the reference also
 * translates its cached copy of the model, which the next model read
 * replaces; the iterate does not move).  The centre of mass follows
 * sp_i3matrix_binary_center_of_mass_shifted (src/linear_alg.c:444-475):
 * float sums in raster order, then truncation to int. */
EXPORT void or_support_centre_image(OrPhaser *p) {
  const int nx = p->nx, ny = p->ny, nz = p->nz;
  if (nz == 1) {
    for (int ix = 0; ix < nx; ++ix)
      if (flag_at(p->flags, nx, ny, ix, (ny - 1) / 2, 0)) {
        translate_support(p, 0, (ny - 1) / 2, 0);
        break;
      }
    for (int iy = 0; iy < ny; ++iy)
      if (flag_at(p->flags, nx, ny, (nx - 1) / 2, iy, 0)) {
        translate_support(p, (nx - 1) / 2, 0, 0);
        break;
      }
  } else {
    /* each break leaves the inner loop only, as in the reference */
    for (int iy = 0; iy < ny; ++iy)
      for (int iz = 0; iz < nz; ++iz)
        if (flag_at(p->flags, nx, ny, (nx - 1) / 2, iy, iz)) {
          translate_support(p, (nx - 1) / 2, 0, 0);
          break;
        }
    for (int ix = 0; ix < nx; ++ix)
      for (int iz = 0; iz < nz; ++iz)
        if (flag_at(p->flags, nx, ny, ix, (ny - 1) / 2, iz)) {
          translate_support(p, 0, (ny - 1) / 2, 0);
          break;
        }
    for (int ix = 0; ix < nx; ++ix)
      for (int iy = 0; iy < ny; ++iy)
        if (flag_at(p->flags, nx, ny, ix, iy, (nz - 1) / 2)) {
          translate_support(p, 0, 0, (nz - 1) / 2);
          break;
        }
  }
  float c[3] = {0.f, 0.f, 0.f};
  int sum = 0;
  long i = 0;
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x, ++i)
        if (p->flags[i] & SUPPORT) {
          c[0] += (float)((x + (nx - 1) / 2) % nx);
          c[1] += (float)((y + (ny - 1) / 2) % ny);
          c[2] += (float)((z + (nz - 1) / 2) % nz);
          sum++;
        }
  if (sum) {
    c[0] /= (float)sum;
    c[1] /= (float)sum;
    c[2] /= (float)sum;
    c[0] -= (float)((nx - 1) / 2);
    c[1] -= (float)((ny - 1) / 2);
    c[2] -= (float)((nz - 1) / 2);
  }
  translate_support(p, (int)(0.f - c[0]), (int)(0.f - c[1]), (int)(0.f - c[2]));
}

static void support_update(OrPhaser *p, const OrSup *s) {
  if (s->type == 5) return;
  if (s->type == 6) {
    or_support_close_update(p, s->close_size);
    return;
  }
  if (s->type == 3) {
    or_support_centre_image(p);
    return;
  }
  const R radius = or_bezier(s->blur_keys, s->blur_vals, s->nblur, (R)p->iteration);
  const R v = or_bezier(s->val_keys, s->val_vals, s->nval, (R)p->iteration);
  if (s->type == 2) or_support_area_update(p, radius, v);
  else if (s->type == 1) or_support_threshold_update(p, radius, v);
  else if (s->type == 4) or_support_template_update(p, s->tmpl_abs, s->tmpl_sorted, v, s->tmpl_area);
}

/* the reference's support scheduler (src/phasing.c:1000-1022) over an array
 * of nsup support algorithms run in order (src/support_update.c:489-494);
 * the period is the first one's */
EXPORT int or_iterate_multi(OrPhaser *p, const OrAlg *alg, const OrSup *sup, int nsup, int iterations) {
  int ret = 0;
  if (sup && nsup > 0 && sup->type != 0) {
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
        for (int k = 0; k < nsup; ++k) support_update(p, &sup[k]);
        p->iteration++;
        iterations -= 1;
      }
    }
  } else if (iterations) {
    ret = run_alg(p, alg, iterations);
  }
  return ret;
}

EXPORT int or_iterate(OrPhaser *p, const OrAlg *alg, const OrSup *sup, int iterations) {
  return or_iterate_multi(p, alg, sup, sup ? 1 : 0, iterations);
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
