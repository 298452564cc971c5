// Host side of the HIP phasing engine plus its small helper kernels
// (initialisation, reductions, radix select of the shrinkwrap threshold).
#include "engine.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <tuple>
#include <mutex>
#include <vector>

namespace spx {

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local char g_err[512] = "";
void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}
const char* last_error() { return g_err; }

bool device_available() {
  static int state = -1;
  if (state < 0) {
    int n = 0;
    state = (hipGetDeviceCount(&n) == hipSuccess && n > 0) ? 1 : 0;
    if (state) state = (hipSetDevice(0) == hipSuccess) ? 1 : 0;
  }
  return state == 1;
}

// ---------------------------------------------------------------------------
// twiddle tables: for stage s >= 1 of plan (L, E), entry
// twoff[s] + (k-1)*Ns + j holds exp(-2 pi i j k / (Ns R)), evaluated in
// double and rounded to float
// ---------------------------------------------------------------------------
const float2* twiddles(int L, int E) {
  static std::mutex mu;
  static std::map<int, float2*> tables;
  std::lock_guard<std::mutex> lock(mu);
  const int key = L * 64 + E;
  auto it = tables.find(key);
  if (it != tables.end()) return it->second;
  const FftPlan p = make_plan(L, E);
  std::vector<float2> h(p.twtotal > 0 ? p.twtotal : 2, make_float2(0.f, 0.f));
  for (int s = 1; s < p.n; ++s) {
    const int R = p.r[s], Ns = p.ns[s];
    for (int j = 0; j < Ns; ++j)
      for (int k = 1; k < R; ++k) {
        const double a = -2.0 * M_PI * (double)j * (double)k / (double)(Ns * R);
        h[p.twoff[s] + (k - 1) * Ns + j] = make_float2((float)cos(a), (float)sin(a));
      }
  }
  float2* d = nullptr;
  if (hipMalloc(&d, h.size() * sizeof(float2)) != hipSuccess) return nullptr;
  if (hipMemcpy(d, h.data(), h.size() * sizeof(float2), hipMemcpyHostToDevice) != hipSuccess) return nullptr;
  tables[key] = d;
  return d;
}

// ---------------------------------------------------------------------------
// plan-variant tuner
// ---------------------------------------------------------------------------
static int variant_index(const LengthEntry* e, int E) {
  for (int i = 0; i < e->nvar; ++i)
    if (e->var[i].E == E) return i;
  return -1;
}

// Milliseconds per launch of `launch` (median-free: best of a few batches).
template <class F>
static float time_launches(hipStream_t s, F&& launch, int warm = 2, int per = 8) {
  hipEvent_t a, b;
  if (hipEventCreate(&a) != hipSuccess) return 1e30f;
  if (hipEventCreate(&b) != hipSuccess) {
    (void)hipEventDestroy(a);
    return 1e30f;
  }
  for (int i = 0; i < warm; ++i) launch();
  float best = 1e30f;
  for (int rep = 0; rep < 3; ++rep) {
    (void)hipEventRecord(a, s);
    for (int i = 0; i < per; ++i) launch();
    (void)hipEventRecord(b, s);
    (void)hipEventSynchronize(b);
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) == hipSuccess && ms / per < best) best = ms / per;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return (hipGetLastError() == hipSuccess) ? best : 1e30f;
}

// Tuner scratch: finite non-zero data (a support band, measured pixels) so
// that the timed passes take the branches and traffic of a real iteration.
__global__ void k_tune_fill(float2* w, float2* g0, float2* g1, float* amp, uint8_t* fl, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const float v = (float)((unsigned)(i * 2654435761u) & 1023u) * (1.f / 1024.f);
    w[i] = make_float2(v, 0.5f - v);
    g0[i] = make_float2(0.25f * v, v);
    g1[i] = make_float2(v, 0.25f * v);
    amp[i] = 1.f + v;
    fl[i] = (uint8_t)(((i & 63) < 8 ? 1 : 0) | 2);
  }
}

static std::mutex tune_mu;

// Variant choices are made per array shape (nx, ny, nz), once: the first
// engine, batch or transform of that shape tunes them and every later one
// of the same shape, whatever its number of reconstructions or layout, uses
// the same variants, so a reconstruction computes the same bits alone and in
// a batch.  Shapes do not share choices: a 2-D 256 x 256 request and a
// 256^3 volume each get the variants that suit their own passes.
struct ShapeVariants {
  int v[3];
};
static std::map<std::tuple<int, int, int>, ShapeVariants> shape_variants;

// variant with E elements per thread (forced by SPX_FFT_E=<E> for every
// pass, or SPX_FFT_E=<row E>,<column E>; else 8) among those usable in the
// layout, else the first usable one
static int default_choice(const LengthEntry* e, bool need_t4, bool column = false) {
  const char* force = getenv("SPX_FFT_E");
  int pref = force ? atoi(force) : 8;
  if (force && column)
    if (const char* comma = strchr(force, ',')) pref = atoi(comma + 1);
  int first = -1;
  for (int i = 0; i < e->nvar; ++i) {
    if (need_t4 && !e->var[i].t4) continue;
    if (e->var[i].E == pref) return i;
    if (first < 0) first = i;
  }
  return first < 0 ? 0 : first;
}

static bool tuning_off() {
  const char* force = getenv("SPX_FFT_E");
  const char* notune = getenv("SPX_NO_TUNE");
  return (force && atoi(force) > 0) || (notune && atoi(notune) == 1);
}

// The specialised passes (k_col_plain, k_row_plain) serve the common case:
// the projection pass for the phase objective without amplitude margins,
// the real-space pass for one algorithm and beta without gp output or idle
// slots.  SPX_NO_PLAIN=1 keeps the general kernels (for A/B runs).  The tuner
// times the passes that will run.
static bool plain_off() {
  static const int off = [] {
    const char* e = getenv("SPX_NO_PLAIN");
    return (e && atoi(e) == 1) ? 1 : 0;
  }();
  return off != 0;
}
static bool plain_projection(int mode, const ColArgs& a) {
  return mode == kColProject && a.objective == 0 && !a.margin && !plain_off();
}

static void launch_projection(const Variant& v, const ColArgs& ca, bool wide, hipStream_t s) {
  if (plain_projection(kColProject, ca) && v.col_project_plain(ca, wide, s)) return;
  if (wide) v.col_project_wide(ca, s);
  else v.col(kColProject, ca, s);
}
static void launch_update(const Variant& v, const UpdateArgs& u, bool tiled, hipStream_t s) {
  if (!plain_off() && !u.gp && v.update_plain(u, tiled, s)) return;
  v.update(u, false, tiled, s);
}

// out = the transpose of each of nrec row-major (rows x cols) arrays,
// itself row-major (cols x rows): out[c * rows + r] = in[r * cols + c].
// 32 x 32 tiles through LDS (one padding column against bank conflicts), so
// both sides move 256-byte row segments of float2.
template <class T>
__global__ __launch_bounds__(256) void k_transpose(const T* __restrict__ in, T* __restrict__ out, int rows, int cols,
                                                  long per_rec) {
  __shared__ T tile[32][33];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
  const T* src = in + (long)blockIdx.z * per_rec;
  T* dst = out + (long)blockIdx.z * per_rec;
#pragma unroll
  for (int k = 0; k < 32; k += 8) {
    const int r = r0 + ty + k, c = c0 + tx;
    if (r < rows && c < cols) tile[ty + k][tx] = src[(long)r * cols + c];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 32; k += 8) {
    const int c = c0 + ty + k, r = r0 + tx;
    if (c < cols && r < rows) dst[(long)c * rows + r] = tile[tx][ty + k];
  }
}

template <class T>
static void enqueue_transpose(const T* in, T* out, int rows, int cols, int nrec, hipStream_t s) {
  const long per = (long)rows * cols;
  for (int r0 = 0; r0 < nrec; r0 += 65535) {  // grid z holds at most 65535 reconstructions
    const dim3 grid((cols + 31) / 32, (rows + 31) / 32, std::min(65535, nrec - r0));
    hipLaunchKernelGGL(k_transpose<T>, grid, dim3(256), 0, s, in + r0 * per, out + r0 * per, rows, cols, per);
  }
}

// The Fourier-space pass of a 2-D shape on its transposed copy: W (ny rows
// of nx) -> Wt (nx columns of ny, each contiguous), the column pass on Wt
// with the amplitudes in the same layout, and back.  Each step streams
// whole rows or columns, where the column pass on W reads 32-byte pieces a
// row apart (one DRAM page each once the arrays outgrow the Infinity Cache).
static void enqueue_fourier_transposed(const Shape& sh, float2* W, float2* Wt, const float* ampt, ColArgs a,
                                       hipStream_t s) {
  enqueue_transpose(W, Wt, sh.ny, sh.nx, sh.nrec, s);
  a.in = Wt;
  a.W = Wt;
  a.ampx = ampt;
  a.geo = sh.gytr;
  a.outer_per_rec = 1;
  // contiguous lines: the row passes' work-groups (several per CU); the
  // column kernel on the same geometry as the fallback for other hints
  if (!sh.ey->line_project(a, s)) sh.ey->col(kColProject, a, s);
  enqueue_transpose(Wt, W, sh.nx, sh.ny, sh.nrec, s);
}

// Single-pass tuning of one axis outside a 2-D pair (1-D and 3-D shapes):
// times the pass's hot kernel on zero-filled scratch of the real geometry
// and returns the variant index.  Caller holds tune_mu.
static int tune_length(const LengthEntry* e, int lines, const LineGeom* col_geo) {
  int choice = default_choice(e, false, col_geo != nullptr);
  if (e->nvar > 1 && !tuning_off()) {
    const long L = e->L;
    long n;
    if (col_geo) n = (long)(col_geo->nouter - 1) * col_geo->outer_stride + (L - 1) * col_geo->axis_stride + col_geo->ncontig;
    else n = (long)lines * L;
    hipStream_t s = nullptr;
    float2 *b0 = nullptr, *b1 = nullptr, *b2 = nullptr;
    float* amp = nullptr;
    uint8_t* fl = nullptr;
    bool ok = hipStreamCreateWithFlags(&s, hipStreamNonBlocking) == hipSuccess &&
              hipMalloc(&b0, sizeof(float2) * n) == hipSuccess && hipMalloc(&b1, sizeof(float2) * n) == hipSuccess &&
              hipMalloc(&b2, sizeof(float2) * n) == hipSuccess && hipMalloc(&amp, sizeof(float) * n) == hipSuccess &&
              hipMalloc(&fl, (size_t)n) == hipSuccess;
    ok = ok && hipMemsetAsync(b0, 0, sizeof(float2) * n, s) == hipSuccess &&
         hipMemsetAsync(b1, 0, sizeof(float2) * n, s) == hipSuccess &&
         hipMemsetAsync(amp, 0, sizeof(float) * n, s) == hipSuccess && hipMemsetAsync(fl, 0, (size_t)n, s) == hipSuccess;
    if (ok) {
      float best = 1e30f;
      for (int i = 0; i < e->nvar; ++i) {
        const Variant& v = e->var[i];
        float ms;
        if (col_geo) {
          ColArgs c{};
          c.in = b0;
          c.W = b0;
          c.geo = *col_geo;
          c.ampx = amp;
          ms = time_launches(s, [&] { launch_projection(v, c, false, s); });
        } else {
          UpdateArgs u{};
          u.W = b0;
          u.gcur = b1;
          u.gnew = b2;
          u.flags = fl;
          u.nlines = lines;
          u.lines_per_rec = lines;
          u.alg = kER;
          u.scale = 1.f;
          u.nsize = 1.f;
          ms = time_launches(s, [&] { launch_update(v, u, false, s); });
        }
        if (getenv("SPX_TUNE_LOG"))
          fprintf(stderr, "tune length %d (%s, %d lines): E=%d  %.2f us\n", (int)L, col_geo ? "column" : "row",
                  col_geo ? col_geo->ncontig * col_geo->nouter : lines, v.E, ms * 1000.f);
        if (ms < best) {
          best = ms;
          choice = i;
        }
      }
    }
    (void)hipStreamSynchronize(s);
    void* bufs[] = {b0, b1, b2, amp, fl};
    for (void* p : bufs)
      if (p) (void)hipFree(p);
    if (s) (void)hipStreamDestroy(s);
  }
  return choice;
}

// Joint tuning of a 2-D shape's row (x) and column (y) passes, timed as one
// iteration runs them: column pass then row pass, in sequence, because a
// variant's cost depends on where (which XCD's cache) the previous pass left
// the data.  Both layouts of the Fourier-space arrays are timed (row-major,
// and 4-column tiled when nx % 4 == 0 and the row variant has a tiled
// kernel).  The variants (*vx, *vy) are chosen at the shape's first use,
// for one reconstruction and for batches together (choose_pair_variants),
// and serve both layouts; the layout is chosen per shape and per class of batch
// size (1, 2-3, 4-7, 8 or more reconstructions: in a batch whole
// reconstructions stay on one XCD, which favours the tiled layout, and more
// of them fill the chip with wider work-groups).  Layouts and work-group
// widths compute the same bits, so these choices never change a result.
// Returns true for the tiled layout.  *wide: in the row-major layout, the
// projection pass runs with the wider work-groups
// (LengthOps::col_project_wide), also timed here with the chosen variants;
// SPX_COL_WIDE=0/1 forces the choice.  *nt: which of the passes' streams
// use the non-temporal hint (bit 0: W loads, 1: W stores, 2: g_k loads,
// 3: g_{k+1} stores).  W crosses XCDs between the passes; one reconstruction
// gains from streaming it past L2 (1024^2: 20.1 against 27.3 us per
// iteration), batches lose (1024^2 x 16: 212 against 201 us, 750^2 x 16:
// 209 against 167 us).  Streaming the iterates helps both, by different
// masks (1024^2: 18.7 us with g_{k+1} stores only; 1024^2 x 16: 184 against
// 201 us with both iterate streams).  The layouts are timed with the
// shape's variants, a default mask per batch class and that mask with W's
// setting flipped, then the best with every W setting (none or both) and
// iterate setting (none, either, both); SPX_NT=<mask> forces it.  Large
// arrays may also run the Fourier-space pass on a transposed copy of W
// (kLayTransposed, enqueue_fourier_transposed).  Caller holds tune_mu.
struct PairChoice {
  bool t4, wide;
  int nt;
  bool tr = false;  // the transposed Fourier-space pass
};
// layouts of a 2-D shape's Fourier-space pass
enum PairLayout { kLayRowMajor = 0, kLayTiled = 1, kLayTransposed = 2 };
static std::map<std::tuple<int, int, int>, PairChoice> tune_layout;
// iterate hints chosen per 3-D shape and batch class (tune_volume_hints)
static std::map<std::tuple<int, int, int, int>, int> volume_hints;

// Tuning records.  With SPX_TUNE_FILE=<path>, every choice the tuner times
// is appended to the file as one line of text, and the file's lines are read
// before the process's first choice, so a later process (a restarted server,
// a profiler run) makes the same choices without timing anything:
//   shape <nx> <ny> <nz> <Ex> <Ey> <Ez>        elements per thread (0: axis of length 1)
//   pair <nx> <ny> <class> <layout> <wide> <hints>   (layout 0 row-major, 1 tiled, 2 transposed pass)
//   volume <nx> <ny> <nz> <class> <hints>
// (class: batch class 0-3, hints: the non-temporal mask).  Lines that name a
// variant this build lacks, or that do not parse, are skipped.  Caller holds
// tune_mu.
static const char* tune_file() {
  const char* p = getenv("SPX_TUNE_FILE");
  return (p && *p) ? p : nullptr;
}

static int variant_of(int L, int E) {
  if (L == 1) return E == 0 ? 0 : -1;
  const LengthEntry* e = find_length(L);
  if (!e) return -1;
  for (int i = 0; i < e->nvar; ++i)
    if (e->var[i].E == E) return i;
  return -1;
}

static int elements_of(int L, int v) {
  if (L == 1) return 0;
  const LengthEntry* e = find_length(L);
  return e && v >= 0 && v < e->nvar ? e->var[v].E : 0;
}

static void load_tune_file() {
  static bool loaded = false;
  if (loaded) return;
  loaded = true;
  const char* path = tune_file();
  FILE* f = path ? fopen(path, "r") : nullptr;
  if (!f) return;
  char line[256];
  while (fgets(line, sizeof line, f)) {
    int a[6];
    if (sscanf(line, "shape %d %d %d %d %d %d", &a[0], &a[1], &a[2], &a[3], &a[4], &a[5]) == 6) {
      ShapeVariants sv{};
      bool ok = a[0] >= 1 && a[1] >= 1 && a[2] >= 1;
      for (int k = 0; ok && k < 3; ++k) {
        sv.v[k] = variant_of(a[k], a[3 + k]);
        ok = sv.v[k] >= 0;
      }
      if (ok) shape_variants[std::make_tuple(a[0], a[1], a[2])] = sv;
    } else if (sscanf(line, "pair %d %d %d %d %d %d", &a[0], &a[1], &a[2], &a[3], &a[4], &a[5]) == 6) {
      tune_layout[std::make_tuple(a[0], a[1], a[2])] =
          PairChoice{a[3] == kLayTiled, a[4] != 0, a[5] & 15, a[3] == kLayTransposed};
    } else if (sscanf(line, "volume %d %d %d %d %d", &a[0], &a[1], &a[2], &a[3], &a[4]) == 5) {
      volume_hints[std::make_tuple(a[0], a[1], a[2], a[3])] = a[4] & 15;
    }
  }
  fclose(f);
}

static void record_tune_line(const char* fmt, ...) {
  const char* path = tune_file();
  if (!path) return;
  FILE* f = fopen(path, "a");
  if (!f) return;
  va_list ap;
  va_start(ap, fmt);
  vfprintf(f, fmt, ap);
  va_end(ap);
  fclose(f);
}

static void record_shape(int nx, int ny, int nz, const ShapeVariants& sv) {
  record_tune_line("shape %d %d %d %d %d %d\n", nx, ny, nz, elements_of(nx, sv.v[0]), elements_of(ny, sv.v[1]),
                   elements_of(nz, sv.v[2]));
}

static int batch_class(int nrec) { return nrec <= 1 ? 0 : nrec < 4 ? 1 : nrec < 8 ? 2 : 3; }

// Tuner scratch for nr reconstructions of a 2-D shape, filled by
// k_tune_fill, on a stream of its own.
struct PairScratch {
  hipStream_t s = nullptr;
  float2 *w = nullptr, *g0 = nullptr, *g1 = nullptr;
  float2* wt = nullptr;  // the transposed pass's copy, on first use
  float* amp = nullptr;
  uint8_t* fl = nullptr;
  long n = 0;
  bool ok = false;
  bool ensure_wt() {
    if (!wt && hipMalloc(&wt, sizeof(float2) * n) != hipSuccess) wt = nullptr;
    return wt != nullptr;
  }
  explicit PairScratch(long n_) : n(n_) {
    ok = hipStreamCreateWithFlags(&s, hipStreamNonBlocking) == hipSuccess &&
         hipMalloc(&w, sizeof(float2) * n) == hipSuccess && hipMalloc(&g0, sizeof(float2) * n) == hipSuccess &&
         hipMalloc(&g1, sizeof(float2) * n) == hipSuccess && hipMalloc(&amp, sizeof(float) * n) == hipSuccess &&
         hipMalloc(&fl, (size_t)n) == hipSuccess;
    if (ok) {
      hipLaunchKernelGGL(k_tune_fill, dim3(1024), dim3(256), 0, s, w, g0, g1, amp, fl, n);
      ok = hipGetLastError() == hipSuccess;
    }
  }
  ~PairScratch() {
    if (s) (void)hipStreamSynchronize(s);
    void* bufs[] = {w, g0, g1, amp, fl, wt};
    for (void* p : bufs)
      if (p) (void)hipFree(p);
    if (s) (void)hipStreamDestroy(s);
  }
};

// Milliseconds per iteration of one configuration of a 2-D shape's passes
// on nr reconstructions: the column pass then the row pass, in sequence,
// the iterates ping-ponging as in a run.
static float time_pair(PairScratch& b, const Variant& vr, const Variant& vc, int nx, int ny, int nr, int lay, bool wd,
                       int ntv) {
  const bool t4 = lay == kLayTiled, trp = lay == kLayTransposed;
  if (trp && !b.ensure_wt()) return 1e30f;
  LineGeom g{nx, nr, (long)nx * ny, t4 ? 4L : (long)nx};
  if (t4) {
    g.tshift = 2;
    g.tstride = 4L * ny;
  }
  LineGeom gt{nx, nr, (long)nx * ny, 1L};
  gt.tshift = 0;
  gt.tstride = ny;
  ColArgs ca{};
  ca.in = b.w;
  ca.W = b.w;
  ca.geo = g;
  ca.ampx = b.amp;
  ca.nt = ntv;
  UpdateArgs u{};
  u.W = b.w;
  u.flags = b.fl;
  u.nlines = nr * ny;
  u.lines_per_rec = ny;
  u.alg = kHIO;
  u.beta = 0.9f;
  u.scale = 1.f;
  u.nsize = 1.f;
  u.nt = ntv;
  float2 *ga = b.g0, *gb = b.g1;
  hipStream_t s = b.s;
  const float ms = time_launches(
      s,
      [&] {
        u.gcur = ga;
        u.gnew = gb;
        if (trp) {
          enqueue_transpose(b.w, b.wt, ny, nx, nr, s);
          ColArgs ct = ca;
          ct.in = b.wt;
          ct.W = b.wt;
          ct.geo = gt;
          if (!vc.line_project(ct, s)) vc.col(kColProject, ct, s);
          enqueue_transpose(b.wt, b.w, nx, ny, nr, s);
        } else {
          launch_projection(vc, ca, wd, s);
        }
        launch_update(vr, u, t4, s);
        std::swap(ga, gb);
      },
      20, 30);
  if (getenv("SPX_TUNE_LOG"))
    fprintf(stderr, "tune %d x %d (x%d) %s, hints %d: row E=%d col E=%d  %.2f us\n", nx, ny, nr,
            t4 ? "tiled" : trp ? "transposed" : (wd ? "row-major wide" : "row-major"), ntv, vr.E, vc.E, ms * 1000.f);
  return ms;
}

// reconstructions a batch class is timed with: the class of 8 or more with
// at least 16, the sizes it mostly serves; the scratch is capped at 16 Mi
// elements (whole reconstructions)
static int timed_recs(int nrec, long N) {
  int nr = batch_class(nrec) == 3 && nrec < 16 ? 16 : nrec;
  if ((long)nr * N > (1L << 24)) nr = (int)((1L << 24) / N);
  return nr < 1 ? 1 : nr;
}

// The variants of a 2-D shape serve every later use of the shape, alone
// and in batches of any size (so a reconstruction computes the same bits
// either way), so they are chosen for both: each pair (row E, column E) is
// timed on one reconstruction and on a batch of 16, in each layout with the
// class's default hints, and the pair with the least cost wins, cost = the
// slowdown against the best pair on one reconstruction + half that in the
// batch.  (900^2: E 8 / 8 is best alone at 23.6 us and E 16 / 16 3.6 %
// slower, but a batch of 16 takes 222 us with 8 / 8 and 163 us with 16 / 16.)
static void choose_pair_variants(const LengthEntry* ex, const LengthEntry* ey, int ny, bool t4_shape, int* rc,
                                 int* cc) {
  const int nx = ex->L, nvx = ex->nvar, nvy = ey->nvar;
  const long N = (long)nx * ny;
  const int nr16 = timed_recs(16, N);
  const int scales[2] = {1, nr16};
  const int nscale = nr16 > 1 ? 2 : 1;
  std::vector<float> t((size_t)2 * nvx * nvy, 1e30f);
  PairScratch b((long)nr16 * N);
  if (!b.ok) return;
  for (int sc = 0; sc < nscale; ++sc) {
    const int nt0 = sc == 0 ? 3 | 8 : 4 | 8;
    for (int lay = 0; lay < 2; ++lay) {
      if (lay == 1 && !t4_shape) continue;
      for (int r = 0; r < nvx; ++r) {
        if (lay == 1 && !ex->var[r].t4) continue;
        for (int c = 0; c < nvy; ++c) {
          const float ms = time_pair(b, ex->var[r], ey->var[c], nx, ny, scales[sc], lay, false, nt0);
          float& slot = t[((size_t)sc * nvx + r) * nvy + c];
          slot = std::min(slot, ms);
        }
      }
    }
  }
  float best[2] = {1e30f, 1e30f};
  for (int sc = 0; sc < nscale; ++sc)
    for (int i = 0; i < nvx * nvy; ++i) best[sc] = std::min(best[sc], t[(size_t)sc * nvx * nvy + i]);
  float cost_best = 1e30f;
  for (int r = 0; r < nvx; ++r)
    for (int c = 0; c < nvy; ++c) {
      float cost = t[(size_t)r * nvy + c] / best[0] - 1.f;
      if (nscale == 2) cost += 0.5f * (t[((size_t)nvx + r) * nvy + c] / best[1] - 1.f);
      if (cost < cost_best) {
        cost_best = cost;
        *rc = r;
        *cc = c;
      }
    }
  if (getenv("SPX_TUNE_LOG"))
    fprintf(stderr, "tune %d x %d variants (alone and x%d) chose row E=%d col E=%d\n", nx, ny, nr16, ex->var[*rc].E,
            ey->var[*cc].E);
}

static int tune_pair(const LengthEntry* ex, const LengthEntry* ey, int ny, int nrec, bool t4_shape, int* vx, int* vy,
                     bool* wide, int* nt) {
  const auto dims = std::make_tuple(ex->L, ny, 1);
  auto known = shape_variants.find(dims);
  const bool need_row = known == shape_variants.end();
  int row_choice = need_row ? default_choice(ex, false) : known->second.v[0];
  int col_choice = need_row ? default_choice(ey, false, true) : known->second.v[1];
  if (need_row && !tuning_off() && (ex->nvar > 1 || ey->nvar > 1))
    choose_pair_variants(ex, ey, ny, t4_shape, &row_choice, &col_choice);
  // SPX_NO_T4=1 / SPX_T4=1 force the layout (when the row variant allows
  // it), SPX_TR=1 / 0 the transposed pass
  const char* not4 = getenv("SPX_NO_T4");
  const char* yes4 = getenv("SPX_T4");
  const char* trenv = getenv("SPX_TR");
  const char* wenv = getenv("SPX_COL_WIDE");
  const char* ntenv = getenv("SPX_NT");
  auto finish = [&](PairChoice pc) -> int {
    *nt = ntenv ? (atoi(ntenv) & 15) : pc.nt;
    if (need_row) {
      shape_variants[dims] = ShapeVariants{{row_choice, col_choice, 0}};
      if (!tuning_off()) record_shape(ex->L, ny, 1, shape_variants[dims]);
    }
    *vx = row_choice;
    *vy = col_choice;
    const bool forced4 = (not4 && atoi(not4) == 1) || (yes4 && atoi(yes4) == 1);
    const bool tr = trenv ? atoi(trenv) == 1 : (pc.tr && !forced4);
    const bool possible = t4_shape && ex->var[row_choice].t4;
    bool t4 = pc.t4 && possible;
    if (not4 && atoi(not4) == 1) t4 = false;
    if (yes4 && atoi(yes4) == 1) t4 = possible;
    t4 = t4 && !tr;
    const bool w = wenv ? atoi(wenv) == 1 : pc.wide;
    *wide = !t4 && !tr && w && ey->var[col_choice].wide;
    return tr ? kLayTransposed : t4 ? kLayTiled : kLayRowMajor;
  };
  const auto key = std::make_tuple(ex->L, ny, batch_class(nrec));
  auto it = tune_layout.find(key);
  if (it != tune_layout.end()) return finish(it->second);
  // the hints the layouts are timed with
  const int nt0 = batch_class(nrec) == 0 ? 3 | 8 : 4 | 8;
  if (tuning_off()) {
    tune_layout[key] = PairChoice{false, false, nt0};
    return finish(tune_layout[key]);
  }
  const int nx = ex->L;
  const int nr = timed_recs(nrec, (long)nx * ny);
  PairScratch b((long)nr * nx * ny);
  const Variant &vr = ex->var[row_choice], &vc = ey->var[col_choice];
  float best = 1e30f;
  int blay = kLayRowMajor, bnt = nt0;
  bool bwide = false;
  // with the shape's variants: 0 row-major, 1 tiled, 2 row-major with the
  // wide projection pass, 3 the transposed pass (from 4 Mi pixels per
  // launch, where the arrays leave the Infinity Cache), each with W's hint
  // as in nt0 and flipped (the hint that helps 1024^2 costs 3072^2 half
  // again: 159 against 243 us per iteration), then 8..15 the best of those
  // with the eight hint masks
  static const int kMasks[8] = {0, 3, 4, 7, 8, 11, 12, 15};
  const bool try_tr = (long)nx * ny * nr >= (1L << 22) || (trenv && atoi(trenv) == 1);
  for (int cand = 0; b.ok && cand < 16; ++cand) {
    const bool flip = cand >= 8;
    const int c4 = cand & 3;
    const int lay = flip ? blay : c4 == 1 ? kLayTiled : c4 == 3 ? kLayTransposed : kLayRowMajor;
    const bool wd = flip ? bwide : c4 == 2;
    const int ntv = flip ? kMasks[cand - 8] : (cand < 4 ? nt0 : nt0 ^ 3);
    if (flip && (ntv == nt0 || ntv == (nt0 ^ 3))) continue;
    if (!flip && c4 == 1 && !(t4_shape && vr.t4)) continue;
    if (!flip && c4 == 2 && !vc.wide) continue;
    if (!flip && c4 == 3 && !try_tr) continue;
    const float ms = time_pair(b, vr, vc, nx, ny, nr, lay, wd, ntv);
    if (ms < best) {
      best = ms;
      blay = lay;
      bwide = wd;
      bnt = ntv;
    }
  }
  tune_layout[key] = PairChoice{blay == kLayTiled, bwide, bnt, blay == kLayTransposed};
  record_tune_line("pair %d %d %d %d %d %d\n", nx, ny, batch_class(nrec), blay, bwide ? 1 : 0, bnt);
  if (getenv("SPX_TUNE_LOG"))
    fprintf(stderr, "tune %d x %d (x%d) chose %s%s, hints %d: row E=%d col E=%d\n", nx, ny, nr,
            blay == kLayTiled ? "tiled" : blay == kLayTransposed ? "transposed" : "row-major", bwide ? " wide" : "",
            bnt, vr.E, vc.E);
  return finish(tune_layout[key]);
}

// ---------------------------------------------------------------------------
// small kernels
// ---------------------------------------------------------------------------
namespace {

constexpr int kRedThreads = 256;

#define GRID_STRIDE(i, n) \
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < (n); i += (long)gridDim.x * blockDim.x)

__global__ void k_fill_c64(float2* p, long n, float2 v) {
  GRID_STRIDE(i, n) p[i] = v;
}
__global__ void k_fill_u32(unsigned* p, long n, unsigned v) {
  GRID_STRIDE(i, n) p[i] = v;
}

// amplitudes as the projection reads them: unmeasured pixels get the NaN
// pattern kUnmeasuredBits, measured NaNs the canonical quiet NaN
// row-major index i of reconstruction-stacked (ny, nx) arrays -> index in
// the 4-column tiled layout
__device__ __forceinline__ long t4_index(long i, int nx, int ny) {
  const long N = (long)nx * ny;
  const long r = i / N, j = i - r * N;
  const int y = (int)(j / nx), x = (int)(j - (long)y * nx);
  return r * N + t4_offset(x, ny) + 4L * y;
}

// encoded amplitudes (unmeasured: kUnmeasuredBits), written row-major
// (t4 == 0) or tiled
__global__ void k_encode_amp(const float* amp, const uint8_t* flags, float* ampx, long n, int nx, int ny, int t4) {
  GRID_STRIDE(i, n) {
    float a = amp[i];
    if (!(flags[i] & kMeasured)) a = __uint_as_float(kUnmeasuredBits);
    else if (isnan(a)) a = __uint_as_float(0x7FC00000u);
    ampx[t4 ? t4_index(i, nx, ny) : i] = a;
  }
}

// phased amplitudes of the SpRecoverAmplitudes objective; unmeasured pixels
// carry kUnmeasuredBits in .x
__global__ void k_encode_phased(const float2* ph, const uint8_t* flags, float2* out, long n, int nx, int ny, int t4) {
  GRID_STRIDE(i, n) {
    const float2 v = (flags[i] & kMeasured) ? ph[i] : make_float2(__uint_as_float(kUnmeasuredBits), 0.f);
    out[t4 ? t4_index(i, nx, ny) : i] = v;
  }
}

__global__ void k_layout_f32(const float* in, float* out, long n, int nx, int ny, int t4) {
  GRID_STRIDE(i, n) out[t4 ? t4_index(i, nx, ny) : i] = in[i];
}

__global__ void k_untile_c64(const float2* in, float2* out, long n, int nx, int ny) {
  GRID_STRIDE(i, n) out[i] = in[t4_index(i, nx, ny)];
}

__global__ void k_scale_copy(const float2* in, float2* out, long n, float s) {
  GRID_STRIDE(i, n) {
    const float2 v = in[i];
    // sp_cmul(v, (s, 0)) as in sp_image_scale (reference src/image_util.c:1967-1970)
    out[i] = make_float2(v.x * s - v.y * 0.f, v.x * 0.f + s * v.y);
  }
}

// per-reconstruction scale factors (random-values initialisation)
__global__ void k_scale_rec(float2* buf, long n_per_rec, long n, const float* s) {
  GRID_STRIDE(i, n) {
    const float f = s[i / n_per_rec];
    buf[i] = make_float2(buf[i].x * f, buf[i].y * f);
  }
}

__global__ void k_sub(const float2* a, const float2* b, float2* out, long n) {
  GRID_STRIDE(i, n) out[i] = make_float2(a[i].x - b[i].x, a[i].y - b[i].y);
}

// counter-based generator: splitmix64 finaliser of (seed, index, stream)
__device__ __forceinline__ double rng_u01(uint64_t seed, uint64_t i, uint64_t stream) {
  uint64_t z = seed + 0x9E3779B97F4A7C15ull * (2 * i + stream + 1);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);  // [0, 1)
}

// |amp| with a random (or zero) phase: sp_image_rephase (reference
// src/image_util.c:1044-1067)
// n_per_rec pixels per reconstruction; reconstruction r of the range draws
// from seed + r at its local pixel index.
__global__ void k_init_phases(const float* amp, float2* out, long n_per_rec, long n, uint64_t seed, int random) {
  GRID_STRIDE(i, n) {
    const float a = fabsf(amp[i]);
    if (random) {
      const long r = i / n_per_rec;
      const float phase = (float)(rng_u01(seed + (uint64_t)r, i - r * n_per_rec, 0) * M_PI * 2.0);
      out[i] = make_float2((float)(cos((double)phase) * a), (float)(sin((double)phase) * a));
    } else {
      out[i] = make_float2(a, 0.f);
    }
  }
}

__global__ void k_random_values(float2* out, long n_per_rec, long n, uint64_t seed) {
  GRID_STRIDE(i, n) {
    const long r = i / n_per_rec, j = i - r * n_per_rec;
    out[i] = make_float2((float)((float)rng_u01(seed + (uint64_t)r, j, 0) - 0.5),
                         (float)((float)rng_u01(seed + (uint64_t)r, j, 1) - 0.5));
  }
}

__global__ void k_amp2(const float* amp, float2* out, long n) {
  GRID_STRIDE(i, n) out[i] = make_float2(amp[i] * amp[i], 0.f);
}

__global__ void k_zero_unmeasured(float2* buf, const uint8_t* flags, long n) {
  GRID_STRIDE(i, n) if (!(flags[i] & kMeasured)) buf[i] = make_float2(0.f, 0.f);
}

// f1 of the difference map (reference src/phasing.c:1260-1274)
__global__ void k_dm_f1(const float2* g, const uint8_t* flags, float gamma1, float2* out, long n) {
  GRID_STRIDE(i, n) {
    const float2 v = g[i];
    out[i] = (flags[i] & kInsideSupport) ? v : make_float2(-gamma1 * v.x, -gamma1 * v.y);
  }
}

// --- per-reconstruction reductions: grid (blocks_per_rec, nrec); each block
// --- strides over its reconstruction and writes NV partial sums; a second
// --- kernel adds the partials of each reconstruction in a fixed order, so
// --- results do not depend on scheduling
#define REC_STRIDE(i, n_per_rec)                                                       \
  const long rec_base_ = (long)blockIdx.y * (n_per_rec);                              \
  for (long i = rec_base_ + blockIdx.x * (long)blockDim.x + threadIdx.x; i < rec_base_ + (n_per_rec); \
       i += (long)gridDim.x * blockDim.x)

template <int NV>
__device__ __forceinline__ void write_partials(double (&acc)[NV], double* partials) {
  __shared__ double red[NV * 32];
  block_sum<NV, kRedThreads>(acc, red);
  if (threadIdx.x == 0)
    for (int q = 0; q < NV; ++q) partials[((long)blockIdx.y * gridDim.x + blockIdx.x) * NV + q] = acc[q];
}

// E_R sums over the model before projection (reference src/phasing.c:826-832)
__global__ __launch_bounds__(kRedThreads) void k_ereal(const float2* gp, const uint8_t* flags, long n,
                                                      double* partials) {
  double acc[2] = {0.0, 0.0};
  REC_STRIDE(i, n) {
    const float2 v = gp[i];
    const double m = (double)sqrtf(v.y * v.y + v.x * v.x);
    if (!(flags[i] & kInsideSupport)) acc[0] += m * m;
    acc[1] += m * m;
  }
  write_partials<2>(acc, partials);
}

__global__ __launch_bounds__(kRedThreads) void k_sumsq_c64(const float2* in, long n, double* partials) {
  double acc[1] = {0.0};
  REC_STRIDE(i, n) {
    const float2 v = in[i];
    acc[0] += (double)(v.y * v.y + v.x * v.x);
  }
  write_partials<1>(acc, partials);
}

__global__ __launch_bounds__(kRedThreads) void k_sumsq_f32(const float* in, long n, double* partials) {
  double acc[1] = {0.0};
  REC_STRIDE(i, n) acc[0] += (double)(in[i] * in[i]);
  write_partials<1>(acc, partials);
}

__global__ __launch_bounds__(kRedThreads) void k_count_support(const uint8_t* flags, long n, double* partials) {
  double acc[1] = {0.0};
  REC_STRIDE(i, n) acc[0] += (flags[i] & kInsideSupport) ? 1.0 : 0.0;
  write_partials<1>(acc, partials);
}

// one block per reconstruction: out[rec*nv + q] = sum of its partials
__global__ __launch_bounds__(256) void k_reduce_rec(const double* partials, int parts, int nv, double* out) {
  __shared__ double sh[256];
  const double* p = partials + (long)blockIdx.x * parts * nv;
  for (int q = 0; q < nv; ++q) {
    double s = 0.0;
    for (int b = threadIdx.x; b < parts; b += blockDim.x) s += p[(long)b * nv + q];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[(long)blockIdx.x * nv + q] = sh[0];
    __syncthreads();
  }
}

// --- exact k-th largest |blur| per reconstruction (descending index k, the
// --- reference's qsort-then-index, src/support_update.c:143-144) by a
// --- three-level radix select over the float bit patterns (non-negative
// --- floats order like their bits): 11 + 11 + 10 bits.
__global__ void k_ctl_init(unsigned* ctl, int nrec, unsigned k) {
  for (int r = threadIdx.x; r < nrec; r += blockDim.x) {
    unsigned* c = ctl + r * kCtlStride;
    c[kCtlK] = k;
    c[kCtlPrefix] = 0;
    c[kCtlThresh] = 0;
    for (int x = 0; x < kHistSlices; ++x) c[kCtlMax + x] = 0;
  }
}

// one block per reconstruction
__global__ __launch_bounds__(1024) void k_select(const unsigned* hist_all, int level, unsigned* ctl_all) {
  __shared__ unsigned scan[1024];
  __shared__ unsigned lvl1[2048];
  const int nbins = (level == 3) ? 1024 : 2048;
  const unsigned* hist =
      hist_all + (long)blockIdx.x * kHistStride + (level == 1 ? 0 : level == 2 ? kHistLevel2 : kHistLevel3);
  unsigned* ctl = ctl_all + blockIdx.x * kCtlStride;
  const int per = nbins / 1024;  // 1 or 2 bins per thread, walked high to low
  const int tid = threadIdx.x;
  if (level == 1) {
    // the per-XCD copies of the level-1 histogram, summed
    for (int b = tid; b < 2048; b += 1024) {
      unsigned c = 0;
      for (int x = 0; x < kHistSlices; ++x) c += hist[x * 2048 + b];
      lvl1[b] = c;
    }
    __syncthreads();
    hist = lvl1;
  }
  const unsigned k = ctl[kCtlK];
  unsigned local = 0;
  for (int q = 0; q < per; ++q) local += hist[nbins - 1 - (tid * per + q)];
  scan[tid] = local;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const unsigned v = (tid >= off) ? scan[tid - off] : 0u;
    __syncthreads();
    scan[tid] += v;
    __syncthreads();
  }
  unsigned before = scan[tid] - local;  // elements in strictly higher bins
  const unsigned prefix = ctl[kCtlPrefix];
  __syncthreads();
  for (int q = 0; q < per; ++q) {
    const int b = nbins - 1 - (tid * per + q);
    const unsigned c = hist[b];
    if (c > 0 && before <= k && k < before + c) {
      ctl[kCtlK] = k - before;
      if (level == 1) ctl[kCtlPrefix] = (unsigned)b;
      else if (level == 2) ctl[kCtlPrefix] = (prefix << 11) | (unsigned)b;
      else ctl[kCtlThresh] = (prefix << 10) | (unsigned)b;
    }
    before += c;
  }
}

// grid (blocks_per_rec, nrec)
__global__ __launch_bounds__(256) void k_hist_level(const float* absb, long n, unsigned* hist_all, const unsigned* ctl_all,
                                                    int level) {
  __shared__ unsigned h[2048];
  const int nb = (level == 2) ? 2048 : 1024;
  unsigned* hist = hist_all + (long)blockIdx.y * kHistStride + (level == 2 ? kHistLevel2 : kHistLevel3);
  for (int i = threadIdx.x; i < nb; i += blockDim.x) h[i] = 0u;
  __syncthreads();
  const unsigned pre = ctl_all[blockIdx.y * kCtlStride + kCtlPrefix];
  REC_STRIDE(i, n) {
    const unsigned b = __float_as_uint(absb[i]);
    if (level == 2) {
      if ((b >> 21) == pre) atomicAdd(&h[(b >> 10) & 2047u], 1u);
    } else {
      if ((b >> 10) == pre) atomicAdd(&h[b & 1023u], 1u);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nb; i += blockDim.x)
    if (h[i]) atomicAdd(&hist[i], h[i]);
}

// support bit = |blur| > threshold (reference src/support_update.c:392-400);
// mode 0: threshold = selected value, mode 1: max * rel
// (src/support_update.c:175).  Grid (blocks_per_rec, nrec).
__global__ void k_mask(const float* absb, uint8_t* flags, long n, unsigned* ctl_all, int mode, float rel) {
  unsigned* ctl = ctl_all + blockIdx.y * kCtlStride;
  unsigned mx = 0;
  for (int x = 0; x < kHistSlices; ++x) mx = (ctl[kCtlMax + x] > mx) ? ctl[kCtlMax + x] : mx;
  const float t = (mode == 0) ? __uint_as_float(ctl[kCtlThresh]) : __uint_as_float(mx) * rel;
  // record the threshold of mode 1 (mode 0 already holds it in kCtlThresh)
  if (mode == 1 && blockIdx.x == 0 && threadIdx.x == 0) ctl[kCtlThresh] = __float_as_uint(t);
  REC_STRIDE(i, n) {
    const uint8_t f = flags[i];
    flags[i] = (absb[i] > t) ? (uint8_t)(f | kInsideSupport) : (uint8_t)(f & ~kInsideSupport);
  }
}

// One pass of the support closing (reference src/support_update.c:213-280):
// a dilation (dilate = 1: inside if the pixel or one of its 6 neighbours is)
// or an erosion (outside if the pixel or one of its neighbours is).
// Neighbours beyond the array's faces do not take part, and reconstructions
// stacked along z do not see each other.  The other flag bits are copied.
__global__ void k_close_pass(const uint8_t* __restrict__ src, uint8_t* __restrict__ dst, int nx, int ny, int nz,
                             long N, long n, int dilate) {
  const long sy = nx, sz = (long)nx * ny;
  GRID_STRIDE(i, n) {
    const long j = i % N;
    const int x = (int)(j % nx);
    const long r = j / nx;
    const int y = (int)(r % ny), z = (int)(r / ny);
    const bool want = dilate != 0;  // the neighbour state that flips the pixel
    bool flip = ((src[i] & kInsideSupport) != 0) == want;
    if (x > 0) flip |= ((src[i - 1] & kInsideSupport) != 0) == want;
    if (x < nx - 1) flip |= ((src[i + 1] & kInsideSupport) != 0) == want;
    if (y > 0) flip |= ((src[i - sy] & kInsideSupport) != 0) == want;
    if (y < ny - 1) flip |= ((src[i + sy] & kInsideSupport) != 0) == want;
    if (z > 0) flip |= ((src[i - sz] & kInsideSupport) != 0) == want;
    if (z < nz - 1) flip |= ((src[i + sz] & kInsideSupport) != 0) == want;
    const bool inside = want ? flip : !flip;
    dst[i] = (uint8_t)((src[i] & ~kInsideSupport) | (inside ? kInsideSupport : 0));
  }
}

// ---------------------------------------------------------------------------
// Centre-image support update (reference src/support_update.c:282-389 with
// sp_pixel_flags_translate_mask, src/image_util.c:3587-3611, and
// sp_i3matrix_binary_center_of_mass_shifted, src/linear_alg.c:444-475).
// The support bit moves cyclically; the other flag bits stay.  Translations
// compose into offsets o (support at p = original support at p - o).
// ---------------------------------------------------------------------------
__device__ __forceinline__ int wrap_to(long v, int m) { return (int)(((v % m) + m) % m); }

// One work-group of 256 threads per reconstruction: the scans for support
// on the middle row / column (2-D) or the middle planes (3-D), in the
// reference's order.  In 3-D each break leaves the inner loop only, so a scan
// may translate more than once, and later checks see the moved support.
__global__ __launch_bounds__(256) void k_centre_offsets(const uint8_t* __restrict__ flags, int nx, int ny, int nz,
                                                        long N, unsigned* ctl_all) {
  const uint8_t* f = flags + (long)blockIdx.x * N;
  const int hx = (nx - 1) / 2, hy = (ny - 1) / 2, hz = (nz - 1) / 2;
  int ox = 0, oy = 0, oz = 0;
  auto inside = [&](int x, int y, int z) {
    return (f[((long)wrap_to(z - oz, nz) * ny + wrap_to(y - oy, ny)) * nx + wrap_to(x - ox, nx)] & kInsideSupport) != 0;
  };
  const int t = threadIdx.x;
  if (nz == 1) {
    bool any = false;
    for (int ix = t; ix < nx; ix += 256) any |= inside(ix, hy, 0);
    if (__syncthreads_or(any)) oy = wrap_to((long)oy + hy, ny);
    any = false;
    for (int iy = t; iy < ny; iy += 256) any |= inside(hx, iy, 0);
    if (__syncthreads_or(any)) ox = wrap_to((long)ox + hx, nx);
  } else {
    for (int iy = 0; iy < ny; ++iy) {
      bool any = false;
      for (int iz = t; iz < nz; iz += 256) any |= inside(hx, iy, iz);
      if (__syncthreads_or(any)) ox = wrap_to((long)ox + hx, nx);
    }
    for (int ix = 0; ix < nx; ++ix) {
      bool any = false;
      for (int iz = t; iz < nz; iz += 256) any |= inside(ix, hy, iz);
      if (__syncthreads_or(any)) oy = wrap_to((long)oy + hy, ny);
    }
    for (int ix = 0; ix < nx; ++ix) {
      bool any = false;
      for (int iy = t; iy < ny; iy += 256) any |= inside(ix, iy, hz);
      if (__syncthreads_or(any)) oz = wrap_to((long)oz + hz, nz);
    }
  }
  if (t == 0) {
    unsigned* c = ctl_all + (long)blockIdx.x * kCtlStride;
    c[kCtlCentre] = (unsigned)ox;
    c[kCtlCentre + 1] = (unsigned)oy;
    c[kCtlCentre + 2] = (unsigned)oz;
  }
}

// out[p] = support bit of src at p - o (mode 0: o = the middle-line offsets,
// out one byte per pixel; mode 1: o = the centre-of-mass offsets, out = the
// flags with their support bit replaced)
__global__ void k_centre_move(const uint8_t* __restrict__ src, uint8_t* __restrict__ out, int nx, int ny, int nz, long N,
                              long n, const unsigned* __restrict__ ctl_all, int mode) {
  GRID_STRIDE(i, n) {
    const long rec = i / N, j = i - rec * N;
    const int x = (int)(j % nx);
    const long r = j / nx;
    const int y = (int)(r % ny), z = (int)(r / ny);
    const unsigned* c = ctl_all + rec * kCtlStride;
    int ox, oy, oz;
    if (mode == 0) {
      ox = (int)c[kCtlCentre];
      oy = (int)c[kCtlCentre + 1];
      oz = (int)c[kCtlCentre + 2];
    } else {
      ox = (int)(c[kCtlCentreShift] & 0xFFFFu);
      oy = (int)(c[kCtlCentreShift] >> 16);
      oz = (int)c[kCtlCentreShift + 1];
    }
    const long sidx = rec * N + ((long)wrap_to(z - oz, nz) * ny + wrap_to(y - oy, ny)) * nx + wrap_to(x - ox, nx);
    const uint8_t bit = src[sidx] & kInsideSupport;
    out[i] = mode == 0 ? bit : (uint8_t)((out[i] & ~kInsideSupport) | bit);
  }
}

// One work-group of 256 threads per reconstruction: the support's centre of
// mass in the translated frame, summed in float in raster order as the
// reference sums it (src/linear_alg.c:444-475), and the translation that
// moves it to the origin.  The sums are of non-negative integers, so float
// partial sums are exact up to 2^24: chunks of 4096 pixels whose exact total
// keeps a sum within 2^24 are added at once; from the first chunk that would
// pass it, that sum is accumulated pixel by pixel in raster order (one
// thread over the chunk's support pixels, compacted into LDS in order).
__global__ __launch_bounds__(256) void k_centre_mass(const uint8_t* __restrict__ sup, int nx, int ny, int nz, long N,
                                                     unsigned* ctl_all) {
  constexpr int PER = 16, CH = 256 * PER;
  __shared__ ushort4 list[CH];
  __shared__ int wsum[4][4];
  __shared__ int wpre[4];
  __shared__ int need_list;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const uint8_t* f = sup + (long)blockIdx.x * N;
  const int hx = (nx - 1) / 2, hy = (ny - 1) / 2, hz = (nz - 1) / 2;
  const long lim = 1L << 24;
  // thread 0's running state
  float s[3] = {0.f, 0.f, 0.f};
  long ex[3] = {0, 0, 0};
  bool seq[3] = {false, false, false};
  long count = 0;
  for (long c0 = 0; c0 < N; c0 += CH) {
    const long p0 = c0 + (long)t * PER;
    int x = (int)(p0 % nx);
    const long r = p0 / nx;
    int y = (int)(r % ny), z = (int)(r / ny);
    int cnt = 0, sx = 0, sy = 0, sz = 0;
    unsigned short cx[PER], cy[PER], cz[PER];
    unsigned in_mask = 0;
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const long p = p0 + j;
      const bool in = p < N && (f[p < N ? p : N - 1] & kInsideSupport);
      cx[j] = (unsigned short)((x + hx) % nx);
      cy[j] = (unsigned short)((y + hy) % ny);
      cz[j] = (unsigned short)((z + hz) % nz);
      if (in) {
        in_mask |= 1u << j;
        ++cnt;
        sx += cx[j];
        sy += cy[j];
        sz += cz[j];
      }
      if (++x == nx) {
        x = 0;
        if (++y == ny) {
          y = 0;
          ++z;
        }
      }
    }
    // per-wave inclusive scan of cnt and sums of the four values
    int inc = cnt;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int v = __shfl_up(inc, off, 64);
      if (lane >= off) inc += v;
    }
    int v4[4] = {cnt, sx, sy, sz};
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) v4[q] += __shfl_xor(v4[q], off, 64);
    if (lane == 0)
#pragma unroll
      for (int q = 0; q < 4; ++q) wsum[w][q] = v4[q];
    __syncthreads();
    if (t == 0) {
      int tot[4] = {0, 0, 0, 0};
      int pre = 0;
      for (int k = 0; k < 4; ++k) {
        wpre[k] = pre;
        pre += wsum[k][0];
        for (int q = 0; q < 4; ++q) tot[q] += wsum[k][q];
      }
      int need = 0;
      for (int q = 0; q < 3; ++q) {
        if (!seq[q] && ex[q] + tot[q + 1] <= lim) {
          ex[q] += tot[q + 1];
        } else {
          if (!seq[q]) {
            seq[q] = true;
            s[q] = (float)ex[q];  // exact: ex[q] <= 2^24
          }
          need = 1;
        }
      }
      count += tot[0];
      wsum[0][0] = tot[0];
      need_list = need;
    }
    __syncthreads();
    if (need_list) {
      int o = wpre[w] + inc - cnt;  // exclusive offset of this thread's pixels
#pragma unroll
      for (int j = 0; j < PER; ++j)
        if (in_mask & (1u << j)) list[o++] = make_ushort4(cx[j], cy[j], cz[j], 0);
      __syncthreads();
      if (t == 0) {
        const int n = wsum[0][0];
        for (int e = 0; e < n; ++e) {
          const ushort4 v = list[e];
          if (seq[0]) s[0] += (float)v.x;
          if (seq[1]) s[1] += (float)v.y;
          if (seq[2]) s[2] += (float)v.z;
        }
      }
    }
    __syncthreads();
  }
  if (t == 0) {
    for (int q = 0; q < 3; ++q)
      if (!seq[q]) s[q] = (float)ex[q];
    int d[3] = {0, 0, 0};
    if (count) {
      const float h[3] = {(float)hx, (float)hy, (float)hz};
      for (int q = 0; q < 3; ++q) {
        float c = s[q] / (float)count;
        c -= h[q];
        d[q] = (int)(0.f - c);
      }
    }
    unsigned* c = ctl_all + (long)blockIdx.x * kCtlStride;
    const int dx = wrap_to(d[0], nx), dy = wrap_to(d[1], ny), dz = wrap_to(d[2], nz);
    c[kCtlCentreShift] = (unsigned)dx | ((unsigned)dy << 16);
    c[kCtlCentreShift + 1] = (unsigned)dz;
  }
}

// support = tmpl[i] > t for every reconstruction (tmpl holds one of N)
// (reference src/support_update.c:196-200 with :392-400)
__global__ void k_template_mask(const float* __restrict__ tmpl, uint8_t* flags, long N, long n, float t) {
  GRID_STRIDE(i, n) {
    const uint8_t f = flags[i];
    flags[i] = (tmpl[i % N] > t) ? (uint8_t)(f | kInsideSupport) : (uint8_t)(f & ~kInsideSupport);
  }
}

inline dim3 ew_grid(long n) {
  long b = (n + 255) / 256;
  if (b > 8192) b = 8192;
  if (b < 1) b = 1;
  return dim3((unsigned)b);
}

// blocks per reconstruction of the per-reconstruction kernels (~4096
// elements each, 8..512)
inline int rec_blocks(long n) {
  long b = (n + 4095) / 4096;
  if (b < 8) b = 8;
  if (b > 512) b = 512;
  return (int)b;
}

}  // namespace

// ---------------------------------------------------------------------------
// shapes and pass sequences shared by the engine and the one-shot transforms
// ---------------------------------------------------------------------------
static int tune_volume_hints(const Shape& sh0);

int make_shape(int nx, int ny, int nz, int nrec, Shape* sh) {
  *sh = Shape{};
  sh->nx = nx;
  sh->ny = ny;
  sh->nz = nz;
  sh->nrec = nrec;
  if (nx < 1 || ny < 1 || nz < 1 || nrec < 1) {
    set_error("invalid shape %d x %d x %d (x %d)", nx, ny, nz, nrec);
    return -11;
  }
  if ((long)nx * ny * nz >= (1L << 31) || (long)nx * ny * nz * nrec >= (1L << 40)) {
    set_error("array of %ld elements exceeds the int range of the phasing API", (long)nx * ny * nz);
    return -11;
  }
  const int dims[3] = {nx, ny, nz};
  AxisOps* ax[3] = {&sh->ex, &sh->ey, &sh->ez};
  const LengthEntry* found[3] = {nullptr, nullptr, nullptr};
  for (int a = 0; a < 3; ++a) {
    if (dims[a] == 1) continue;
    found[a] = find_length(dims[a]);
    if (!found[a]) {
      set_error("axis length %d is not supported (supported: products of 2, 3, 5 and 7 up to 8192)", dims[a]);
      return -11;
    }
    ax[a]->e = found[a];
  }
  // y lines: (rec, z) outer blocks of nx adjacent lines; z lines: rec outer
  // blocks of nx*ny adjacent lines
  sh->gy = LineGeom{nx, nz * nrec, (long)nx * ny, (long)nx};
  sh->gz = LineGeom{nx * ny, nrec, (long)nx * ny * nz, (long)nx * ny};
  // 2-D shapes may keep their Fourier-space arrays in the 4-column tiled
  // layout (nx % 4 == 0 and a tiled row kernel for nx); the tuner decides
  sh->gyt = LineGeom{nx, nrec, (long)nx * ny, 4L};
  sh->gyt.tshift = 2;
  sh->gyt.tstride = 4L * ny;
  // columns of the transposed copy: column x of reconstruction r at r N + x ny
  sh->gytr = LineGeom{nx, nrec, (long)nx * ny, 1L};
  sh->gytr.tshift = 0;
  sh->gytr.tstride = ny;
  std::lock_guard<std::mutex> lock(tune_mu);
  load_tune_file();
  if (found[0] && found[1] && !found[2]) {
    const int lay = tune_pair(found[0], found[1], ny, nrec, nx % 4 == 0, &sh->ex.v, &sh->ey.v, &sh->colw, &sh->nt);
    sh->t4 = lay == kLayTiled;
    sh->tr = lay == kLayTransposed;
    return 0;
  }
  const auto key = std::make_tuple(nx, ny, nz);
  auto known = shape_variants.find(key);
  if (known != shape_variants.end()) {
    for (int a = 0; a < 3; ++a) ax[a]->v = known->second.v[a];
  } else {
    if (found[0]) sh->ex.v = tune_length(found[0], ny * nz * nrec, nullptr);
    if (found[1]) sh->ey.v = tune_length(found[1], 0, &sh->gy);
    if (found[2]) sh->ez.v = tune_length(found[2], 0, &sh->gz);
    shape_variants[key] = ShapeVariants{{sh->ex.v, sh->ey.v, sh->ez.v}};
    if (!tuning_off()) record_shape(nx, ny, nz, shape_variants[key]);
  }
  if (found[0] && found[1] && found[2]) sh->nt = tune_volume_hints(*sh);
  return 0;
}

int shape_variant_elements(int nx, int ny, int nz, int* e) {
  Shape sh;
  if (int rc = make_shape(nx, ny, nz, 1, &sh)) return rc;
  const AxisOps* ax[3] = {&sh.ex, &sh.ey, &sh.ez};
  for (int a = 0; a < 3; ++a) e[a] = *ax[a] ? (*ax[a])->var().E : 0;
  return 0;
}

static RowArgs row_args(const Shape& sh, const float2* in, float2* out) {
  RowArgs r{};
  r.in = in;
  r.out = out;
  r.nlines = sh.ny * sh.nz * sh.nrec;
  r.lines_per_rec = sh.ny * sh.nz;
  r.L = sh.nx;
  return r;
}

// FFT over all non-trivial axes: x by a row pass, then y and z by column passes
static void enqueue_fft_impl(const Shape& sh, const float2* in, float2* out, int dir, bool abs_in, hipStream_t s) {
  const float2* src = in;
  if (sh.ex) {
    sh.ex->row(abs_in ? kRowAbsFwd : (dir < 0 ? kRowFwd : kRowInv), row_args(sh, in, out), s);
    src = out;
  }
  ColArgs c{};
  if (sh.ey) {
    c.in = src;
    c.W = out;
    c.geo = sh.gy;
    sh.ey->col(dir < 0 ? kColFwd : kColInv, c, s);
    src = out;
  }
  if (sh.ez) {
    c.in = src;
    c.W = out;
    c.geo = sh.gz;
    sh.ez->col(dir < 0 ? kColFwd : kColInv, c, s);
    src = out;
  }
  if (src == in && in != out)
    (void)hipMemcpyAsync(out, in, sizeof(float2) * (size_t)sh.nx * sh.ny * sh.nz * sh.nrec, hipMemcpyDeviceToDevice,
                         s);
}

void enqueue_fft(const Shape& sh, const float2* in, float2* out, int dir, hipStream_t s) {
  enqueue_fft_impl(sh, in, out, dir, false, s);
}

void enqueue_abs_fft(const Shape& sh, const float2* in, float2* out, hipStream_t s) {
  enqueue_fft_impl(sh, in, out, -1, true, s);
}

// Column work after the x pass: the last non-trivial axis gets `mode`; the
// middle axis (y of a 3-D array) is transformed forward before and inverse
// after it.  For kColEFourier nothing is written back.  `tiled`: the arrays
// are in the shape's 4-column tiled layout (the iteration's W, Wf and
// Fourier-space inputs when sh.t4); scratch transforms pass false.
static void enqueue_fourier(const Shape& sh, int mode, const float2* in, float2* out, ColArgs base, hipStream_t s,
                            bool tiled) {
  const bool three = sh.ez && sh.ey;
  ColArgs a = base;
  // y lines of a 3-D shape have outer index (rec, z); every other
  // geometry has outer index rec
  if (three) {
    a.in = in;
    a.W = out;
    a.geo = sh.gy;
    a.outer_per_rec = sh.nz;
    sh.ey->col(kColFwd, a, s);
    a = base;
    a.in = out;
    a.W = out;
    a.geo = sh.gz;
    a.outer_per_rec = 1;
    if (!(plain_projection(mode, a) && sh.ez->col_project_plain(a, false, s))) sh.ez->col(mode, a, s);
    if (mode != kColEFourier) {
      a = base;
      a.in = out;
      a.W = out;
      a.geo = sh.gy;
      a.outer_per_rec = sh.nz;
      sh.ey->col(kColInv, a, s);
    }
  } else if (sh.ez) {
    a.in = in;
    a.W = out;
    a.geo = sh.gz;
    a.outer_per_rec = 1;
    if (!(plain_projection(mode, a) && sh.ez->col_project_plain(a, false, s))) sh.ez->col(mode, a, s);
  } else {
    a.in = in;
    a.W = out;
    a.geo = tiled ? sh.gyt : sh.gy;
    a.outer_per_rec = 1;
    const bool wide = mode == kColProject && !tiled && sh.colw;
    if (plain_projection(mode, a) && sh.ey->col_project_plain(a, wide, s)) return;
    if (wide) sh.ey->col_project_wide(a, s);
    else sh.ey->col(mode, a, s);
  }
}

// Non-temporal hints for the iterates of a 3-D shape's real-space pass
// (UpdateArgs::nt bits 2 and 3; W keeps ordinary accesses in 3-D), timed
// over whole iterations of the shape (x pass, y forward, z projection, y
// inverse) on scratch, per shape and class of batch size.  Streaming both
// iterates took 256^3 from 269 to 247 us per HIO iteration and 96^3 x 32
// from 546 to 500 us, but 64^3 x 32 from 122 to 129 us.  SPX_NT=<mask>
// forces the iterate bits.  Caller holds tune_mu.
static int tune_volume_hints(const Shape& sh0) {
  const char* ntenv = getenv("SPX_NT");
  if (ntenv) return atoi(ntenv) & 12;
  const auto key = std::make_tuple(sh0.nx, sh0.ny, sh0.nz, batch_class(sh0.nrec));
  auto it = volume_hints.find(key);
  if (it != volume_hints.end()) return it->second;
  if (tuning_off()) return volume_hints[key] = 0;
  const long N = (long)sh0.nx * sh0.ny * sh0.nz;
  int nr = batch_class(sh0.nrec) == 3 && sh0.nrec < 16 ? 16 : sh0.nrec;
  if ((long)nr * N > (1L << 24)) nr = (int)((1L << 24) / N);
  if (nr < 1) nr = 1;
  Shape sh = sh0;
  sh.nrec = nr;
  sh.gy = LineGeom{sh.nx, sh.nz * nr, (long)sh.nx * sh.ny, (long)sh.nx};
  sh.gz = LineGeom{sh.nx * sh.ny, nr, N, (long)sh.nx * sh.ny};
  const long n = (long)nr * N;
  hipStream_t s = nullptr;
  float2 *w = nullptr, *g0 = nullptr, *g1 = nullptr;
  float* amp = nullptr;
  uint8_t* fl = nullptr;
  bool ok = hipStreamCreateWithFlags(&s, hipStreamNonBlocking) == hipSuccess &&
            hipMalloc(&w, sizeof(float2) * n) == hipSuccess && hipMalloc(&g0, sizeof(float2) * n) == hipSuccess &&
            hipMalloc(&g1, sizeof(float2) * n) == hipSuccess && hipMalloc(&amp, sizeof(float) * n) == hipSuccess &&
            hipMalloc(&fl, (size_t)n) == hipSuccess;
  if (ok) {
    hipLaunchKernelGGL(k_tune_fill, dim3(1024), dim3(256), 0, s, w, g0, g1, amp, fl, n);
    ok = hipGetLastError() == hipSuccess;
  }
  int best_mask = 0;
  float best = 1e30f;
  for (int mask = 0; ok && mask <= 12; mask += 4) {
    ColArgs ca{};
    ca.ampx = amp;
    UpdateArgs u{};
    u.W = w;
    u.flags = fl;
    u.nlines = sh.ny * sh.nz * nr;
    u.lines_per_rec = sh.ny * sh.nz;
    u.alg = kHIO;
    u.beta = 0.9f;
    u.scale = 1.f;
    u.nsize = 1.f;
    u.nt = mask;
    float2 *ga = g0, *gb = g1;
    const float ms = time_launches(
        s,
        [&] {
          enqueue_fourier(sh, kColProject, w, w, ca, s, false);
          u.gcur = ga;
          u.gnew = gb;
          launch_update(sh.ex->var(), u, false, s);
          std::swap(ga, gb);
        },
        2, 5);
    if (getenv("SPX_TUNE_LOG"))
      fprintf(stderr, "tune %d x %d x %d (x%d) iterate hints %d: %.2f us\n", sh.nx, sh.ny, sh.nz, nr, mask,
              ms * 1000.f);
    if (ms < best) {
      best = ms;
      best_mask = mask;
    }
  }
  (void)hipStreamSynchronize(s);
  void* bufs[] = {w, g0, g1, amp, fl};
  for (void* p : bufs)
    if (p) (void)hipFree(p);
  if (s) (void)hipStreamDestroy(s);
  record_tune_line("volume %d %d %d %d %d\n", sh0.nx, sh0.ny, sh0.nz, batch_class(sh0.nrec), best_mask);
  return volume_hints[key] = best_mask;
}

// column tiles per reconstruction of the last axis (the efourier partials)
static int last_axis_tiles_per_rec(const Shape& sh) {
  if (sh.ez) {
    LineGeom g = sh.gz;
    g.nouter = 1;  // one reconstruction
    return sh.ez->col_blocks(g);
  }
  LineGeom g = sh.t4 ? sh.gyt : sh.gy;
  g.nouter = 1;
  return sh.ey->col_blocks(g);
}

// ---------------------------------------------------------------------------
// Transforms of any axis length up to kBluesteinMaxL (Bluestein's chirp-z
// method), for the Image-level and device-buffer transforms: FFTW plans any
// length (reference src/fft.c:79,95,163,185), the Stockham passes only
// products of 2, 3, 5 and 7.  With w[n] = exp(i pi n^2 / L),
//   X[k] = conj(w[k]) sum_n (x[n] conj(w[n])) w[k - n],
// a circular convolution of length M >= 2L - 1 (a power of two, so the
// Stockham passes transform it): zero-padded a = x conj(w), B = FFT_M(b)/M
// of the chirp filter b (w at 0..L-1 and at M-L+1..M-1), conv = IFFT_M(FFT_M(a) B).
// The inverse transform is the conjugate of the forward one of conj(x).
// w and B are computed in double on the host once per length and kept.
// ---------------------------------------------------------------------------
namespace {
struct Chirp {
  int M = 0;
  float2* w = nullptr;  // L entries
  float2* B = nullptr;  // M entries, FFT_M(b) / M
};
std::map<int, Chirp> g_chirps;
std::mutex g_chirp_mu;

void host_fft_pow2(std::vector<double>& re, std::vector<double>& im, int sign) {
  const int n = (int)re.size();
  for (int i = 1, j = 0; i < n; ++i) {
    int bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) {
      std::swap(re[i], re[j]);
      std::swap(im[i], im[j]);
    }
  }
  for (int len = 2; len <= n; len <<= 1)
    for (int i = 0; i < n; i += len)
      for (int j = 0; j < len / 2; ++j) {
        const double an = sign * 2.0 * M_PI * (double)j / (double)len;
        const double c = cos(an), sn = sin(an);
        const int a = i + j, b = a + len / 2;
        const double xr = re[b] * c - im[b] * sn, xi = re[b] * sn + im[b] * c;
        re[b] = re[a] - xr;
        im[b] = im[a] - xi;
        re[a] += xr;
        im[a] += xi;
      }
}

__global__ void k_bs_pre(const float2* __restrict__ in, float2* __restrict__ sc, long outer, int L, int M, long inner,
                         const float2* __restrict__ w, int conj_in) {
  const long n = outer * M * inner;
  GRID_STRIDE(i, n) {
    const long ii = i % inner, r = i / inner;
    const int m = (int)(r % M);
    const long o = r / M;
    float2 v = make_float2(0.f, 0.f);
    if (m < L) {
      float2 x = in[(o * L + m) * inner + ii];
      if (conj_in) x.y = -x.y;
      const float2 c = w[m];
      v = make_float2(x.x * c.x + x.y * c.y, x.y * c.x - x.x * c.y);  // x conj(w)
    }
    sc[i] = v;
  }
}

__global__ void k_bs_mul(float2* sc, long outer, int M, long inner, const float2* __restrict__ B) {
  const long n = outer * M * inner;
  GRID_STRIDE(i, n) {
    const int m = (int)((i / inner) % M);
    sc[i] = cmul(sc[i], B[m]);
  }
}

__global__ void k_bs_post(const float2* __restrict__ sc, float2* __restrict__ out, long outer, int L, int M, long inner,
                          const float2* __restrict__ w, int conj_out) {
  const long n = outer * L * inner;
  GRID_STRIDE(i, n) {
    const long ii = i % inner, r = i / inner;
    const int k = (int)(r % L);
    const long o = r / L;
    const float2 c = sc[(o * M + k) * inner + ii], wk = w[k];
    float2 x = make_float2(c.x * wk.x + c.y * wk.y, c.y * wk.x - c.x * wk.y);  // conj(w) c
    if (conj_out) x.y = -x.y;
    out[i] = x;
  }
}

// the Gaussian transfer function of kColBlur (src/image_filter.c:80-91) as
// an element-wise step, for shapes transformed by fft_any
__global__ void k_blur_transfer(float2* d, int nx, int ny, int nz, float sigma) {
  const long n = (long)nx * ny * nz;
  const double NN = (double)((float)n);
  const double k0 = -2.0 * (3.14159265358979323846 * 3.14159265358979323846);
  const double s2 = (double)sigma * (double)sigma;
  GRID_STRIDE(i, n) {
    const int ix = (int)(i % nx);
    const long rest = i / nx;
    const int iy = (int)(rest % ny), iz = (int)(rest / ny);
    const float dx = (nx - ix < ix) ? (float)(nx - ix) : (float)ix;
    const float dy = (ny - iy < iy) ? (float)(ny - iy) : (float)iy;
    const float dz = (nz - iz < iz) ? (float)(nz - iz) : (float)iz;
    const float dist = (float)sqrt((double)dx * dx + (double)dy * dy + (double)dz * dz);
    const float r = dist / (float)nx;
    const float h = (float)(exp(k0 * ((double)r * (double)r) * s2) / NN);
    d[i] = make_float2(d[i].x * h, d[i].y * h);
  }
}

// --- Engines of shapes with an axis outside the Stockham lengths run their
// --- iterations unfused: full transforms (fft_any) and these element-wise
// --- steps, with the per-pixel arithmetic of the fused passes.

// the projection pass's work on a full transform (src/phasing.c:1061-1109)
__global__ void k_project_any(float2* d, ColArgs a, long n) {
  const bool plain = a.objective == 0 && !a.margin;
  GRID_STRIDE(i, n) {
    float2 G = d[i];
    if (plain) {
      G = project_phase(G, a.ampx[i], a.centro != 0);
    } else {
      if (a.centro) G.y = 0.f;
      if (a.objective == 0) {
        G = project_pixel(G, a.ampx[i], a, i);
      } else {
        const float2 P = a.phased_amp[i];
        if (__float_as_uint(P.x) != kUnmeasuredBits) G = P;
      }
    }
    d[i] = G;
  }
}

// the real-space update on the inverse transform P (and Pf, the difference
// map's transform of f1), unscaled (src/phasing.c:1111-1258)
__global__ void k_update_any(const float2* __restrict__ P, const float2* __restrict__ Pf, UpdateArgs a, long n, int dm) {
  const bool er = !(a.alg == kHIO || a.alg == kRAAR);
  const float beta = a.beta;
  const float cp = a.alg == kHIO ? -beta : 1.f - 2.f * beta;
  const float cg = a.alg == kHIO ? 1.f : beta;
  const SupportConstraints sc(a.constraints);
  GRID_STRIDE(i, n) {
    const float2 g0 = a.gcur[i];
    const bool in = a.flags[i] & kInsideSupport;
    const float2 Pv = P[i];
    float2 g1;
    if (!dm) {
      const float2 gp = make_float2(Pv.x * a.scale, Pv.y * a.scale);
      if (a.gp) a.gp[i] = gp;
      float2 o = make_float2(fmaf(cp, gp.x, cg * g0.x), fmaf(cp, gp.y, cg * g0.y));
      o = er ? make_float2(0.f, 0.f) : o;
      const float2 gi = sc.apply(gp);
      g1 = make_float2(in ? gi.x : o.x, in ? gi.y : o.y);
    } else {
      if (a.gp) a.gp[i] = Pv;
      const float2 Pfv = Pf[i];
      const float2 gin = make_float2(g0.x + beta * ((1.f + a.gamma2) * Pv.x / a.nsize - a.gamma2 * g0.x),
                                     g0.y + beta * ((1.f + a.gamma2) * Pv.y / a.nsize - a.gamma2 * g0.y));
      g1 = in ? gin : g0;
      g1.x -= beta * Pfv.x / a.nsize;
      g1.y -= beta * Pfv.y / a.nsize;
      const float2 gc = sc.apply(g1);
      g1 = in ? gc : g1;
    }
    a.gnew[i] = g1;
  }
}

// E_F sums over a full transform (src/phasing.c:879-888), as kColEFourier
__global__ __launch_bounds__(kRedThreads) void k_efourier_any(const float2* F, const float* ampx, long n,
                                                              double* partials) {
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  REC_STRIDE(i, n) {
    const float af = cabsf_ref(F[i]);
    const float ab = ampx[i];
    if (__float_as_uint(ab) != kUnmeasuredBits) {
      const double aa = fabs((double)ab);
      const double d = (double)af - aa;
      acc[0] += d * d;
      acc[1] += aa * aa;
      acc[2] += aa;
      acc[3] += (double)af;
    } else {
      acc[1] += (double)af * (double)af;
    }
  }
  write_partials<4>(acc, partials);
}

// |S| into absb with the level-1 histogram (slice = block % 8) or the
// running maximum per slice, as kRowBlurOut
__global__ __launch_bounds__(256) void k_blur_abs_any(const float2* S, float* absb, unsigned* hist, unsigned* ctl, long n) {
  float mx = 0.f;
  const int slice = blockIdx.x & (kHistSlices - 1);
  GRID_STRIDE(i, n) {
    const float m = cabsf_ref(S[i]);
    absb[i] = m;
    mx = fmaxf(mx, m);
    if (hist) atomicAdd(&hist[slice * 2048 + (__float_as_uint(m) >> 21)], 1u);
  }
  if (ctl) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(&ctl[kCtlMax + slice], __float_as_uint(mx));
  }
}

__global__ void k_abs_c64(const float2* in, float2* out, long n) {
  GRID_STRIDE(i, n) out[i] = make_float2(cabsf_ref(in[i]), 0.f);
}
}  // namespace

static int next_pow2(int v) {
  int p = 1;
  while (p < v) p <<= 1;
  return p;
}

bool fft_length_supported(int L) {
  return L >= 1 && (L == 1 || find_length(L) != nullptr || (L <= kBluesteinMaxL && find_length(next_pow2(2 * L - 1))));
}

static const Chirp* chirp_for(int L) {
  std::lock_guard<std::mutex> lock(g_chirp_mu);
  auto it = g_chirps.find(L);
  if (it != g_chirps.end()) return &it->second;
  Chirp c;
  c.M = next_pow2(2 * L - 1);
  std::vector<float2> w(L);
  std::vector<double> br(c.M, 0.0), bi(c.M, 0.0);
  for (int k = 0; k < L; ++k) {
    const long q = ((long)k * k) % (2L * L);  // exp(i pi k^2 / L) has period 2L in k^2
    const double an = M_PI * (double)q / (double)L;
    w[k] = make_float2((float)cos(an), (float)sin(an));
    br[k] = cos(an);
    bi[k] = sin(an);
    if (k) {
      br[c.M - k] = cos(an);
      bi[c.M - k] = sin(an);
    }
  }
  host_fft_pow2(br, bi, -1);
  std::vector<float2> B(c.M);
  for (int m = 0; m < c.M; ++m) B[m] = make_float2((float)(br[m] / c.M), (float)(bi[m] / c.M));
  if (hipMalloc(&c.w, sizeof(float2) * L) != hipSuccess || hipMalloc(&c.B, sizeof(float2) * c.M) != hipSuccess ||
      hipMemcpy(c.w, w.data(), sizeof(float2) * L, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c.B, B.data(), sizeof(float2) * c.M, hipMemcpyHostToDevice) != hipSuccess) {
    set_error("could not set up the Bluestein tables of length %d", L);
    return nullptr;
  }
  return &(g_chirps[L] = c);
}

// Scratch elements fft_any needs for (nx, ny, nz): the largest padded axis.
static long fft_any_scratch(int nx, int ny, int nz) {
  const int dims[3] = {nx, ny, nz};
  const long n = (long)nx * ny * nz;
  long need = 0;
  for (int a = 0; a < 3; ++a)
    if (dims[a] > 1 && !find_length(dims[a])) need = std::max(need, n / dims[a] * next_pow2(2 * dims[a] - 1));
  return need;
}

// The unnormalised transform of every axis: Stockham passes where the
// length has a plan (default variants), Bluestein's method elsewhere.
static int enqueue_fft_any(const float2* in, float2* out, int nx, int ny, int nz, int dir, float2* sc, hipStream_t s) {
  const int dims[3] = {nx, ny, nz};
  const float2* src = in;
  for (int a = 0; a < 3; ++a) {
    const int L = dims[a];
    if (L == 1) continue;
    long inner = 1, outer = 1;
    for (int b = 0; b < a; ++b) inner *= dims[b];
    for (int b = a + 1; b < 3; ++b) outer *= dims[b];
    const LengthEntry* e = find_length(L);
    if (e) {
      const Variant& v = e->var[default_choice(e, false)];
      if (a == 0) {
        RowArgs r{};
        r.in = src;
        r.out = out;
        r.nlines = (int)outer;
        r.lines_per_rec = (int)outer;
        r.L = L;
        v.row(dir < 0 ? kRowFwd : kRowInv, r, s);
      } else {
        ColArgs c{};
        c.in = src;
        c.W = out;
        c.geo = LineGeom{(int)inner, (int)outer, (long)L * inner, inner};
        v.col(dir < 0 ? kColFwd : kColInv, c, s);
      }
    } else {
      const Chirp* ch = chirp_for(L);
      if (!ch) return -20;
      const int M = ch->M;
      const LengthEntry* em = find_length(M);
      const Variant& v = em->var[default_choice(em, false)];
      const long np = outer * M * inner;
      hipLaunchKernelGGL(k_bs_pre, ew_grid(np), dim3(256), 0, s, src, sc, outer, L, M, inner, ch->w, dir > 0 ? 1 : 0);
      if (a == 0) {
        RowArgs r{};
        r.in = sc;
        r.out = sc;
        r.nlines = (int)outer;
        r.lines_per_rec = (int)outer;
        r.L = M;
        v.row(kRowFwd, r, s);
        hipLaunchKernelGGL(k_bs_mul, ew_grid(np), dim3(256), 0, s, sc, outer, M, inner, ch->B);
        v.row(kRowInv, r, s);
      } else {
        ColArgs c{};
        c.in = sc;
        c.W = sc;
        c.geo = LineGeom{(int)inner, (int)outer, (long)M * inner, inner};
        v.col(kColFwd, c, s);
        hipLaunchKernelGGL(k_bs_mul, ew_grid(np), dim3(256), 0, s, sc, outer, M, inner, ch->B);
        v.col(kColInv, c, s);
      }
      hipLaunchKernelGGL(k_bs_post, ew_grid(outer * L * inner), dim3(256), 0, s, sc, out, outer, L, M, inner, ch->w,
                         dir > 0 ? 1 : 0);
    }
    src = out;
  }
  if (src == in && in != out)
    SPX_HIP(hipMemcpyAsync(out, in, sizeof(float2) * (size_t)nx * ny * nz, hipMemcpyDeviceToDevice, s));
  return hipGetLastError() == hipSuccess ? 0 : -20;
}

// The shape's axes all have Stockham plans (the engine's own path), or not
// but each is within fft_any's reach (1: use fft_any); -11 otherwise.
static int any_path(int nx, int ny, int nz) {
  const int dims[3] = {nx, ny, nz};
  bool native = true;
  for (int a = 0; a < 3; ++a) {
    if (dims[a] < 1) {
      set_error("invalid shape %d x %d x %d", nx, ny, nz);
      return -11;
    }
    if (dims[a] > 1 && !find_length(dims[a])) {
      native = false;
      if (!fft_length_supported(dims[a])) {
        set_error("axis length %d is not supported (products of 2, 3, 5 and 7 up to 8192, any other length up to %d)",
                  dims[a], kBluesteinMaxL);
        return -11;
      }
    }
  }
  if ((long)nx * ny * nz >= (1L << 31)) {
    set_error("array of %ld elements exceeds the int range of the phasing API", (long)nx * ny * nz);
    return -11;
  }
  return native ? 0 : 1;
}

// ---------------------------------------------------------------------------
// Engine
// ---------------------------------------------------------------------------
Engine* Engine::create(int nx, int ny, int nz, int nrec) {
  if (!device_available()) {
    set_error("no HIP device available");
    return nullptr;
  }
  const int path = any_path(nx, ny, nz);
  if (path < 0) return nullptr;
  if (nx < 2 || (ny < 2 && nz < 2)) {
    set_error("the HIP engine phases 2-D and 3-D arrays with nx > 1 (got %d x %d x %d)", nx, ny, nz);
    return nullptr;
  }
  Shape sh, one;
  if (path == 1) {
    // an axis without a Stockham plan: one reconstruction, unfused iterations
    if (nrec != 1) {
      set_error("batches need axis lengths that are products of 2, 3, 5 and 7 up to 8192 (got %d x %d x %d)", nx, ny,
                nz);
      return nullptr;
    }
    sh.nx = one.nx = nx;
    sh.ny = one.ny = ny;
    sh.nz = one.nz = nz;
    sh.nrec = one.nrec = 1;
  } else if (make_shape(nx, ny, nz, nrec, &sh) || make_shape(nx, ny, nz, 1, &one)) {
    return nullptr;
  }
  Engine* e = new Engine();
  e->any_mode = path == 1;
  e->sh = sh;
  e->one = one;
  e->nx = nx;
  e->ny = ny;
  e->nz = nz;
  e->nrec = nrec;
  e->N = (long)nx * ny * nz;
  e->NT = e->N * nrec;
  if (e->alloc_all()) {
    delete e;
    return nullptr;
  }
  return e;
}

int Engine::alloc_all() {
  SPX_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  SPX_HIP(hipEventCreate(&ev0));
  SPX_HIP(hipEventCreate(&ev1));
  const size_t c = sizeof(float2) * (size_t)NT, f = sizeof(float) * (size_t)NT;
  SPX_HIP(hipMalloc(&g[0], c));
  SPX_HIP(hipMalloc(&g[1], c));
  SPX_HIP(hipMalloc(&W, c));
  SPX_HIP(hipMalloc(&gp, c));
  SPX_HIP(hipMalloc(&S1, c));
  SPX_HIP(hipMalloc(&ampx, f));
  SPX_HIP(hipMalloc(&absb, f));
  SPX_HIP(hipMalloc(&flags, (size_t)NT));
  int parts = any_mode ? rec_blocks(N) : last_axis_tiles_per_rec(sh);
  if (parts < rec_blocks(N)) parts = rec_blocks(N);
  if (any_mode) SPX_HIP(hipMalloc(&Sany, sizeof(float2) * (size_t)std::max(1L, fft_any_scratch(nx, ny, nz))));
  SPX_HIP(hipMalloc(&partials, sizeof(double) * 4 * (size_t)parts * nrec));
  SPX_HIP(hipMalloc(&red, sizeof(double) * 4 * (size_t)nrec));
  SPX_HIP(hipMalloc(&hist, sizeof(unsigned) * (size_t)kHistStride * nrec));
  SPX_HIP(hipMalloc(&ctl, sizeof(unsigned) * (size_t)kCtlStride * nrec));
  hipLaunchKernelGGL(k_fill_c64, ew_grid(NT), dim3(256), 0, s, g[0], NT, make_float2(0.f, 0.f));
  hipLaunchKernelGGL(k_fill_c64, ew_grid(NT), dim3(256), 0, s, g[1], NT, make_float2(0.f, 0.f));
  hipLaunchKernelGGL(k_fill_c64, ew_grid(NT), dim3(256), 0, s, gp, NT, make_float2(0.f, 0.f));
  SPX_HIP(hipMemsetAsync(flags, 0, (size_t)NT, s));
  SPX_HIP(hipMemsetAsync(absb, 0, f, s));
  SPX_HIP(hipMemsetAsync(ctl, 0, sizeof(unsigned) * (size_t)kCtlStride * nrec, s));
  hipLaunchKernelGGL(k_encode_amp, ew_grid(NT), dim3(256), 0, s, absb, flags, ampx, NT, nx, ny, sh.t4 ? 1 : 0);
  // no wait: everything later on the stream is ordered behind these
  SPX_HIP(hipGetLastError());
  return 0;
}

Engine::~Engine() {
  // teardown: errors here have nowhere to go
  if (s) (void)hipStreamSynchronize(s);
  drop_graphs();
  void* ptrs[] = {g[0], g[1], W,        Wf,  gp,   S1,   ampx,    amp_min,  amp_max,   absb,
                  phased, flags, partials, red, hist, ctl, alg_rec, beta_rec, rec_active, Sany, Wt, ampt};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (auto& e : templates) (void)hipFree(e.second);
  if (red_host) (void)hipHostFree(red_host);
  if (ev0) (void)hipEventDestroy(ev0);
  if (ev1) (void)hipEventDestroy(ev1);
  if (s) (void)hipStreamDestroy(s);
}

// ---------------------------------------------------------------------------
// Host-to-device uploads that return at once.  A host array is copied into a
// ring of pinned host memory shared by every engine of the process, and its
// DMA is queued on the engine's stream with an event recorded behind it; a
// region of the ring is reused once the events of the uploads that held it
// have completed.  The caller may reuse its array on return, and the upload
// stays ordered with the stream's kernels.  Waiting on the stream instead
// would also wait for other streams' work queued ahead on the same hardware
// queue (milliseconds when many batches are served at once).  Arrays above a
// quarter of the ring are copied from pageable memory with a wait.
// ---------------------------------------------------------------------------
namespace {
constexpr size_t kRingBytes = 64u << 20;

struct UploadRing {
  std::mutex mu;
  uint8_t* base = nullptr;
  size_t head = 0;
  // an event is recorded on streams of the device it was created on
  struct Ev {
    hipEvent_t ev;
    int dev;
  };
  struct Span {
    size_t off, len;
    Ev ev;
  };
  std::vector<Span> live;
  std::vector<Ev> spare;
};
UploadRing g_ring;
}  // namespace

int upload_h2d(void* dst, const void* src, size_t bytes, hipStream_t s) {
  if (bytes == 0) return 0;
  if (bytes > kRingBytes / 4) {
    SPX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    SPX_HIP(hipStreamSynchronize(s));
    return 0;
  }
  UploadRing& g = g_ring;
  int dev = 0;
  SPX_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g.mu);
  if (!g.base) SPX_HIP(hipHostMalloc((void**)&g.base, kRingBytes, hipHostMallocDefault));
  const size_t len = (bytes + 255) & ~(size_t)255;
  size_t off = g.head;
  if (off + len > kRingBytes) off = 0;
  // wait for the uploads still reading [off, off + len); drop finished ones
  for (size_t i = 0; i < g.live.size();) {
    UploadRing::Span& sp = g.live[i];
    const bool overlap = sp.off < off + len && off < sp.off + sp.len;
    bool done = false;
    if (overlap) {
      SPX_HIP(hipEventSynchronize(sp.ev.ev));
      done = true;
    } else {
      const hipError_t q = hipEventQuery(sp.ev.ev);
      if (q != hipSuccess && q != hipErrorNotReady) SPX_HIP(q);
      done = q == hipSuccess;
    }
    if (done) {
      g.spare.push_back(sp.ev);
      sp = g.live.back();
      g.live.pop_back();
    } else {
      ++i;
    }
  }
  UploadRing::Ev ev{nullptr, dev};
  for (size_t i = 0; i < g.spare.size(); ++i)
    if (g.spare[i].dev == dev) {
      ev = g.spare[i];
      g.spare[i] = g.spare.back();
      g.spare.pop_back();
      break;
    }
  if (!ev.ev) SPX_HIP(hipEventCreateWithFlags(&ev.ev, hipEventDisableTiming));
  memcpy(g.base + off, src, bytes);
  const hipError_t e1 = hipMemcpyAsync(dst, g.base + off, bytes, hipMemcpyHostToDevice, s);
  const hipError_t e2 = e1 == hipSuccess ? hipEventRecord(ev.ev, s) : e1;
  if (e2 != hipSuccess) {
    // the region stays free: nothing was queued, or the queued copy is
    // waited for here since no event guards it
    if (e1 == hipSuccess) (void)hipStreamSynchronize(s);
    g.spare.push_back(ev);
    SPX_HIP(e2);
  }
  g.live.push_back({off, len, ev});
  g.head = off + len;
  return 0;
}

// encoded amplitudes: the raw values pass through `absb` (scratch) once
int Engine::upload_amplitudes(const float* amp, const uint8_t* fl, int rec) {
  const long off = rec < 0 ? 0 : rec * N, n = rec < 0 ? NT : N;
  if (int rc = upload_h2d(absb + off, amp, sizeof(float) * n, s)) return rc;
  if (int rc = upload_h2d(flags + off, fl, (size_t)n, s)) return rc;
  hipLaunchKernelGGL(k_encode_amp, ew_grid(n), dim3(256), 0, s, absb + off, flags + off, ampx + off, n, nx, ny,
                     sh.t4 ? 1 : 0);
  SPX_HIP(hipGetLastError());
  ++flags_version;
  ++amp_version;
  return 0;
}

int Engine::upload_flags(const uint8_t* fl, int rec) {
  const long off = rec < 0 ? 0 : rec * N, n = rec < 0 ? NT : N;
  if (int rc = upload_h2d(flags + off, fl, (size_t)n, s)) return rc;
  ++flags_version;
  return 0;
}

int Engine::download_flags(uint8_t* fl, int rec) {
  const long off = rec < 0 ? 0 : rec * N, n = rec < 0 ? NT : N;
  SPX_HIP(hipMemcpyAsync(fl, flags + off, (size_t)n, hipMemcpyDeviceToHost, s));
  SPX_HIP(hipStreamSynchronize(s));
  return 0;
}

// Buffers that appear on first use change the kernel arguments a captured
// graph holds, so those setters drop the graphs.
int Engine::upload_margins(const float* lo, const float* hi) {
  drop_graphs();
  if (!amp_min) SPX_HIP(hipMalloc(&amp_min, sizeof(float) * NT));
  if (!amp_max) SPX_HIP(hipMalloc(&amp_max, sizeof(float) * NT));
  // staged through the |blur| scratch into the Fourier-space layout
  SPX_HIP(hipMemcpyAsync(absb, lo, sizeof(float) * NT, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_layout_f32, ew_grid(NT), dim3(256), 0, s, absb, amp_min, NT, nx, ny, sh.t4 ? 1 : 0);
  SPX_HIP(hipMemcpyAsync(absb, hi, sizeof(float) * NT, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_layout_f32, ew_grid(NT), dim3(256), 0, s, absb, amp_max, NT, nx, ny, sh.t4 ? 1 : 0);
  SPX_HIP(hipGetLastError());
  SPX_HIP(hipStreamSynchronize(s));
  return 0;
}

int Engine::upload_phased(const float2* ph, const uint8_t* fl) {
  drop_graphs();
  if (!phased) SPX_HIP(hipMalloc(&phased, sizeof(float2) * NT));
  SPX_HIP(hipMemcpyAsync(S1, ph, sizeof(float2) * NT, hipMemcpyHostToDevice, s));
  SPX_HIP(hipMemcpyAsync(flags, fl, (size_t)NT, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_encode_phased, ew_grid(NT), dim3(256), 0, s, S1, flags, phased, NT, nx, ny, sh.t4 ? 1 : 0);
  SPX_HIP(hipGetLastError());
  SPX_HIP(hipStreamSynchronize(s));
  ++flags_version;
  return 0;
}

int Engine::upload_model(const float2* model, int rec) {
  const long off = rec < 0 ? 0 : rec * N, n = rec < 0 ? NT : N;
  SPX_HIP(hipMemcpyAsync(g[cur] + off, model, sizeof(float2) * n, hipMemcpyHostToDevice, s));
  SPX_HIP(hipStreamSynchronize(s));
  invalidate_R(rec);
  return 0;
}

int Engine::download(const float2* dev, float2* host, int rec) {
  const long off = rec < 0 ? 0 : rec * N, n = rec < 0 ? NT : N;
  SPX_HIP(hipMemcpyAsync(host, dev + off, sizeof(float2) * n, hipMemcpyDeviceToHost, s));
  SPX_HIP(hipStreamSynchronize(s));
  return 0;
}

// Captured graphs read the tables at run time, so new contents need no
// recapture; a table's first allocation or its removal changes the kernel
// arguments.  A null table is freed, so iterations go back to the shared
// algorithm type and beta schedule.
// The tables and flags below are uploaded in stream order (upload_h2d):
// iterations queued before a call use the old values, those queued after it
// the new ones.
int Engine::set_rec_params(const int* alg, const float* beta) {
  if ((alg && !alg_rec) || (beta && !beta_rec) || (!alg && alg_rec) || (!beta && beta_rec)) drop_graphs();
  // queued iterations may still read a table freed below
  if ((!alg && alg_rec) || (!beta && beta_rec)) SPX_HIP(hipStreamSynchronize(s));
  if (alg) {
    if (!alg_rec) SPX_HIP(hipMalloc(&alg_rec, sizeof(int) * nrec));
    if (int rc = upload_h2d(alg_rec, alg, sizeof(int) * nrec, s)) return rc;
  } else if (alg_rec) {
    SPX_HIP(hipFree(alg_rec));
    alg_rec = nullptr;
  }
  if (beta) {
    if (!beta_rec) SPX_HIP(hipMalloc(&beta_rec, sizeof(float) * nrec));
    if (int rc = upload_h2d(beta_rec, beta, sizeof(float) * nrec, s)) return rc;
  } else if (beta_rec) {
    SPX_HIP(hipFree(beta_rec));
    beta_rec = nullptr;
  }
  return 0;
}

// active[r]: 0 idle, 2 vacant, any other value active
int Engine::set_rec_active(const int* active) {
  bool idle = false;
  std::vector<uint8_t> h((size_t)nrec, kSlotActive);
  if (active)
    for (int r = 0; r < nrec; ++r) {
      h[r] = active[r] == 2 ? kSlotVacant : active[r] ? kSlotActive : kSlotIdle;
      if (h[r] != kSlotActive) idle = true;
    }
  if (idle) {
    if (!rec_active) {
      drop_graphs();  // a new kernel argument
      SPX_HIP(hipMalloc(&rec_active, (size_t)nrec));
    }
    if (int rc = upload_h2d(rec_active, h.data(), (size_t)nrec, s)) return rc;
  }
  if (idle != some_idle) drop_graphs();
  some_idle = idle;
  return 0;
}

int Engine::fft_full(const Shape& shp, const float2* in, float2* out, int dir) {
  if (any_mode) return enqueue_fft_any(in, out, nx, ny, nz, dir, Sany, s);
  enqueue_fft(shp, in, out, dir, s);
  return 0;
}

int Engine::fft_nd(const float2* in, float2* out, int dir) {
  if (int rc = fft_full(sh, in, out, dir)) return rc;
  SPX_HIP(hipGetLastError());
  return 0;
}

int Engine::reduce_per_rec(int parts, int nv, double* host_out, int count) {
  const int nr = count < 0 ? nrec : count;
  hipLaunchKernelGGL(k_reduce_rec, dim3(nr), dim3(256), 0, s, partials, parts, nv, red);
  SPX_HIP(hipGetLastError());
  SPX_HIP(hipMemcpyAsync(host_out, red, sizeof(double) * nv * nr, hipMemcpyDeviceToHost, s));
  SPX_HIP(hipStreamSynchronize(s));
  return 0;
}

int Engine::init_model_from_amplitudes(const float* amp_raw, int fl, uint64_t sd, int rec) {
  const int cnt = rec < 0 ? nrec : 1;
  const long off = rec < 0 ? 0 : rec * N, n = cnt * N;
  const Shape& shp = rec < 0 ? sh : one;
  if (int rc = upload_h2d(absb + off, amp_raw, sizeof(float) * n, s)) return rc;
  const float inv_n = (float)(1.0 / (double)N);
  if (fl & 1 /* SpModelRandomPhases */ || fl & 2 /* SpModelZeroPhases */) {
    hipLaunchKernelGGL(k_init_phases, ew_grid(n), dim3(256), 0, s, absb + off, S1 + off, N, n, sd,
                       (fl & 1) ? 1 : 0);
    if (int rc = fft_full(shp, S1 + off, S1 + off, +1)) return rc;
    hipLaunchKernelGGL(k_scale_copy, ew_grid(n), dim3(256), 0, s, S1 + off, g[cur] + off, n, inv_n);
  } else if (fl & 4 /* SpModelRandomValues */) {
    // reference src/phasing.c:666-679: scale so that sum |g|^2 = sum amp^2 / N
    const int bpr = rec_blocks(N);
    std::vector<double> amp2((size_t)cnt), mod2((size_t)cnt);
    hipLaunchKernelGGL(k_sumsq_f32, dim3(bpr, cnt), dim3(kRedThreads), 0, s, absb + off, N, partials);
    if (reduce_per_rec(bpr, 1, amp2.data(), cnt)) return -20;
    hipLaunchKernelGGL(k_random_values, ew_grid(n), dim3(256), 0, s, g[cur] + off, N, n, sd);
    hipLaunchKernelGGL(k_sumsq_c64, dim3(bpr, cnt), dim3(kRedThreads), 0, s, g[cur] + off, N, partials);
    if (reduce_per_rec(bpr, 1, mod2.data(), cnt)) return -20;
    std::vector<float> sc((size_t)cnt);
    for (int r = 0; r < cnt; ++r) sc[r] = (float)sqrt((amp2[r] / (double)N) / mod2[r]);
    float* dsc = (float*)partials;  // free again: the reductions above are done
    SPX_HIP(hipMemcpyAsync(dsc, sc.data(), sizeof(float) * cnt, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_scale_rec, ew_grid(n), dim3(256), 0, s, g[cur] + off, N, n, dsc);
    SPX_HIP(hipStreamSynchronize(s));
  } else {
    return -3;
  }
  SPX_HIP(hipGetLastError());
  return 0;
}

int Engine::init_model_user(const float2* model, int) {
  SPX_HIP(hipMemcpyAsync(g[cur], model, sizeof(float2) * NT, hipMemcpyHostToDevice, s));
  return 0;
}

int Engine::finish_model_init(int fl, int rec) {
  const long off = rec < 0 ? 0 : rec * N, n = rec < 0 ? NT : N;
  const Shape& shp = rec < 0 ? sh : one;
  if (fl & 256 /* SpModelMaskedOutZeroed */) {
    if (int rc = fft_full(shp, g[cur] + off, S1 + off, -1)) return rc;
    hipLaunchKernelGGL(k_zero_unmeasured, ew_grid(n), dim3(256), 0, s, S1 + off, flags + off, n);
    if (int rc = fft_full(shp, S1 + off, S1 + off, +1)) return rc;
    hipLaunchKernelGGL(k_scale_copy, ew_grid(n), dim3(256), 0, s, S1 + off, g[cur] + off, n,
                       (float)(1.0 / (double)N));
  }
  hipLaunchKernelGGL(k_fill_c64, ew_grid(n), dim3(256), 0, s, g[cur ^ 1] + off, n, make_float2(0.f, 0.f));
  hipLaunchKernelGGL(k_fill_c64, ew_grid(n), dim3(256), 0, s, gp + off, n, make_float2(0.f, 0.f));
  SPX_HIP(hipGetLastError());
  invalidate_R(rec);
  return 0;
}

int Engine::ensure_transposed() {
  if (!Wt) SPX_HIP(hipMalloc(&Wt, sizeof(float2) * NT));
  if (!ampt) SPX_HIP(hipMalloc(&ampt, sizeof(float) * NT));
  if (ampt_version != amp_version) {
    enqueue_transpose(ampx, ampt, ny, nx, nrec, s);
    SPX_HIP(hipGetLastError());
    ampt_version = amp_version;
  }
  return 0;
}

int Engine::ensure_R() {
  if (any_mode) return 0;  // unfused iterations transform the iterate in full
  if (!R_valid) {
    if (sh.t4) sh.ex->row_t4(row_args(sh, g[cur], W), s);
    else sh.ex->row(kRowFwd, row_args(sh, g[cur], W), s);
    R_valid = true;
    R_stale.clear();
  } else if (!R_stale.empty()) {
    // one reconstruction's lines at a time (its own tiled block when t4)
    std::sort(R_stale.begin(), R_stale.end());
    R_stale.erase(std::unique(R_stale.begin(), R_stale.end()), R_stale.end());
    for (int r : R_stale) {
      const long off = (long)r * N;
      if (sh.t4) one.ex->row_t4(row_args(one, g[cur] + off, W + off), s);
      else one.ex->row(kRowFwd, row_args(one, g[cur] + off, W + off), s);
    }
    R_stale.clear();
  }
  SPX_HIP(hipGetLastError());
  return 0;
}

int Engine::ensure_Rf(float gamma1) {
  if (any_mode) return 0;  // f1 is formed and transformed in every unfused iteration
  if (!Wf) SPX_HIP(hipMalloc(&Wf, sizeof(float2) * NT));
  if (Rf_valid && Rf_version == flags_version && Rf_gamma1 == gamma1) return 0;
  if (sh.t4) {
    // the tiled transform cannot run in place: f1 goes through the scratch
    hipLaunchKernelGGL(k_dm_f1, ew_grid(NT), dim3(256), 0, s, g[cur], flags, gamma1, S1, NT);
    sh.ex->row_t4(row_args(sh, S1, Wf), s);
  } else {
    hipLaunchKernelGGL(k_dm_f1, ew_grid(NT), dim3(256), 0, s, g[cur], flags, gamma1, Wf, NT);
    sh.ex->row(kRowFwd, row_args(sh, Wf, Wf), s);
  }
  SPX_HIP(hipGetLastError());
  Rf_valid = true;
  Rf_version = flags_version;
  Rf_gamma1 = gamma1;
  return 0;
}

int Engine::iterate(const IterParams& p, const float* betas, int n, bool want_gp) {
  return iterate_impl(p, betas, n, want_gp, nullptr);
}

// `marks`, when given, receives 2n+1 events: before the Fourier-space passes
// of iteration i (2i), between them and the real-space pass (2i+1), and after
// the last pass (2n).
// Unfused iterations of the any-length mode: per iteration the full
// transform of g_k, the projection, the inverse transform (and the same for
// the difference map's f1, in W), then the update, each step element-wise
// with the fused passes' arithmetic.  Marks as iterate_impl's.
int Engine::iterate_any(const IterParams& p, const float* betas, int n, bool want_gp, hipEvent_t* marks) {
  const bool dm = (p.alg == kDiffMap);
  ColArgs ca{};
  ca.ampx = ampx;
  ca.amp_min = amp_min;
  ca.amp_max = amp_max;
  ca.phased_amp = phased;
  ca.centro = (p.constraints & kCentrosymmetricObject) ? 1 : 0;
  ca.margin = (p.constraints & kAmplitudeErrorMargin) ? 1 : 0;
  ca.objective = dm ? 0 : p.objective;
  ColArgs caf = ca;
  caf.centro = 0;  // the difference map constrains FFT(g0) only (src/phasing.c:1232-1235)
  UpdateArgs u{};
  u.flags = flags;
  u.alg = p.alg;
  u.constraints = p.constraints;
  u.gamma1 = p.gamma1;
  u.gamma2 = p.gamma2;
  u.scale = (float)(1.0 / (double)N);
  u.nsize = (float)N;
  for (int i = 0; i < n; ++i) {
    if (marks) SPX_HIP(hipEventRecord(marks[2 * i], s));
    if (int rc = fft_full(sh, g[cur], S1, -1)) return rc;
    hipLaunchKernelGGL(k_project_any, ew_grid(N), dim3(256), 0, s, S1, ca, N);
    if (int rc = fft_full(sh, S1, S1, +1)) return rc;
    if (dm) {
      hipLaunchKernelGGL(k_dm_f1, ew_grid(N), dim3(256), 0, s, g[cur], flags, p.gamma1, W, N);
      if (int rc = fft_full(sh, W, W, -1)) return rc;
      hipLaunchKernelGGL(k_project_any, ew_grid(N), dim3(256), 0, s, W, caf, N);
      if (int rc = fft_full(sh, W, W, +1)) return rc;
    }
    if (marks) SPX_HIP(hipEventRecord(marks[2 * i + 1], s));
    u.gcur = g[cur];
    u.gnew = g[cur ^ 1];
    u.gp = (want_gp && i == n - 1) ? gp : nullptr;
    u.beta = betas ? betas[i] : 0.f;
    hipLaunchKernelGGL(k_update_any, ew_grid(N), dim3(256), 0, s, S1, W, u, N, dm ? 1 : 0);
    cur ^= 1;
  }
  if (marks) SPX_HIP(hipEventRecord(marks[2 * n], s));
  SPX_HIP(hipGetLastError());
  return 0;
}

int Engine::iterate_impl(const IterParams& p, const float* betas, int n, bool want_gp, hipEvent_t* marks) {
  if (n <= 0) return 0;
  if (any_mode) return iterate_any(p, betas, n, want_gp, marks);
  const bool dm = (p.alg == kDiffMap);
  if (int r = ensure_R()) return r;
  if (dm)
    if (int r = ensure_Rf(p.gamma1)) return r;
  ColArgs ca{};
  ca.ampx = ampx;
  ca.amp_min = amp_min;
  ca.amp_max = amp_max;
  ca.phased_amp = phased;
  ca.centro = (p.constraints & kCentrosymmetricObject) ? 1 : 0;
  ca.margin = (p.constraints & kAmplitudeErrorMargin) ? 1 : 0;
  ca.objective = dm ? 0 : p.objective;
  ca.nt = sh.nt;  // 3-D shapes set only the iterate bits, which column passes ignore
  ca.rec_active = some_idle ? rec_active : nullptr;
  ColArgs caf = ca;
  caf.centro = 0;  // the difference map constrains FFT(g0) only (src/phasing.c:1232-1235)
  UpdateArgs u{};
  u.W = W;
  u.Wf = Wf;
  u.flags = flags;
  u.beta_rec = beta_rec;
  u.alg_rec = dm ? nullptr : alg_rec;
  u.nlines = ny * nz * nrec;
  u.lines_per_rec = ny * nz;
  u.alg = p.alg;
  u.constraints = p.constraints;
  u.gamma1 = p.gamma1;
  u.gamma2 = p.gamma2;
  u.scale = (float)(1.0 / (double)N);
  u.nsize = (float)N;
  u.nt = sh.nt;
  u.rec_active = some_idle ? rec_active : nullptr;
  // the transposed Fourier-space pass serves the plain projection only
  const bool trp = sh.tr && !dm && plain_projection(kColProject, ca);
  if (trp)
    if (int r = ensure_transposed()) return r;
  for (int i = 0; i < n; ++i) {
    if (marks) SPX_HIP(hipEventRecord(marks[2 * i], s));
    if (trp) enqueue_fourier_transposed(sh, W, Wt, ampt, ca, s);
    else enqueue_fourier(sh, kColProject, W, W, ca, s, sh.t4);
    if (dm) enqueue_fourier(sh, kColProject, Wf, Wf, caf, s, sh.t4);
    if (marks) SPX_HIP(hipEventRecord(marks[2 * i + 1], s));
    u.gcur = g[cur];
    u.gnew = g[cur ^ 1];
    u.gp = (want_gp && i == n - 1) ? gp : nullptr;
    u.beta = betas ? betas[i] : 0.f;
    // the specialised pass: ER / HIO / RAAR without gp (per-reconstruction
    // tables and idle slots in its TAB form)
    const bool plain = !dm && !u.gp && !plain_off();
    if (!(plain && sh.ex->update_plain(u, sh.t4, s))) sh.ex->update(u, dm, sh.t4, s);
    cur ^= 1;
  }
  if (marks) SPX_HIP(hipEventRecord(marks[2 * n], s));
  SPX_HIP(hipGetLastError());
  if (dm) {
    Rf_version = flags_version;
    Rf_gamma1 = p.gamma1;
  } else {
    Rf_valid = false;
  }
  return 0;
}

void Engine::drop_graphs() {
  for (auto& x : gexec)
    if (x) {
      (void)hipGraphExecDestroy(x);
      x = nullptr;
    }
  gvalid = false;
}

int Engine::iterate_graph(const IterParams& p, float beta, int n) {
  if (any_mode) {
    std::vector<float> b((size_t)std::max(n, 1), beta);
    return iterate_impl(p, b.data(), n, false, nullptr);
  }
  if (n < 0) return 0;  // n == 0 captures without running
  if (int r = ensure_R()) return r;
  if (p.alg == kDiffMap)
    if (int r = ensure_Rf(p.gamma1)) return r;
  // allocations and the amplitudes' transpose stay out of the capture
  if (sh.tr)
    if (int r = ensure_transposed()) return r;
  const bool same = gvalid && gkey.alg == p.alg && gkey.constraints == p.constraints &&
                    gkey.objective == p.objective && gkey.gamma1 == p.gamma1 && gkey.gamma2 == p.gamma2 &&
                    gbeta == beta;
  if (!same) {
    drop_graphs();
    const int cur0 = cur;
    for (int parity = 0; parity < 2; ++parity) {
      cur = parity;
      hipGraph_t graph = nullptr;
      SPX_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      const int rc = iterate_impl(p, &beta, 1, false, nullptr);
      const hipError_t ec = hipStreamEndCapture(s, &graph);
      if (rc || ec != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        cur = cur0;
        set_error("graph capture failed");
        return rc ? rc : -20;
      }
      const hipError_t ei = hipGraphInstantiate(&gexec[parity], graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (ei != hipSuccess) {
        cur = cur0;
        set_error("graph instantiation failed: %s", hipGetErrorString(ei));
        return -20;
      }
    }
    cur = cur0;
    gkey = p;
    gbeta = beta;
    gvalid = true;
  }
  for (int i = 0; i < n; ++i) {
    SPX_HIP(hipGraphLaunch(gexec[cur], s));
    cur ^= 1;
  }
  if (n > 0 && p.alg != kDiffMap) Rf_valid = false;
  return 0;
}

int Engine::time_iterations(const IterParams& p, const float* betas, int n, float* ms) {
  if (int r = ensure_R()) return r;
  if (p.alg == kDiffMap)
    if (int r = ensure_Rf(p.gamma1)) return r;
  SPX_HIP(hipEventRecord(ev0, s));
  if (int r = iterate(p, betas, n, false)) return r;
  SPX_HIP(hipEventRecord(ev1, s));
  SPX_HIP(hipEventSynchronize(ev1));
  SPX_HIP(hipEventElapsedTime(ms, ev0, ev1));
  return 0;
}

int Engine::time_passes(const IterParams& p, const float* betas, int n, float* ms_fourier, float* ms_real) {
  if (n <= 0) return 0;
  if (int r = ensure_R()) return r;
  if (p.alg == kDiffMap)
    if (int r = ensure_Rf(p.gamma1)) return r;
  std::vector<hipEvent_t> ev((size_t)(2 * n + 1), nullptr);
  int rc = 0;
  for (auto& e : ev)
    if (hipEventCreate(&e) != hipSuccess) rc = -20;
  if (!rc) rc = iterate_impl(p, betas, n, false, ev.data());
  if (!rc && hipEventSynchronize(ev[2 * n]) != hipSuccess) rc = -20;
  double f = 0.0, r = 0.0;
  for (int i = 0; i < n && !rc; ++i) {
    float a = 0.f, b = 0.f;
    if (hipEventElapsedTime(&a, ev[2 * i], ev[2 * i + 1]) != hipSuccess ||
        hipEventElapsedTime(&b, ev[2 * i + 1], ev[2 * i + 2]) != hipSuccess)
      rc = -20;
    f += a;
    r += b;
  }
  for (auto& e : ev)
    if (e) (void)hipEventDestroy(e);
  if (rc) {
    set_error("event timing failed");
    return rc;
  }
  *ms_fourier = (float)f;
  *ms_real = (float)r;
  return 0;
}

int Engine::efourier(double* out, float2* fout) {
  if (any_mode) {
    const int bpr = rec_blocks(N);
    if (int rc = fft_full(sh, g[cur], S1, -1)) return rc;
    if (fout) SPX_HIP(hipMemcpyAsync(fout, S1, sizeof(float2) * (size_t)N, hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(k_efourier_any, dim3(bpr, 1), dim3(kRedThreads), 0, s, S1, ampx, N, partials);
    SPX_HIP(hipGetLastError());
    return reduce_per_rec(bpr, 4, out);
  }
  if (int r = ensure_R()) return r;
  ColArgs ca{};
  ca.ampx = ampx;
  ca.partials = partials;
  ca.fout = fout;
  // the 3-D path needs somewhere to put the y-transformed lines: S1
  enqueue_fourier(sh, kColEFourier, W, S1, ca, s, sh.t4);
  SPX_HIP(hipGetLastError());
  return reduce_per_rec(last_axis_tiles_per_rec(sh), 4, out);
}

int Engine::efourier_enqueue() {
  if (int r = ensure_R()) return r;
  if (!red_host) SPX_HIP(hipHostMalloc((void**)&red_host, sizeof(double) * 4 * (size_t)nrec, hipHostMallocDefault));
  if (any_mode) {
    const int bpr = rec_blocks(N);
    if (int rc = fft_full(sh, g[cur], S1, -1)) return rc;
    hipLaunchKernelGGL(k_efourier_any, dim3(bpr, 1), dim3(kRedThreads), 0, s, S1, ampx, N, partials);
    hipLaunchKernelGGL(k_reduce_rec, dim3(1), dim3(256), 0, s, partials, bpr, 4, red);
    SPX_HIP(hipGetLastError());
    SPX_HIP(hipMemcpyAsync(red_host, red, sizeof(double) * 4, hipMemcpyDeviceToHost, s));
    red_host_pending = true;
    return 0;
  }
  ColArgs ca{};
  ca.ampx = ampx;
  ca.partials = partials;
  enqueue_fourier(sh, kColEFourier, W, S1, ca, s, sh.t4);
  SPX_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_reduce_rec, dim3(nrec), dim3(256), 0, s, partials, last_axis_tiles_per_rec(sh), 4, red);
  SPX_HIP(hipGetLastError());
  SPX_HIP(hipMemcpyAsync(red_host, red, sizeof(double) * 4 * (size_t)nrec, hipMemcpyDeviceToHost, s));
  red_host_pending = true;
  return 0;
}

int Engine::efourier_fetch(double* out) {
  if (!red_host_pending) {
    set_error("no Fourier error was enqueued");
    return -1;
  }
  SPX_HIP(hipStreamSynchronize(s));
  memcpy(out, red_host, sizeof(double) * 4 * (size_t)nrec);
  red_host_pending = false;
  return 0;
}

int Engine::fmodel(float2* fout) {
  std::vector<double> dummy((size_t)4 * nrec);
  if (!sh.t4) return efourier(dummy.data(), fout);
  // the transform comes out in the tiled layout: untile it into `fout`
  if (int r = efourier(dummy.data(), S1)) return r;
  hipLaunchKernelGGL(k_untile_c64, ew_grid(NT), dim3(256), 0, s, S1, fout, NT, nx, ny);
  SPX_HIP(hipGetLastError());
  return 0;
}

int Engine::ereal(double* out) {
  const int bpr = rec_blocks(N);
  hipLaunchKernelGGL(k_ereal, dim3(bpr, nrec), dim3(kRedThreads), 0, s, gp, flags, N, partials);
  SPX_HIP(hipGetLastError());
  return reduce_per_rec(bpr, 2, out);
}

int Engine::support_count(double* out) {
  const int bpr = rec_blocks(N);
  hipLaunchKernelGGL(k_count_support, dim3(bpr, nrec), dim3(kRedThreads), 0, s, flags, N, partials);
  SPX_HIP(hipGetLastError());
  return reduce_per_rec(bpr, 1, out);
}

int Engine::model_change(float2* out) {
  hipLaunchKernelGGL(k_sub, ew_grid(NT), dim3(256), 0, s, g[cur], g[cur ^ 1], out, NT);
  SPX_HIP(hipGetLastError());
  return 0;
}

// |blur(|g1|, sigma)| into absb, with the level-1 histograms and the running
// maxima (reference src/support_update.c:138-141,169-172)
int Engine::blur_to_abs(float sigma, bool with_hist) {
  if (any_mode) {
    hipLaunchKernelGGL(k_abs_c64, ew_grid(N), dim3(256), 0, s, g[cur], S1, N);
    if (int rc = fft_full(sh, S1, S1, -1)) return rc;
    hipLaunchKernelGGL(k_blur_transfer, ew_grid(N), dim3(256), 0, s, S1, nx, ny, nz, sigma);
    if (int rc = fft_full(sh, S1, S1, +1)) return rc;
    hipLaunchKernelGGL(k_blur_abs_any, ew_grid(N), dim3(256), 0, s, S1, absb, with_hist ? hist : nullptr,
                       with_hist ? nullptr : ctl, N);
    SPX_HIP(hipGetLastError());
    return 0;
  }
  sh.ex->row(kRowAbsFwd, row_args(sh, g[cur], S1), s);
  ColArgs cb{};
  cb.nx = nx;
  cb.ny = ny;
  cb.nz = nz;
  cb.sigma = sigma;
  enqueue_fourier(sh, kColBlur, S1, S1, cb, s, false);
  RowArgs r = row_args(sh, S1, nullptr);
  r.absb = absb;
  r.hist = with_hist ? hist : nullptr;
  r.ctl = with_hist ? nullptr : ctl;  // the maximum serves the threshold update only
  sh.ex->row(kRowBlurOut, r, s);
  SPX_HIP(hipGetLastError());
  return 0;
}

int Engine::shrinkwrap_area(float sigma, float area) {
  // k = (int)(N * area) with N converted to float first (src/support_update.c:144)
  long k = (long)(int)((float)N * area);
  if (k >= N) k = N - 1;
  if (k < 0) k = 0;
  const int bpr = rec_blocks(N);
  const long nh = (long)kHistStride * nrec;
  hipLaunchKernelGGL(k_fill_u32, ew_grid(nh), dim3(256), 0, s, hist, nh, 0u);
  hipLaunchKernelGGL(k_ctl_init, dim3(1), dim3(256), 0, s, ctl, nrec, (unsigned)k);
  if (int rc = blur_to_abs(sigma, true)) return rc;
  hipLaunchKernelGGL(k_select, dim3(nrec), dim3(1024), 0, s, hist, 1, ctl);
  hipLaunchKernelGGL(k_hist_level, dim3(bpr, nrec), dim3(256), 0, s, absb, N, hist, ctl, 2);
  hipLaunchKernelGGL(k_select, dim3(nrec), dim3(1024), 0, s, hist, 2, ctl);
  hipLaunchKernelGGL(k_hist_level, dim3(bpr, nrec), dim3(256), 0, s, absb, N, hist, ctl, 3);
  hipLaunchKernelGGL(k_select, dim3(nrec), dim3(1024), 0, s, hist, 3, ctl);
  hipLaunchKernelGGL(k_mask, dim3(bpr, nrec), dim3(256), 0, s, absb, flags, N, ctl, 0, 0.f);
  SPX_HIP(hipGetLastError());
  ++flags_version;
  return 0;
}

int Engine::last_threshold(float* t) {
  std::vector<unsigned> c((size_t)kCtlStride * nrec);
  SPX_HIP(hipMemcpyAsync(c.data(), ctl, sizeof(unsigned) * c.size(), hipMemcpyDeviceToHost, s));
  SPX_HIP(hipStreamSynchronize(s));
  for (int r = 0; r < nrec; ++r) memcpy(&t[r], &c[(size_t)r * kCtlStride + kCtlThresh], sizeof(float));
  return 0;
}

int Engine::shrinkwrap_threshold(float sigma, float rel) {
  hipLaunchKernelGGL(k_ctl_init, dim3(1), dim3(256), 0, s, ctl, nrec, 0u);
  if (int rc = blur_to_abs(sigma, false)) return rc;
  hipLaunchKernelGGL(k_mask, dim3(rec_blocks(N), nrec), dim3(256), 0, s, absb, flags, N, ctl, 1, rel);
  SPX_HIP(hipGetLastError());
  ++flags_version;
  return 0;
}

// The passes ping-pong between the flags and the byte view of `absb`
// (4 NT bytes); 2 size passes end back in the flags.
int Engine::support_close(int size) {
  if (size <= 0) return 0;
  uint8_t* tmp = reinterpret_cast<uint8_t*>(absb);
  uint8_t *src = flags, *dst = tmp;
  for (int k = 0; k < 2 * size; ++k) {
    hipLaunchKernelGGL(k_close_pass, ew_grid(NT), dim3(256), 0, s, src, dst, nx, ny, nz, N, NT, k < size ? 1 : 0);
    std::swap(src, dst);
  }
  SPX_HIP(hipGetLastError());
  ++flags_version;
  return 0;
}

// The centre-image update on the device: the middle-line scans, the
// support moved by them into the byte view of `absb`, its centre of mass,
// and the flags' support bits replaced by that copy moved to the origin.
int Engine::support_centre_image() {
  if (nx > 65535 || ny > 65535) {
    set_error("centre image: axis lengths above 65535");
    return -1;
  }
  uint8_t* tmp = reinterpret_cast<uint8_t*>(absb);
  hipLaunchKernelGGL(k_centre_offsets, dim3(nrec), dim3(256), 0, s, flags, nx, ny, nz, N, ctl);
  hipLaunchKernelGGL(k_centre_move, ew_grid(NT), dim3(256), 0, s, flags, tmp, nx, ny, nz, N, NT, ctl, 0);
  hipLaunchKernelGGL(k_centre_mass, dim3(nrec), dim3(256), 0, s, tmp, nx, ny, nz, N, ctl);
  hipLaunchKernelGGL(k_centre_move, ew_grid(NT), dim3(256), 0, s, tmp, flags, nx, ny, nz, N, NT, ctl, 1);
  SPX_HIP(hipGetLastError());
  ++flags_version;
  return 0;
}

int Engine::support_template(unsigned long long key, const float* tmpl_abs, float t) {
  float* d = nullptr;
  for (auto& e : templates)
    if (e.first == key) d = e.second;
  if (!d) {
    SPX_HIP(hipMalloc(&d, sizeof(float) * N));
    templates.emplace_back(key, d);
    SPX_HIP(hipMemcpyAsync(d, tmpl_abs, sizeof(float) * N, hipMemcpyHostToDevice, s));
    SPX_HIP(hipStreamSynchronize(s));
  }
  hipLaunchKernelGGL(k_template_mask, ew_grid(NT), dim3(256), 0, s, d, flags, N, NT, t);
  SPX_HIP(hipGetLastError());
  ++flags_version;
  return 0;
}

// support from the thresholded Patterson map |IFFT(amp^2)| > max * value
// (reference src/phasing.c:744-763)
int Engine::patterson_support(const float* amp_raw, float value) {
  SPX_HIP(hipMemcpyAsync(absb, amp_raw, sizeof(float) * NT, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_amp2, ew_grid(NT), dim3(256), 0, s, absb, S1, NT);
  if (any_mode) {
    if (int rc = fft_full(sh, S1, S1, +1)) return rc;
    hipLaunchKernelGGL(k_ctl_init, dim3(1), dim3(256), 0, s, ctl, nrec, 0u);
    hipLaunchKernelGGL(k_blur_abs_any, ew_grid(N), dim3(256), 0, s, S1, absb, nullptr, ctl, N);
    hipLaunchKernelGGL(k_mask, dim3(rec_blocks(N), nrec), dim3(256), 0, s, absb, flags, N, ctl, 1, value);
    SPX_HIP(hipGetLastError());
    SPX_HIP(hipStreamSynchronize(s));
    ++flags_version;
    return 0;
  }
  ColArgs c{};
  if (sh.ey) {
    c.in = S1;
    c.W = S1;
    c.geo = sh.gy;
    sh.ey->col(kColInv, c, s);
  }
  if (sh.ez) {
    c.in = S1;
    c.W = S1;
    c.geo = sh.gz;
    sh.ez->col(kColInv, c, s);
  }
  hipLaunchKernelGGL(k_ctl_init, dim3(1), dim3(256), 0, s, ctl, nrec, 0u);
  RowArgs r = row_args(sh, S1, nullptr);
  r.absb = absb;
  r.ctl = ctl;
  sh.ex->row(kRowBlurOut, r, s);
  hipLaunchKernelGGL(k_mask, dim3(rec_blocks(N), nrec), dim3(256), 0, s, absb, flags, N, ctl, 1, value);
  SPX_HIP(hipGetLastError());
  SPX_HIP(hipStreamSynchronize(s));
  ++flags_version;
  return 0;
}

// ---------------------------------------------------------------------------
// one-shot transforms for the Image-level API
// ---------------------------------------------------------------------------
static std::mutex g_oneshot_mu;

static int oneshot(const float2* in, float2* out, int nx, int ny, int nz, int dir, float sigma, bool is_blur) {
  std::lock_guard<std::mutex> lock(g_oneshot_mu);
  if (!device_available()) {
    set_error("no HIP device available");
    return -10;
  }
  const int path = any_path(nx, ny, nz);
  if (path < 0) return path;
  if (is_blur && (nx < 2 || (ny < 2 && nz < 2))) {
    set_error("blur needs a 2-D or 3-D array with nx > 1");
    return -11;
  }
  Shape sh;
  if (path == 0)
    if (int rc = make_shape(nx, ny, nz, 1, &sh)) return rc;
  static hipStream_t s = nullptr;
  if (!s) SPX_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  const size_t bytes = sizeof(float2) * (size_t)nx * ny * nz;
  float2* d = nullptr;
  float2* sc = nullptr;
  SPX_HIP(hipMalloc(&d, bytes));
  if (path == 1 && hipMalloc(&sc, sizeof(float2) * (size_t)fft_any_scratch(nx, ny, nz)) != hipSuccess) {
    (void)hipFree(d);
    set_error("out of device memory for the transform scratch");
    return -20;
  }
  int rc = 0;
  if (hipMemcpyAsync(d, in, bytes, hipMemcpyHostToDevice, s) != hipSuccess) rc = -20;
  if (!rc && path == 1) {
    // any length: transform, the transfer function element-wise, inverse
    rc = enqueue_fft_any(d, d, nx, ny, nz, is_blur ? -1 : dir, sc, s);
    if (!rc && is_blur) {
      hipLaunchKernelGGL(k_blur_transfer, ew_grid((long)nx * ny * nz), dim3(256), 0, s, d, nx, ny, nz, sigma);
      rc = enqueue_fft_any(d, d, nx, ny, nz, +1, sc, s);
    }
  } else if (!rc) {
    if (!is_blur) {
      enqueue_fft(sh, d, d, dir, s);
    } else {
      sh.ex->row(kRowFwd, row_args(sh, d, d), s);
      ColArgs cb{};
      cb.nx = nx;
      cb.ny = ny;
      cb.nz = nz;
      cb.sigma = sigma;
      enqueue_fourier(sh, kColBlur, d, d, cb, s, false);
      sh.ex->row(kRowInv, row_args(sh, d, d), s);
    }
  }
  if (!rc && hipGetLastError() != hipSuccess) rc = -20;
  if (!rc && hipMemcpyAsync(out, d, bytes, hipMemcpyDeviceToHost, s) != hipSuccess) rc = -20;
  if (hipStreamSynchronize(s) != hipSuccess && !rc) rc = -20;
  (void)hipFree(d);
  if (sc) (void)hipFree(sc);
  if (rc == -20) set_error("HIP error in one-shot transform");
  return rc;
}

int fft_host(const float2* in, float2* out, int nx, int ny, int nz, int dir) {
  return oneshot(in, out, nx, ny, nz, dir, 0.f, false);
}

int blur_host(const float2* in, float2* out, int nx, int ny, int nz, float sigma) {
  return oneshot(in, out, nx, ny, nz, 0, sigma, true);
}

int fft_device(const float2* in, float2* out, int nx, int ny, int nz, int dir) {
  std::lock_guard<std::mutex> lock(g_oneshot_mu);
  if (!device_available()) {
    set_error("no HIP device available");
    return -10;
  }
  const int path = any_path(nx, ny, nz);
  if (path < 0) return path;
  static hipStream_t s = nullptr;
  if (!s) SPX_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  if (path == 1) {
    float2* sc = nullptr;
    SPX_HIP(hipMalloc(&sc, sizeof(float2) * (size_t)fft_any_scratch(nx, ny, nz)));
    int rc = enqueue_fft_any(in, out, nx, ny, nz, dir, sc, s);
    if (hipStreamSynchronize(s) != hipSuccess && !rc) rc = -20;
    (void)hipFree(sc);
    if (rc) set_error("HIP error in the transform");
    return rc;
  }
  Shape sh;
  if (int rc = make_shape(nx, ny, nz, 1, &sh)) return rc;
  enqueue_fft(sh, in, out, dir, s);
  SPX_HIP(hipGetLastError());
  SPX_HIP(hipStreamSynchronize(s));
  return 0;
}


// ---------------------------------------------------------------------------
// transform that allocates nothing (no counterpart in the reference)
// ---------------------------------------------------------------------------
// Callers that may not allocate device memory stage their array through
// this fixed array instead of the per-call buffer of fft_host: kFixedElems
// complex values of data followed by three twiddle tables (one per axis) of
// up to 4096 entries.  Every device the library runs on holds its own copy
// for the life of the process; calls on one device share it, so calls that
// overlap on one device race.
static __device__ __align__(256) float2 scratch[kFixedElems + 3 * 4096];

namespace {
struct FixedDev {
  float2* ptr = nullptr;  // device address of `scratch` on this device
  int twL[3] = {0, 0, 0}, twE[3] = {0, 0, 0};  // tables resident in the twiddle slots
};
std::mutex g_fixed_mu;
FixedDev g_fixed[64];
float2 g_fixed_tw_host[4096];
}  // namespace

// The shape's passes with the default variants and no tuning (the tuner
// times candidate variants on buffers it allocates).
static int make_shape_fixed(int nx, int ny, int nz, Shape* sh) {
  *sh = Shape{};
  sh->nx = nx;
  sh->ny = ny;
  sh->nz = nz;
  sh->nrec = 1;
  const int dims[3] = {nx, ny, nz};
  AxisOps* ax[3] = {&sh->ex, &sh->ey, &sh->ez};
  for (int a = 0; a < 3; ++a) {
    if (dims[a] == 1) continue;
    const LengthEntry* e = find_length(dims[a]);
    // the fixed array keeps 4096 twiddles per axis: plans of longer lines
    // need more
    if (!e || dims[a] > 4096) {
      set_error("axis length %d is not supported by the fixed-array transform (products of 2, 3, 5 and 7 up to "
                "4096)", dims[a]);
      return -11;
    }
    ax[a]->e = e;
    ax[a]->v = default_choice(e, false);
  }
  sh->gy = LineGeom{nx, nz, (long)nx * ny, (long)nx};
  sh->gz = LineGeom{nx * ny, 1, (long)nx * ny * nz, (long)nx * ny};
  return 0;
}

int fft_fixed(const float2* in, float2* out, int nx, int ny, int nz, int dir) {
  if (nx < 1 || ny < 1 || nz < 1) {
    set_error("invalid shape %d x %d x %d", nx, ny, nz);
    return -11;
  }
  const long n = (long)nx * ny * nz;
  if (n > kFixedElems) {
    set_error("%d x %d x %d = %ld elements exceed the %ld of the fixed device array", nx, ny, nz, n,
              (long)kFixedElems);
    return -12;
  }
  std::lock_guard<std::mutex> lock(g_fixed_mu);
  if (!device_available()) {
    set_error("no HIP device available");
    return -10;
  }
  Shape sh;
  if (int rc = make_shape_fixed(nx, ny, nz, &sh)) return rc;
  int dev = 0;
  SPX_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) {
    set_error("device %d out of range", dev);
    return -20;
  }
  FixedDev& fd = g_fixed[dev];
  if (!fd.ptr) SPX_HIP(hipGetSymbolAddress((void**)&fd.ptr, HIP_SYMBOL(scratch)));
  float2* data = fd.ptr;
  const hipStream_t s = nullptr;  // the null stream: no stream to create
  // twiddle tables of the axes' plans in the three slots after the data
  const float2* tw[3] = {nullptr, nullptr, nullptr};
  const AxisOps* ax[3] = {&sh.ex, &sh.ey, &sh.ez};
  for (int a = 0; a < 3; ++a) {
    if (!*ax[a]) continue;
    const int L = ax[a]->e->L, E = ax[a]->var().E;
    float2* slot = fd.ptr + kFixedElems + 4096L * a;
    tw[a] = slot;
    if (fd.twL[a] == L && fd.twE[a] == E) continue;
    const FftPlan p = make_plan(L, E);
    for (int st = 1; st < p.n; ++st) {
      const int R = p.r[st], Ns = p.ns[st];
      for (int j = 0; j < Ns; ++j)
        for (int k = 1; k < R; ++k) {
          const double an = -2.0 * M_PI * (double)j * (double)k / (double)(Ns * R);
          g_fixed_tw_host[p.twoff[st] + (k - 1) * Ns + j] = make_float2((float)cos(an), (float)sin(an));
        }
    }
    if (p.twtotal > 0)
      SPX_HIP(hipMemcpyAsync(slot, g_fixed_tw_host, sizeof(float2) * p.twtotal, hipMemcpyHostToDevice, s));
    SPX_HIP(hipStreamSynchronize(s));  // the host table is reused for the next axis
    fd.twL[a] = L;
    fd.twE[a] = E;
  }
  SPX_HIP(hipMemcpyAsync(data, in, sizeof(float2) * n, hipMemcpyHostToDevice, s));
  if (sh.ex) {
    RowArgs r = row_args(sh, data, data);
    r.tw = tw[0];
    sh.ex->row(dir < 0 ? kRowFwd : kRowInv, r, s);
  }
  ColArgs c{};
  c.in = data;
  c.W = data;
  if (sh.ey) {
    c.geo = sh.gy;
    c.tw = tw[1];
    sh.ey->col(dir < 0 ? kColFwd : kColInv, c, s);
  }
  if (sh.ez) {
    c.geo = sh.gz;
    c.tw = tw[2];
    sh.ez->col(dir < 0 ? kColFwd : kColInv, c, s);
  }
  SPX_HIP(hipGetLastError());
  SPX_HIP(hipMemcpyAsync(out, data, sizeof(float2) * n, hipMemcpyDeviceToHost, s));
  SPX_HIP(hipStreamSynchronize(s));
  return 0;
}

}  // namespace spx
