// Pass kernels of the HIP phasing engine.
//
// Every array is (nz, ny, nx) row-major with x fastest, like the reference's
// sp_c3matrix (include/spimage/linear_alg.h:436-438).  One work-group owns
// whole lines along one axis, reads them once into LDS, transforms them there
// and writes them back, so every pass may run in place.
//
//   row passes    lines along x (contiguous), M lines per work-group
//   column passes lines along y or z (strided), C adjacent lines per
//                 work-group so that every global access moves C*8
//                 contiguous bytes
//
// The Fourier-space work of an iteration (constraints + modulus projection,
// reference src/phasing.c:1061-1109) runs in the column pass of the last axis
// between its forward and inverse transforms; the real-space update (ER/HIO/
// RAAR/DM, src/phasing.c:1111-1258), the 1/N scale and the support
// constraints (src/phasing.c:1026-1059) run in the row pass between the
// inverse x-transform of this iteration and the forward x-transform of the
// next one.
#pragma once
#include "fft_core.h"

namespace spx {

// pixel flag bits, same values as SpPhasingPixelFlags (reference
// include/spimage/phasing.h:19)
constexpr uint8_t kInsideSupport = 1;
constexpr uint8_t kMeasured = 2;

// Device amplitude encoding: unmeasured pixels carry this NaN pattern so the
// projection needs no separate flag read.  A measured NaN amplitude is stored
// as the canonical quiet NaN, whose pattern differs.
constexpr uint32_t kUnmeasuredBits = 0xFFFFFFFFu;

// Phasing algorithm ids, same values as SpPhasingAlgorithmType
// (reference include/spimage/phasing.h:16).
enum : int { kHIO = 1, kRAAR = 2, kDiffMap = 3, kER = 4 };
// SpPhasingConstraints bits (reference include/spimage/phasing.h:17)
enum : int {
  kRealObject = 1, kPositiveRealObject = 2, kPositiveComplexObject = 4,
  kPositivityFlipping = 8, kCentrosymmetricObject = 16, kAmplitudeErrorMargin = 32
};

// In 2-D iterations the row-transformed iterate W crosses XCDs between the
// two passes (rows are spread over XCDs in one pass, columns in the other).
// A single reconstruction's passes load (nt bit 0) and store (bit 1) it with
// the non-temporal hint: it streams through L2 and leaves fewer dirty lines
// for the end-of-kernel write-back.  The iterates g_k and g_{k+1} of the
// real-space pass may stream too (bits 2 and 3).  Which streams gain depends
// on the shape and the batch size (batches lose with W streamed: partial
// lines of neighbouring column groups are then not merged); the tuner
// decides per shape and batch class.  3-D passes use ordinary accesses.
typedef float spx_v2f __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void st_nt(float2* p, float2 v) {
  const spx_v2f x = {v.x, v.y};
  __builtin_nontemporal_store(x, reinterpret_cast<spx_v2f*>(p));
}
__device__ __forceinline__ float2 ld_nt(const float2* p) {
  const spx_v2f x = __builtin_nontemporal_load(reinterpret_cast<const spx_v2f*>(p));
  return make_float2(x.x, x.y);
}
__device__ __forceinline__ float2 ld_line(const float2* p, int nt) { return (nt & 1) ? ld_nt(p) : *p; }
__device__ __forceinline__ void st_line(float2* p, float2 v, int nt) {
  if (nt & 2) st_nt(p, v);
  else *p = v;
}

// Map hardware work-group order (round-robin over the 8 XCDs) onto tiles so
// that each XCD receives a contiguous range: neighbouring column groups share
// 128-byte lines and then share one L2.
__device__ __forceinline__ int xcd_tile(int bid, int nb) {
  const int q = nb >> 3, r = nb & 7;
  const int x = bid & 7, k = bid >> 3;
  return x * q + (x < r ? x : r) + k;
}

// ---------------------------------------------------------------------------
// Argument blocks (passed by value).
// ---------------------------------------------------------------------------
struct LineGeom {
  // line element i of line (outer, c) lives at
  //   outer*outer_stride + (c >> tshift)*tstride + (c & (2^tshift - 1)) + i*axis_stride
  // natural layout: tshift large (one tile); 4-column tiled layout: tshift 2
  int ncontig = 0;       // lines of one outer index (x extent for column passes)
  int nouter = 0;        // number of outer indices
  long outer_stride = 0;
  long axis_stride = 0;
  int tshift = 30;
  long tstride = 0;
};

// 4-column tiled layout of a 2-D array with `ny` rows of L elements (one
// reconstruction after another): element (y, x) sits at
// (x / 4) * 4 ny + 4 y + x % 4, so a column pass reads each group of 4
// columns as one contiguous block and a row pass that holds 4 consecutive
// rows writes whole 128-byte lines.  The engine keeps its Fourier-space
// arrays (the row-transformed iterate, amplitudes, margins, phased
// amplitudes) in this layout for 2-D shapes with nx % 4 == 0.
__device__ __forceinline__ long t4_offset(int x, int ny) { return (long)(x >> 2) * (4L * ny) + (x & 3); }

enum RowMode : int {
  kRowFwd = 0,        // out = FFT_x(in)
  kRowInv = 1,        // out = IFFT_x(in)
  kRowAbsFwd = 2,     // out = FFT_x(|in|)         (blur input, src/support_update.c:139-141)
  kRowBlurOut = 3,    // absb = |IFFT_x(in)| plus max / level-1 histogram
};

// Per-reconstruction select scratch: 3 histograms of 2048 bins and a small
// control block (k, prefix, threshold bits, max bits) for each reconstruction.
// Per-reconstruction select scratch.  The level-1 histogram (top 11 bits of
// |blur|) is counted by a row pass with thousands of work-groups, whose
// global atomics would all meet on the few bins the values fall in; it is
// kept in kHistSlices copies, one per XCD (work-group x lands on XCD x % 8),
// which the select sums.  Levels 2 and 3 follow.  The running maximum has
// one slot per XCD in the same way.
constexpr int kHistSlices = 8;
constexpr int kHistLevel2 = kHistSlices * 2048, kHistLevel3 = kHistLevel2 + 2048;
constexpr int kHistStride = kHistLevel3 + 2048;
constexpr int kCtlStride = 16;
// The centre-image update keeps its offsets there too: the middle-line
// translation (x, y, z in slots 3..5) and the centre-of-mass translation
// (x | y << 16 in slot 6, z in slot 7), each wrapped into [0, n).
enum CtlSlot {
  kCtlK = 0, kCtlPrefix = 1, kCtlThresh = 2, kCtlCentre = 3 /* .. 5 */, kCtlCentreShift = 6 /* .. 7 */,
  kCtlMax = 8 /* .. 15, one per XCD */
};

struct RowArgs {
  const float2* in;
  float2* out;
  float* absb;             // kRowBlurOut: |blur| output (f32)
  unsigned* hist;          // kRowBlurOut: per-reconstruction histograms, or null
  unsigned* ctl;           // kRowBlurOut: per-reconstruction control (running max), or null
  const float2* tw;
  int nlines;
  int lines_per_rec;       // lines of one reconstruction (ny * nz)
  int L;
};

// States of a batch slot in the per-reconstruction activity table
// (spx_batch_set_active): active slots iterate; idle ones keep their state
// unchanged across iterations; vacant ones keep nothing (their passes do
// no work) until the slot is set again.
constexpr uint8_t kSlotIdle = 0, kSlotActive = 1, kSlotVacant = 2;

struct UpdateArgs {
  float2* W;               // in: column-pass output C; out: FFT_x of the new iterate
  float2* Wf;              // diff map: same for f1
  const float2* gcur;      // iterate g_k (the reference's g0 in this iteration)
  float2* gnew;            // iterate g_{k+1}
  float2* gp;              // model before projection, or null
  const uint8_t* flags;
  const float2* tw;
  const float* beta_rec;   // per-reconstruction beta, or null for `beta`
  const int* alg_rec;      // per-reconstruction algorithm, or null for `alg`
  const uint8_t* rec_active;  // per-reconstruction slot state (kSlot*), or null for all active
  int nlines;
  int lines_per_rec;
  int alg;
  int constraints;
  float beta;
  float gamma1, gamma2;
  float scale;             // 1/N as float
  float nsize;             // N as float (diff map divides by it)
  int nt;                  // non-temporal hint: W / Wf loads (bit 0) and stores (bit 1), g_k loads
                           // (bit 2), g_{k+1} stores (bit 3)
};

enum ColMode : int {
  kColFwd = 0,            // in place FFT along the axis
  kColInv = 1,            // in place IFFT along the axis
  kColProject = 2,        // FFT, Fourier constraints + projection, IFFT
  kColEFourier = 3,       // FFT, error-metric partial sums, optional write
  kColBlur = 4,           // FFT, Gaussian transfer function, IFFT
};

struct ColArgs {
  const float2* in;          // lines are read from here ...
  float2* W;                 // ... and written here (may alias `in`)
  const float2* tw;
  LineGeom geo;
  // projection
  const float* ampx;         // encoded amplitudes (kUnmeasuredBits for unmeasured)
  const float* amp_min;
  const float* amp_max;
  const float2* phased_amp;  // SpRecoverAmplitudes objective (unmeasured: kUnmeasuredBits in .x)
  int centro;                // apply the centrosymmetric Fourier constraint
  int margin;                // SpAmplitudeErrorMargin
  int objective;             // 0 phases, 1 amplitudes
  // efourier
  double* partials;          // 4 doubles per work-group
  float2* fout;              // optional copy of the transform (fmodel)
  // blur
  int nx, ny, nz;
  float sigma;
  int nt;                    // line data loaded (bit 0) and stored (bit 1) with the non-temporal hint
  // per-reconstruction slot state (kSlot*; the work-groups of idle and
  // vacant slots return at once), or null; outer index / outer_per_rec is
  // the reconstruction
  const uint8_t* rec_active;
  int outer_per_rec;
};

// ---------------------------------------------------------------------------
// Shared pieces.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float cabsf_ref(float2 a) {
  // sp_cabs: sqrt(im*im + re*re) evaluated in float (linear_alg.h:163-165)
  return sqrtf(a.y * a.y + a.x * a.x);
}

// Modulus projection of one Fourier-space pixel (src/phasing.c:1070-1100).
__device__ __forceinline__ float2 project_pixel(float2 g, float ampbits_f, const ColArgs& a, long idx) {
  const uint32_t bits = __float_as_uint(ampbits_f);
  if (bits == kUnmeasuredBits) return g;
  const float amp = ampbits_f;
  float m = 1.f;
  if (!a.margin) {
    m = amp / cabsf_ref(g);
  } else {
    const float aa = cabsf_ref(g);
    const float lo = a.amp_min[idx], hi = a.amp_max[idx];
    if (aa < lo) m = lo / aa;
    else if (aa > hi) m = hi / aa;
  }
  if (isfinite(m)) return make_float2(g.x * m, g.y * m);
  return make_float2(amp, 0.f);
}

// The modulus projection of project_pixel without margins, branch-free,
// after the centrosymmetric constraint (src/phasing.c:1061-1100).
// m = amp / |G| is computed as amp * rsqrt(|G|^2): 1-2 ulp from the
// correctly rounded quotient, a third of its instructions.  A denormal
// |G|^2 is scaled by 2^64 first (exactly), as the reciprocal square root
// flushes denormal inputs; |G|^2 = 0 still gives an infinite m.
__device__ __forceinline__ float2 project_phase(float2 G, float amp, bool centro) {
  G.y = centro ? 0.f : G.y;
  const float sq = G.y * G.y + G.x * G.x;
  const bool tiny = sq < 1.17549435e-38f;
  const float m = amp * rsqrtf(tiny ? sq * 0x1p64f : sq) * (tiny ? 0x1p32f : 1.f);
  const bool fin = isfinite(m);
  const float2 P = make_float2(fin ? G.x * m : amp, fin ? G.y * m : 0.f);
  const bool unm = __float_as_uint(amp) == kUnmeasuredBits;
  return make_float2(unm ? G.x : P.x, unm ? G.y : P.y);
}

// Real-space support constraints (src/phasing.c:1026-1059), inside support only.
__device__ __forceinline__ float2 apply_constraints(float2 g, int c) {
  if (c & kRealObject) {
    g.y = 0.f;
  } else if (c & kPositiveRealObject) {
    if (g.x < 0.f) g.x = (c & kPositivityFlipping) ? fabsf(g.x) : 0.f;
    g.y = 0.f;
  } else if (c & kPositiveComplexObject) {
    if (g.x < 0.f) g.x = (c & kPositivityFlipping) ? fabsf(g.x) : 0.f;
    if (g.y < 0.f) g.y = (c & kPositivityFlipping) ? fabsf(g.y) : 0.f;
  }
  return g;
}

// apply_constraints with the choice among the constraint sets made once per
// thread (the flags are uniform), so the per-pixel code has no branches
struct SupportConstraints {
  bool zero_im, clamp_re, clamp_im, flip;
  __device__ explicit SupportConstraints(int c) {
    const bool real = c & kRealObject;
    const bool posr = !real && (c & kPositiveRealObject);
    const bool posc = !real && !posr && (c & kPositiveComplexObject);
    zero_im = real || posr;
    clamp_re = posr || posc;
    clamp_im = posc;
    flip = c & kPositivityFlipping;
  }
  __device__ __forceinline__ float2 apply(float2 g) const {
    const float x = (clamp_re && g.x < 0.f) ? (flip ? fabsf(g.x) : 0.f) : g.x;
    const float y = zero_im ? 0.f : ((clamp_im && g.y < 0.f) ? (flip ? fabsf(g.y) : 0.f) : g.y);
    return make_float2(x, y);
  }
};

// Block-wide sum of NV doubles; result valid in thread 0.
template <int NV, int NT>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* scratch /* NV*32 */) {
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    double x = v[q];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    v[q] = x;
  }
  constexpr int NW = (NT + 63) / 64;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) scratch[q * 32 + w] = v[q];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      double s = 0.0;
      for (int i = 0; i < NW; ++i) s += scratch[q * 32 + i];
      v[q] = s;
    }
  }
}


// ---------------------------------------------------------------------------
// Row kernels: M lines along x per work-group, T threads per line.
// ---------------------------------------------------------------------------
template <int L, int E>
struct RowShape {
  static constexpr int T = PlanOf<L, E>::T;
  // about 64 threads per work-group, 128 for lines of up to 128 points
  // (3-D volumes up to 128^3): short lines give more, smaller work-groups.
  // Against about 256 threads: 256^2 alone 9.6 -> 8.7 us, 720^2 alone 21.7
  // -> 20.6 us, 720^2 x 16 113 -> 107 us per iteration; 96^3 x 32 prefers
  // 120 threads to 60 (475 against 488 us).
  static constexpr int M = (T >= 64) ? 1 : (L > 128 ? 64 / T : 128 / T);
  static constexpr int NT = T * M;
};

template <int L, int E, int MODE>
__global__ __launch_bounds__((RowShape<L, E>::NT)) void k_row(RowArgs a) {
  using PL = PlanOf<L, E>;
  constexpr int T = RowShape<L, E>::T, M = RowShape<L, E>::M, NT = RowShape<L, E>::NT;
  constexpr int NR = PL::NREG;
  using View = LdsView<L, M, false>;
  __shared__ float2 shA[View::SIZE];
  __shared__ float2 shB[SingleBuf<L>::value ? 1 : View::SIZE];
  __shared__ unsigned hsh[(MODE == kRowBlurOut) ? 2048 : 1];
  __shared__ float wmax[(MODE == kRowBlurOut) ? (NT + 63) / 64 : 1];
  const int tid = threadIdx.x;
  const int g = tid / T, t = tid - g * T;
  const int blk = xcd_tile(blockIdx.x, gridDim.x);
  const int line = blk * M + g;
  const bool act = line < a.nlines;
  const View va{shA, g}, vb{shB, g};
  // idle lanes of a partial last work-group address the last line (loads
  // stay in bounds and need no predicate); only their stores are skipped
  const long base = (long)(act ? line : a.nlines - 1) * L;

  float2 v[NR];
#pragma unroll
  for (int q = 0; q < NR; ++q) {
    float2 x = a.in[base + in_pos<L, E>(t, q)];
    if constexpr (MODE == kRowAbsFwd) x = make_float2(cabsf_ref(x), 0.f);
    v[q] = x;
  }
  TwRegs<L, E> twr;
  load_tw<L, E>(a.tw, t, twr);
  if constexpr (MODE == kRowBlurOut) {
    if (a.hist)
      for (int i = tid; i < 2048; i += NT) hsh[i] = 0u;
  }
  if constexpr (MODE == kRowInv || MODE == kRowBlurOut) fft_line<L, E, +1>(v, va, vb, t, twr);
  else fft_line<L, E, -1>(v, va, vb, t, twr);

  if constexpr (MODE == kRowBlurOut) {
    // the work-group's lines belong to one reconstruction unless lines of
    // several small reconstructions share it; then counts go straight to
    // global memory
    const int rec = line / a.lines_per_rec;
    const int rec_first = (blk * M) / a.lines_per_rec;
    const int last_line = (blk * M + M - 1) < a.nlines ? (blk * M + M - 1) : (a.nlines - 1);
    const bool one_rec = rec_first == last_line / a.lines_per_rec;
    const int slice = blockIdx.x & (kHistSlices - 1);
    float mx = 0.f;
    if (a.hist) __syncthreads();  // histogram zeroing done
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      if (act) {
        const float m = cabsf_ref(v[q]);
        a.absb[base + out_pos<L, E>(t, q)] = m;
        mx = fmaxf(mx, m);
        if (a.hist) {
          if (one_rec) atomicAdd(&hsh[__float_as_uint(m) >> 21], 1u);
          else atomicAdd(&a.hist[(long)rec * kHistStride + slice * 2048 + (__float_as_uint(m) >> 21)], 1u);
        }
      }
    }
    if (a.ctl) {
      // one atomic per work-group, into its XCD's slot
      if (one_rec) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, 64));
        if ((tid & 63) == 0) wmax[tid >> 6] = mx;
        __syncthreads();
        if (tid == 0) {
          for (int w = 1; w < (NT + 63) / 64; ++w) mx = fmaxf(mx, wmax[w]);
          atomicMax(&a.ctl[rec_first * kCtlStride + kCtlMax + slice], __float_as_uint(mx));
        }
      } else if (act) {
        atomicMax(&a.ctl[rec * kCtlStride + kCtlMax + slice], __float_as_uint(mx));
      }
    }
    if (a.hist && one_rec) {
      __syncthreads();
      unsigned* h = a.hist + (long)rec_first * kHistStride + slice * 2048;
      for (int i = tid; i < 2048; i += NT)
        if (hsh[i]) atomicAdd(&h[i], hsh[i]);
    }
  } else {
#pragma unroll
    for (int q = 0; q < NR; ++q)
      if (act) a.out[base + out_pos<L, E>(t, q)] = v[q];
  }
}

// The fused real-space pass of one iteration: inverse x-transform of the
// column-pass output, 1/N scale, algorithm update, support constraints, store
// of the new iterate, then forward x-transform of the new iterate (and, for
// the difference map, of its f1) for the next iteration's column pass.  The
// update runs on registers at the inverse transform's output positions; its
// other operands (g_k, flags) are loaded before the inverse transform.
// Thread layout of row work-groups.  Natural layout: M lines of T threads
// each (thread = line * T + t).  4-column tiled layout of W: 4 consecutive
// lines whose lanes interleave as 16 positions x 4 lines, so that one
// wavefront load covers 16 consecutive x of 4 rows: four whole 128-byte
// lines of a row-major array and four whole 4 x 4 tiles of the tiled one.
template <int L, int E, bool T4>
struct RowGeo {
  static constexpr int T = PlanOf<L, E>::T;
  static constexpr int M = T4 ? 4 : RowShape<L, E>::M;
  static constexpr int NT = T * M;
  __device__ static __forceinline__ int line_of(int tid) { return T4 ? ((tid >> 4) & 3) : tid / T; }
  __device__ static __forceinline__ int pos_of(int tid) {
    return T4 ? (((tid >> 6) << 4) | (tid & 15)) : tid - (tid / T) * T;
  }
};
// the tiled row kernels need T % 16 == 0, 4 T <= 1024 threads and two LDS
// buffers of 4 lines within the 160 KiB of LDS a gfx950 work-group may use
template <int L, int E>
struct T4Ok {
  static constexpr int T = PlanOf<L, E>::T;
  static constexpr bool value = (T % 16 == 0) && (4 * T <= 1024) && (SingleBuf<L>::buffers * LdsView<L, 4, false>::SIZE * 8 <= kLdsBytes);
};

template <int L, int E, bool DM, bool T4>
__global__ __launch_bounds__((RowGeo<L, E, T4>::NT)) void k_row_update(UpdateArgs a) {
  using PL = PlanOf<L, E>;
  using RG = RowGeo<L, E, T4>;
  constexpr int M = RG::M;
  constexpr int NR = PL::NREG;
  using View = LdsView<L, M, false>;
  __shared__ float2 shA[View::SIZE];
  __shared__ float2 shB[SingleBuf<L>::value ? 1 : View::SIZE];
  const int tid = threadIdx.x;
  const int g = RG::line_of(tid), t = RG::pos_of(tid);
  const int blk = xcd_tile(blockIdx.x, gridDim.x);
  const int line = blk * M + g;
  bool act = line < a.nlines;
  bool idle_line = false;
  if (a.rec_active) {
    // kSlotIdle reconstructions keep their state: W stays in place and the
    // iterate is carried over to the other ping-pong buffer; kSlotVacant
    // ones keep nothing, so their lines are skipped.  A work-group without
    // an active line does only those copies and leaves before its first
    // barrier
    bool any = false;
    for (int k = 0; k < M; ++k) {
      const int l = blk * M + k;
      if (l < a.nlines && a.rec_active[l / a.lines_per_rec] == kSlotActive) any = true;
    }
    if (!any) {
      for (int i = tid; i < M * L; i += RG::NT) {
        const int l = blk * M + i / L;
        if (l < a.nlines && a.rec_active[l / a.lines_per_rec] == kSlotIdle)
          a.gnew[(long)l * L + i % L] = a.gcur[(long)l * L + i % L];
      }
      return;
    }
    const uint8_t st = act ? a.rec_active[line / a.lines_per_rec] : kSlotActive;
    idle_line = act && st == kSlotIdle;
    act = act && st == kSlotActive;
  }
  const View va{shA, g}, vb{shB, g};
  // idle lanes of a partial last work-group address the last line (loads
  // stay in bounds and need no predicate); only their stores are skipped
  const int lc = line < a.nlines ? line : a.nlines - 1;
  const long base = (long)lc * L;
  // W (and Wf): natural rows, or the 4-column tiled layout
  long wbase = base;
  if constexpr (T4) {
    const int rec = lc / a.lines_per_rec, y = lc - rec * a.lines_per_rec;
    wbase = (long)rec * L * a.lines_per_rec + 4L * y;
  }
  auto wpos = [&](int x) -> long {
    if constexpr (T4) return t4_offset(x, a.lines_per_rec);
    else return x;
  };

  // Operands in the order they are needed: the twiddles and W for the
  // inverse transform, then g_k, the flags and the per-reconstruction
  // parameters for the update, which arrive during the transform.  Uniform
  // choices (cache hints, optional outputs) are branched on outside the
  // per-element loops.
  TwRegs<L, E> twr;
  load_tw<L, E>(a.tw, t, twr);
  float2 v[NR], vf[DM ? NR : 1], g0v[NR];
  uint8_t fl[NR];
  if (a.nt & 1) {
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      v[q] = ld_nt(&a.W[wbase + wpos(in_pos<L, E>(t, q))]);
      if constexpr (DM) vf[q] = ld_nt(&a.Wf[wbase + wpos(in_pos<L, E>(t, q))]);
    }
  } else {
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      v[q] = a.W[wbase + wpos(in_pos<L, E>(t, q))];
      if constexpr (DM) vf[q] = a.Wf[wbase + wpos(in_pos<L, E>(t, q))];
    }
  }
  if (a.nt & 4) {
#pragma unroll
    for (int q = 0; q < NR; ++q) g0v[q] = ld_nt(&a.gcur[base + out_pos<L, E>(t, q)]);
  } else {
#pragma unroll
    for (int q = 0; q < NR; ++q) g0v[q] = a.gcur[base + out_pos<L, E>(t, q)];
  }
#pragma unroll
  for (int q = 0; q < NR; ++q) fl[q] = a.flags[base + out_pos<L, E>(t, q)];
  const int rec = act ? line / a.lines_per_rec : 0;
  const float beta = a.beta_rec ? a.beta_rec[rec] : a.beta;
  const int alg = a.alg_rec ? a.alg_rec[rec] : a.alg;

  fft_line<L, E, +1>(v, va, vb, t, twr);
  if constexpr (DM) {
    __syncthreads();
    fft_line<L, E, +1>(vf, va, vb, t, twr);
  }

  // outside the support g1 = cp * gp + cg * g0: HIO g0 - beta gp
  // (src/phasing.c:1160-1168), RAAR beta g0 + (1 - 2 beta) gp
  // (src/phasing.c:1194-1215), ER 0 (src/phasing.c:1128-1136); inside, gp
  const bool er = !(alg == kHIO || alg == kRAAR);
  const float cp = alg == kHIO ? -beta : 1.f - 2.f * beta;
  const float cg = alg == kHIO ? 1.f : beta;
  const SupportConstraints sc(a.constraints);
  float2 g1v[NR];
#pragma unroll
  for (int q = 0; q < NR; ++q) {
    const float2 P = v[q];
    const float2 g0 = g0v[q];
    const bool in = fl[q] & kInsideSupport;
    float2 g1;
    if constexpr (!DM) {
      const float2 gp = make_float2(P.x * a.scale, P.y * a.scale);
      float2 o = make_float2(fmaf(cp, gp.x, cg * g0.x), fmaf(cp, gp.y, cg * g0.y));
      o = er ? make_float2(0.f, 0.f) : o;
      const float2 gi = sc.apply(gp);
      g1 = make_float2(in ? gi.x : o.x, in ? gi.y : o.y);
    } else {
      // src/phasing.c:1243-1252; gp stays unscaled like the CPU engine
      const float2 Pf = vf[q];
      const float2 gin = make_float2(g0.x + beta * ((1.f + a.gamma2) * P.x / a.nsize - a.gamma2 * g0.x),
                                     g0.y + beta * ((1.f + a.gamma2) * P.y / a.nsize - a.gamma2 * g0.y));
      g1 = in ? gin : g0;
      g1.x -= beta * Pf.x / a.nsize;
      g1.y -= beta * Pf.y / a.nsize;
      const float2 gc = sc.apply(g1);
      g1 = in ? gc : g1;
      // f1 of the next iteration (src/phasing.c:1260-1274)
      vf[q] = in ? g1 : make_float2(-a.gamma1 * g1.x, -a.gamma1 * g1.y);
    }
    g1v[q] = g1;
  }
  if (a.gp && act) {
#pragma unroll
    for (int q = 0; q < NR; ++q)
      a.gp[base + out_pos<L, E>(t, q)] = DM ? v[q] : make_float2(v[q].x * a.scale, v[q].y * a.scale);
  }
  if (act) {
    if (a.nt & 8) {
#pragma unroll
      for (int q = 0; q < NR; ++q) st_nt(&a.gnew[base + out_pos<L, E>(t, q)], g1v[q]);
    } else {
#pragma unroll
      for (int q = 0; q < NR; ++q) a.gnew[base + out_pos<L, E>(t, q)] = g1v[q];
    }
  } else if (idle_line) {
#pragma unroll
    for (int q = 0; q < NR; ++q) a.gnew[base + out_pos<L, E>(t, q)] = g0v[q];
  }
#pragma unroll
  for (int q = 0; q < NR; ++q) v[q] = g1v[q];
  out_to_in<L, E>(v, va, t);
  if constexpr (DM) out_to_in<L, E>(vf, va, t);
  __syncthreads();
  fft_line<L, E, -1>(v, va, vb, t, twr);
  if constexpr (DM) {
    __syncthreads();
    fft_line<L, E, -1>(vf, va, vb, t, twr);
  }
  if (act) {
    if (a.nt & 2) {
#pragma unroll
      for (int q = 0; q < NR; ++q) {
        st_nt(&a.W[wbase + wpos(out_pos<L, E>(t, q))], v[q]);
        if constexpr (DM) st_nt(&a.Wf[wbase + wpos(out_pos<L, E>(t, q))], vf[q]);
      }
    } else {
#pragma unroll
      for (int q = 0; q < NR; ++q) {
        a.W[wbase + wpos(out_pos<L, E>(t, q))] = v[q];
        if constexpr (DM) a.Wf[wbase + wpos(out_pos<L, E>(t, q))] = vf[q];
      }
    }
  }
}

// The real-space pass for the common case, specialised: ER / HIO / RAAR
// without the model-before-projection output.  W's cache hint is a template
// argument (NTW 0 or 3: loads and stores alike); the iterates' hints (nt
// bits 2, 3) stay arguments.  TAB adds what batches of mixed requests need:
// per-reconstruction algorithm and beta tables and idle reconstructions,
// handled as k_row_update handles them.  The same arithmetic as
// k_row_update, without its other optional paths: kept apart, those cost
// this pass about 0.4 us at 1024^2.
template <int L, int E, bool T4, int NTW, bool TAB>
__global__ __launch_bounds__((RowGeo<L, E, T4>::NT)) void k_row_plain(UpdateArgs a) {
  using PL = PlanOf<L, E>;
  using RG = RowGeo<L, E, T4>;
  constexpr int M = RG::M;
  constexpr int NR = PL::NREG;
  using View = LdsView<L, M, false>;
  __shared__ float2 shA[View::SIZE];
  __shared__ float2 shB[SingleBuf<L>::value ? 1 : View::SIZE];
  const int tid = threadIdx.x;
  const int g = RG::line_of(tid), t = RG::pos_of(tid);
  const int blk = xcd_tile(blockIdx.x, gridDim.x);
  const int line = blk * M + g;
  bool act = line < a.nlines;
  bool idle_line = false;
  if constexpr (TAB) {
    if (a.rec_active) {
      // kSlotIdle reconstructions keep their state (the iterate is carried
      // over to the other ping-pong buffer); kSlotVacant ones are skipped
      bool any = false;
      for (int k = 0; k < M; ++k) {
        const int l = blk * M + k;
        if (l < a.nlines && a.rec_active[l / a.lines_per_rec] == kSlotActive) any = true;
      }
      if (!any) {
        for (int i = tid; i < M * L; i += RG::NT) {
          const int l = blk * M + i / L;
          if (l < a.nlines && a.rec_active[l / a.lines_per_rec] == kSlotIdle)
            a.gnew[(long)l * L + i % L] = a.gcur[(long)l * L + i % L];
        }
        return;
      }
      const uint8_t st = act ? a.rec_active[line / a.lines_per_rec] : kSlotActive;
      idle_line = act && st == kSlotIdle;
      act = act && st == kSlotActive;
    }
  }
  const View va{shA, g}, vb{shB, g};
  const int lc = line < a.nlines ? line : a.nlines - 1;
  const long base = (long)lc * L;
  long wbase = base;
  if constexpr (T4) {
    const int rec = lc / a.lines_per_rec, y = lc - rec * a.lines_per_rec;
    wbase = (long)rec * L * a.lines_per_rec + 4L * y;
  }
  auto wpos = [&](int x) -> long {
    if constexpr (T4) return t4_offset(x, a.lines_per_rec);
    else return x;
  };
  float2 v[NR], g0v[NR];
  uint8_t fl[NR];
#pragma unroll
  for (int q = 0; q < NR; ++q) {
    const float2* p = &a.W[wbase + wpos(in_pos<L, E>(t, q))];
    v[q] = NTW ? ld_nt(p) : *p;
  }
  if (a.nt & 4) {
#pragma unroll
    for (int q = 0; q < NR; ++q) g0v[q] = ld_nt(&a.gcur[base + out_pos<L, E>(t, q)]);
  } else {
#pragma unroll
    for (int q = 0; q < NR; ++q) g0v[q] = a.gcur[base + out_pos<L, E>(t, q)];
  }
#pragma unroll
  for (int q = 0; q < NR; ++q) fl[q] = a.flags[base + out_pos<L, E>(t, q)];
  TwRegs<L, E> twr;
  load_tw<L, E>(a.tw, t, twr);
  fft_line<L, E, +1>(v, va, vb, t, twr);
  // src/phasing.c:1128-1136 (ER), 1160-1168 (HIO), 1194-1215 (RAAR)
  int alg = a.alg;
  float beta = a.beta;
  if constexpr (TAB) {
    const int rec = act ? line / a.lines_per_rec : 0;
    if (a.beta_rec) beta = a.beta_rec[rec];
    if (a.alg_rec) alg = a.alg_rec[rec];
  }
  const bool er = !(alg == kHIO || alg == kRAAR);
  const float cp = alg == kHIO ? -beta : 1.f - 2.f * beta;
  const float cg = alg == kHIO ? 1.f : beta;
  const SupportConstraints sc(a.constraints);
#pragma unroll
  for (int q = 0; q < NR; ++q) {
    const float2 gp = make_float2(v[q].x * a.scale, v[q].y * a.scale);
    const float2 g0 = g0v[q];
    float2 o = make_float2(fmaf(cp, gp.x, cg * g0.x), fmaf(cp, gp.y, cg * g0.y));
    o = er ? make_float2(0.f, 0.f) : o;
    const float2 gi = sc.apply(gp);
    const bool in = fl[q] & kInsideSupport;
    v[q] = make_float2(in ? gi.x : o.x, in ? gi.y : o.y);
  }
  if (act) {
    if (a.nt & 8) {
#pragma unroll
      for (int q = 0; q < NR; ++q) st_nt(&a.gnew[base + out_pos<L, E>(t, q)], v[q]);
    } else {
#pragma unroll
      for (int q = 0; q < NR; ++q) a.gnew[base + out_pos<L, E>(t, q)] = v[q];
    }
  } else if (TAB && idle_line) {
#pragma unroll
    for (int q = 0; q < NR; ++q) a.gnew[base + out_pos<L, E>(t, q)] = g0v[q];
  }
  out_to_in<L, E>(v, va, t);
  __syncthreads();
  fft_line<L, E, -1>(v, va, vb, t, twr);
  if (act) {
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      float2* p = &a.W[wbase + wpos(out_pos<L, E>(t, q))];
      if constexpr (NTW != 0) st_nt(p, v[q]);
      else *p = v[q];
    }
  }
}

// Forward x-transform of row-major `in` into the 4-column tiled `out`
// (in != out): the row-transformed iterate W of a 2-D shape.
template <int L, int E>
__global__ __launch_bounds__((RowGeo<L, E, true>::NT)) void k_row_fwd_t4(RowArgs a) {
  using RG = RowGeo<L, E, true>;
  constexpr int NR = PlanOf<L, E>::NREG;
  using View = LdsView<L, 4, false>;
  __shared__ float2 shA[View::SIZE];
  __shared__ float2 shB[SingleBuf<L>::value ? 1 : View::SIZE];
  const int tid = threadIdx.x;
  const int g = RG::line_of(tid), t = RG::pos_of(tid);
  const int blk = xcd_tile(blockIdx.x, gridDim.x);
  const int line = blk * 4 + g;
  const bool act = line < a.nlines;
  const View va{shA, g}, vb{shB, g};
  const int lc = act ? line : a.nlines - 1;
  const long base = (long)lc * L;
  const int rec = lc / a.lines_per_rec, y = lc - rec * a.lines_per_rec;
  const long wbase = (long)rec * L * a.lines_per_rec + 4L * y;
  float2 v[NR];
#pragma unroll
  for (int q = 0; q < NR; ++q) v[q] = a.in[base + in_pos<L, E>(t, q)];
  TwRegs<L, E> twr;
  load_tw<L, E>(a.tw, t, twr);
  fft_line<L, E, -1>(v, va, vb, t, twr);
  if (act) {
#pragma unroll
    for (int q = 0; q < NR; ++q) st_nt(&a.out[wbase + t4_offset(out_pos<L, E>(t, q), a.lines_per_rec)], v[q]);
  }
}

// ---------------------------------------------------------------------------
// Column kernels: C adjacent lines (contiguous in memory) per work-group.
// ---------------------------------------------------------------------------
// WIDE: at least 8 adjacent lines (64-byte segments).  In the row-major
// layout 32-byte segments waste much of each 128-byte line fetched, so a
// batch with enough column groups to fill the chip runs its projection pass
// wide; a single reconstruction keeps 4 lines per work-group and twice the
// work-groups (the tuner times both).
template <int L, int E, bool WIDE = false>
struct ColShape {
  static constexpr int T = PlanOf<L, E>::T;
  // C >= 4 adjacent lines (32-byte segments) and about 256 threads, within
  // 1024 threads and the work-group's LDS
  static constexpr int LS = L + (L >> 4);
  static constexpr int Cthr = (1024 / T) >= 1 ? (1024 / T) : 1;
  static constexpr int NBUF = SingleBuf<L>::buffers;
  static constexpr int Cmem = (kLdsBytes / (NBUF * LS * 8)) >= 1 ? (kLdsBytes / (NBUF * LS * 8)) : 1;
  static constexpr int Cmin = WIDE ? 8 : 4;
  // a power of two, so that column groups start on 128-byte lines (96^3 x
  // 32: 21 columns -> 16, 501 -> 479 us per iteration)
  static constexpr int pow2_floor(int v) {
    int p = 1;
    while (p * 2 <= v) p *= 2;
    return p;
  }
  static constexpr int Cw = pow2_floor((256 / T) > Cmin ? (256 / T) : Cmin);
  static constexpr int C1 = Cw < 64 ? Cw : 64;
  static constexpr int C2 = C1 < Cthr ? C1 : Cthr;
  static constexpr int C = C2 < Cmem ? C2 : Cmem;
  static constexpr int NT = T * C;
};

template <int L, int E, int MODE, bool WIDE = false>
__global__ __launch_bounds__((ColShape<L, E, WIDE>::NT)) void k_col(ColArgs a) {
  using PL = PlanOf<L, E>;
  constexpr int T = ColShape<L, E, WIDE>::T, C = ColShape<L, E, WIDE>::C, NT = ColShape<L, E, WIDE>::NT;
  constexpr int NR = PL::NREG;
  using View = LdsView<L, C, true>;
  __shared__ float2 shA[View::SIZE];
  __shared__ float2 shB[SingleBuf<L>::value ? 1 : View::SIZE];
  __shared__ double red[(MODE == kColEFourier) ? 4 * 32 : 1];
  const int tid = threadIdx.x;
  const int c = tid % C, t = tid / C;
  const int ngrp = (a.geo.ncontig + C - 1) / C;
  const int blk = xcd_tile(blockIdx.x, gridDim.x);
  const int outer = blk / ngrp;
  if (MODE != kColEFourier && a.rec_active && a.rec_active[outer / a.outer_per_rec] != kSlotActive) return;
  const int x = (blk - outer * ngrp) * C + c;
  const bool act = x < a.geo.ncontig;
  // idle lanes address the last line of the group (in bounds, no predicate)
  const int xc = act ? x : a.geo.ncontig - 1;
  const long base = (long)outer * a.geo.outer_stride + (long)(xc >> a.geo.tshift) * a.geo.tstride +
                    (xc & ((1 << a.geo.tshift) - 1));
  const long st = a.geo.axis_stride;
  const View va{shA, c}, vb{shB, c};

  // twiddles and line data first, then the amplitudes of the projection,
  // which arrive during the forward transform
  TwRegs<L, E> twr;
  load_tw<L, E>(a.tw, t, twr);
  float2 v[NR];
  if (a.nt & 1) {
#pragma unroll
    for (int q = 0; q < NR; ++q) v[q] = ld_nt(&a.in[base + in_pos<L, E>(t, q) * st]);
  } else {
#pragma unroll
    for (int q = 0; q < NR; ++q) v[q] = a.in[base + in_pos<L, E>(t, q) * st];
  }
  float ampv[(MODE == kColProject || MODE == kColEFourier) ? NR : 1];
  if constexpr (MODE == kColProject || MODE == kColEFourier) {
    if (MODE == kColEFourier || a.objective == 0) {
#pragma unroll
      for (int q = 0; q < NR; ++q) ampv[q] = a.ampx[base + out_pos<L, E>(t, q) * st];
    }
  }
  if constexpr (MODE == kColInv) fft_line<L, E, +1>(v, va, vb, t, twr);
  else fft_line<L, E, -1>(v, va, vb, t, twr);

  if constexpr (MODE == kColProject) {
    if (a.objective == 0 && !a.margin) {
      // the common case, branch-free: the modulus projection of
      // project_pixel without margins (src/phasing.c:1070-1100)
      const bool centro = a.centro != 0;
#pragma unroll
      for (int q = 0; q < NR; ++q) v[q] = project_phase(v[q], ampv[q], centro);
    } else {
#pragma unroll
      for (int q = 0; q < NR; ++q) {
        if (!act) continue;
        const long idx = base + out_pos<L, E>(t, q) * st;
        float2 G = v[q];
        if (a.centro) G.y = 0.f;  // src/phasing.c:1061-1068
        if (a.objective == 0) {
          G = project_pixel(G, ampv[q], a, idx);
        } else {
          // src/phasing.c:1103-1109: measured pixels take the phased
          // amplitude; unmeasured entries carry kUnmeasuredBits in .x
          const float2 P = a.phased_amp[idx];
          if (__float_as_uint(P.x) != kUnmeasuredBits) G = P;
        }
        v[q] = G;
      }
    }
    out_to_in<L, E>(v, va, t);
    __syncthreads();
    fft_line<L, E, +1>(v, va, vb, t, twr);
  } else if constexpr (MODE == kColBlur) {
    // Gaussian transfer function of sp_gaussian_blur (src/image_filter.c:80-91)
    // with dist_to_top_left (src/image_util.c:1012-1032), computed in double
    // and rounded to float like the reference.
    const double NN = (double)((float)((long)a.nx * a.ny * a.nz));
    const double k0 = -2.0 * (3.14159265358979323846 * 3.14159265358979323846);
    const double s2 = (double)a.sigma * (double)a.sigma;
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      if (!act) continue;
      const long idx = base + out_pos<L, E>(t, q) * st;
      const int ix = (int)(idx % a.nx);
      const long rest = idx / a.nx;
      const int iy = (int)(rest % a.ny);
      const int iz = (int)((rest / a.ny) % a.nz);  // reconstructions stack along z
      const float dx = (a.nx - ix < ix) ? (float)(a.nx - ix) : (float)ix;
      const float dy = (a.ny - iy < iy) ? (float)(a.ny - iy) : (float)iy;
      const float dz = (a.nz - iz < iz) ? (float)(a.nz - iz) : (float)iz;
      const float dist = (float)sqrt((double)dx * dx + (double)dy * dy + (double)dz * dz);
      const float r = dist / (float)a.nx;
      const float h = (float)(exp(k0 * ((double)r * (double)r) * s2) / NN);
      v[q] = make_float2(v[q].x * h, v[q].y * h);
    }
    out_to_in<L, E>(v, va, t);
    __syncthreads();
    fft_line<L, E, +1>(v, va, vb, t, twr);
  } else if constexpr (MODE == kColEFourier) {
    // src/phasing.c:879-888: measured pixels add (|F|-|amp|)^2 and amp^2,
    // unmeasured ones |F|^2; FcFo sums |amp| and |F| over measured pixels.
    double acc[4] = {0.0, 0.0, 0.0, 0.0};  // nom, den, fcfo_num, fcfo_den
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      if (!act) continue;
      const long idx = base + out_pos<L, E>(t, q) * st;
      const float2 F = v[q];
      if (a.fout) a.fout[idx] = F;
      const float af = cabsf_ref(F);
      const float ab = ampv[q];
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
    block_sum<4, NT>(acc, red);
    // indexed by tile: the tiles of one reconstruction are contiguous
    if (tid == 0) {
#pragma unroll
      for (int qq = 0; qq < 4; ++qq) a.partials[(long)blk * 4 + qq] = acc[qq];
    }
    return;
  }

  if (act) {
    if (a.nt & 2) {
#pragma unroll
      for (int q = 0; q < NR; ++q) st_nt(&a.W[base + out_pos<L, E>(t, q) * st], v[q]);
    } else {
#pragma unroll
      for (int q = 0; q < NR; ++q) a.W[base + out_pos<L, E>(t, q) * st] = v[q];
    }
  }
}

// The Fourier-space pass for the common case, specialised: the phase
// objective without amplitude margins (the branch-free projection of
// k_col), W's cache hint a template argument (NTW 0 or 3).  Idle
// reconstructions of a batch still leave at once.
template <int L, int E, bool WIDE, int NTW>
__global__ __launch_bounds__((ColShape<L, E, WIDE>::NT)) void k_col_plain(ColArgs a) {
  using PL = PlanOf<L, E>;
  constexpr int C = ColShape<L, E, WIDE>::C;
  constexpr int NR = PL::NREG;
  using View = LdsView<L, C, true>;
  __shared__ float2 shA[View::SIZE];
  __shared__ float2 shB[SingleBuf<L>::value ? 1 : View::SIZE];
  const int tid = threadIdx.x;
  const int c = tid % C, t = tid / C;
  const int ngrp = (a.geo.ncontig + C - 1) / C;
  const int blk = xcd_tile(blockIdx.x, gridDim.x);
  const int outer = blk / ngrp;
  if (a.rec_active && a.rec_active[outer / a.outer_per_rec] != kSlotActive) return;
  const int x = (blk - outer * ngrp) * C + c;
  const bool act = x < a.geo.ncontig;
  const int xc = act ? x : a.geo.ncontig - 1;
  const long base = (long)outer * a.geo.outer_stride + (long)(xc >> a.geo.tshift) * a.geo.tstride +
                    (xc & ((1 << a.geo.tshift) - 1));
  const long st = a.geo.axis_stride;
  const View va{shA, c}, vb{shB, c};
  float2 v[NR];
  float ampv[NR];
#pragma unroll
  for (int q = 0; q < NR; ++q) {
    const float2* p = &a.in[base + in_pos<L, E>(t, q) * st];
    v[q] = NTW ? ld_nt(p) : *p;
  }
#pragma unroll
  for (int q = 0; q < NR; ++q) ampv[q] = a.ampx[base + out_pos<L, E>(t, q) * st];
  TwRegs<L, E> twr;
  load_tw<L, E>(a.tw, t, twr);
  fft_line<L, E, -1>(v, va, vb, t, twr);
  const bool centro = a.centro != 0;
#pragma unroll
  for (int q = 0; q < NR; ++q) v[q] = project_phase(v[q], ampv[q], centro);
  out_to_in<L, E>(v, va, t);
  __syncthreads();
  fft_line<L, E, +1>(v, va, vb, t, twr);
  if (act) {
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      float2* p = &a.W[base + out_pos<L, E>(t, q) * st];
      if constexpr (NTW != 0) st_nt(p, v[q]);
      else *p = v[q];
    }
  }
}

// The Fourier-space pass on contiguous lines (the transposed copy of W of a
// 2-D shape, whose columns are rows of L = ny elements): k_col_plain's
// arithmetic with the row passes' geometry, RowShape<L, E>::M lines per
// work-group, so that several work-groups share a CU and overlap their
// phases.  Lines l = 0 .. ncontig * nouter - 1 start at l * L; line l
// belongs to reconstruction l / ncontig; idle reconstructions' lines are
// transformed but not stored.
template <int L, int E, int NTW>
__global__ __launch_bounds__((RowShape<L, E>::NT)) void k_line_project(ColArgs a) {
  using PL = PlanOf<L, E>;
  using RG = RowGeo<L, E, false>;
  constexpr int M = RG::M;
  constexpr int NR = PL::NREG;
  using View = LdsView<L, M, false>;
  __shared__ float2 shA[View::SIZE];
  __shared__ float2 shB[SingleBuf<L>::value ? 1 : View::SIZE];
  const int tid = threadIdx.x;
  const int g = RG::line_of(tid), t = RG::pos_of(tid);
  const int blk = xcd_tile(blockIdx.x, gridDim.x);
  const int nlines = a.geo.ncontig * a.geo.nouter;
  const int line = blk * M + g;
  const int lc = line < nlines ? line : nlines - 1;
  const bool act = line < nlines && !(a.rec_active && a.rec_active[lc / a.geo.ncontig] != kSlotActive);
  const long base = (long)lc * L;
  const View va{shA, g}, vb{shB, g};
  float2 v[NR];
  float ampv[NR];
#pragma unroll
  for (int q = 0; q < NR; ++q) {
    const float2* p = &a.in[base + in_pos<L, E>(t, q)];
    v[q] = NTW ? ld_nt(p) : *p;
  }
#pragma unroll
  for (int q = 0; q < NR; ++q) ampv[q] = a.ampx[base + out_pos<L, E>(t, q)];
  TwRegs<L, E> twr;
  load_tw<L, E>(a.tw, t, twr);
  fft_line<L, E, -1>(v, va, vb, t, twr);
  const bool centro = a.centro != 0;
#pragma unroll
  for (int q = 0; q < NR; ++q) v[q] = project_phase(v[q], ampv[q], centro);
  out_to_in<L, E>(v, va, t);
  __syncthreads();
  fft_line<L, E, +1>(v, va, vb, t, twr);
  if (act) {
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      float2* p = &a.W[base + out_pos<L, E>(t, q)];
      if constexpr (NTW != 0) st_nt(p, v[q]);
      else *p = v[q];
    }
  }
}

// ---------------------------------------------------------------------------
// Launch helpers and the per-length dispatch table.
// ---------------------------------------------------------------------------
// device twiddle table of plan (L, E), built once per process (engine.hip)
const float2* twiddles(int L, int E);

template <int L, int E>
struct LengthOps {
  static const float2* tw() {
    static const float2* t = twiddles(L, E);
    return t;
  }
  // a twiddle table in the arguments (the no-allocation transform keeps its
  // own) takes precedence over the shared one
  static void row(int mode, const RowArgs& a0, hipStream_t s) {
    RowArgs a = a0;
    a.tw = a0.tw ? a0.tw : tw();
    const dim3 grid((a.nlines + RowShape<L, E>::M - 1) / RowShape<L, E>::M), blk(RowShape<L, E>::NT);
    switch (mode) {
      case kRowFwd: hipLaunchKernelGGL((k_row<L, E, kRowFwd>), grid, blk, 0, s, a); break;
      case kRowInv: hipLaunchKernelGGL((k_row<L, E, kRowInv>), grid, blk, 0, s, a); break;
      case kRowAbsFwd: hipLaunchKernelGGL((k_row<L, E, kRowAbsFwd>), grid, blk, 0, s, a); break;
      case kRowBlurOut: hipLaunchKernelGGL((k_row<L, E, kRowBlurOut>), grid, blk, 0, s, a); break;
    }
  }
  static constexpr bool t4 = T4Ok<L, E>::value;
  static void update(const UpdateArgs& a0, bool dm, bool tiled, hipStream_t s) {
    UpdateArgs a = a0;
    a.tw = tw();
    if (tiled) {
      if constexpr (t4) {
        const dim3 grid((a.nlines + 3) / 4), blk(RowGeo<L, E, true>::NT);
        if (dm) hipLaunchKernelGGL((k_row_update<L, E, true, true>), grid, blk, 0, s, a);
        else hipLaunchKernelGGL((k_row_update<L, E, false, true>), grid, blk, 0, s, a);
      }
      return;
    }
    const dim3 grid((a.nlines + RowShape<L, E>::M - 1) / RowShape<L, E>::M), blk(RowShape<L, E>::NT);
    if (dm) hipLaunchKernelGGL((k_row_update<L, E, true, false>), grid, blk, 0, s, a);
    else hipLaunchKernelGGL((k_row_update<L, E, false, false>), grid, blk, 0, s, a);
  }
  // the specialised passes; false when the arguments need the general ones
  // (a W hint other than none or loads and stores both)
  template <bool T4K, bool TAB>
  static void row_plain_launch(const UpdateArgs& a, int ntw, dim3 grid, dim3 blk, hipStream_t s) {
    if (ntw) hipLaunchKernelGGL((k_row_plain<L, E, T4K, 3, TAB>), grid, blk, 0, s, a);
    else hipLaunchKernelGGL((k_row_plain<L, E, T4K, 0, TAB>), grid, blk, 0, s, a);
  }
  static bool update_plain(const UpdateArgs& a0, bool tiled, hipStream_t s) {
    UpdateArgs a = a0;
    a.tw = tw();
    const int ntw = a.nt & 3;
    if (ntw != 0 && ntw != 3) return false;
    const bool tab = a.alg_rec || a.beta_rec || a.rec_active;
    if (tiled) {
      if constexpr (t4) {
        const dim3 grid((a.nlines + 3) / 4), blk(RowGeo<L, E, true>::NT);
        if (tab) row_plain_launch<true, true>(a, ntw, grid, blk, s);
        else row_plain_launch<true, false>(a, ntw, grid, blk, s);
        return true;
      }
      return false;
    }
    const dim3 grid((a.nlines + RowShape<L, E>::M - 1) / RowShape<L, E>::M), blk(RowShape<L, E>::NT);
    if (tab) row_plain_launch<false, true>(a, ntw, grid, blk, s);
    else row_plain_launch<false, false>(a, ntw, grid, blk, s);
    return true;
  }
  // the projection pass on contiguous lines (k_line_project); false when
  // W's hint is neither none nor loads and stores both
  static bool line_project(const ColArgs& a0, hipStream_t s) {
    ColArgs a = a0;
    a.tw = a0.tw ? a0.tw : tw();
    const int ntw = a.nt & 3;
    if (ntw != 0 && ntw != 3) return false;
    const int nlines = a.geo.ncontig * a.geo.nouter;
    const dim3 grid((nlines + RowShape<L, E>::M - 1) / RowShape<L, E>::M), blk(RowShape<L, E>::NT);
    if (ntw) hipLaunchKernelGGL((k_line_project<L, E, 3>), grid, blk, 0, s, a);
    else hipLaunchKernelGGL((k_line_project<L, E, 0>), grid, blk, 0, s, a);
    return true;
  }
  template <bool WIDE>
  static bool col_plain_launch(const ColArgs& a0, hipStream_t s) {
    ColArgs a = a0;
    a.tw = tw();
    const int ntw = a.nt & 3;
    if (ntw != 0 && ntw != 3) return false;
    const dim3 grid(((a.geo.ncontig + ColShape<L, E, WIDE>::C - 1) / ColShape<L, E, WIDE>::C) * a.geo.nouter),
        blk(ColShape<L, E, WIDE>::NT);
    if (ntw) hipLaunchKernelGGL((k_col_plain<L, E, WIDE, 3>), grid, blk, 0, s, a);
    else hipLaunchKernelGGL((k_col_plain<L, E, WIDE, 0>), grid, blk, 0, s, a);
    return true;
  }
  static bool col_project_plain(const ColArgs& a0, bool wide, hipStream_t s) {
    if (wide) {
      if constexpr (ColShape<L, E, true>::C > ColShape<L, E>::C) return col_plain_launch<true>(a0, s);
      return false;
    }
    return col_plain_launch<false>(a0, s);
  }
  static void row_t4(const RowArgs& a0, hipStream_t s) {
    if constexpr (t4) {
      RowArgs a = a0;
      a.tw = tw();
      hipLaunchKernelGGL((k_row_fwd_t4<L, E>), dim3((a.nlines + 3) / 4), dim3(RowGeo<L, E, true>::NT), 0, s, a);
    }
  }
  static int col_blocks(const LineGeom& g) {
    return ((g.ncontig + ColShape<L, E>::C - 1) / ColShape<L, E>::C) * g.nouter;
  }
  static void col(int mode, const ColArgs& a0, hipStream_t s) {
    ColArgs a = a0;
    a.tw = a0.tw ? a0.tw : tw();
    const dim3 grid(col_blocks(a.geo)), blk(ColShape<L, E>::NT);
    switch (mode) {
      case kColFwd: hipLaunchKernelGGL((k_col<L, E, kColFwd>), grid, blk, 0, s, a); break;
      case kColInv: hipLaunchKernelGGL((k_col<L, E, kColInv>), grid, blk, 0, s, a); break;
      case kColProject: hipLaunchKernelGGL((k_col<L, E, kColProject>), grid, blk, 0, s, a); break;
      case kColEFourier: hipLaunchKernelGGL((k_col<L, E, kColEFourier>), grid, blk, 0, s, a); break;
      case kColBlur: hipLaunchKernelGGL((k_col<L, E, kColBlur>), grid, blk, 0, s, a); break;
    }
  }
  // the projection pass with at least 8 lines per work-group, where that
  // differs from col()'s work-group
  static constexpr bool wide = ColShape<L, E, true>::C > ColShape<L, E>::C;
  static int colw_blocks(const LineGeom& g) {
    return ((g.ncontig + ColShape<L, E, true>::C - 1) / ColShape<L, E, true>::C) * g.nouter;
  }
  static void col_project_wide(const ColArgs& a0, hipStream_t s) {
    if constexpr (wide) {
      ColArgs a = a0;
      a.tw = tw();
      hipLaunchKernelGGL((k_col<L, E, kColProject, true>), dim3(colw_blocks(a.geo)), dim3(ColShape<L, E, true>::NT), 0,
                         s, a);
    }
  }
};

// One compiled variant (plan with E elements per thread at the edges).
struct Variant {
  int E;
  bool t4;  // has the tiled-layout row kernels
  void (*row)(int, const RowArgs&, hipStream_t);
  void (*update)(const UpdateArgs&, bool, bool, hipStream_t);
  void (*row_t4)(const RowArgs&, hipStream_t);
  void (*col)(int, const ColArgs&, hipStream_t);
  int (*col_blocks)(const LineGeom&);
  bool wide;  // has a wider projection pass (col_project_wide)
  void (*col_project_wide)(const ColArgs&, hipStream_t);
  bool (*update_plain)(const UpdateArgs&, bool, hipStream_t);
  bool (*col_project_plain)(const ColArgs&, bool, hipStream_t);
  bool (*line_project)(const ColArgs&, hipStream_t);
};

template <int L, int E>
constexpr Variant variant() {
  static_assert(PlanOf<L, E>::T <= 1024, "a line of this plan needs more than 1024 threads");
  return Variant{E,
                 LengthOps<L, E>::t4,
                 &LengthOps<L, E>::row,
                 &LengthOps<L, E>::update,
                 &LengthOps<L, E>::row_t4,
                 &LengthOps<L, E>::col,
                 &LengthOps<L, E>::col_blocks,
                 LengthOps<L, E>::wide,
                 &LengthOps<L, E>::col_project_wide,
                 &LengthOps<L, E>::update_plain,
                 &LengthOps<L, E>::col_project_plain,
                 &LengthOps<L, E>::line_project};
}

// Dispatch table entry for one length: its compiled variants.
struct LengthEntry {
  int L;
  int nvar;
  Variant var[3];
};

// The passes of one axis of a shape: the length's entry and the variant the
// tuner chose for that shape (the value behaves like a pointer: `ax->row(..)`,
// and false when the axis is trivial).
struct AxisOps {
  const LengthEntry* e = nullptr;
  int v = 0;
  explicit operator bool() const { return e != nullptr; }
  const AxisOps* operator->() const { return this; }
  const Variant& var() const { return e->var[v]; }
  void row(int mode, const RowArgs& a, hipStream_t s) const { var().row(mode, a, s); }
  void update(const UpdateArgs& a, bool dm, bool tiled, hipStream_t s) const { var().update(a, dm, tiled, s); }
  void row_t4(const RowArgs& a, hipStream_t s) const { var().row_t4(a, s); }
  void col(int mode, const ColArgs& a, hipStream_t s) const { var().col(mode, a, s); }
  void col_project_wide(const ColArgs& a, hipStream_t s) const { var().col_project_wide(a, s); }
  bool update_plain(const UpdateArgs& a, bool tiled, hipStream_t s) const { return var().update_plain(a, tiled, s); }
  bool line_project(const ColArgs& a, hipStream_t s) const { return var().line_project(a, s); }
  bool col_project_plain(const ColArgs& a, bool wide, hipStream_t s) const {
    return var().col_project_plain(a, wide, s);
  }
  int col_blocks(const LineGeom& g) const { return var().col_blocks(g); }
};

// Registry implemented by the instantiation units.
LengthEntry* find_length(int L);

}  // namespace spx
