// Device building blocks of the Stockham FFT used by every pass kernel of the
// phasing engine.
//
// Conventions follow the reference's FFTW3 calls (src/fft.c:79,95,163,185):
// unnormalised complex-to-complex transforms, forward sign -1
// (X_k = sum_n x_n e^{-2 pi i k n / L}), inverse sign +1, row-major with x
// fastest.
//
// A length-L line is split into radix stages (2, 3, 4, 5, 6, 7, 8, 9, 10,
// 12, 15, 16) chosen at compile time.  T threads cooperate on a line.  Stage 0 reads its inputs from
// registers (loaded straight from global memory: butterfly b takes the
// elements b + k L/R0, so consecutive threads read consecutive addresses),
// the middle stages go through LDS (two buffers, one barrier per stage) and
// the last stage leaves its outputs in registers at b + r L/R_last.  Plans
// start and end with the same radix whenever that does not cost an extra
// stage, so the output positions of one transform are the input positions of
// the next: an element-wise step between an inverse and a forward transform
// (projection, algorithm update) never touches LDS.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace spx {

// ---------------------------------------------------------------------------
// Compile-time plan for one transform length.
// ---------------------------------------------------------------------------
struct FftPlan {
  int L;          // transform length
  int n;          // number of radix stages
  int r[16];      // radix of each stage (stage 0 first)
  int ns[16];     // Ns of each stage: product of the radices before it
  int twoff[16];  // offset of the stage's twiddles in the per-length table
  int twtotal;    // total twiddle entries (complex) for this length
  int maxr;       // largest radix
  int T;          // threads cooperating on one line
  bool ok;        // L factors into the supported radices
};

constexpr int ilog2c(int v) {
  int e = 0;
  while ((1 << e) < v) ++e;
  return e;
}
constexpr int gcd_c(int a, int b) {
  while (b) {
    const int t = a % b;
    a = b;
    b = t;
  }
  return a;
}
constexpr int pow2_stages(int e) { return (e + 3) / 4; }

// Appends the radices of 2^m (each at most 2^kmax), as even as possible,
// larger first: 2^6 with kmax 2 gives 4,4,4; 2^5 gives 4,4,2.
constexpr void push_pow2(FftPlan& p, int m, int kmax) {
  if (m <= 0) return;
  const int k = (m + kmax - 1) / kmax;
  int left = m;
  for (int i = 0; i < k; ++i) {
    const int take = (left + (k - i) - 1) / (k - i);
    p.r[p.n++] = 1 << take;
    left -= take;
  }
}

// Radices a stage can have: powers of two up to 16, 3, 5, 7, and the
// products 6, 9, 10, 12, 15 done in registers by one Cooley-Tukey step
// (dft_ct).
constexpr int kMixRadices[12] = {16, 15, 12, 10, 9, 8, 7, 6, 5, 4, 3, 2};

// The fewest radices (at most `depth`) from kMixRadices whose product is m,
// largest first, into out; returns their number, or 99 if there are none.
constexpr int split_radices(int m, int depth, int* out) {
  if (m == 1) return 0;
  int cap = 1;
  for (int d = 0; d < depth && cap < m; ++d) cap *= 16;
  if (depth == 0 || cap < m) return 99;
  int best = 99;
  int bestv[8] = {};
  for (int i = 0; i < 12; ++i) {
    const int r = kMixRadices[i];
    if (m % r != 0) continue;
    int sub[8] = {};
    const int c = split_radices(m / r, depth - 1, sub);
    if (c + 1 < best) {
      best = c + 1;
      bestv[0] = r;
      for (int k = 0; k < c; ++k) bestv[k + 1] = sub[k];
    }
  }
  if (best < 99)
    for (int k = 0; k < best; ++k) out[k] = bestv[k];
  return best;
}

// Palindromic plan e, (fewest middle radices), e with e^2 | L: among the
// edge radices within a factor 2 of E the one with the fewest stages, ties
// going to the one closest to E; if none is that close, the closest.  False
// when no radix e >= 2 has e^2 | L.
constexpr bool mixed_plan(int L, int E, FftPlan& p) {
  int best_e = 0, best_c = 99, best_key = 1 << 30;
  int best_mid[8] = {};
  for (int i = 0; i < 12; ++i) {
    const int e = kMixRadices[i];
    if (L % (e * e) != 0) continue;
    int mid[8] = {};
    int c = 99;
    for (int d = 0; d <= 6 && c == 99; ++d) c = split_radices(L / (e * e), d, mid);
    if (c == 99) continue;
    const int ratio = e >= E ? (e * 1024) / E : (E * 1024) / e;  // 1024 = equal
    const int key = ratio <= 2048 ? c * 4096 + ratio : (1 << 20) + ratio;
    if (key < best_key) {
      best_e = e;
      best_c = c;
      best_key = key;
      for (int k = 0; k < c; ++k) best_mid[k] = mid[k];
    }
  }
  if (!best_e) return false;
  p.n = 0;
  p.r[p.n++] = best_e;
  for (int k = 0; k < best_c; ++k) p.r[p.n++] = best_mid[k];
  p.r[p.n++] = best_e;
  return true;
}

// Plan of a length-L line whose threads hold E elements at the edges
// (E = 4, 8 or 16; also the largest power-of-two radix used).  The edge
// radix 2^ke appears first and last so the transform is palindromic.  When
// L's power-of-two part is too small for edges of E (a line of 2^a 3^b 5^c
// 7^d with 2^a < E^2), the plan of mixed radices with edges closest to E
// replaces it if it has fewer stages: 1000 = 2 2 5 5 5 2 becomes 10 10 10.
constexpr FftPlan make_plan(int L, int E) {
  FftPlan p{};
  p.L = L;
  p.n = 0;
  p.ok = L >= 2;
  int rem = L, n3 = 0, n5 = 0, n7 = 0;
  while (rem % 7 == 0) { ++n7; rem /= 7; }
  while (rem % 5 == 0) { ++n5; rem /= 5; }
  while (rem % 3 == 0) { ++n3; rem /= 3; }
  if (!p.ok || (rem & (rem - 1)) != 0) {
    p.ok = false;
    return p;
  }
  const int a2 = ilog2c(rem), kE = ilog2c(E);
  if (L == 3 || L == 5 || L == 7 || (L == rem && L <= E)) {
    p.r[p.n++] = L;
  } else {
    const int ke = (a2 / 2 < kE) ? a2 / 2 : kE;
    if (ke >= 1) {
      p.r[p.n++] = 1 << ke;
      push_pow2(p, a2 - 2 * ke, kE);
      for (int i = 0; i < n3; ++i) p.r[p.n++] = 3;
      for (int i = 0; i < n5; ++i) p.r[p.n++] = 5;
      for (int i = 0; i < n7; ++i) p.r[p.n++] = 7;
      p.r[p.n++] = 1 << ke;
    } else {
      push_pow2(p, a2, kE);
      for (int i = 0; i < n3; ++i) p.r[p.n++] = 3;
      for (int i = 0; i < n5; ++i) p.r[p.n++] = 5;
      for (int i = 0; i < n7; ++i) p.r[p.n++] = 7;
    }
    if (n3 + n5 + n7 > 0 && ke < kE) {
      FftPlan q = p;
      if (mixed_plan(L, E, q) && q.n < p.n) {
        p.n = q.n;
        for (int s = 0; s < q.n; ++s) p.r[s] = q.r[s];
      }
    }
  }
  int ns = 1, off = 0, mr = 1;
  for (int s = 0; s < p.n; ++s) {
    p.ns[s] = ns;
    p.twoff[s] = off;
    if (s > 0) off += ns * (p.r[s] - 1);
    ns *= p.r[s];
    if (p.r[s] > mr) mr = p.r[s];
  }
  p.twtotal = off;
  p.maxr = mr;
  const int e0 = p.r[0], el = p.r[p.n - 1];
  p.T = L / (e0 / gcd_c(e0, el) * el);
  if (ns != L) p.ok = false;
  return p;
}

template <int L, int E>
struct PlanOf {
  static constexpr FftPlan P = make_plan(L, E);
  static_assert(P.ok, "transform length does not factor into the supported radices");
  static constexpr int T = P.T;
  static constexpr int E0 = P.r[0];
  static constexpr int EL = P.r[P.n - 1];
  static constexpr int NREG = L / T;       // elements each thread holds at the edges
  static constexpr int B0 = (L / E0) / T;  // stage-0 butterflies per thread
  static constexpr int BL = (L / EL) / T;  // last-stage butterflies per thread
  static constexpr bool PAL = (E0 == EL);  // output positions == input positions
  static_assert(B0 * E0 == NREG && BL * EL == NREG, "edge stages must divide evenly over the threads");
  // butterflies of stage s one thread computes
  static constexpr int bpt(int s) { return ((L / P.r[s]) + T - 1) / T; }
  // first twiddle register of stage s >= 1: stages 1..s-1 each hold
  // bpt * (R - 1) values
  static constexpr int twreg(int s) {
    int o = 0;
    for (int q = 1; q < s; ++q) o += bpt(q) * (P.r[q] - 1);
    return o;
  }
  static constexpr int NTW = twreg(P.n);  // twiddle registers per thread
};

// the edge registers of one thread
template <int L, int E>
using RegLine = float2[PlanOf<L, E>::NREG];

// The twiddles one thread applies, for every stage after the first, read
// from the global table once at kernel start (the table is a few KiB and
// shared by every work-group, so these reads hit in the caches) and used by
// both the inverse and the forward transform of a pass.
template <int L, int E>
struct TwRegs {
  float2 w[PlanOf<L, E>::NTW > 0 ? PlanOf<L, E>::NTW : 1];
};

// ---------------------------------------------------------------------------
// Complex helpers on float2 (x = re, y = im).
// ---------------------------------------------------------------------------
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// multiply by (DIR * i): DIR = -1 gives -i, DIR = +1 gives +i
template <int DIR>
__device__ __forceinline__ float2 rot(float2 a) {
  return DIR < 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
}
// twiddle w stored for the forward transform; the inverse uses conj(w)
template <int DIR>
__device__ __forceinline__ float2 twmul(float2 a, float2 w) {
  return DIR < 0 ? cmul(a, w) : make_float2(a.x * w.x + a.y * w.y, a.y * w.x - a.x * w.y);
}

// ---------------------------------------------------------------------------
// Small in-register DFTs, X_r = sum_k x_k exp(DIR 2 pi i r k / R).
// ---------------------------------------------------------------------------
template <int DIR>
__device__ __forceinline__ void dft2(float2& a, float2& b) {
  const float2 t = a;
  a = cadd(t, b);
  b = csub(t, b);
}

template <int DIR>
__device__ __forceinline__ void dft4(float2& x0, float2& x1, float2& x2, float2& x3) {
  const float2 t0 = cadd(x0, x2), t1 = csub(x0, x2);
  const float2 t2 = cadd(x1, x3), t3 = rot<DIR>(csub(x1, x3));
  x0 = cadd(t0, t2);
  x2 = csub(t0, t2);
  x1 = cadd(t1, t3);
  x3 = csub(t1, t3);
}

template <int DIR>
__device__ __forceinline__ void dft8(float2* x) {
  constexpr float c = 0.70710678118654752440f;
  float2 e0 = x[0], e1 = x[2], e2 = x[4], e3 = x[6];
  float2 o0 = x[1], o1 = x[3], o2 = x[5], o3 = x[7];
  dft4<DIR>(e0, e1, e2, e3);
  dft4<DIR>(o0, o1, o2, o3);
  // W8^1 = (c, DIR c), W8^2 = DIR i, W8^3 = (-c, DIR c)
  const float2 w1 = make_float2(c * (o1.x - DIR * o1.y), c * (o1.y + DIR * o1.x));
  const float2 w2 = rot<DIR>(o2);
  const float2 w3 = make_float2(c * (-o3.x - DIR * o3.y), c * (-o3.y + DIR * o3.x));
  x[0] = cadd(e0, o0); x[4] = csub(e0, o0);
  x[1] = cadd(e1, w1); x[5] = csub(e1, w1);
  x[2] = cadd(e2, w2); x[6] = csub(e2, w2);
  x[3] = cadd(e3, w3); x[7] = csub(e3, w3);
}

template <int DIR>
__device__ __forceinline__ void dft16(float2* x) {
  float2 e[8], o[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) { e[k] = x[2 * k]; o[k] = x[2 * k + 1]; }
  dft8<DIR>(e);
  dft8<DIR>(o);
  // W16^r = (cos(2 pi r/16), DIR sin(2 pi r/16))
  constexpr float C1 = 0.92387953251128675613f, S1 = 0.38268343236508977173f;
  constexpr float C2 = 0.70710678118654752440f;
  const float cs[8] = {1.f, C1, C2, S1, 0.f, -S1, -C2, -C1};
  const float sn[8] = {0.f, S1, C2, C1, 1.f, C1, C2, S1};
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    float2 w;
    if (r == 0) w = o[0];
    else if (r == 4) w = rot<DIR>(o[4]);
    else w = make_float2(o[r].x * cs[r] - DIR * o[r].y * sn[r], o[r].y * cs[r] + DIR * o[r].x * sn[r]);
    x[r] = cadd(e[r], w);
    x[r + 8] = csub(e[r], w);
  }
}

template <int DIR>
__device__ __forceinline__ void dft3(float2* x) {
  constexpr float h = 0.86602540378443864676f;  // sqrt(3)/2
  const float2 s = cadd(x[1], x[2]);
  const float2 d = rot<DIR>(csub(x[1], x[2]));
  const float2 m = make_float2(x[0].x - 0.5f * s.x, x[0].y - 0.5f * s.y);
  x[0] = cadd(x[0], s);
  x[1] = make_float2(m.x + h * d.x, m.y + h * d.y);
  x[2] = make_float2(m.x - h * d.x, m.y - h * d.y);
}

template <int DIR>
__device__ __forceinline__ void dft5(float2* x) {
  constexpr float c1 = 0.30901699437494742410f;   // cos(2 pi/5)
  constexpr float c2 = -0.80901699437494742410f;  // cos(4 pi/5)
  constexpr float s1 = 0.95105651629515357212f;   // sin(2 pi/5)
  constexpr float s2 = 0.58778525229247312917f;   // sin(4 pi/5)
  const float2 a1 = cadd(x[1], x[4]), b1 = csub(x[1], x[4]);
  const float2 a2 = cadd(x[2], x[3]), b2 = csub(x[2], x[3]);
  const float2 x0 = x[0];
  const float2 m1 = make_float2(x0.x + c1 * a1.x + c2 * a2.x, x0.y + c1 * a1.y + c2 * a2.y);
  const float2 m2 = make_float2(x0.x + c2 * a1.x + c1 * a2.x, x0.y + c2 * a1.y + c1 * a2.y);
  const float2 n1 = rot<DIR>(make_float2(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y));
  const float2 n2 = rot<DIR>(make_float2(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y));
  x[0] = make_float2(x0.x + a1.x + a2.x, x0.y + a1.y + a2.y);
  x[1] = cadd(m1, n1);
  x[4] = csub(m1, n1);
  x[2] = cadd(m2, n2);
  x[3] = csub(m2, n2);
}

// X_m = x_0 + sum_k cos(2 pi m k/7) (x_k + x_{7-k})
//           + DIR i sum_k sin(2 pi m k/7) (x_k - x_{7-k}),  k = 1..3,
// and X_{7-m} the same with the sine terms negated.
template <int DIR>
__device__ __forceinline__ void dft7(float2* x) {
  constexpr float c1 = 0.62348980185873353053f;   // cos(2 pi/7)
  constexpr float c2 = -0.22252093395631440429f;  // cos(4 pi/7)
  constexpr float c3 = -0.90096886790241912624f;  // cos(6 pi/7)
  constexpr float s1 = 0.78183148246802980871f;   // sin(2 pi/7)
  constexpr float s2 = 0.97492791218182360702f;   // sin(4 pi/7)
  constexpr float s3 = 0.43388373911755812048f;   // sin(6 pi/7)
  const float2 a1 = cadd(x[1], x[6]), b1 = csub(x[1], x[6]);
  const float2 a2 = cadd(x[2], x[5]), b2 = csub(x[2], x[5]);
  const float2 a3 = cadd(x[3], x[4]), b3 = csub(x[3], x[4]);
  const float2 x0 = x[0];
  const float2 m1 = make_float2(x0.x + c1 * a1.x + c2 * a2.x + c3 * a3.x, x0.y + c1 * a1.y + c2 * a2.y + c3 * a3.y);
  const float2 m2 = make_float2(x0.x + c2 * a1.x + c3 * a2.x + c1 * a3.x, x0.y + c2 * a1.y + c3 * a2.y + c1 * a3.y);
  const float2 m3 = make_float2(x0.x + c3 * a1.x + c1 * a2.x + c2 * a3.x, x0.y + c3 * a1.y + c1 * a2.y + c2 * a3.y);
  const float2 n1 =
      rot<DIR>(make_float2(s1 * b1.x + s2 * b2.x + s3 * b3.x, s1 * b1.y + s2 * b2.y + s3 * b3.y));
  const float2 n2 =
      rot<DIR>(make_float2(s2 * b1.x - s3 * b2.x - s1 * b3.x, s2 * b1.y - s3 * b2.y - s1 * b3.y));
  const float2 n3 =
      rot<DIR>(make_float2(s3 * b1.x - s1 * b2.x + s2 * b3.x, s3 * b1.y - s1 * b2.y + s2 * b3.y));
  x[0] = make_float2(x0.x + a1.x + a2.x + a3.x, x0.y + a1.y + a2.y + a3.y);
  x[1] = cadd(m1, n1);
  x[6] = csub(m1, n1);
  x[2] = cadd(m2, n2);
  x[5] = csub(m2, n2);
  x[3] = cadd(m3, n3);
  x[4] = csub(m3, n3);
}

template <int R, int DIR>
__device__ __forceinline__ void dft(float2* x);

// (cos, sin) of 2 pi m / R, 0 < m < R, for the composite radices
template <int R>
__device__ __forceinline__ float2 unit_root(int m) {
  if constexpr (R == 6) {
    constexpr float c[6] = {1.f, 0.5f, -0.5f, -1.f, -0.5f, 0.5f};
    constexpr float s[6] = {0.f, 0.86602540378443864676f, 0.86602540378443864676f, 0.f, -0.86602540378443864676f,
                            -0.86602540378443864676f};
    return make_float2(c[m], s[m]);
  } else if constexpr (R == 9) {
    constexpr float c[9] = {1.f, 0.76604444311897803520f, 0.17364817766693034885f, -0.5f, -0.93969262078590838405f,
                            -0.93969262078590838405f, -0.5f, 0.17364817766693034885f, 0.76604444311897803520f};
    constexpr float s[9] = {0.f, 0.64278760968653932632f, 0.98480775301220805936f, 0.86602540378443864676f,
                            0.34202014332566873304f, -0.34202014332566873304f, -0.86602540378443864676f,
                            -0.98480775301220805936f, -0.64278760968653932632f};
    return make_float2(c[m], s[m]);
  } else if constexpr (R == 10) {
    constexpr float c[10] = {1.f, 0.80901699437494742410f, 0.30901699437494742410f, -0.30901699437494742410f,
                             -0.80901699437494742410f, -1.f, -0.80901699437494742410f, -0.30901699437494742410f,
                             0.30901699437494742410f, 0.80901699437494742410f};
    constexpr float s[10] = {0.f, 0.58778525229247312917f, 0.95105651629515357212f, 0.95105651629515357212f,
                             0.58778525229247312917f, 0.f, -0.58778525229247312917f, -0.95105651629515357212f,
                             -0.95105651629515357212f, -0.58778525229247312917f};
    return make_float2(c[m], s[m]);
  } else if constexpr (R == 12) {
    constexpr float c[12] = {1.f, 0.86602540378443864676f, 0.5f, 0.f, -0.5f, -0.86602540378443864676f, -1.f,
                             -0.86602540378443864676f, -0.5f, 0.f, 0.5f, 0.86602540378443864676f};
    constexpr float s[12] = {0.f, 0.5f, 0.86602540378443864676f, 1.f, 0.86602540378443864676f, 0.5f, 0.f, -0.5f,
                             -0.86602540378443864676f, -1.f, -0.86602540378443864676f, -0.5f};
    return make_float2(c[m], s[m]);
  } else {
    static_assert(R == 15, "no table for this radix");
    constexpr float c[15] = {1.f, 0.91354545764260089550f, 0.66913060635885821383f, 0.30901699437494742410f,
                             -0.10452846326765347140f, -0.5f, -0.80901699437494742410f, -0.97814760073380563793f,
                             -0.97814760073380563793f, -0.80901699437494742410f, -0.5f, -0.10452846326765347140f,
                             0.30901699437494742410f, 0.66913060635885821383f, 0.91354545764260089550f};
    constexpr float s[15] = {0.f, 0.40673664307580020775f, 0.74314482547739423501f, 0.95105651629515357212f,
                             0.99452189536827333692f, 0.86602540378443864676f, 0.58778525229247312917f,
                             0.20791169081775933710f, -0.20791169081775933710f, -0.58778525229247312917f,
                             -0.86602540378443864676f, -0.99452189536827333692f, -0.95105651629515357212f,
                             -0.74314482547739423501f, -0.40673664307580020775f};
    return make_float2(c[m], s[m]);
  }
}

// DFT of R = R1 R2 points in registers by one Cooley-Tukey step: R2-point
// DFTs of the R1 subsequences x[n1 + R1 n2], the twiddles
// exp(DIR 2 pi i n1 k2 / R), then R1-point DFTs; X[R2 k1 + k2] comes out in
// natural order.
template <int R1, int R2, int DIR>
__device__ __forceinline__ void dft_ct(float2* x) {
  constexpr int R = R1 * R2;
  float2 y[R];
#pragma unroll
  for (int n1 = 0; n1 < R1; ++n1) {
    float2 c[R2];
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) c[n2] = x[n1 + R1 * n2];
    dft<R2, DIR>(c);
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) {
      const int m = n1 * k2;
      if (m == 0) {
        y[n1 * R2 + k2] = c[k2];
      } else {
        const float2 w = unit_root<R>(m);
        y[n1 * R2 + k2] = make_float2(c[k2].x * w.x - DIR * c[k2].y * w.y, c[k2].y * w.x + DIR * c[k2].x * w.y);
      }
    }
  }
#pragma unroll
  for (int k2 = 0; k2 < R2; ++k2) {
    float2 r[R1];
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) r[n1] = y[n1 * R2 + k2];
    dft<R1, DIR>(r);
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) x[R2 * k1 + k2] = r[k1];
  }
}

template <int R, int DIR>
__device__ __forceinline__ void dft(float2* x) {
  if constexpr (R == 2) dft2<DIR>(x[0], x[1]);
  else if constexpr (R == 3) dft3<DIR>(x);
  else if constexpr (R == 4) dft4<DIR>(x[0], x[1], x[2], x[3]);
  else if constexpr (R == 5) dft5<DIR>(x);
  else if constexpr (R == 6) dft_ct<2, 3, DIR>(x);
  else if constexpr (R == 7) dft7<DIR>(x);
  else if constexpr (R == 8) dft8<DIR>(x);
  else if constexpr (R == 9) dft_ct<3, 3, DIR>(x);
  else if constexpr (R == 10) dft_ct<2, 5, DIR>(x);
  else if constexpr (R == 12) dft_ct<4, 3, DIR>(x);
  else if constexpr (R == 15) dft_ct<3, 5, DIR>(x);
  else if constexpr (R == 16) dft16<DIR>(x);
  else static_assert(R == 2, "unsupported radix");
}

// ---------------------------------------------------------------------------
// LDS views.  A work-group holds G lines of length L.  Row passes keep each
// line contiguous (index = g*LP + pos, LP = L + pad); column passes interleave
// the lines (index = pos*G + g) so that lanes of one wavefront, which walk g
// fastest, touch consecutive banks.
// ---------------------------------------------------------------------------
// One padding slot after every 16 positions: the stage-0 writes (stride =
// radix, up to 16) and the early middle-stage writes then spread over the
// banks instead of piling onto one.
// LDS for the line buffers of a work-group: gfx950 gives a work-group up to
// 160 KiB; 4 KiB stay for the kernels' reduction scratch
constexpr int kLdsBytes = 156 * 1024;

// Lines too long for two buffers of 4 lines each run their stages in one
// buffer: a stage loads all its inputs, synchronises, then stores (two
// barriers per stage instead of one), which halves the LDS and keeps 4 lines
// per work-group for the column passes and the tiled layout.
template <int L>
struct SingleBuf {
  static constexpr int LS = L + (L >> 4);
  static constexpr bool value = 2 * 4 * (LS + 1) * 8 > kLdsBytes;
  static constexpr int buffers = value ? 1 : 2;
};

template <int L, int G, bool INTERLEAVED>
struct LdsView {
  static constexpr int slot(int pos) { return pos + (pos >> 4); }
  static constexpr int LS = L + (L >> 4);            // padded line length
  static constexpr int LP = LS + ((LS % 2) ? 0 : 1);  // odd pitch between lines
  static constexpr int SIZE = INTERLEAVED ? LS * G : LP * G;
  float2* base;
  int g;
  __device__ __forceinline__ float2& at(int pos) const {
    if constexpr (INTERLEAVED) return base[slot(pos) * G + g];
    else return base[g * LP + slot(pos)];
  }
};

// Global twiddle table of plan (L, E): stage s >= 1 keeps exp(-2 pi i j k /
// (Ns R)) at twoff[s] + (k-1) Ns + j, so lanes with consecutive butterflies
// read consecutive entries.
template <int L, int E, int S>
__device__ __forceinline__ void load_tw_stage(const float2* __restrict__ tw, int t, TwRegs<L, E>& r) {
  constexpr FftPlan P = PlanOf<L, E>::P;
  if constexpr (S < P.n) {
    constexpr int R = P.r[S], NS = P.ns[S], NB = L / R, T = P.T;
    constexpr int BPT = PlanOf<L, E>::bpt(S), O = PlanOf<L, E>::twreg(S);
#pragma unroll
    for (int i = 0; i < BPT; ++i) {
      const int b = t + i * T;
      if (BPT * T == NB || b < NB) {
        const int j = b % NS;
#pragma unroll
        for (int k = 1; k < R; ++k) r.w[O + i * (R - 1) + (k - 1)] = tw[P.twoff[S] + (k - 1) * NS + j];
      }
    }
    load_tw_stage<L, E, S + 1>(tw, t, r);
  }
}

template <int L, int E>
__device__ __forceinline__ void load_tw(const float2* __restrict__ tw, int t, TwRegs<L, E>& r) {
  load_tw_stage<L, E, 1>(tw, t, r);
}

// Middle stage S: LDS `in` -> LDS `out` (no barrier inside).
template <int L, int E, int S, int DIR, class View>
__device__ __forceinline__ void mid_stage(const View& in, const View& out, int t, const TwRegs<L, E>& tr) {
  constexpr FftPlan P = PlanOf<L, E>::P;
  constexpr int R = P.r[S], NS = P.ns[S], T = P.T;
  constexpr int NB = L / R;
  constexpr int BPT = PlanOf<L, E>::bpt(S), O = PlanOf<L, E>::twreg(S);
#pragma unroll
  for (int i = 0; i < BPT; ++i) {
    const int b = t + i * T;
    if (BPT * T == NB || b < NB) {
      float2 x[R];
#pragma unroll
      for (int k = 0; k < R; ++k) x[k] = in.at(b + k * NB);
      const int j = b % NS;
#pragma unroll
      for (int k = 1; k < R; ++k) x[k] = twmul<DIR>(x[k], tr.w[O + i * (R - 1) + (k - 1)]);
      dft<R, DIR>(x);
      const int base = (b - j) * R + j;
#pragma unroll
      for (int r = 0; r < R; ++r) out.at(base + r * NS) = x[r];
    }
  }
}

// Middle stage S in one buffer: every input of the thread's butterflies is
// read, the work-group synchronises, then the outputs are written over them.
template <int L, int E, int S, int DIR, class View>
__device__ __forceinline__ void mid_stage_inplace(const View& buf, int t, const TwRegs<L, E>& tr) {
  constexpr FftPlan P = PlanOf<L, E>::P;
  constexpr int R = P.r[S], NS = P.ns[S], T = P.T;
  constexpr int NB = L / R;
  constexpr int BPT = PlanOf<L, E>::bpt(S), O = PlanOf<L, E>::twreg(S);
  float2 x[BPT][R];
#pragma unroll
  for (int i = 0; i < BPT; ++i) {
    const int b = t + i * T;
    if (BPT * T == NB || b < NB) {
#pragma unroll
      for (int k = 0; k < R; ++k) x[i][k] = buf.at(b + k * NB);
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < BPT; ++i) {
    const int b = t + i * T;
    if (BPT * T == NB || b < NB) {
      const int j = b % NS;
#pragma unroll
      for (int k = 1; k < R; ++k) x[i][k] = twmul<DIR>(x[i][k], tr.w[O + i * (R - 1) + (k - 1)]);
      dft<R, DIR>(x[i]);
      const int base = (b - j) * R + j;
#pragma unroll
      for (int r = 0; r < R; ++r) buf.at(base + r * NS) = x[i][r];
    }
  }
}

template <int L, int E, int S, int DIR, class View>
__device__ __forceinline__ void mid_stages(const View& a, const View& b, int t, const TwRegs<L, E>& tr) {
  constexpr FftPlan P = PlanOf<L, E>::P;
  if constexpr (SingleBuf<L>::value) {
    if constexpr (S < P.n - 1) {
      mid_stage_inplace<L, E, S, DIR>(a, t, tr);
      __syncthreads();
      mid_stages<L, E, S + 1, DIR>(a, b, t, tr);
    }
  } else if constexpr (S < P.n - 1) {
    // stage S reads the buffer stage S-1 wrote: a for odd S, b for even S
    if constexpr (S % 2 == 1) mid_stage<L, E, S, DIR>(a, b, t, tr);
    else mid_stage<L, E, S, DIR>(b, a, t, tr);
    __syncthreads();
    mid_stages<L, E, S + 1, DIR>(a, b, t, tr);
  }
}

// Full transform of one line.  On entry v[i*E0 + k] holds the element at
// in_pos(t, i, k) = (t + i T) + k L/E0; on exit v[i*EL + r] holds the output
// at out_pos(t, i, r) = (t + i T) + r L/EL.  Uses LDS buffers a and b; ends
// without a barrier (the last stage only reads LDS), so a caller that writes
// LDS next must synchronise first.
template <int L, int E, int DIR, class View>
__device__ __forceinline__ void fft_line(RegLine<L, E>& v, const View& a, const View& b, int t,
                                         const TwRegs<L, E>& tr) {
  constexpr FftPlan P = PlanOf<L, E>::P;
  constexpr int T = P.T;
  if constexpr (P.n == 1) {
    dft<L, DIR>(v);
  } else {
    constexpr int E0 = P.r[0], B0 = PlanOf<L, E>::B0;
    // stage 0: no twiddles (Ns = 1); outputs land at b*E0 + r
#pragma unroll
    for (int i = 0; i < B0; ++i) {
      dft<E0, DIR>(v + i * E0);
      const int bb = t + i * T;
#pragma unroll
      for (int r = 0; r < E0; ++r) a.at(bb * E0 + r) = v[i * E0 + r];
    }
    __syncthreads();
    mid_stages<L, E, 1, DIR>(a, b, t, tr);
    // last stage reads the buffer written by stage n-2
    constexpr int SL = P.n - 1;
    constexpr int EL = P.r[SL], NS = P.ns[SL], BL = PlanOf<L, E>::BL;
    const View& src = (SingleBuf<L>::value || SL % 2 == 1) ? a : b;
#pragma unroll
    for (int i = 0; i < BL; ++i) {
      const int bb = t + i * T;  // bb < NS: j = bb
      float2* x = v + i * EL;
#pragma unroll
      for (int k = 0; k < EL; ++k) x[k] = src.at(bb + k * NS);
      constexpr int O = PlanOf<L, E>::twreg(SL);
#pragma unroll
      for (int k = 1; k < EL; ++k) x[k] = twmul<DIR>(x[k], tr.w[O + i * (EL - 1) + (k - 1)]);
      dft<EL, DIR>(x);
    }
  }
}

// Moves the outputs of one transform to the input positions of the next
// when the plan's edge radices differ (a no-op for palindromic plans).
template <int L, int E, class View>
__device__ __forceinline__ void out_to_in(RegLine<L, E>& v, const View& a, int t) {
  if constexpr (!PlanOf<L, E>::PAL) {
    constexpr int T = PlanOf<L, E>::T, E0 = PlanOf<L, E>::E0, EL = PlanOf<L, E>::EL;
    constexpr int B0 = PlanOf<L, E>::B0, BL = PlanOf<L, E>::BL;
    __syncthreads();
#pragma unroll
    for (int i = 0; i < BL; ++i)
#pragma unroll
      for (int r = 0; r < EL; ++r) a.at((t + i * T) + r * (L / EL)) = v[i * EL + r];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < B0; ++i)
#pragma unroll
      for (int k = 0; k < E0; ++k) v[i * E0 + k] = a.at((t + i * T) + k * (L / E0));
  }
}

// positions of the edge registers
template <int L, int E>
__device__ __forceinline__ int in_pos(int t, int q) {
  constexpr int E0 = PlanOf<L, E>::E0, T = PlanOf<L, E>::T;
  return (t + (q / E0) * T) + (q % E0) * (L / E0);
}
template <int L, int E>
__device__ __forceinline__ int out_pos(int t, int q) {
  constexpr int EL = PlanOf<L, E>::EL, T = PlanOf<L, E>::T;
  return (t + (q / EL) * T) + (q % EL) * (L / EL);
}

}  // namespace spx
