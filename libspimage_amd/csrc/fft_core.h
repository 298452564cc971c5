// Device building blocks of the LDS-resident Stockham FFT used by every pass
// kernel of the phasing engine.
//
// Conventions follow the reference's FFTW3 calls (src/fft.c:79,95,163,185):
// unnormalised complex-to-complex transforms, forward sign -1
// (X_k = sum_n x_n e^{-2 pi i k n / L}), inverse sign +1, row-major with x
// fastest.  A length-L transform is split into radix stages (16, 8, 4, 2, 3, 5)
// chosen at compile time; each stage reads and writes the line through LDS in
// Stockham (self-sorting) order, so no bit-reversal pass exists.
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
  int T;          // threads cooperating on one line = L / maxr
  bool ok;        // L factors into the supported radices
};

constexpr int ilog2c(int v) {
  int e = 0;
  while ((1 << e) < v) ++e;
  return e;
}

constexpr FftPlan make_plan(int L) {
  FftPlan p{};
  p.L = L;
  p.n = 0;
  p.ok = L >= 2;
  int rem = L;
  while (rem % 5 == 0) { p.r[p.n++] = 5; rem /= 5; }
  while (rem % 3 == 0) { p.r[p.n++] = 3; rem /= 3; }
  if ((rem & (rem - 1)) != 0) p.ok = false;
  int e = ilog2c(rem);
  // split the power-of-two part into radix 16/8/4/2 stages, largest first,
  // never leaving a lone radix-2 behind a radix-16 when two radix-8 fit.
  while (e > 0) {
    int take = 0;
    if (e >= 10 || e == 4 || e == 7 || e == 8) take = 4;   // 2^10 = 16*8*8
    else if (e >= 3) take = 3;                              // 2^9 = 8*8*8, 2^5 = 8*4
    else take = e;
    p.r[p.n++] = 1 << take;
    e -= take;
  }
  int ns = 1, off = 0, mr = 1;
  for (int s = 0; s < p.n; ++s) {
    p.ns[s] = ns;
    p.twoff[s] = off;
    off += ns * (p.r[s] - 1);
    ns *= p.r[s];
    if (p.r[s] > mr) mr = p.r[s];
  }
  p.twtotal = off;
  p.maxr = mr;
  p.T = L / mr;
  if (ns != L) p.ok = false;
  return p;
}

template <int L>
struct PlanOf {
  static constexpr FftPlan P = make_plan(L);
  static_assert(P.ok, "transform length does not factor into radices 2,3,4,5,8,16");
  static constexpr int T = P.T;
};

// ---------------------------------------------------------------------------
// Complex helpers on float2 (x = re, y = im).
// ---------------------------------------------------------------------------
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cscl(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
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
  float2 t = a;
  a = cadd(t, b);
  b = csub(t, b);
}

template <int DIR>
__device__ __forceinline__ void dft4(float2& x0, float2& x1, float2& x2, float2& x3) {
  float2 t0 = cadd(x0, x2), t1 = csub(x0, x2);
  float2 t2 = cadd(x1, x3), t3 = rot<DIR>(csub(x1, x3));
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
  float2 w1 = make_float2(c * (o1.x - DIR * o1.y), c * (o1.y + DIR * o1.x));
  float2 w2 = rot<DIR>(o2);
  float2 w3 = make_float2(c * (-o3.x - DIR * o3.y), c * (-o3.y + DIR * o3.x));
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
  float2 s = cadd(x[1], x[2]);
  float2 d = rot<DIR>(csub(x[1], x[2]));
  float2 m = make_float2(x[0].x - 0.5f * s.x, x[0].y - 0.5f * s.y);
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
  float2 a1 = cadd(x[1], x[4]), b1 = csub(x[1], x[4]);
  float2 a2 = cadd(x[2], x[3]), b2 = csub(x[2], x[3]);
  float2 x0 = x[0];
  float2 m1 = make_float2(x0.x + c1 * a1.x + c2 * a2.x, x0.y + c1 * a1.y + c2 * a2.y);
  float2 m2 = make_float2(x0.x + c2 * a1.x + c1 * a2.x, x0.y + c2 * a1.y + c1 * a2.y);
  float2 n1 = rot<DIR>(make_float2(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y));
  float2 n2 = rot<DIR>(make_float2(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y));
  x[0] = make_float2(x0.x + a1.x + a2.x, x0.y + a1.y + a2.y);
  x[1] = cadd(m1, n1);
  x[4] = csub(m1, n1);
  x[2] = cadd(m2, n2);
  x[3] = csub(m2, n2);
}

template <int R, int DIR>
__device__ __forceinline__ void dft(float2* x) {
  if constexpr (R == 2) dft2<DIR>(x[0], x[1]);
  else if constexpr (R == 3) dft3<DIR>(x);
  else if constexpr (R == 4) dft4<DIR>(x[0], x[1], x[2], x[3]);
  else if constexpr (R == 5) dft5<DIR>(x);
  else if constexpr (R == 8) dft8<DIR>(x);
  else if constexpr (R == 16) dft16<DIR>(x);
  else static_assert(R == 2, "unsupported radix");
}

// ---------------------------------------------------------------------------
// LDS views.  A work-group holds G lines of length L.  Row passes keep each
// line contiguous (index = g*LP + pos, LP = L + pad); column passes interleave
// the lines (index = pos*G + g) so that lanes of one wavefront, which walk g
// fastest, touch consecutive banks.
// ---------------------------------------------------------------------------
template <int L, int G, bool INTERLEAVED>
struct LdsView {
  static constexpr int LP = L + ((L >= 32) ? 1 : 0);  // odd pitch breaks power-of-two bank strides
  static constexpr int SIZE = INTERLEAVED ? L * G : LP * G;
  float2* base;
  int g;
  __device__ __forceinline__ float2& at(int pos) const {
    if constexpr (INTERLEAVED) return base[pos * G + g];
    else return base[g * LP + pos];
  }
};

// One Stockham radix-R stage over a line held in LDS.  `t` is the thread's
// index among the T threads of its line.  Reads happen before the barrier,
// writes after it, so the stage runs in place.
template <int L, int R, int NS, int T, int DIR, int TWOFF, class View>
__device__ __forceinline__ void stockham_stage(const View& v, int t, const float2* __restrict__ tw) {
  constexpr int NB = L / R;
  constexpr int BPT = (NB + T - 1) / T;
  float2 x[BPT][R];
#pragma unroll
  for (int i = 0; i < BPT; ++i) {
    const int b = t + i * T;
    if (BPT * T == NB || b < NB) {
#pragma unroll
      for (int k = 0; k < R; ++k) x[i][k] = v.at(b + k * NB);
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < BPT; ++i) {
    const int b = t + i * T;
    if (BPT * T == NB || b < NB) {
      const int j = (NS == 1) ? 0 : (b % NS);
      if constexpr (NS > 1) {
        const float2* w = tw + TWOFF + j * (R - 1);
#pragma unroll
        for (int k = 1; k < R; ++k) x[i][k] = twmul<DIR>(x[i][k], w[k - 1]);
      }
      dft<R, DIR>(x[i]);
      const int base = (b - j) * R + j;
#pragma unroll
      for (int r = 0; r < R; ++r) v.at(base + r * NS) = x[i][r];
    }
  }
  __syncthreads();
}

template <int L, int DIR, int S, class View>
__device__ __forceinline__ void run_stages(const View& v, int t, const float2* __restrict__ tw) {
  constexpr FftPlan P = PlanOf<L>::P;
  if constexpr (S < P.n) {
    stockham_stage<L, P.r[S], P.ns[S], P.T, DIR, P.twoff[S]>(v, t, tw);
    run_stages<L, DIR, S + 1>(v, t, tw);
  }
}

// Full length-L transform of the line in LDS (natural order in, natural order
// out).  Caller must have synchronised after filling LDS; the last stage ends
// with a barrier so the result may be read immediately.
template <int L, int DIR, class View>
__device__ __forceinline__ void fft_line(const View& v, int t, const float2* __restrict__ tw) {
  run_stages<L, DIR, 0>(v, t, tw);
}

}  // namespace spx
