// Phase-retrieval transfer function over many reconstructions, on the
// device: the computation of the reference's prtf() (src/_spimage_prtf.py:
// 17-101) with sp_image_superimpose (src/image_util.c:2973-3005),
// sp_image_phase_match (:3160-3186) and the unnormalised forward transform.
//
// Each reconstruction after the first is aligned against the running sum
// of the aligned ones, so the loop over reconstructions is sequential; each
// step is a handful of kernels on one stream (two transforms of the
// magnitudes, the correlation and convolution products and their inverse
// transforms, a first-index argmax, a gather that shifts and reflects, the
// phase match, the running sums), with no host round trip inside the loop:
// the shift and reflection the argmax picks stay in device memory and the
// gather reads them there.
#include <limits.h>
#include <math.h>
#include <string.h>

#include <vector>

#include "engine.h"
#include "sp_prtf.h"

namespace spx {
namespace {

constexpr int kThreads = 256;

inline dim3 grid_for(long n) {
  long b = (n + kThreads - 1) / kThreads;
  if (b > 2048) b = 2048;
  if (b < 1) b = 1;
  return dim3((unsigned)b);
}

#define PRTF_STRIDE(i, n) \
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < (n); i += (long)gridDim.x * blockDim.x)

// |in| as a real complex value (sp_image_dephase, src/image_util.c:1035-1042)
__global__ void k_magnitude(const float2* in, float2* out, long n) {
  PRTF_STRIDE(i, n) out[i] = make_float2(cabsf_ref(in[i]), 0.f);
}

// corr = A conj(B) (sp_image_cross_correlate, src/image_util.c:1356-1358);
// with conv, B becomes A B (sp_image_convolute, :1524-1526)
__global__ void k_corr_conv(const float2* A, float2* B, float2* corr, long n, int conv) {
  PRTF_STRIDE(i, n) {
    const float2 a = A[i], b = B[i];
    corr[i] = make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
    if (conv) B[i] = make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
  }
}

// larger magnitude wins; equal magnitudes go to the lower index, so the
// result is the first maximum in raster order (sp_c3matrix_max,
// include/spimage/linear_alg.h:1974-1988)
__device__ __forceinline__ void keep_best(float& v, int& i, float v2, int i2) {
  if (v2 > v || (v2 == v && i2 < i)) {
    v = v2;
    i = i2;
  }
}

__global__ __launch_bounds__(kThreads) void k_argmax_part(const float2* c, long n, float* pv, int* pi) {
  float v = -1.f;
  int idx = INT_MAX;
  PRTF_STRIDE(i, n) keep_best(v, idx, cabsf_ref(c[i]), (int)i);
  for (int off = 32; off > 0; off >>= 1) {
    const float v2 = __shfl_xor(v, off, 64);
    const int i2 = __shfl_xor(idx, off, 64);
    keep_best(v, idx, v2, i2);
  }
  __shared__ float sv[kThreads / 64];
  __shared__ int si[kThreads / 64];
  if ((threadIdx.x & 63) == 0) {
    sv[threadIdx.x >> 6] = v;
    si[threadIdx.x >> 6] = idx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < kThreads / 64; ++k) keep_best(v, idx, sv[k], si[k]);
    pv[blockIdx.x] = v;
    pi[blockIdx.x] = idx;
  }
}

// The superposition decision (src/image_util.c:2983-3004): the shift is the
// coordinates of the correlation peak, or of the convolution peak plus one
// when that is strictly larger, which also reflects.  dec = {x, y, z, reflect}.
__global__ __launch_bounds__(kThreads) void k_decide(const float* pv1, const int* pi1, const float* pv2,
                                                    const int* pi2, int nparts, int enantio, int nx, int ny,
                                                    int* dec) {
  // (value, index) pairs under keep_best form a total order, so the tree
  // below finds the same first maximum as a scan in raster order
  __shared__ float sv[2][kThreads];
  __shared__ int si[2][kThreads];
  const int tid = threadIdx.x;
  float v1 = -1.f, v2 = -1.f;
  int i1 = INT_MAX, i2 = INT_MAX;
  for (int k = tid; k < nparts; k += kThreads) {
    keep_best(v1, i1, pv1[k], pi1[k]);
    if (enantio) keep_best(v2, i2, pv2[k], pi2[k]);
  }
  sv[0][tid] = v1;
  si[0][tid] = i1;
  sv[1][tid] = v2;
  si[1][tid] = i2;
  __syncthreads();
  for (int off = kThreads / 2; off > 0; off >>= 1) {
    if (tid < off) {
      keep_best(sv[0][tid], si[0][tid], sv[0][tid + off], si[0][tid + off]);
      keep_best(sv[1][tid], si[1][tid], sv[1][tid + off], si[1][tid + off]);
    }
    __syncthreads();
  }
  if (tid != 0) return;
  int idx = si[0][0], add = 0;
  if (enantio && sv[1][0] > sv[0][0]) {
    idx = si[1][0];
    add = 1;
  }
  dec[0] = idx % nx + add;
  dec[1] = (idx / nx) % ny + add;
  dec[2] = idx / (nx * ny) + add;
  dec[3] = add;
}

__device__ __forceinline__ int pmod(int v, int m) { return ((v % m) + m) % m; }

// out(q) = reflect(img)(q - d), mout(q) = mask(q - d), cyclic
// (sp_image_translate with SP_TRANSLATE_WRAP_AROUND, src/image_util.c:3188-
// 3215, after sp_image_reflect(.., SP_ORIGO), :161-175, which moves the
// image but not the mask).  In 3-D the reflection leaves the middle plane
// of an odd number of planes in place, as the reference's in-place swap loop
// does (src/image_util.c:290-306).
// also counts, where cnt is given, the pixels outside the aligned mask
// (the k_count_outside step of a reconstruction with a mask)
__global__ void k_align(const float2* img, const int* mask, float2* out, int* mout, float* cnt, int nx, int ny, int nz,
                        const int* dec) {
  const int dx = dec[0], dy = dec[1], dz = dec[2], refl = dec[3];
  const long n = (long)nx * ny * nz;
  PRTF_STRIDE(q, n) {
    const int x = (int)(q % nx);
    const long r = q / nx;
    const int y = (int)(r % ny), z = (int)(r / ny);
    const int px = pmod(x - dx, nx), py = pmod(y - dy, ny), pz = pmod(z - dz, nz);
    const long p = ((long)pz * ny + py) * nx + px;
    const int mv = mask ? mask[p] : 0;
    if (mout) mout[q] = mv;
    if (cnt && mv == 0) cnt[q] += 1.f;
    long s = p;
    if (refl && !(nz > 1 && (nz & 1) && pz == nz / 2)) s = ((long)(nz - 1 - pz) * ny + (ny - 1 - py)) * nx + (nx - 1 - px);
    out[q] = img[s];
  }
}

// sum over pixels of |a|^2 exp(i (arg a - arg b)) (sp_image_phase_match with
// weighted = 2, src/image_util.c:3160-3177), in double
__global__ __launch_bounds__(kThreads) void k_phase_part(const float2* a, const float2* b, long n, double2* part) {
  double sx = 0.0, sy = 0.0;
  PRTF_STRIDE(i, n) {
    const float2 av = a[i], bv = b[i];
    const float ang = atan2f(av.y, av.x) - atan2f(bv.y, bv.x);
    const double m = (double)(av.y * av.y + av.x * av.x);
    sx += cos((double)ang) * m;
    sy += sin((double)ang) * m;
  }
  for (int off = 32; off > 0; off >>= 1) {
    sx += __shfl_xor(sx, off, 64);
    sy += __shfl_xor(sy, off, 64);
  }
  __shared__ double s[2][kThreads / 64];
  if ((threadIdx.x & 63) == 0) {
    s[0][threadIdx.x >> 6] = sx;
    s[1][threadIdx.x >> 6] = sy;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < kThreads / 64; ++k) {
      sx += s[0][k];
      sy += s[1][k];
    }
    part[blockIdx.x] = make_double2(sx, sy);
  }
}

// b = |b| exp(i (arg b + phi)), phi the angle of the summed partials
// (src/image_util.c:3178-3184)
// also adds the result to the running sum avg (the k_add_c64 step)
__global__ __launch_bounds__(kThreads) void k_phase_apply(const float2* b, float2* out, float2* avg, long n,
                                                          const double2* part, int nparts) {
  // every block sums the partials in the same fixed order
  __shared__ double sx[kThreads], sy[kThreads];
  const int tid = threadIdx.x;
  double x = 0.0, y = 0.0;
  for (int k = tid; k < nparts; k += kThreads) {
    x += part[k].x;
    y += part[k].y;
  }
  sx[tid] = x;
  sy[tid] = y;
  __syncthreads();
  for (int off = kThreads / 2; off > 0; off >>= 1) {
    if (tid < off) {
      sx[tid] += sx[tid + off];
      sy[tid] += sy[tid + off];
    }
    __syncthreads();
  }
  const float phi = (float)atan2(sy[0], sx[0]);
  PRTF_STRIDE(i, n) {
    const float2 v = b[i];
    const float ang = atan2f(v.y, v.x) + phi;
    const double m = (double)cabsf_ref(v);
    const float2 r = make_float2((float)(cos((double)ang) * m), (float)(sin((double)ang) * m));
    out[i] = r;
    avg[i] = make_float2(avg[i].x + r.x, avg[i].y + r.y);
  }
}

__global__ void k_add_c64(float2* acc, const float2* b, long n) {
  PRTF_STRIDE(i, n) acc[i] = make_float2(acc[i].x + b[i].x, acc[i].y + b[i].y);
}

// running sum of F / |F| where F != 0, and a count of the zeros
// (src/_spimage_prtf.py:80-86)
__global__ void k_unit_accum(const float2* F, float2* sum, int* zeros, long n, int first) {
  PRTF_STRIDE(i, n) {
    const float2 f = F[i];
    float2 u = make_float2(0.f, 0.f);
    int z = 0;
    if (f.x != 0.f || f.y != 0.f) {
      const float m = hypotf(f.x, f.y);
      u = make_float2(f.x / m, f.y / m);
    } else {
      z = 1;
    }
    if (first) {
      sum[i] = u;
      zeros[i] = z;
    } else {
      sum[i] = make_float2(sum[i].x + u.x, sum[i].y + u.y);
      zeros[i] += z;
    }
  }
}

// pixels outside this reconstruction's mask (src/_spimage_prtf.py:72-73)
__global__ void k_count_outside(const int* m, float* cnt, long n) {
  PRTF_STRIDE(i, n) if (m[i] == 0) cnt[i] += 1.f;
}

// |sum / n|, zero where any transform was zero (src/_spimage_prtf.py:91-94)
__global__ void k_prtf_final(const float2* sum, const int* zeros, float* out, long n, float nimg) {
  PRTF_STRIDE(i, n) {
    const float2 s = make_float2(sum[i].x / nimg, sum[i].y / nimg);
    out[i] = zeros[i] > 0 ? 0.f : cabsf_ref(s);
  }
}

__global__ void k_fill_f32(float* p, long n, float v) { PRTF_STRIDE(i, n) p[i] = v; }

// device buffers of one call, released on every path
struct Buffers {
  std::vector<void*> ptrs;
  hipStream_t s = nullptr;
  ~Buffers() {
    if (s) (void)hipStreamSynchronize(s);
    for (void* p : ptrs) (void)hipFree(p);
    if (s) (void)hipStreamDestroy(s);
  }
  template <class T>
  int alloc(T** p, size_t count) {
    *p = nullptr;
    if (hipMalloc(p, sizeof(T) * (count ? count : 1)) != hipSuccess) {
      set_error("spx_prtf: device allocation of %zu bytes failed", sizeof(T) * count);
      return -20;
    }
    ptrs.push_back(*p);
    return 0;
  }
};

int prtf_impl(const float* images, const int* masks, int nx, int ny, int nz, int n, int translate, int enantio,
              float* prtf, float* super_image, float* aligned, int* aligned_masks, float* mask_count) {
  if (!images || !prtf || n < 1 || nx < 1 || ny < 1 || nz < 1) {
    set_error("spx_prtf: images, prtf, n >= 1 and positive sizes are required");
    return -1;
  }
  if (!device_available()) {
    set_error("no HIP device available");
    return -10;
  }
  Shape sh;
  if (int rc = make_shape(nx, ny, nz, 1, &sh)) return rc;
  const long N = (long)nx * ny * nz;
  Buffers b;
  SPX_HIP(hipStreamCreateWithFlags(&b.s, hipStreamNonBlocking));
  hipStream_t s = b.s;
  const int nparts = (int)grid_for(N).x;
  float2 *imgs = nullptr, *avg = nullptr, *t1 = nullptr, *t2 = nullptr, *t3 = nullptr, *sum = nullptr;
  int *msk = nullptr, *mout = nullptr, *zeros = nullptr, *pi = nullptr, *dec = nullptr;
  float *cnt = nullptr, *pv = nullptr;
  double2* ph = nullptr;
  if (b.alloc(&imgs, (size_t)n * N) || b.alloc(&avg, N) || b.alloc(&t1, N) || b.alloc(&t2, N) ||
      b.alloc(&t3, N) || b.alloc(&sum, N) || b.alloc(&mout, N) || b.alloc(&zeros, N) || b.alloc(&cnt, N) ||
      b.alloc(&pv, 2 * nparts) || b.alloc(&pi, 2 * nparts) || b.alloc(&dec, 4) || b.alloc(&ph, nparts))
    return -20;
  if (masks && b.alloc(&msk, (size_t)n * N)) return -20;
  SPX_HIP(hipMemcpyAsync(imgs, images, sizeof(float2) * n * N, hipMemcpyHostToDevice, s));
  if (masks) SPX_HIP(hipMemcpyAsync(msk, masks, sizeof(int) * n * N, hipMemcpyHostToDevice, s));
  // the mask sums decide which reconstructions count in mask_count; a
  // translation keeps the sum
  std::vector<char> has_mask((size_t)n, 0);
  if (masks)
    for (int i = 0; i < n; ++i) {
      long tot = 0;
      for (long k = 0; k < N; ++k) tot += masks[(long)i * N + k];
      has_mask[i] = tot > 0;
    }
  const dim3 g = grid_for(N), blk(kThreads);
  hipLaunchKernelGGL(k_fill_f32, g, blk, 0, s, cnt, N, 0.f);
  SPX_HIP(hipMemcpyAsync(avg, imgs, sizeof(float2) * N, hipMemcpyDeviceToDevice, s));
  enqueue_fft(sh, imgs, t1, -1, s);
  hipLaunchKernelGGL(k_unit_accum, g, blk, 0, s, t1, sum, zeros, N, 1);
  for (int i = 1; i < n; ++i) {
    float2* img = imgs + (long)i * N;
    int* im = msk ? msk + (long)i * N : nullptr;
    const float2* cur = img;
    const int* curm = im;
    if (translate) {
      if (sh.ex) {
        // the magnitudes taken in the transforms' x pass
        enqueue_abs_fft(sh, avg, t1, s);
        enqueue_abs_fft(sh, img, t2, s);
      } else {
        hipLaunchKernelGGL(k_magnitude, g, blk, 0, s, avg, t1, N);
        enqueue_fft(sh, t1, t1, -1, s);
        hipLaunchKernelGGL(k_magnitude, g, blk, 0, s, img, t2, N);
        enqueue_fft(sh, t2, t2, -1, s);
      }
      hipLaunchKernelGGL(k_corr_conv, g, blk, 0, s, t1, t2, t3, N, enantio ? 1 : 0);
      enqueue_fft(sh, t3, t3, +1, s);
      hipLaunchKernelGGL(k_argmax_part, g, blk, 0, s, t3, N, pv, pi);
      if (enantio) {
        enqueue_fft(sh, t2, t2, +1, s);
        hipLaunchKernelGGL(k_argmax_part, g, blk, 0, s, t2, N, pv + nparts, pi + nparts);
      }
      hipLaunchKernelGGL(k_decide, dim3(1), dim3(kThreads), 0, s, pv, pi, pv + nparts, pi + nparts, nparts,
                         enantio ? 1 : 0, nx, ny, dec);
      hipLaunchKernelGGL(k_align, g, blk, 0, s, img, im, t1, mout, (im && has_mask[i]) ? cnt : nullptr, nx, ny, nz,
                         dec);
      hipLaunchKernelGGL(k_phase_part, g, blk, 0, s, avg, t1, N, ph);
      // the aligned, phase-matched reconstruction replaces the input in
      // place and joins the running sum
      hipLaunchKernelGGL(k_phase_apply, g, blk, 0, s, t1, img, avg, N, ph, nparts);
      if (im) SPX_HIP(hipMemcpyAsync(im, mout, sizeof(int) * N, hipMemcpyDeviceToDevice, s));
    } else {
      hipLaunchKernelGGL(k_add_c64, g, blk, 0, s, avg, cur, N);
      if (curm && has_mask[i]) hipLaunchKernelGGL(k_count_outside, g, blk, 0, s, curm, cnt, N);
    }
    enqueue_fft(sh, cur, t2, -1, s);
    hipLaunchKernelGGL(k_unit_accum, g, blk, 0, s, t2, sum, zeros, N, 0);
  }
  float* dprtf = reinterpret_cast<float*>(t3);
  hipLaunchKernelGGL(k_prtf_final, g, blk, 0, s, sum, zeros, dprtf, N, (float)n);
  SPX_HIP(hipGetLastError());
  SPX_HIP(hipMemcpyAsync(prtf, dprtf, sizeof(float) * N, hipMemcpyDeviceToHost, s));
  if (super_image) SPX_HIP(hipMemcpyAsync(super_image, avg, sizeof(float2) * N, hipMemcpyDeviceToHost, s));
  if (aligned) SPX_HIP(hipMemcpyAsync(aligned, imgs, sizeof(float2) * n * N, hipMemcpyDeviceToHost, s));
  if (aligned_masks) {
    if (msk) SPX_HIP(hipMemcpyAsync(aligned_masks, msk, sizeof(int) * n * N, hipMemcpyDeviceToHost, s));
    else memset(aligned_masks, 0, sizeof(int) * n * N);
  }
  if (mask_count) SPX_HIP(hipMemcpyAsync(mask_count, cnt, sizeof(float) * N, hipMemcpyDeviceToHost, s));
  SPX_HIP(hipStreamSynchronize(s));
  return 0;
}

}  // namespace
}  // namespace spx

extern "C" int spx_prtf(const float* images, const int* masks, int nx, int ny, int nz, int n, int translate,
                        int enantio, float* prtf, float* super_image, float* aligned, int* aligned_masks,
                        float* mask_count) {
  return spx::prtf_impl(images, masks, nx, ny, nz, n, translate, enantio, prtf, super_image, aligned, aligned_masks,
                        mask_count);
}
