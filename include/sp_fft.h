/* Image-level FFT and Gaussian blur, computed on the GPU by the engine's own
 * Stockham transforms.
 *
 * The reference maps sp_image_fft/ifft(_fast) to FFTW wrappers
 * (include/spimage/fft.h:25-28,87-115; src/fft.c:53-99,152-189) and offers
 * host-to-device one-shots sp_image_cuda_fft/ifft (include/spimage/fft.h:129-132,
 * src/fft.cu:7-43).  All of those names are exported here and share one
 * implementation: unnormalised c2c over (z,y,x), forward sign -1.
 * Supported lengths per axis: products of 2, 3, 5 and 7 up to 8192 (Stockham
 * passes; 5250, 6250, 7350, 7560 and 7840 excepted), and any other length up
 * to 4096 (Bluestein's method over a power-of-two Stockham transform);
 * beyond that the functions print an error and return NULL / leave the
 * output untouched.
 */
#ifndef SPX_SP_FFT_H
#define SPX_SP_FFT_H 1

#include "sp_core.h"

#ifdef __cplusplus
extern "C" {
#endif

/* new image = FFT(img); result has shifted=1, phased=1, zero mask and
 * image_center = n/2 (reference src/fft.c:152-175) */
spimage_EXPORT Image *sp_image_fft(const Image *img);
/* new image = unnormalised inverse FFT(img); aborts unless img is phased and
 * shifted, like the reference (src/fft.c:53-85) */
spimage_EXPORT Image *sp_image_ifft(const Image *img);
/* out = FFT(in) / IFFT(in), raw data only (reference src/fft.c:87-99,177-189) */
spimage_EXPORT void sp_image_fft_fast(const Image *in, Image *out);
spimage_EXPORT void sp_image_ifft_fast(const Image *in, Image *out);

/* binary aliases of the reference's macro targets (include/spimage/fft.h:25-28)
 * and of its CUDA one-shots (include/spimage/fft.h:129-132) */
spimage_EXPORT Image *sp_image_fftw3(const Image *img);
spimage_EXPORT Image *sp_image_ifftw3(const Image *img);
spimage_EXPORT void sp_image_fftw3_fast(const Image *in, Image *out);
spimage_EXPORT void sp_image_ifftw3_fast(const Image *in, Image *out);
spimage_EXPORT Image *sp_image_cuda_fft(const Image *img);
spimage_EXPORT Image *sp_image_cuda_ifft(const Image *img);

/* Gaussian blur by multiplication with exp(-2 pi^2 (r/nx)^2 radius^2)/N in
 * Fourier space (reference src/image_filter.c:80-91) */
spimage_EXPORT Image *sp_gaussian_blur(Image *in, real radius);

#ifdef __cplusplus
}
#endif
#endif
