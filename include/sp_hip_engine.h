/* Plain-pointer C ABI of the HIP phasing engine (device buffers, sizes,
 * return codes; no library types beyond SpPhaser).  These entry points stand
 * where the reference's CUDA glue sits:
 *
 *   spx_device_available      sp_cuda_get_device_type
 *                             (reference include/spimage/cuda_util.h:101-109, src/cuda_util.c:10-29)
 *   spx_fft_device            the cufftPlan/cufftExecC2C call sites
 *                             (reference src/phasing.cu:47-221, src/phasing.c:716-720, src/fft.cu:14-38)
 *   spx_phaser_time_iterations  the reference's speed tests
 *                             (reference tests/PhasingTests.c:1068-1217), timed with HIP events
 *                             on the engine's stream
 *
 * Every function returns 0 on success and a negative code on error; the text
 * of the last error is available from spx_last_error().
 */
#ifndef SPX_SP_HIP_ENGINE_H
#define SPX_SP_HIP_ENGINE_H 1

#include "sp_phasing.h"

#ifdef __cplusplus
extern "C" {
#endif

/* 1 when a HIP device is usable, 0 otherwise */
spimage_EXPORT int spx_device_available(void);
/* Free and total bytes of device memory on the library's device
 * (hipMemGetInfo). */
spimage_EXPORT int spx_device_memory(unsigned long long *free_bytes, unsigned long long *total_bytes);
spimage_EXPORT const char *spx_last_error(void);
/* Page-locked host memory (hipHostMalloc) for staging large downloads,
 * which then move at the full link rate; NULL when it cannot be allocated.
 * No counterpart in the reference. */
spimage_EXPORT void *spx_host_alloc(unsigned long long bytes);
spimage_EXPORT void spx_host_free(void *p);
/* 1 when a line of length L can be transformed by the engine */
spimage_EXPORT int spx_supported_length(int L);
/* Tunes arrays of shape (nz, ny, nx) stacked nrec times as the first engine
 * or batch of that shape and size class would (variants per shape; layout
 * and column-pass width per class of nrec: 1, 2-3, 4-7, 8 or more), so that
 * later engines and batches start without timing runs.  A server calls it
 * for the shapes and batch sizes it expects.  With the environment variable
 * SPX_TUNE_FILE=<path>, every timed choice is appended to that file as a
 * line of text, and a process reads the file before its first choice, so
 * that a restarted server (or a profiler run) makes the same choices
 * without timing anything.  0, or -11 for an unsupported shape, -20 without
 * a device. */
spimage_EXPORT int spx_tune_shape(int nx, int ny, int nz, int nrec);
/* Elements per thread (4, 8 or 16) of the transform variants the tuner chose
 * for L x L arrays in row passes and in column passes (tuning the shape if
 * no engine has used it yet).  One choice per shape serves both layouts of
 * the Fourier-space arrays; `tiled` is kept for callers that ask per layout. */
spimage_EXPORT int spx_fft_variant(int L, int tiled, int *row_elements, int *col_elements);

/* 1 when the phaser's engine keeps its Fourier-space work arrays (the
 * row-transformed iterate, amplitudes, margins, phased amplitudes) in the
 * 4-column tiled layout, 0 when row-major, 2 when they stay row-major and
 * the Fourier-space pass of plain iterations runs on a transposed copy
 * (large 2-D arrays), negative without an engine.  The iterates, supports
 * and every array crossing this ABI stay row-major. */
spimage_EXPORT int spx_phaser_tiled(SpPhaser *ph);

/* Which streams of the phaser's passes carry the non-temporal hint, as the
 * tuner chose them for its shape and batch class (or SPX_NT forced them):
 * bit 0 Fourier-space work-array loads, bit 1 its stores, bit 2 loads of
 * the current iterate, bit 3 stores of the next one.  Negative without an
 * engine.  The hints change timings only, never results. */
spimage_EXPORT int spx_phaser_nt_hints(SpPhaser *ph);

/* Unnormalised c2c transform of an (nz, ny, nx) x-fastest complex64 array in
 * device memory; direction -1 forward, +1 inverse; d_in == d_out allowed.
 * Runs on the library's stream and returns after completion. */
spimage_EXPORT int spx_fft_device(void *d_in, void *d_out, int nx, int ny, int nz, int direction);

/* The unnormalised transform of sp_image_fft_fast / sp_image_ifft_fast
 * (direction -1 / +1) on a host array of nx*ny*nz interleaved complex
 * floats, for callers that cannot allocate device memory (harnesses that
 * forbid hipMalloc, serving loops that must not allocate).  No counterpart
 * in the reference.  It allocates nothing: the array is staged through a
 * fixed device array of 2^21 complex values (16 MiB, plus 96 KiB of twiddle
 * tables) that each device holds for the life of the process, so larger
 * arrays are rejected with -12; it uses the default transform plans, not
 * tuned ones, and takes axis lengths that are products of 2, 3, 5 and 7 up
 * to 4096 (-11 otherwise: the Bluestein tables of other lengths are
 * allocated).
 * It synchronises before returning.  Calls on one device share the fixed
 * array: calls that overlap on one device race. */
spimage_EXPORT int spx_fft_noalloc(const float *in, float *out, int nx, int ny, int nz, int direction);

/* Run `iterations` phasing iterations of the phaser's current algorithm
 * (no support updates) and report the elapsed device time in milliseconds,
 * measured with HIP events on the engine's stream. */
spimage_EXPORT int spx_phaser_time_iterations(SpPhaser *ph, int iterations, float *ms);

/* The same iterations with HIP events around each pass: total milliseconds
 * spent in the Fourier-space passes (column transforms with the modulus
 * projection) and in the real-space passes (row transforms with the
 * ER/HIO/RAAR/DM update). */
spimage_EXPORT int spx_phaser_time_passes(SpPhaser *ph, int iterations, float *ms_fourier, float *ms_real);

/* Device addresses of the phaser's current iterate g1, previous iterate g0
 * and model before projection gp (complex64, image_size elements each). */
spimage_EXPORT int spx_phaser_device_arrays(SpPhaser *ph, void **g1, void **g0, void **gp);

/* Threshold applied by the phaser's last area / threshold / Patterson
 * support update (the value the reference computes at
 * src/support_update.c:144,175 and src/phasing.c:754). */
spimage_EXPORT int spx_phaser_last_threshold(SpPhaser *ph, float *threshold);

/* Seed of the engine's counter-based generator used by
 * sp_phaser_init_model(..., SpModelRandomPhases / SpModelRandomValues). */
spimage_EXPORT void spx_phaser_set_seed(SpPhaser *ph, unsigned long long seed);

#ifdef __cplusplus
}
#endif
#endif
