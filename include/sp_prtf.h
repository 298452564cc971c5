/* Phase-retrieval transfer function (PRTF) over many reconstructions of one
 * array shape, on the GPU.
 *
 * spx_prtf computes what the reference's Python prtf() computes
 * (reference src/_spimage_prtf.py:17-101) with its libspimage calls
 * (sp_image_superimpose, src/image_util.c:2973-3005; sp_image_phase_match,
 * :3160-3186; sp_image_fftw3): reconstruction 0 starts the running sum;
 * each later one is moved onto the running sum (integer cyclic shift from
 * the peak of the magnitudes' cross-correlation or, when larger, of their
 * convolution, which also reflects it through the origin), given the
 * constant phase that matches it to the sum (weights |sum|^2), added to the
 * sum, and its Fourier transform's unit phasors are accumulated.  The PRTF is
 * |sum of unit phasors| / n, zero wherever one of the transforms is zero.
 *
 * Host arrays are (nz, ny, nx) row-major with x fastest; complex values are
 * interleaved float pairs.  Returns 0, or a negative code with
 * spx_last_error() set.  The work stays on the device from the upload of the
 * inputs to the download of the results.
 */
#ifndef SPX_SP_PRTF_H
#define SPX_SP_PRTF_H 1

#include "sp_core.h"

#ifdef __cplusplus
extern "C" {
#endif

/* images: n * N complex; masks: n * N ints (the supports) or NULL;
 * translate: align (superimpose + phase match) as prtf(translate=True);
 * enantio: allow the reflected match, as prtf(enantio=True).
 * Outputs (NULL to skip any but prtf):
 *   prtf          N floats, not fftshifted (prtf() returns it fftshifted)
 *   super_image   N complex: the sum of the aligned reconstructions
 *   aligned       n * N complex: the aligned reconstructions ("images")
 *   aligned_masks n * N ints: their masks, moved with them ("masks")
 *   mask_count    N floats: over reconstructions 1..n-1 with a non-empty
 *                 mask, how many have the pixel outside it ("super_mask"
 *                 is mask_count > 0) */
spimage_EXPORT int spx_prtf(const float *images, const int *masks, int nx, int ny, int nz, int n, int translate,
                            int enantio, float *prtf, float *super_image, float *aligned, int *aligned_masks,
                            float *mask_count);

#ifdef __cplusplus
}
#endif

#endif
