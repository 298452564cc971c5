/* Umbrella header of the HIP phasing engine: the libspimage phasing boundary
 * (reference include/spimage.h, which includes phasing.h, support_update.h,
 * fft.h, image.h, linear_alg.h and map.h among others). */
#ifndef SPX_SPIMAGE_H
#define SPX_SPIMAGE_H 1
#include "sp_core.h"
#include "sp_phasing.h"
#include "sp_fft.h"
#include "sp_hip_engine.h"
#include "sp_cuda_util.h"
#endif
