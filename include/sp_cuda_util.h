/* The device queries of libspimage's CUDA glue that stay public without CUDA.
 *
 *   SpCUDADeviceType            reference include/spimage/cuda_util.h:101
 *   sp_cuda_get_device_type     reference include/spimage/cuda_util.h:105, src/cuda_util.c:10-29
 *   sp_cuda_launch_parameters   reference include/spimage/cuda_util.h:109, src/cuda_util.c:39-60
 *
 * The reference declares these outside its #ifdef _USE_CUDA block, so the
 * SWIG module (reference src/spimage_pybackend.i:165) and the C tests
 * (reference tests/PhasingTests.c:1049,1084,1128,1185,1208,1342) use them in
 * every build.  Here the "CUDA device" is the HIP device the engine runs on.
 */
#ifndef SPX_SP_CUDA_UTIL_H
#define SPX_SP_CUDA_UTIL_H 1

#include <stdint.h>

#include "sp_core.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum { SpCUDANoDevice = 0, SpCUDAEmulatedDevice = 1, SpCUDAHardwareDevice = 2 } SpCUDADeviceType;

/* SpCUDAHardwareDevice when a HIP device is usable, SpCUDANoDevice
 * otherwise (there is no emulated device on this platform). */
spimage_EXPORT SpCUDADeviceType sp_cuda_get_device_type(void);

/* Grid and block size of a one-element-per-thread launch over `size`
 * elements: the block starts at 64 threads and doubles while the grid would
 * exceed the device's limit (2^31 - 1 blocks); grid = ceil(size / block).
 * A size beyond what 1024-thread blocks cover is fatal, as in the reference. */
spimage_EXPORT void sp_cuda_launch_parameters(int64_t size, int *gridSize, int *blockSize);

#ifdef __cplusplus
}
#endif
#endif
