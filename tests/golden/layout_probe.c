/* Prints the sizes, field offsets and enum values of the phasing boundary's
 * types as one JSON object.
 *
 * Built two ways from this one file:
 *   -DPROBE_REFERENCE against the reference's headers (make_layout.py, in a
 *     checkout that has them): writes tests/golden/reference_layout.json;
 *   against this repository's include/spimage.h (tests/test_c_abi.py): the
 *     result must equal that fixture, so a caller compiled against either
 *     header set sees the same memory.
 * SpPhaser is compared up to `gp`; the reference's CUDA members after it are
 * replaced by one opaque engine pointer here. */
#include <stddef.h>
#include <stdio.h>

#ifdef PROBE_REFERENCE
/* the reference's own headers, in the order its umbrella header includes
 * them (its fft.h is left out: it needs FFTW's header) */
#include "spimage/image.h"
#include "spimage/sperror.h"
#include "spimage/list.h"
#include "spimage/map.h"
#include "spimage/phasing.h"
#include "spimage/support_update.h"
#include "spimage/cuda_util.h"
#include "spimage/image_util.h"
#else
#include "spimage.h"
#endif

static int first = 1;
static void item(const char *key, long v) {
  printf("%s\n  \"%s\": %ld", first ? "{" : ",", key, v);
  first = 0;
}
#define SZ(T) item("sizeof " #T, (long)sizeof(T))
#define OF(T, f) item(#T "." #f, (long)offsetof(T, f))
#define EV(e) item("enum " #e, (long)(e))

int main(void) {
  SZ(real);
  SZ(Complex);
  SZ(sp_3matrix); OF(sp_3matrix, x); OF(sp_3matrix, y); OF(sp_3matrix, z); OF(sp_3matrix, data);
  SZ(sp_i3matrix); OF(sp_i3matrix, data);
  SZ(sp_c3matrix); OF(sp_c3matrix, data);
  SZ(Detector); OF(Detector, image_center); OF(Detector, pixel_size); OF(Detector, detector_distance);
  OF(Detector, wavelength); OF(Detector, orientation);
  SZ(Image); OF(Image, phased); OF(Image, scaled); OF(Image, image); OF(Image, mask); OF(Image, detector);
  OF(Image, shifted); OF(Image, rec_coords); OF(Image, num_dimensions);
  SZ(sp_list); OF(sp_list, data); OF(sp_list, used); OF(sp_list, size);
  SZ(sp_smap); OF(sp_smap, keys); OF(sp_smap, values);
  SZ(SpPhasingAlgorithm); OF(SpPhasingAlgorithm, type); OF(SpPhasingAlgorithm, params);
  SZ(SpPhasingHIOParameters); OF(SpPhasingHIOParameters, beta); OF(SpPhasingHIOParameters, constraints);
  SZ(SpPhasingERParameters); OF(SpPhasingERParameters, constraints);
  SZ(SpPhasingDiffMapParameters); OF(SpPhasingDiffMapParameters, beta); OF(SpPhasingDiffMapParameters, gamma1);
  OF(SpPhasingDiffMapParameters, gamma2); OF(SpPhasingDiffMapParameters, constraints);
  SZ(SpSupportAlgorithm); OF(SpSupportAlgorithm, type); OF(SpSupportAlgorithm, params);
  OF(SpSupportAlgorithm, function);
  SZ(SpSupportArray); OF(SpSupportArray, size); OF(SpSupportArray, algorithms); OF(SpSupportArray, update_period);
  SZ(SpSupportThresholdParameters); OF(SpSupportThresholdParameters, blur_radius_map);
  OF(SpSupportThresholdParameters, threshold);
  SZ(SpSupportAreaParameters); OF(SpSupportAreaParameters, blur_radius_map); OF(SpSupportAreaParameters, area);
  SZ(SpSupportTemplateParameters); OF(SpSupportTemplateParameters, blured); OF(SpSupportTemplateParameters, sorted);
  OF(SpSupportTemplateParameters, area); OF(SpSupportTemplateParameters, original_area);
  SZ(SpSupportCloseParameters); OF(SpSupportCloseParameters, size);
  OF(SpPhaser, amplitudes); OF(SpPhaser, amplitudes_min); OF(SpPhaser, amplitudes_max);
  OF(SpPhaser, phased_amplitudes); OF(SpPhaser, pixel_flags); OF(SpPhaser, phasing_objective);
  OF(SpPhaser, algorithm); OF(SpPhaser, sup_algorithm); OF(SpPhaser, iteration); OF(SpPhaser, image_size);
  OF(SpPhaser, nx); OF(SpPhaser, ny); OF(SpPhaser, nz); OF(SpPhaser, model); OF(SpPhaser, model_iteration);
  OF(SpPhaser, old_model); OF(SpPhaser, old_model_iteration); OF(SpPhaser, model_change);
  OF(SpPhaser, model_change_iteration); OF(SpPhaser, support); OF(SpPhaser, support_iteration);
  OF(SpPhaser, amplitudes_image); OF(SpPhaser, fmodel); OF(SpPhaser, fmodel_iteration);
  OF(SpPhaser, model_before_projection); OF(SpPhaser, model_before_projection_iteration);
  OF(SpPhaser, engine); OF(SpPhaser, g0); OF(SpPhaser, g1); OF(SpPhaser, gp);
  EV(SpModelRandomPhases); EV(SpModelZeroPhases); EV(SpModelRandomValues); EV(SpModelMaskedOutZeroed);
  EV(SpSupportFromPatterson);
  EV(SpHIO); EV(SpRAAR); EV(SpDiffMap); EV(SpER);
  EV(SpNoConstraints); EV(SpRealObject); EV(SpPositiveRealObject); EV(SpPositiveComplexObject);
  EV(SpPositivityFlipping); EV(SpCentrosymmetricObject); EV(SpAmplitudeErrorMargin);
  EV(SpEngineAutomatic); EV(SpEngineCPU); EV(SpEngineCUDA);
  EV(SpPixelInsideSupport); EV(SpPixelMeasuredAmplitude);
  EV(SpRecoverPhases); EV(SpRecoverAmplitudes);
  EV(SpSupportThreshold); EV(SpSupportArea); EV(SpSupportCentreImage); EV(SpSupportTemplate);
  EV(SpSupportStatic); EV(SpSupportClose);
  EV(SpCUDANoDevice); EV(SpCUDAEmulatedDevice); EV(SpCUDAHardwareDevice);
  EV(SP_COPY_DATA); EV(SP_COPY_MASK); EV(SP_COPY_ALL);
  printf("\n}\n");
  return 0;
}
