/* Phasing API of libspimage served by the HIP engine.
 *
 * Every declaration keeps the reference's name, signature and return codes;
 * each cites the reference interface it replaces.  The CUDA-typed members of
 * SpPhaser (reference include/spimage/phasing.h:92-107) are replaced by one
 * opaque engine pointer; the fields that callers read or write (algorithm,
 * sup_algorithm, model, iteration, nx/ny/nz, image_size) keep their names.
 */
#ifndef SPX_SP_PHASING_H
#define SPX_SP_PHASING_H 1

#include "sp_core.h"

#ifdef __cplusplus
extern "C" {
#endif

/* enums: reference include/spimage/phasing.h:14-20 */
typedef enum { SpModelRandomPhases = 1, SpModelZeroPhases = 2, SpModelRandomValues = 4, SpModelMaskedOutZeroed = 256 } SpModelInitialization;
typedef enum { SpSupportFromPatterson = 1 } SpSupportInitialization;
typedef enum { SpHIO = 1, SpRAAR, SpDiffMap, SpER } SpPhasingAlgorithmType;
typedef enum {
  SpNoConstraints = 0, SpRealObject = 1, SpPositiveRealObject = 2, SpPositiveComplexObject = 4,
  SpPositivityFlipping = 8, SpCentrosymmetricObject = 16, SpAmplitudeErrorMargin = 32
} SpPhasingConstraints;
/* SpEngineCUDA keeps value 2 because callers pass it for "the GPU engine"
 * (reference src/_spimage_reconstructor.py:541-546); here it selects HIP. */
typedef enum { SpEngineAutomatic = 0, SpEngineCPU = 1, SpEngineCUDA = 2, SpEngineHIP = 2 } SpPhasingEngine;
typedef enum { SpPixelInsideSupport = 1, SpPixelMeasuredAmplitude = 2 } SpPhasingPixelFlags;
typedef enum { SpRecoverPhases = 0, SpRecoverAmplitudes = 1 } SpPhasingObjective;

/* parameter blocks: reference include/spimage/phasing.h:22-46 */
typedef struct { sp_smap *beta; SpPhasingConstraints constraints; } SpPhasingHIOParameters;
typedef struct { SpPhasingConstraints constraints; } SpPhasingERParameters;
typedef SpPhasingHIOParameters SpPhasingRAARParameters;
typedef struct { sp_smap *beta; real gamma1; real gamma2; SpPhasingConstraints constraints; } SpPhasingDiffMapParameters;
typedef struct { SpPhasingAlgorithmType type; void *params; } SpPhasingAlgorithm;

#include "sp_support.h"

/* phaser state: reference include/spimage/phasing.h:49-108 */
typedef struct {
  sp_3matrix *amplitudes;
  sp_3matrix *amplitudes_min;
  sp_3matrix *amplitudes_max;
  sp_c3matrix *phased_amplitudes;
  sp_i3matrix *pixel_flags;
  SpPhasingObjective phasing_objective;
  SpPhasingAlgorithm *algorithm;
  SpSupportArray *sup_algorithm;
  int iteration;
  int image_size;
  int nx;
  int ny;
  int nz;
  Image *model;
  int model_iteration;
  Image *old_model;
  int old_model_iteration;
  Image *model_change;
  int model_change_iteration;
  Image *support;
  int support_iteration;
  Image *amplitudes_image;
  Image *fmodel;
  int fmodel_iteration;
  Image *model_before_projection;
  int model_before_projection_iteration;
  SpPhasingEngine engine;
  Image *g0; /* host images of the CPU engine: always NULL here */
  Image *g1;
  Image *gp;
  void *hip_engine; /* device state: iterates, amplitudes, flags, work buffers */
} SpPhaser;

/* algorithm allocators: reference include/spimage/phasing.h:111-114, src/phasing.c:16-76 */
spimage_EXPORT SpPhasingAlgorithm *sp_phasing_hio_alloc(sp_smap *beta, SpPhasingConstraints constraints);
spimage_EXPORT SpPhasingAlgorithm *sp_phasing_raar_alloc(sp_smap *beta, SpPhasingConstraints constraints);
spimage_EXPORT SpPhasingAlgorithm *sp_phasing_diff_map_alloc(sp_smap *beta, real gamma1, real gamma2, SpPhasingConstraints constraints);
spimage_EXPORT SpPhasingAlgorithm *sp_phasing_er_alloc(SpPhasingConstraints constraints);

/* lifetime: reference include/spimage/phasing.h:116-117, src/phasing.c:78-186 */
spimage_EXPORT SpPhaser *sp_phaser_alloc(void);
spimage_EXPORT void sp_phaser_free(SpPhaser *ph);

/* getters (phaser-owned images, cached per iteration):
 * reference include/spimage/phasing.h:119-135, src/phasing.c:365-554,780-801 */
spimage_EXPORT const Image *sp_phaser_model(SpPhaser *ph);
spimage_EXPORT const Image *sp_phaser_model_with_support(SpPhaser *ph);
spimage_EXPORT const Image *sp_phaser_model_before_projection(SpPhaser *ph);
spimage_EXPORT const Image *sp_phaser_old_model(SpPhaser *ph);
spimage_EXPORT const Image *sp_phaser_fmodel(SpPhaser *ph);
spimage_EXPORT const Image *sp_phaser_fmodel_with_mask(SpPhaser *ph);
spimage_EXPORT Image *sp_phaser_model_change(SpPhaser *ph);
spimage_EXPORT const Image *sp_phaser_support(SpPhaser *ph);
spimage_EXPORT const Image *sp_phaser_amplitudes(SpPhaser *ph);

/* metrics: reference include/spimage/phasing.h:126-127,142, src/phasing.c:803-893
 * (FcFo = sum amp / sum |F| over measured pixels, the CPU engine's ratio) */
spimage_EXPORT real sp_phaser_ereal(SpPhaser *ph);
spimage_EXPORT real sp_phaser_efourier(SpPhaser *ph, real *FcFo);
spimage_EXPORT real sp_phaser_support_fraction(SpPhaser *ph);

/* setters (data is copied in): reference include/spimage/phasing.h:128-132,141,
 * src/phasing.c:93-95,189-363 */
spimage_EXPORT void sp_phaser_set_model(SpPhaser *ph, const Image *model);
spimage_EXPORT void sp_phaser_set_support(SpPhaser *ph, const Image *support);
spimage_EXPORT void sp_phaser_set_phased_amplitudes(SpPhaser *ph, const Image *phased_amplitudes);
spimage_EXPORT void sp_phaser_set_amplitudes(SpPhaser *ph, const Image *amplitudes);
spimage_EXPORT void sp_phaser_set_amplitudes_margins(SpPhaser *ph, const Image *amplitudes_min, const Image *amplitudes_max);
spimage_EXPORT void sp_phaser_set_objective(SpPhaser *ph, SpPhasingObjective obj);

/* control: reference include/spimage/phasing.h:136-140, src/phasing.c:558-778,895-1024.
 * Return codes are the reference's; in addition sp_phaser_init returns
 *   -9  when SpEngineCPU is requested (this library has no CPU engine)
 *   -10 when no HIP device is present (the reference would fall back to CPU)
 * and sp_phaser_iterate returns -11 when the HIP engine cannot serve the
 * array shape (an axis length above 4096 without a Stockham plan, or a 1-D
 * array). */
spimage_EXPORT int sp_phaser_init(SpPhaser *ph, SpPhasingAlgorithm *alg, SpSupportArray *sup_alg, SpPhasingEngine engine);
spimage_EXPORT int sp_phaser_set_phasing_algorithm(SpPhaser *ph, SpPhasingAlgorithm *alg);
/* defined but not declared by the reference (src/phasing.c:597-604) */
spimage_EXPORT int sp_phaser_set_support_algorithm(SpPhaser *ph, SpSupportArray *alg);
spimage_EXPORT int sp_phaser_init_model(SpPhaser *ph, const Image *model, int flags);
spimage_EXPORT int sp_phaser_init_support(SpPhaser *ph, const Image *support, int flags, real value);
spimage_EXPORT int sp_phaser_iterate(SpPhaser *ph, int iterations);

/* support scheduling entry point: reference include/spimage/phasing.h:173,
 * src/support_update.c:489-494 */
spimage_EXPORT int sp_support_array_update(SpSupportArray *array, SpPhaser *ph);

/* per-type support updates used as SpSupportAlgorithm.function:
 * reference include/spimage/phasing.h:161-171, src/support_update.c:124-389 */
spimage_EXPORT int sp_support_area_update_support(SpSupportAlgorithm *alg, SpPhaser *ph);
spimage_EXPORT int sp_support_threshold_update_support(SpSupportAlgorithm *alg, SpPhaser *ph);
spimage_EXPORT int sp_support_template_update_support(SpSupportAlgorithm *alg, SpPhaser *ph);
spimage_EXPORT int sp_support_static_update_support(SpSupportAlgorithm *alg, SpPhaser *ph);
spimage_EXPORT int sp_support_close_update_support(SpSupportAlgorithm *alg, SpPhaser *ph);
spimage_EXPORT int sp_support_centre_image(SpSupportAlgorithm *alg, SpPhaser *ph);

#ifdef __cplusplus
}
#endif
#endif
