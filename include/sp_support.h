/* Support-update (shrinkwrap) types and allocators.
 * Reference include/spimage/support_update.h:9-69, src/support_update.c:7-122,456-487.
 */
#ifndef SPX_SP_SUPPORT_H
#define SPX_SP_SUPPORT_H 1

#include "sp_core.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum { SpSupportThreshold = 1, SpSupportArea, SpSupportCentreImage, SpSupportTemplate, SpSupportStatic, SpSupportClose } SpSupportAlgorithmType;

typedef struct { sp_smap *blur_radius_map; sp_smap *threshold; } SpSupportThresholdParameters;
typedef struct { sp_smap *blur_radius_map; sp_smap *area; } SpSupportAreaParameters;
typedef struct { void *dummy; } SpSupportCentreImageParameters;
typedef struct { Image *blured; Image *sorted; sp_smap *area; real original_area; } SpSupportTemplateParameters;
typedef struct { void *dummy; } SpSupportStaticParameters;
typedef struct { int size; } SpSupportCloseParameters;

/* function is called as int (*)(SpSupportAlgorithm *, SpPhaser *)
 * (reference src/support_update.c:491) */
typedef struct {
  SpSupportAlgorithmType type;
  void *params;
  void *function;
} SpSupportAlgorithm;

typedef struct {
  int size;
  SpSupportAlgorithm **algorithms;
  int update_period;
} SpSupportArray;

/* reference include/spimage/support_update.h:59-64 */
spimage_EXPORT SpSupportAlgorithm *sp_support_threshold_alloc(sp_smap *blur_radius, sp_smap *threshold);
spimage_EXPORT SpSupportAlgorithm *sp_support_area_alloc(sp_smap *blur_radius, sp_smap *area);
spimage_EXPORT SpSupportAlgorithm *sp_support_template_alloc(Image *initial_support, real blur_radius, sp_smap *area);
spimage_EXPORT SpSupportAlgorithm *sp_support_static_alloc(void);
spimage_EXPORT SpSupportAlgorithm *sp_support_close_alloc(int size);
spimage_EXPORT SpSupportAlgorithm *sp_support_centre_image_alloc(void);

/* reference include/spimage/support_update.h:66-69 */
spimage_EXPORT SpSupportArray *sp_support_array_alloc(int size, int update_period);
spimage_EXPORT SpSupportArray *sp_support_array_init(SpSupportAlgorithm *algorithm, int update_period);
spimage_EXPORT void sp_support_array_set(SpSupportArray *array, int i, SpSupportAlgorithm *algorithm);
spimage_EXPORT void sp_support_array_append(SpSupportArray *array, SpSupportAlgorithm *algorithm);

#ifdef __cplusplus
}
#endif
#endif
