/* Core types of the libspimage phasing boundary, layout-compatible with the
 * reference so that callers compiled against the reference headers (Hawk, the
 * SWIG wrapper, the C tests) see the same structs.
 *
 *   real / Complex            reference include/spimage/linear_alg.h:39-57
 *   sp_3matrix/i3/c3matrix    reference include/spimage/linear_alg.h:95-114
 *   Detector / Image          reference include/spimage/image.h:33-84
 *   sp_list / sp_smap         reference include/spimage/list.h, map.h:16-141
 *
 * Only FP32 builds exist (the reference's DOUBLE_PRECISION=OFF default,
 * reference CMakeLists.txt:45,108-112).
 */
#ifndef SPX_SP_CORE_H
#define SPX_SP_CORE_H 1

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef __cplusplus
extern "C" {
#endif

#define spimage_EXPORT __attribute__((visibility("default")))

typedef float real;
typedef struct {
  float re, im;
} Complex;
#define sp_real(a) (a).re
#define sp_imag(a) (a).im
#define REAL_MAX FLT_MAX
#define REAL_EPSILON FLT_EPSILON

typedef struct { unsigned int size; real *data; } sp_vector;
typedef struct { unsigned int rows; unsigned int cols; real *data; } sp_matrix;
typedef sp_matrix SpRotation;

typedef struct { unsigned int x, y, z; real *data; } sp_3matrix;
typedef struct { unsigned int x, y, z; int *data; } sp_i3matrix;
typedef struct { unsigned int x, y, z; Complex *data; } sp_c3matrix;

typedef enum { SP_1D = 1, SP_2D = 2, SP_3D = 3 } Dimensions;

typedef struct {
  real image_center[3];
  real pixel_size[3];
  real detector_distance;
  real wavelength;
  SpRotation *orientation;
} Detector;

typedef struct {
  int phased;
  int scaled;
  sp_c3matrix *image;
  sp_i3matrix *mask;
  Detector *detector;
  int shifted;
  sp_3matrix *rec_coords;
  Dimensions num_dimensions;
} Image;

/* duplicate flags (reference include/spimage/image_util.h:47-50) */
#define SP_COPY_DATA 1
#define SP_COPY_DETECTOR 0
#define SP_COPY_MASK 2
#define SP_COPY_ALL 3
#define SP_ZERO_PHASE 1
#define SP_RANDOM_PHASE 2

/* allocation is plain malloc/free (the reference's sp_malloc without
 * SP_MEM_DEBUG, reference include/spimage/mem_util.h) */
#define sp_malloc(s) malloc(s)
#define sp_calloc(n, s) calloc(n, s)
#define sp_realloc(p, s) realloc(p, s)
#define sp_free(p) free(p)

/* ---- complex helpers (reference include/spimage/linear_alg.h:153-201) ---- */
static inline Complex sp_cinit(real a, real b) { Complex r = {a, b}; return r; }
static inline real sp_cabs(Complex a) { return sqrt(sp_imag(a) * sp_imag(a) + sp_real(a) * sp_real(a)); }
static inline real sp_cabs2(Complex a) { return sp_imag(a) * sp_imag(a) + sp_real(a) * sp_real(a); }
static inline Complex sp_cadd(Complex a, Complex b) { sp_real(a) += sp_real(b); sp_imag(a) += sp_imag(b); return a; }
static inline Complex sp_csub(Complex a, Complex b) { sp_real(a) -= sp_real(b); sp_imag(a) -= sp_imag(b); return a; }
static inline Complex sp_cscale(Complex a, real b) { sp_real(a) *= b; sp_imag(a) *= b; return a; }
static inline Complex sp_cmul(Complex a, Complex b) {
  Complex r;
  sp_real(r) = sp_real(a) * sp_real(b) - sp_imag(a) * sp_imag(b);
  sp_imag(r) = sp_real(a) * sp_imag(b) + sp_real(b) * sp_imag(a);
  return r;
}
static inline real sp_min(real a, real b) { return a < b ? a : b; }
static inline real sp_max(real a, real b) { return a > b ? a : b; }

/* ---- 3-D matrices (reference src/linear_alg.c:9-90; calloc'ed data) ---- */
spimage_EXPORT sp_3matrix *sp_3matrix_alloc(unsigned int nx, unsigned int ny, unsigned int nz);
spimage_EXPORT sp_i3matrix *sp_i3matrix_alloc(unsigned int nx, unsigned int ny, unsigned int nz);
spimage_EXPORT sp_c3matrix *sp_c3matrix_alloc(unsigned int nx, unsigned int ny, unsigned int nz);
spimage_EXPORT void sp_3matrix_free(sp_3matrix *m);
spimage_EXPORT void sp_i3matrix_free(sp_i3matrix *m);
spimage_EXPORT void sp_c3matrix_free(sp_c3matrix *m);

/* ---- Image lifetime (reference src/image_util.c:1223-1300; the reference
 * header maps sp_image_alloc/free/duplicate to these underscored symbols,
 * reference include/spimage/image_util.h:120-147) ---- */
spimage_EXPORT Image *_sp_image_alloc(int x, int y, int z, const char *file, int line);
spimage_EXPORT void _sp_image_free(Image *in, const char *file, int line);
spimage_EXPORT Image *_sp_image_duplicate(const Image *in, int flags, const char *file, int line);
#define sp_image_alloc(x, y, z) _sp_image_alloc(x, y, z, __FILE__, __LINE__)
#define sp_image_free(in) _sp_image_free(in, __FILE__, __LINE__)
#define sp_image_duplicate(in, flags) _sp_image_duplicate(in, flags, __FILE__, __LINE__)

static inline int sp_image_x(const Image *a) { return (int)a->image->x; }
static inline int sp_image_y(const Image *a) { return (int)a->image->y; }
static inline int sp_image_z(const Image *a) { return (int)a->image->z; }
static inline long long sp_image_size(const Image *a) {
  return (long long)a->image->x * a->image->y * a->image->z;
}

/* ---- parameter schedules: sorted map with cubic-Bezier interpolation
 * (reference include/spimage/list.h, include/spimage/map.h:16-141).  The
 * library exports out-of-line copies of every function for FFI callers
 * (ctypes); C callers get these inline versions. ---- */
typedef struct {
  real *data;
  unsigned int used;
  unsigned int size;
} sp_list;

typedef struct {
  sp_list *keys;
  sp_list *values;
} sp_smap;

spimage_EXPORT sp_smap *spx_smap_alloc(unsigned int init_size);
spimage_EXPORT void spx_smap_free(sp_smap *m);
spimage_EXPORT void spx_smap_insert(sp_smap *m, real key, real value);
spimage_EXPORT sp_smap *spx_smap_create_from_pair(real key, real value);
spimage_EXPORT real spx_smap_interpolate(sp_smap *m, real x);
spimage_EXPORT real spx_smap_max(sp_smap *m);
spimage_EXPORT real spx_smap_min(sp_smap *m);

static inline sp_smap *sp_smap_alloc(unsigned int init_size) { return spx_smap_alloc(init_size); }
static inline void sp_smap_free(sp_smap *m) { spx_smap_free(m); }
static inline void sp_smap_insert(sp_smap *m, real key, real value) { spx_smap_insert(m, key, value); }
static inline sp_smap *sp_smap_create_from_pair(real key, real value) { return spx_smap_create_from_pair(key, value); }
static inline real sp_smap_interpolate(sp_smap *m, real x) { return spx_smap_interpolate(m, x); }
static inline real sp_smap_max(sp_smap *m) { return spx_smap_max(m); }
static inline real sp_smap_min(sp_smap *m) { return spx_smap_min(m); }
static inline sp_list *sp_smap_get_keys(sp_smap *m) { return m->keys; }
static inline sp_list *sp_smap_get_values(sp_smap *m) { return m->values; }
static inline unsigned int sp_list_size(sp_list *l) { return l->used; }
static inline real sp_list_get(sp_list *l, unsigned int n) { return l->data[n]; }

/* p_drand48 (reference src/image_util.c:124-127) and its seed */
spimage_EXPORT real p_drand48(void);
spimage_EXPORT void sp_srand(int i);

#ifdef __cplusplus
}
#endif
#endif
