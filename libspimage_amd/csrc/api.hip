// The libspimage phasing API (reference include/spimage/phasing.h,
// support_update.h, fft.h) implemented over the HIP engine.
//
// Host-side control (argument checks, return codes, the support scheduler,
// parameter schedules, getter caching) follows the reference's CPU code
// (src/phasing.c, src/support_update.c); every array operation runs on the
// device through spx::Engine.  There is no CPU engine in this library.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "engine.h"
#include "spimage.h"

using spx::Engine;

namespace {

// Per-phaser host bookkeeping hung off SpPhaser.hip_engine.
struct HostState {
  Engine* eng = nullptr;
  unsigned long long seed = 0;
  bool device_flags_newer = false;  // a support update changed the device flags
  bool have_margins = false;
  bool have_phased = false;
};

HostState* hs(SpPhaser* ph) {
  if (!ph->hip_engine) ph->hip_engine = new HostState();
  return (HostState*)ph->hip_engine;
}

Engine* engine_of(SpPhaser* ph) { return ph->hip_engine ? ((HostState*)ph->hip_engine)->eng : nullptr; }

void report(const char* where) { fprintf(stderr, "%s: %s\n", where, spx::last_error()); }

// Creates the device state once the shape is known (first setter).
Engine* ensure_engine(SpPhaser* ph) {
  HostState* h = hs(ph);
  if (h->eng) return h->eng;
  if (ph->engine != SpEngineHIP || !ph->nx) return nullptr;
  h->eng = Engine::create(ph->nx, ph->ny, ph->nz);
  if (!h->eng) {
    report("libspimage_hip");
    return nullptr;
  }
  h->eng->seed = h->seed;
  return h->eng;
}

std::vector<uint8_t> flags_u8(const SpPhaser* ph) {
  std::vector<uint8_t> f((size_t)ph->image_size);
  for (int i = 0; i < ph->image_size; ++i) f[i] = (uint8_t)(ph->pixel_flags->data[i] & 3);
  return f;
}

// Bring the host copy of the flags up to date after device-side support updates.
void pull_flags(SpPhaser* ph) {
  HostState* h = hs(ph);
  if (!h->device_flags_newer || !h->eng || !ph->pixel_flags) return;
  std::vector<uint8_t> f((size_t)ph->image_size);
  if (h->eng->download_flags(f.data())) {
    report("sp_phaser flags download");
    return;
  }
  for (int i = 0; i < ph->image_size; ++i)
    ph->pixel_flags->data[i] = (ph->pixel_flags->data[i] & ~1) | (f[i] & 1);
  h->device_flags_newer = false;
}

void push_flags(SpPhaser* ph) {
  Engine* e = ensure_engine(ph);
  if (!e) return;
  std::vector<uint8_t> f = flags_u8(ph);
  if (ph->amplitudes) {
    // amplitudes are re-encoded whenever the measured bits may have moved
    if (e->upload_amplitudes(ph->amplitudes->data, f.data())) report("sp_phaser amplitudes upload");
  } else if (e->upload_flags(f.data())) {
    report("sp_phaser flags upload");
  }
}

// reference src/phasing.c:271-292
void phaser_check_dimensions(SpPhaser* ph, const Image* a) {
  if (!ph->nx) {
    ph->nx = sp_image_x(a);
    ph->ny = sp_image_y(a);
    ph->nz = sp_image_z(a);
    ph->image_size = ph->nx * ph->ny * ph->nz;
  }
  if (ph->nx != sp_image_x(a) || ph->ny != sp_image_y(a) || ph->nz != sp_image_z(a)) abort();
}

void ensure_pixel_flags(SpPhaser* ph) {
  if (!ph->pixel_flags) ph->pixel_flags = sp_i3matrix_alloc(ph->nx, ph->ny, ph->nz);
}

Image* alloc_like(SpPhaser* ph) { return sp_image_alloc(ph->nx, ph->ny, ph->nz); }

int download_into(SpPhaser* ph, Image* img, const float2* dev) {
  Engine* e = engine_of(ph);
  if (!e) return -1;
  if (e->download(dev, (float2*)img->image->data)) {
    report("sp_phaser download");
    return -1;
  }
  return 0;
}

// cubic-Bezier schedule; x-control points (p0, m, m, p3) with m the midpoint
// and y-control points (y0, y0, y3, y3), the curve of the reference's
// sp_smap_interpolate (include/spimage/map.h:103-141).  With
// u = (x - p0) / (p3 - p0) the curve is u = t^3 - 1.5 t^2 + 1.5 t; t = s + 1/2
// turns it into s^3 + 0.75 s + (0.5 - u) = 0, solved by Cardano's formula.
double bezier_value(double p0x, double p0y, double p3x, double p3y, double x) {
  const double u = (x - p0x) / (p3x - p0x);
  const double q = 0.5 - u;
  const double d = sqrt(q * q / 4.0 + 1.0 / 64.0);
  const double t = cbrt(-q / 2.0 + d) + cbrt(-q / 2.0 - d) + 0.5;
  const double w = 1.0 - t;
  return p0y * w * w * w + 3.0 * p0y * t * w * w + 3.0 * p3y * t * t * w + p3y * t * t * t;
}

float beta_at(SpPhasingAlgorithm* alg, int it) {
  sp_smap* m = nullptr;
  if (alg->type == SpHIO || alg->type == SpRAAR) m = ((SpPhasingHIOParameters*)alg->params)->beta;
  else if (alg->type == SpDiffMap) m = ((SpPhasingDiffMapParameters*)alg->params)->beta;
  return m ? spx_smap_interpolate(m, (real)it) : 0.f;
}

int run_phasing(SpPhaser* ph, int n) {
  if (n <= 0) return 0;
  Engine* e = engine_of(ph);
  if (!e) {
    fprintf(stderr, "sp_phaser_iterate: HIP engine not initialised (%s)\n", spx::last_error());
    return -11;
  }
  spx::IterParams p{};
  p.alg = ph->algorithm->type;
  p.objective = ph->phasing_objective;
  if (p.alg == SpER) {
    p.constraints = ((SpPhasingERParameters*)ph->algorithm->params)->constraints;
  } else if (p.alg == SpDiffMap) {
    SpPhasingDiffMapParameters* d = (SpPhasingDiffMapParameters*)ph->algorithm->params;
    p.constraints = d->constraints;
    p.gamma1 = d->gamma1;
    p.gamma2 = d->gamma2;
  } else {
    p.constraints = ((SpPhasingHIOParameters*)ph->algorithm->params)->constraints;
  }
  if ((p.constraints & SpAmplitudeErrorMargin) && !hs(ph)->have_margins &&
      (p.objective == SpRecoverPhases || p.alg == SpDiffMap)) {
    // reference src/phasing.c:1071-1073 (sp_error_fatal)
    fprintf(stderr, "Amplitude margin map is a null pointer.\n");
    abort();
  }
  std::vector<float> betas((size_t)n);
  for (int i = 0; i < n; ++i) betas[i] = beta_at(ph->algorithm, ph->iteration + i);
  // eager launches: on MI355X a replayed HIP graph of the two pass kernels
  // measured about 2.4 us slower per kernel than back-to-back launches
  const int rc = e->iterate(p, betas.data(), n, true);
  if (rc) {
    report("sp_phaser_iterate");
    return rc;
  }
  ph->iteration += n;
  return 0;
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------------------
// matrices and images (reference src/linear_alg.c:9-90, src/image_util.c:1223-1300)
// ---------------------------------------------------------------------------
sp_3matrix* sp_3matrix_alloc(unsigned int nx, unsigned int ny, unsigned int nz) {
  sp_3matrix* m = (sp_3matrix*)malloc(sizeof(sp_3matrix));
  m->x = nx; m->y = ny; m->z = nz;
  m->data = (real*)calloc((size_t)nx * ny * nz, sizeof(real));
  return m;
}
sp_i3matrix* sp_i3matrix_alloc(unsigned int nx, unsigned int ny, unsigned int nz) {
  sp_i3matrix* m = (sp_i3matrix*)malloc(sizeof(sp_i3matrix));
  m->x = nx; m->y = ny; m->z = nz;
  m->data = (int*)calloc((size_t)nx * ny * nz, sizeof(int));
  return m;
}
sp_c3matrix* sp_c3matrix_alloc(unsigned int nx, unsigned int ny, unsigned int nz) {
  sp_c3matrix* m = (sp_c3matrix*)malloc(sizeof(sp_c3matrix));
  m->x = nx; m->y = ny; m->z = nz;
  m->data = (Complex*)calloc((size_t)nx * ny * nz, sizeof(Complex));
  return m;
}
void sp_3matrix_free(sp_3matrix* m) { if (m) { free(m->data); free(m); } }
void sp_i3matrix_free(sp_i3matrix* m) { if (m) { free(m->data); free(m); } }
void sp_c3matrix_free(sp_c3matrix* m) { if (m) { free(m->data); free(m); } }

Image* _sp_image_alloc(int x, int y, int z, const char*, int) {
  if (x < 1 || y < 1 || z < 1) {
    perror("Dimensions must be positive integers");
    abort();
  }
  Image* r = (Image*)calloc(1, sizeof(Image));
  r->detector = (Detector*)calloc(1, sizeof(Detector));
  r->detector->image_center[0] = (x - 1) / 2.0;
  r->detector->image_center[1] = (y - 1) / 2.0;
  r->detector->image_center[2] = (z - 1) / 2.0;
  r->mask = sp_i3matrix_alloc(x, y, z);
  r->image = sp_c3matrix_alloc(x, y, z);
  r->num_dimensions = (z == 1) ? SP_2D : SP_3D;
  return r;
}

void _sp_image_free(Image* in, const char*, int) {
  if (!in) return;
  sp_c3matrix_free(in->image);
  sp_i3matrix_free(in->mask);
  free(in->detector);
  free(in);
}

Image* _sp_image_duplicate(const Image* in, int flags, const char*, int) {
  Image* r = (Image*)malloc(sizeof(Image));
  memcpy(r, in, sizeof(Image));
  r->detector = (Detector*)malloc(sizeof(Detector));
  memcpy(r->detector, in->detector, sizeof(Detector));
  r->image = sp_c3matrix_alloc(in->image->x, in->image->y, in->image->z);
  if (flags & SP_COPY_DATA) memcpy(r->image->data, in->image->data, sizeof(Complex) * (size_t)sp_image_size(in));
  r->mask = sp_i3matrix_alloc(in->image->x, in->image->y, in->image->z);
  if (flags & SP_COPY_MASK) memcpy(r->mask->data, in->mask->data, sizeof(int) * (size_t)sp_image_size(in));
  return r;
}

real p_drand48(void) { return (real)(((double)rand()) / (double)RAND_MAX); }
void sp_srand(int i) { srand(i); }

// ---------------------------------------------------------------------------
// sorted map (reference include/spimage/map.h:16-141, list.h)
// ---------------------------------------------------------------------------
static sp_list* list_alloc(unsigned int n) {
  sp_list* l = (sp_list*)malloc(sizeof(sp_list));
  l->size = n ? n : 1;
  l->used = 0;
  l->data = (real*)malloc(sizeof(real) * l->size);
  return l;
}
static void list_insert(sp_list* l, unsigned int at, real v) {
  if (l->used == l->size) {
    l->size = l->size * 2 + 4;
    l->data = (real*)realloc(l->data, sizeof(real) * l->size);
  }
  memmove(l->data + at + 1, l->data + at, sizeof(real) * (l->used - at));
  l->data[at] = v;
  l->used++;
}

sp_smap* spx_smap_alloc(unsigned int init_size) {
  sp_smap* m = (sp_smap*)malloc(sizeof(sp_smap));
  m->keys = list_alloc(init_size);
  m->values = list_alloc(init_size);
  return m;
}
void spx_smap_free(sp_smap* m) {
  if (!m) return;
  free(m->keys->data); free(m->keys);
  free(m->values->data); free(m->values);
  free(m);
}
void spx_smap_insert(sp_smap* m, real key, real value) {
  unsigned int at = 0;
  while (at < m->keys->used && !(m->keys->data[at] > key)) ++at;
  list_insert(m->keys, at, key);
  list_insert(m->values, at, value);
}
sp_smap* spx_smap_create_from_pair(real key, real value) {
  sp_smap* m = spx_smap_alloc(1);
  spx_smap_insert(m, key, value);
  return m;
}
real spx_smap_interpolate(sp_smap* m, real x) {
  const unsigned int n = m->keys->used;
  unsigned int idx = 0;
  while (idx < n && !(x < m->keys->data[idx])) ++idx;
  if (idx == 0) return m->values->data[0];
  if (idx == n) return m->values->data[n - 1];
  return (real)bezier_value(m->keys->data[idx - 1], m->values->data[idx - 1], m->keys->data[idx],
                            m->values->data[idx], x);
}
real spx_smap_max(sp_smap* m) {
  real v = m->values->data[0];
  for (unsigned int i = 1; i < m->values->used; ++i)
    if (m->values->data[i] > v) v = m->values->data[i];
  return v;
}
real spx_smap_min(sp_smap* m) {
  real v = m->values->data[0];
  for (unsigned int i = 1; i < m->values->used; ++i)
    if (m->values->data[i] < v) v = m->values->data[i];
  return v;
}

// ---------------------------------------------------------------------------
// algorithms (reference src/phasing.c:16-76)
// ---------------------------------------------------------------------------
SpPhasingAlgorithm* sp_phasing_diff_map_alloc(sp_smap* beta, real gamma1, real gamma2, SpPhasingConstraints c) {
  SpPhasingAlgorithm* r = (SpPhasingAlgorithm*)malloc(sizeof(SpPhasingAlgorithm));
  r->type = SpDiffMap;
  SpPhasingDiffMapParameters* p = (SpPhasingDiffMapParameters*)malloc(sizeof(SpPhasingDiffMapParameters));
  p->beta = beta;
  // Elser's optimal gammas (V. Elser 2003) at beta(0) when gamma1 is infinite;
  // like the reference, gamma2 is also chosen by testing gamma1
  const real b = spx_smap_interpolate(beta, 0);
  const real sigma = 0.1f;
  p->gamma1 = isinf(gamma1) ? -(4 + (2 + b) * sigma + b * sigma * sigma) / (b * (4 - sigma + sigma * sigma)) : gamma1;
  p->gamma2 = isinf(gamma1) ? (3 - b) / (2 * b) : gamma2;
  p->constraints = c;
  r->params = p;
  return r;
}

static SpPhasingAlgorithm* beta_alg(SpPhasingAlgorithmType t, sp_smap* beta, SpPhasingConstraints c) {
  SpPhasingAlgorithm* r = (SpPhasingAlgorithm*)malloc(sizeof(SpPhasingAlgorithm));
  r->type = t;
  SpPhasingHIOParameters* p = (SpPhasingHIOParameters*)malloc(sizeof(SpPhasingHIOParameters));
  p->beta = beta;
  p->constraints = c;
  r->params = p;
  return r;
}
SpPhasingAlgorithm* sp_phasing_raar_alloc(sp_smap* beta, SpPhasingConstraints c) { return beta_alg(SpRAAR, beta, c); }
SpPhasingAlgorithm* sp_phasing_hio_alloc(sp_smap* beta, SpPhasingConstraints c) { return beta_alg(SpHIO, beta, c); }
SpPhasingAlgorithm* sp_phasing_er_alloc(SpPhasingConstraints c) {
  SpPhasingAlgorithm* r = (SpPhasingAlgorithm*)malloc(sizeof(SpPhasingAlgorithm));
  r->type = SpER;
  SpPhasingERParameters* p = (SpPhasingERParameters*)malloc(sizeof(SpPhasingERParameters));
  p->constraints = c;
  r->params = p;
  return r;
}

// ---------------------------------------------------------------------------
// phaser lifetime (reference src/phasing.c:78-186,558-604)
// ---------------------------------------------------------------------------
SpPhaser* sp_phaser_alloc(void) {
  SpPhaser* r = (SpPhaser*)calloc(1, sizeof(SpPhaser));
  r->model_iteration = -1;
  // the reference leaves old_model_iteration at 0, so its old_model getter
  // returns NULL before the first iteration; -1 makes it return g0 instead
  r->old_model_iteration = -1;
  r->fmodel_iteration = -1;
  r->model_before_projection_iteration = -1;
  r->model_change_iteration = -1;
  r->support_iteration = -1;
  r->phasing_objective = SpRecoverPhases;
  return r;
}

void sp_phaser_set_objective(SpPhaser* ph, SpPhasingObjective obj) { ph->phasing_objective = obj; }

void sp_phaser_free(SpPhaser* ph) {
  if (!ph) return;
  Image* imgs[] = {ph->model, ph->model_before_projection, ph->model_change, ph->amplitudes_image,
                   ph->old_model, ph->support, ph->fmodel};
  for (Image* i : imgs)
    if (i) sp_image_free(i);
  sp_3matrix_free(ph->amplitudes);
  sp_3matrix_free(ph->amplitudes_min);
  sp_3matrix_free(ph->amplitudes_max);
  sp_i3matrix_free(ph->pixel_flags);
  sp_c3matrix_free(ph->phased_amplitudes);
  if (ph->hip_engine) {
    HostState* h = (HostState*)ph->hip_engine;
    delete h->eng;
    delete h;
  }
  free(ph);
}

int sp_phaser_init(SpPhaser* ph, SpPhasingAlgorithm* alg, SpSupportArray* sup_alg, SpPhasingEngine engine) {
  if (!ph) {
    fprintf(stderr, "Phaser is NULL!\n");
    return -1;
  }
  if (!alg) {
    fprintf(stderr, "Algorithm is NULL!\n");
    return -2;
  }
  ph->algorithm = alg;
  ph->sup_algorithm = sup_alg;
  ph->iteration = 0;
  if (engine == SpEngineCPU) {
    fprintf(stderr, "libspimage_hip provides the HIP engine only; SpEngineCPU is served by the reference library.\n");
    return -9;
  }
  if (!spx::device_available()) {
    fprintf(stderr, "Requested engine unavailable: no HIP device.\n");
    return -10;
  }
  ph->engine = SpEngineHIP;
  // the reference seeds its device generator from rand() (src/phasing.c:578)
  hs(ph)->seed = (unsigned long long)rand();
  if (Engine* e = engine_of(ph)) e->seed = hs(ph)->seed;
  return 0;
}

int sp_phaser_set_phasing_algorithm(SpPhaser* ph, SpPhasingAlgorithm* alg) {
  if (!alg) {
    fprintf(stderr, "Algorithm is NULL!\n");
    return -2;
  }
  ph->algorithm = alg;
  return 0;
}

int sp_phaser_set_support_algorithm(SpPhaser* ph, SpSupportArray* alg) {
  if (!alg) {
    fprintf(stderr, "Phasing algorithm is NULL!\n");
    return -2;
  }
  ph->sup_algorithm = alg;
  return 0;
}

// ---------------------------------------------------------------------------
// setters (reference src/phasing.c:189-363)
// ---------------------------------------------------------------------------
void sp_phaser_set_model(SpPhaser* ph, const Image* model) {
  phaser_check_dimensions(ph, model);
  ph->model_iteration = -1;
  Engine* e = ensure_engine(ph);
  if (e && e->upload_model((const float2*)model->image->data)) report("sp_phaser_set_model");
}

void sp_phaser_set_support(SpPhaser* ph, const Image* support) {
  phaser_check_dimensions(ph, support);
  ph->support_iteration = -1;
  ensure_pixel_flags(ph);
  pull_flags(ph);
  for (int i = 0; i < ph->image_size; i++) {
    if (sp_real(support->image->data[i])) ph->pixel_flags->data[i] |= SpPixelInsideSupport;
    else ph->pixel_flags->data[i] &= ~SpPixelInsideSupport;
  }
  push_flags(ph);
}

void sp_phaser_set_phased_amplitudes(SpPhaser* ph, const Image* pa) {
  phaser_check_dimensions(ph, pa);
  if (!ph->phased_amplitudes) ph->phased_amplitudes = sp_c3matrix_alloc(ph->nx, ph->ny, ph->nz);
  ensure_pixel_flags(ph);
  pull_flags(ph);
  for (int i = 0; i < ph->image_size; i++) {
    ph->phased_amplitudes->data[i] = pa->image->data[i];
    if (pa->mask->data[i]) ph->pixel_flags->data[i] |= SpPixelMeasuredAmplitude;
    else ph->pixel_flags->data[i] &= ~SpPixelMeasuredAmplitude;
  }
  Engine* e = ensure_engine(ph);
  if (!e) return;
  std::vector<uint8_t> f = flags_u8(ph);
  if (e->upload_phased((const float2*)ph->phased_amplitudes->data, f.data())) report("sp_phaser_set_phased_amplitudes");
  hs(ph)->have_phased = true;
  push_flags(ph);
}

void sp_phaser_set_amplitudes(SpPhaser* ph, const Image* amplitudes) {
  phaser_check_dimensions(ph, amplitudes);
  if (!ph->amplitudes) ph->amplitudes = sp_3matrix_alloc(ph->nx, ph->ny, ph->nz);
  ensure_pixel_flags(ph);
  pull_flags(ph);
  int masked = 0;
  for (int i = 0; i < ph->image_size; i++) {
    ph->amplitudes->data[i] = sp_real(amplitudes->image->data[i]);
    if (amplitudes->mask->data[i]) {
      ph->pixel_flags->data[i] |= SpPixelMeasuredAmplitude;
      masked++;
    } else {
      ph->pixel_flags->data[i] &= ~SpPixelMeasuredAmplitude;
    }
  }
  if (masked == 0) fprintf(stderr, "Amplitudes mask is all zeros!\n");
  push_flags(ph);
}

void sp_phaser_set_amplitudes_margins(SpPhaser* ph, const Image* amin, const Image* amax) {
  phaser_check_dimensions(ph, amin);
  phaser_check_dimensions(ph, amax);
  if (!ph->amplitudes_min) ph->amplitudes_min = sp_3matrix_alloc(ph->nx, ph->ny, ph->nz);
  if (!ph->amplitudes_max) ph->amplitudes_max = sp_3matrix_alloc(ph->nx, ph->ny, ph->nz);
  for (int i = 0; i < ph->image_size; i++) {
    ph->amplitudes_min->data[i] = sp_real(amin->image->data[i]);
    ph->amplitudes_max->data[i] = sp_real(amax->image->data[i]);
  }
  Engine* e = ensure_engine(ph);
  if (!e) return;
  if (e->upload_margins(ph->amplitudes_min->data, ph->amplitudes_max->data)) report("sp_phaser_set_amplitudes_margins");
  else hs(ph)->have_margins = true;
}

// ---------------------------------------------------------------------------
// initialisation (reference src/phasing.c:606-778)
// ---------------------------------------------------------------------------
int sp_phaser_init_model(SpPhaser* ph, const Image* user_model, int flags) {
  if (!ph) return -1;
  if (user_model) phaser_check_dimensions(ph, user_model);
  if (!user_model && !(flags & (SpModelRandomPhases | SpModelZeroPhases | SpModelRandomValues))) return -3;
  if (!user_model && !ph->amplitudes) {
    fprintf(stderr, "sp_phaser_init_model: amplitudes must be set before a model is derived from them\n");
    return -4;
  }
  Engine* e = ensure_engine(ph);
  if (!e) {
    fprintf(stderr, "sp_phaser_init_model: HIP engine unavailable (%s)\n", spx::last_error());
    return -10;
  }
  if (ph->model) sp_image_free(ph->model);
  if (ph->model_change) sp_image_free(ph->model_change);
  ph->model = nullptr;
  ph->model_change = nullptr;
  int rc;
  if (user_model) {
    ph->model = sp_image_duplicate(user_model, SP_COPY_ALL);
    rc = e->init_model_user((const float2*)user_model->image->data, flags);
  } else {
    ph->model = alloc_like(ph);
    rc = e->init_model_from_amplitudes(ph->amplitudes->data, flags, hs(ph)->seed);
  }
  if (!rc) rc = e->finish_model_init(flags);
  if (rc) {
    report("sp_phaser_init_model");
    return rc == -3 ? -3 : -10;
  }
  if (e->download(e->g1(), (float2*)ph->model->image->data)) report("sp_phaser_init_model");
  ph->model->phased = 1;
  ph->model_change = sp_image_alloc(sp_image_x(ph->model), sp_image_y(ph->model), sp_image_z(ph->model));
  return 0;
}

int sp_phaser_init_support(SpPhaser* ph, const Image* support, int flags, real value) {
  if (support) {
    phaser_check_dimensions(ph, support);
    ensure_pixel_flags(ph);
    pull_flags(ph);
    for (int i = 0; i < ph->image_size; i++) {
      const Complex v = support->image->data[i];
      if (sp_real(v)) ph->pixel_flags->data[i] |= SpPixelInsideSupport;
      else ph->pixel_flags->data[i] &= ~SpPixelInsideSupport;
    }
    push_flags(ph);
    return engine_of(ph) ? 0 : -1;
  }
  if (flags & SpSupportFromPatterson) {
    if (!ph->amplitudes) return -1;
    ensure_pixel_flags(ph);
    pull_flags(ph);
    Engine* e = ensure_engine(ph);
    if (!e) return -1;
    push_flags(ph);
    if (e->patterson_support(ph->amplitudes->data, value)) {
      report("sp_phaser_init_support");
      return -1;
    }
    hs(ph)->device_flags_newer = true;
    pull_flags(ph);
    return 0;
  }
  return -1;
}

// ---------------------------------------------------------------------------
// iteration and support scheduling (reference src/phasing.c:895-1024)
// ---------------------------------------------------------------------------
int sp_phaser_iterate(SpPhaser* ph, int iterations) {
  if (!ph) return -1;
  if (!ph->algorithm) return -2;
  if (!ph->model) return -3;
  if ((!ph->amplitudes && ph->phasing_objective == SpRecoverPhases) ||
      (!ph->phased_amplitudes && ph->phasing_objective == SpRecoverAmplitudes))
    return -4;
  if (!ph->pixel_flags) return -5;
  if (!ph->model_change) return -6;
  const int t = ph->algorithm->type;
  if (t != SpHIO && t != SpRAAR && t != SpDiffMap && t != SpER) return -7;
  if (!engine_of(ph)) return -8;
  int ret = 0;
  SpSupportArray* sup = ph->sup_algorithm;
  if (sup && sup->size > 0 && sup->algorithms[0]->function) {
    if (sup->update_period < 1) {
      fprintf(stderr, "sp_phaser_iterate: support update_period must be >= 1\n");
      return -12;
    }
    while (iterations) {
      int to_support_update = sup->update_period - 1 - ph->iteration % sup->update_period;
      const int to_iterate = iterations < to_support_update ? iterations : to_support_update;
      ret = run_phasing(ph, to_iterate);
      if (ret) return ret;
      iterations -= to_iterate;
      to_support_update -= to_iterate;
      if (to_support_update == 0 && iterations > 0) {
        sp_support_array_update(sup, ph);
        ph->iteration++;
        iterations -= 1;
      }
    }
  } else if (iterations) {
    ret = run_phasing(ph, iterations);
  }
  return ret;
}

// ---------------------------------------------------------------------------
// getters (reference src/phasing.c:365-554,780-801)
// ---------------------------------------------------------------------------
const Image* sp_phaser_amplitudes(SpPhaser* ph) {
  if (!ph->amplitudes_image) {
    ph->amplitudes_image = alloc_like(ph);
    pull_flags(ph);
    for (int i = 0; i < ph->image_size; i++) {
      ph->amplitudes_image->image->data[i] = sp_cinit(ph->amplitudes->data[i], 0);
      ph->amplitudes_image->mask->data[i] = (ph->pixel_flags->data[i] & SpPixelMeasuredAmplitude) ? 1 : 0;
    }
  }
  return ph->amplitudes_image;
}

static Image* model_nc(SpPhaser* ph) {
  if (ph->model_iteration != ph->iteration) {
    if (!ph->model) ph->model = alloc_like(ph);
    ph->model_iteration = ph->iteration;
    Engine* e = engine_of(ph);
    if (!e || download_into(ph, ph->model, e->g1())) return nullptr;
  }
  return ph->model;
}

const Image* sp_phaser_model(SpPhaser* ph) { return model_nc(ph); }

const Image* sp_phaser_model_before_projection(SpPhaser* ph) {
  if (ph->model_before_projection_iteration != ph->iteration) {
    if (!ph->model_before_projection) ph->model_before_projection = alloc_like(ph);
    ph->model_before_projection_iteration = ph->iteration;
    Engine* e = engine_of(ph);
    if (!e || download_into(ph, ph->model_before_projection, e->gp_buf())) return nullptr;
  }
  return ph->model_before_projection;
}

const Image* sp_phaser_model_with_support(SpPhaser* ph) {
  Image* model = model_nc(ph);
  const Image* support = sp_phaser_support(ph);
  if (!model || !support) return nullptr;
  for (long long i = 0; i < sp_image_size(model); i++) {
    const Complex v = support->image->data[i];
    model->mask->data[i] = (sp_real(v) == 0.0f && sp_imag(v) == 0.0f) ? 0 : 1;
  }
  return model;
}

static Image* fmodel_nc(SpPhaser* ph) {
  if (ph->fmodel_iteration != ph->iteration) {
    if (!ph->fmodel) ph->fmodel = alloc_like(ph);
    ph->fmodel_iteration = ph->iteration;
    Engine* e = engine_of(ph);
    if (!e) return nullptr;
    // the efourier pass writes the transform of g1 into a temporary buffer
    float2* dev = nullptr;
    if (hipMalloc(&dev, sizeof(float2) * (size_t)ph->image_size) != hipSuccess) return nullptr;
    int rc = e->fmodel(dev);
    if (!rc) rc = e->download(dev, (float2*)ph->fmodel->image->data);
    (void)hipFree(dev);
    if (rc) {
      report("sp_phaser_fmodel");
      return nullptr;
    }
  }
  ph->fmodel->phased = 1;
  return ph->fmodel;
}

const Image* sp_phaser_fmodel(SpPhaser* ph) { return fmodel_nc(ph); }

const Image* sp_phaser_fmodel_with_mask(SpPhaser* ph) {
  Image* f = fmodel_nc(ph);
  const Image* a = sp_phaser_amplitudes(ph);
  if (!f || !a) return nullptr;
  memcpy(f->mask->data, a->mask->data, sizeof(int) * (size_t)ph->image_size);
  return f;
}

const Image* sp_phaser_old_model(SpPhaser* ph) {
  if (ph->old_model_iteration != ph->iteration) {
    if (!ph->old_model) ph->old_model = alloc_like(ph);
    ph->old_model_iteration = ph->iteration;
    Engine* e = engine_of(ph);
    if (!e || download_into(ph, ph->old_model, e->g0())) return nullptr;
  }
  return ph->old_model;
}

Image* sp_phaser_model_change(SpPhaser* ph) {
  if (ph->model_iteration != ph->iteration) sp_phaser_model(ph);
  if (ph->model_change_iteration != ph->iteration) {
    ph->model_change_iteration = ph->iteration;
    Engine* e = engine_of(ph);
    if (!e) return nullptr;
    float2* dev = nullptr;
    if (hipMalloc(&dev, sizeof(float2) * (size_t)ph->image_size) != hipSuccess) return nullptr;
    int rc = e->model_change(dev);
    if (!rc) rc = e->download(dev, (float2*)ph->model_change->image->data);
    (void)hipFree(dev);
    if (rc) {
      report("sp_phaser_model_change");
      return nullptr;
    }
  }
  return ph->model_change;
}

const Image* sp_phaser_support(SpPhaser* ph) {
  if (!ph->support) ph->support = alloc_like(ph);
  if (ph->iteration != ph->support_iteration) {
    ph->support_iteration = ph->iteration;
    Engine* e = engine_of(ph);
    if (!e || !ph->pixel_flags) return nullptr;
    hs(ph)->device_flags_newer = true;
    pull_flags(ph);
    for (int i = 0; i < ph->image_size; i++)
      ph->support->image->data[i] = sp_cinit((ph->pixel_flags->data[i] & SpPixelInsideSupport) ? 1.f : 0.f, 0.f);
  }
  return ph->support;
}

// ---------------------------------------------------------------------------
// metrics (reference src/phasing.c:803-893)
// ---------------------------------------------------------------------------
real sp_phaser_ereal(SpPhaser* ph) {
  Engine* e = engine_of(ph);
  double s[2];
  if (!e || e->ereal(s)) return -3;
  return (real)sqrt(s[0] / s[1]);
}

real sp_phaser_support_fraction(SpPhaser* ph) {
  Engine* e = engine_of(ph);
  double c;
  if (!e || e->support_count(&c)) return -3;
  return (real)c / (real)ph->image_size;
}

real sp_phaser_efourier(SpPhaser* ph, real* FcFo) {
  Engine* e = engine_of(ph);
  double s[4];
  if (!e || e->efourier(s, nullptr)) return -3;
  if (FcFo) *FcFo = (real)(s[2] / s[3]);
  return (real)sqrt(s[0] / s[1]);
}

// ---------------------------------------------------------------------------
// support algorithms (reference src/support_update.c)
// ---------------------------------------------------------------------------
SpSupportAlgorithm* sp_support_threshold_alloc(sp_smap* blur_radius, sp_smap* threshold) {
  SpSupportAlgorithm* r = (SpSupportAlgorithm*)malloc(sizeof(SpSupportAlgorithm));
  r->type = SpSupportThreshold;
  SpSupportThresholdParameters* p = (SpSupportThresholdParameters*)malloc(sizeof(SpSupportThresholdParameters));
  p->blur_radius_map = blur_radius;
  p->threshold = threshold;
  r->params = p;
  r->function = (void*)sp_support_threshold_update_support;
  return r;
}

SpSupportAlgorithm* sp_support_area_alloc(sp_smap* blur_radius, sp_smap* area) {
  SpSupportAlgorithm* r = (SpSupportAlgorithm*)malloc(sizeof(SpSupportAlgorithm));
  r->type = SpSupportArea;
  SpSupportAreaParameters* p = (SpSupportAreaParameters*)malloc(sizeof(SpSupportAreaParameters));
  p->blur_radius_map = blur_radius;
  p->area = area;
  r->params = p;
  r->function = (void*)sp_support_area_update_support;
  return r;
}

SpSupportAlgorithm* sp_support_static_alloc(void) {
  SpSupportAlgorithm* r = (SpSupportAlgorithm*)malloc(sizeof(SpSupportAlgorithm));
  r->type = SpSupportStatic;
  r->params = malloc(sizeof(SpSupportStaticParameters));
  r->function = (void*)sp_support_static_update_support;
  return r;
}

SpSupportAlgorithm* sp_support_close_alloc(int size) {
  SpSupportAlgorithm* r = (SpSupportAlgorithm*)malloc(sizeof(SpSupportAlgorithm));
  r->type = SpSupportClose;
  SpSupportCloseParameters* p = (SpSupportCloseParameters*)malloc(sizeof(SpSupportCloseParameters));
  p->size = size;
  r->params = p;
  r->function = (void*)sp_support_close_update_support;
  return r;
}

SpSupportAlgorithm* sp_support_centre_image_alloc(void) {
  SpSupportAlgorithm* r = (SpSupportAlgorithm*)malloc(sizeof(SpSupportAlgorithm));
  r->type = SpSupportCentreImage;
  r->params = nullptr;
  r->function = (void*)sp_support_centre_image;
  return r;
}

namespace {
// sp_cabs (include/spimage/linear_alg.h:163-165): the sum in float, then the
// square root rounded to float
inline float cabs_host(Complex c) { return sqrtf(sp_imag(c) * sp_imag(c) + sp_real(c) * sp_real(c)); }

// The template parameters as callers see them (SpSupportTemplateParameters
// first), then this library's own: a key naming the device copy of the
// template and the template's magnitudes |blured| on the host.
struct TemplateParams {
  SpSupportTemplateParameters base;
  unsigned long long key;
  float* abs_blured;
};
unsigned long long g_template_key = 0;

long template_index(long n, real area, real fraction) {
  // (int)(size * area * fraction) in float (src/support_update.c:72,195);
  // clamped to the array
  long k = (long)(int)((real)n * area * fraction);
  if (k >= n) k = n - 1;
  if (k < 0) k = 0;
  return k;
}
}  // namespace

// reference src/support_update.c:47-85
SpSupportAlgorithm* sp_support_template_alloc(Image* initial_support, real blur_radius, sp_smap* area) {
  if (!initial_support || !area) return nullptr;
  Image* blured = sp_gaussian_blur(initial_support, blur_radius);
  if (!blured) return nullptr;
  TemplateParams* x = (TemplateParams*)calloc(1, sizeof(TemplateParams));
  SpSupportTemplateParameters* p = &x->base;
  p->blured = blured;
  p->area = area;
  const long n = sp_image_size(initial_support);
  p->original_area = 0;
  for (long i = 0; i < n; i++)
    if (sp_real(initial_support->image->data[i]) != 0 || sp_imag(initial_support->image->data[i]) != 0)
      p->original_area += 1.0;
  p->original_area /= (real)n;
  // the blurred template sorted by descending magnitude; equal magnitudes
  // may come in any order, as with the reference's qsort
  x->abs_blured = (float*)malloc(sizeof(float) * n);
  std::vector<long> order((size_t)n);
  for (long i = 0; i < n; i++) {
    x->abs_blured[i] = cabs_host(blured->image->data[i]);
    order[i] = i;
  }
  std::sort(order.begin(), order.end(), [&](long a, long b) { return x->abs_blured[a] > x->abs_blured[b]; });
  p->sorted = sp_image_duplicate(blured, SP_COPY_DATA);
  for (long i = 0; i < n; i++) p->sorted->image->data[i] = blured->image->data[order[i]];
  const real max_threshold =
      cabs_host(p->sorted->image->data[template_index(n, spx_smap_max(area), p->original_area)]);
  const real min_threshold =
      cabs_host(p->sorted->image->data[template_index(n, spx_smap_min(area), p->original_area)]);
  if (max_threshold < 1e-4) fprintf(stderr, "Dangerously large template area with this threshold!\n");
  if (1.0 - min_threshold < 1e-4) fprintf(stderr, "Dangerously small template area with this threshold!\n");
  x->key = __atomic_add_fetch(&g_template_key, 1, __ATOMIC_RELAXED);
  SpSupportAlgorithm* r = (SpSupportAlgorithm*)malloc(sizeof(SpSupportAlgorithm));
  r->type = SpSupportTemplate;
  r->params = x;
  r->function = (void*)sp_support_template_update_support;
  return r;
}

SpSupportArray* sp_support_array_alloc(int size, int update_period) {
  SpSupportArray* r = (SpSupportArray*)malloc(sizeof(SpSupportArray));
  r->size = size;
  r->algorithms = (SpSupportAlgorithm**)calloc(size > 0 ? size : 1, sizeof(SpSupportAlgorithm*));
  r->update_period = update_period;
  return r;
}
void sp_support_array_set(SpSupportArray* a, int i, SpSupportAlgorithm* alg) { a->algorithms[i] = alg; }
SpSupportArray* sp_support_array_init(SpSupportAlgorithm* alg, int update_period) {
  SpSupportArray* r = sp_support_array_alloc(1, update_period);
  sp_support_array_set(r, 0, alg);
  return r;
}
void sp_support_array_append(SpSupportArray* a, SpSupportAlgorithm* alg) {
  SpSupportAlgorithm** n = (SpSupportAlgorithm**)malloc((a->size + 1) * sizeof(SpSupportAlgorithm*));
  if (a->size) memcpy(n, a->algorithms, a->size * sizeof(SpSupportAlgorithm*));
  n[a->size] = alg;
  free(a->algorithms);
  a->algorithms = n;
  a->size += 1;
}

int sp_support_array_update(SpSupportArray* a, SpPhaser* ph) {
  for (int i = 0; i < a->size; i++)
    ((int (*)(SpSupportAlgorithm*, SpPhaser*))a->algorithms[i]->function)(a->algorithms[i], ph);
  return 0;
}

int sp_support_area_update_support(SpSupportAlgorithm* alg, SpPhaser* ph) {
  SpSupportAreaParameters* p = (SpSupportAreaParameters*)alg->params;
  Engine* e = engine_of(ph);
  if (!e) return -1;
  const real radius = spx_smap_interpolate(p->blur_radius_map, (real)ph->iteration);
  const real area = spx_smap_interpolate(p->area, (real)ph->iteration);
  if (e->shrinkwrap_area(radius, area)) {
    report("sp_support_area_update_support");
    return -1;
  }
  hs(ph)->device_flags_newer = true;
  return 0;
}

int sp_support_threshold_update_support(SpSupportAlgorithm* alg, SpPhaser* ph) {
  SpSupportThresholdParameters* p = (SpSupportThresholdParameters*)alg->params;
  Engine* e = engine_of(ph);
  if (!e) return -1;
  const real radius = spx_smap_interpolate(p->blur_radius_map, (real)ph->iteration);
  const real rel = spx_smap_interpolate(p->threshold, (real)ph->iteration);
  if (e->shrinkwrap_threshold(radius, rel)) {
    report("sp_support_threshold_update_support");
    return -1;
  }
  hs(ph)->device_flags_newer = true;
  return 0;
}

int sp_support_static_update_support(SpSupportAlgorithm*, SpPhaser*) { return 0; }

// reference src/support_update.c:181-207: the support is the blurred
// template thresholded at the magnitude of its sorted entry
// (int)(N * area(iteration) * original_area)
int sp_support_template_update_support(SpSupportAlgorithm* alg, SpPhaser* ph) {
  TemplateParams* x = (TemplateParams*)alg->params;
  SpSupportTemplateParameters* p = &x->base;
  Engine* e = engine_of(ph);
  if (!e) return -1;
  const long n = sp_image_size(p->sorted);
  if (n != ph->image_size) {
    fprintf(stderr, "sp_support_template_update_support: the template has %ld pixels, the phaser %d\n", n,
            ph->image_size);
    return -1;
  }
  const real area = spx_smap_interpolate(p->area, (real)ph->iteration);
  const real t = cabs_host(p->sorted->image->data[template_index(n, area, p->original_area)]);
  if (e->support_template(x->key, x->abs_blured, t)) {
    report("sp_support_template_update_support");
    return -1;
  }
  hs(ph)->device_flags_newer = true;
  return 0;
}

// reference src/support_update.c:213-280, on the device
int sp_support_close_update_support(SpSupportAlgorithm* alg, SpPhaser* ph) {
  SpSupportCloseParameters* p = (SpSupportCloseParameters*)alg->params;
  Engine* e = engine_of(ph);
  if (!e) return -1;
  if (e->support_close(p->size)) {
    report("sp_support_close_update_support");
    return -1;
  }
  hs(ph)->device_flags_newer = true;
  return 0;
}

// reference src/support_update.c:282-389 (sp_support_centre_image_cpu) with
// sp_pixel_flags_translate_mask (src/image_util.c:3587-3611) and
// sp_i3matrix_binary_center_of_mass_shifted (src/linear_alg.c:444-475), on
// the device (Engine::support_centre_image): no download or upload.
//
// The support bit moves cyclically; the other flag bits stay in place.  The
// reference also translates its cached copy of the model (ph->model), which
// the next model read replaces because the update consumes an iteration;
// the iterate itself does not move, so it does not move here either.
int sp_support_centre_image(SpSupportAlgorithm*, SpPhaser* ph) {
  Engine* e = engine_of(ph);
  if (!e) return -1;
  if (e->support_centre_image()) {
    report("sp_support_centre_image");
    return -1;
  }
  hs(ph)->device_flags_newer = true;
  return 0;
}

// ---------------------------------------------------------------------------
// Image-level FFT and blur (reference src/fft.c:53-99,152-189; image_filter.c:80-91)
// ---------------------------------------------------------------------------
static void set_center_half(Image* r) {
  r->detector->image_center[0] = sp_image_x(r) / 2.0;
  r->detector->image_center[1] = sp_image_y(r) / 2.0;
  r->detector->image_center[2] = sp_image_z(r) / 2.0;
}

Image* sp_image_fft(const Image* img) {
  Image* r = sp_image_duplicate(img, SP_COPY_DETECTOR);
  r->phased = 1;
  if (spx::fft_host((const float2*)img->image->data, (float2*)r->image->data, sp_image_x(img), sp_image_y(img),
                    sp_image_z(img), -1)) {
    report("sp_image_fft");
    sp_image_free(r);
    return nullptr;
  }
  r->shifted = 1;
  set_center_half(r);
  return r;
}

Image* sp_image_ifft(const Image* img) {
  if (!img->phased) {
    fprintf(stderr, "Error: Trying reverse fft an unphased image!\n");
    abort();
  }
  if (!img->shifted) {
    fprintf(stderr, "Error: Trying to rev_fft an unshifted image!\n");
    abort();
  }
  Image* r = sp_image_duplicate(img, SP_COPY_DETECTOR);
  if (spx::fft_host((const float2*)img->image->data, (float2*)r->image->data, sp_image_x(img), sp_image_y(img),
                    sp_image_z(img), +1)) {
    report("sp_image_ifft");
    sp_image_free(r);
    return nullptr;
  }
  r->shifted = 0;
  return r;
}

void sp_image_fft_fast(const Image* in, Image* out) {
  if (spx::fft_host((const float2*)in->image->data, (float2*)out->image->data, sp_image_x(in), sp_image_y(in),
                    sp_image_z(in), -1))
    report("sp_image_fft_fast");
}

void sp_image_ifft_fast(const Image* in, Image* out) {
  if (spx::fft_host((const float2*)in->image->data, (float2*)out->image->data, sp_image_x(in), sp_image_y(in),
                    sp_image_z(in), +1))
    report("sp_image_ifft_fast");
}

Image* sp_image_fftw3(const Image* img) { return sp_image_fft(img); }
Image* sp_image_ifftw3(const Image* img) { return sp_image_ifft(img); }
void sp_image_fftw3_fast(const Image* in, Image* out) { sp_image_fft_fast(in, out); }
void sp_image_ifftw3_fast(const Image* in, Image* out) { sp_image_ifft_fast(in, out); }
Image* sp_image_cuda_fft(const Image* img) { return sp_image_fft(img); }
Image* sp_image_cuda_ifft(const Image* img) { return sp_image_ifft(img); }

Image* sp_gaussian_blur(Image* in, real radius) {
  Image* r = sp_image_duplicate(in, SP_COPY_DETECTOR);
  if (spx::blur_host((const float2*)in->image->data, (float2*)r->image->data, sp_image_x(in), sp_image_y(in),
                     sp_image_z(in), radius)) {
    report("sp_gaussian_blur");
    sp_image_free(r);
    return nullptr;
  }
  // metadata of sp_image_ifft(sp_image_fft(in))
  r->phased = 1;
  r->shifted = 0;
  set_center_half(r);
  return r;
}

// ---------------------------------------------------------------------------
// plain-pointer engine ABI (include/sp_hip_engine.h)
// ---------------------------------------------------------------------------
int spx_device_available(void) { return spx::device_available() ? 1 : 0; }
const char* spx_last_error(void) { return spx::last_error(); }
int spx_device_memory(unsigned long long* free_b, unsigned long long* total_b) {
  if (!spx::device_available()) {
    spx::set_error("no HIP device available");
    return -10;
  }
  size_t f = 0, t = 0;
  const hipError_t e = hipMemGetInfo(&f, &t);
  if (e != hipSuccess) {
    spx::set_error("hipMemGetInfo: %s", hipGetErrorString(e));
    return -20;
  }
  if (free_b) *free_b = f;
  if (total_b) *total_b = t;
  return 0;
}

void* spx_host_alloc(unsigned long long bytes) {
  if (!spx::device_available()) {
    spx::set_error("no HIP device available");
    return nullptr;
  }
  void* p = nullptr;
  const hipError_t e = hipHostMalloc(&p, (size_t)(bytes ? bytes : 1), hipHostMallocDefault);
  if (e != hipSuccess) {
    spx::set_error("hipHostMalloc(%llu): %s", bytes, hipGetErrorString(e));
    return nullptr;
  }
  return p;
}

void spx_host_free(void* p) {
  if (p) (void)hipHostFree(p);
}

int spx_supported_length(int L) { return spx::find_length(L) != nullptr ? 1 : 0; }

// ---------------------------------------------------------------------------
// device queries of the reference's CUDA glue (include/sp_cuda_util.h;
// reference src/cuda_util.c:10-29,39-60)
// ---------------------------------------------------------------------------
SpCUDADeviceType sp_cuda_get_device_type(void) {
  return spx::device_available() ? SpCUDAHardwareDevice : SpCUDANoDevice;
}

void sp_cuda_launch_parameters(int64_t size, int* gridSize, int* blockSize) {
  // HIP's limits: 1024 threads per block, 2^31 - 1 blocks along x
  const int64_t max_block = 1024, max_grid = 2147483647;
  int64_t block = 64;
  while (size > max_grid * block) block *= 2;
  if (block > max_block) {
    fprintf(stderr, "sp_cuda_launch_parameters: %lld elements exceed %lld blocks of %lld threads\n",
            (long long)size, (long long)max_grid, (long long)max_block);
    abort();
  }
  *blockSize = (int)block;
  *gridSize = (int)((size + block - 1) / block);
}

int spx_tune_shape(int nx, int ny, int nz, int nrec) {
  if (!spx::device_available()) {
    spx::set_error("no HIP device");
    return -20;
  }
  spx::Shape sh;
  return spx::make_shape(nx, ny, nz, nrec, &sh);
}

int spx_fft_variant(int L, int tiled, int* row_e, int* col_e) {
  if (!spx::find_length(L)) {
    spx::set_error("length %d is not supported", L);
    return -11;
  }
  (void)tiled;  // one choice per shape serves both layouts
  int e[3];
  if (int rc = spx::shape_variant_elements(L, L, 1, e)) return rc;
  if (row_e) *row_e = e[0];
  if (col_e) *col_e = e[1];
  return 0;
}

int spx_phaser_tiled(SpPhaser* ph) {
  Engine* e = engine_of(ph);
  if (!e) {
    spx::set_error("phaser has no HIP engine");
    return -1;
  }
  return e->layout();
}

int spx_phaser_nt_hints(SpPhaser* ph) {
  Engine* e = engine_of(ph);
  if (!e) {
    spx::set_error("phaser has no HIP engine");
    return -1;
  }
  return e->nt_hints();
}

int spx_fft_device(void* d_in, void* d_out, int nx, int ny, int nz, int direction) {
  if (direction != -1 && direction != 1) {
    spx::set_error("direction must be -1 or +1");
    return -1;
  }
  return spx::fft_device((const float2*)d_in, (float2*)d_out, nx, ny, nz, direction);
}

int spx_fft_noalloc(const float* in, float* out, int nx, int ny, int nz, int direction) {
  if (direction != -1 && direction != 1) {
    spx::set_error("direction must be -1 or +1");
    return -1;
  }
  if (!in || !out) {
    spx::set_error("null array");
    return -1;
  }
  return spx::fft_fixed((const float2*)in, (float2*)out, nx, ny, nz, direction);
}

static spx::IterParams iter_params(SpPhaser* ph) {
  spx::IterParams p{};
  p.alg = ph->algorithm->type;
  p.objective = ph->phasing_objective;
  if (p.alg == SpER) {
    p.constraints = ((SpPhasingERParameters*)ph->algorithm->params)->constraints;
  } else if (p.alg == SpDiffMap) {
    SpPhasingDiffMapParameters* d = (SpPhasingDiffMapParameters*)ph->algorithm->params;
    p.constraints = d->constraints;
    p.gamma1 = d->gamma1;
    p.gamma2 = d->gamma2;
  } else {
    p.constraints = ((SpPhasingHIOParameters*)ph->algorithm->params)->constraints;
  }
  return p;
}

int spx_phaser_time_iterations(SpPhaser* ph, int iterations, float* ms) {
  Engine* e = engine_of(ph);
  if (!e || !ph->algorithm || !ms) {
    spx::set_error("phaser has no HIP engine or algorithm");
    return -8;
  }
  std::vector<float> betas((size_t)(iterations > 0 ? iterations : 1));
  for (int i = 0; i < iterations; ++i) betas[i] = beta_at(ph->algorithm, ph->iteration + i);
  const int rc = e->time_iterations(iter_params(ph), betas.data(), iterations, ms);
  if (!rc) ph->iteration += iterations;
  return rc;
}

int spx_phaser_time_passes(SpPhaser* ph, int iterations, float* ms_fourier, float* ms_real) {
  Engine* e = engine_of(ph);
  if (!e || !ph->algorithm || !ms_fourier || !ms_real) {
    spx::set_error("phaser has no HIP engine or algorithm");
    return -8;
  }
  std::vector<float> betas((size_t)(iterations > 0 ? iterations : 1));
  for (int i = 0; i < iterations; ++i) betas[i] = beta_at(ph->algorithm, ph->iteration + i);
  const int rc = e->time_passes(iter_params(ph), betas.data(), iterations, ms_fourier, ms_real);
  if (!rc) ph->iteration += iterations;
  return rc;
}

int spx_phaser_device_arrays(SpPhaser* ph, void** g1, void** g0, void** gp) {
  Engine* e = engine_of(ph);
  if (!e) return -8;
  if (g1) *g1 = e->g1();
  if (g0) *g0 = e->g0();
  if (gp) *gp = e->gp_buf();
  return 0;
}

int spx_phaser_last_threshold(SpPhaser* ph, float* t) {
  Engine* e = engine_of(ph);
  if (!e || !t) return -8;
  return e->last_threshold(t);
}

void spx_phaser_set_seed(SpPhaser* ph, unsigned long long seed) {
  hs(ph)->seed = seed;
  if (Engine* e = engine_of(ph)) e->seed = seed;
}

}  // extern "C"
