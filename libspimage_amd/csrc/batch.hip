// Batched phasing (include/sp_batch.h): nrec same-shape reconstructions in
// one spx::Engine, advanced with one launch per pass for the whole batch.
// The iteration counter, algorithm and support schedule are shared and follow
// the reference's sp_phaser_iterate scheduler (src/phasing.c:1000-1022).
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <memory>
#include <vector>

#include "engine.h"
#include "sp_batch.h"
#include "spimage.h"

struct SpxBatch {
  spx::Engine* e = nullptr;
  int nx = 0, ny = 0, nz = 0, nrec = 0;
  long N = 0;
  // host copy of the raw amplitudes (nrec * N), zero until set; calloc's
  // pages stay untouched (and free to map) until a reconstruction is set
  std::unique_ptr<float, decltype(&free)> amp{nullptr, &free};
  SpPhasingAlgorithm* alg = nullptr;
  SpSupportArray* sup = nullptr;
  bool rec_betas = false;  // per-reconstruction betas replace the schedule
  bool graphs = false;
  int iteration = 0;
};

namespace {

int fail(int code, const char* msg) {
  spx::set_error("%s", msg);
  return code;
}

bool check(SpxBatch* b, int rec) { return b && b->e && rec < b->nrec; }

spx::IterParams params_of(const SpxBatch* b) {
  spx::IterParams p{};
  const SpPhasingAlgorithm* a = b->alg;
  p.alg = a->type;
  p.objective = SpRecoverPhases;
  if (a->type == SpER) {
    p.constraints = ((SpPhasingERParameters*)a->params)->constraints;
  } else if (a->type == SpDiffMap) {
    SpPhasingDiffMapParameters* d = (SpPhasingDiffMapParameters*)a->params;
    p.constraints = d->constraints;
    p.gamma1 = d->gamma1;
    p.gamma2 = d->gamma2;
  } else {
    p.constraints = ((SpPhasingHIOParameters*)a->params)->constraints;
  }
  return p;
}

float beta_of(const SpxBatch* b, int it) {
  const SpPhasingAlgorithm* a = b->alg;
  sp_smap* m = nullptr;
  if (a->type == SpHIO || a->type == SpRAAR) m = ((SpPhasingHIOParameters*)a->params)->beta;
  else if (a->type == SpDiffMap) m = ((SpPhasingDiffMapParameters*)a->params)->beta;
  return m ? spx_smap_interpolate(m, (real)it) : 0.f;
}

int run_phasing(SpxBatch* b, int n) {
  if (n <= 0) return 0;
  const spx::IterParams p = params_of(b);
  std::vector<float> betas((size_t)n);
  bool constant = true;
  for (int i = 0; i < n; ++i) {
    betas[i] = beta_of(b, b->iteration + i);
    if (betas[i] != betas[0]) constant = false;
  }
  int rc;
  // a graph replays one beta; per-reconstruction betas are read from the
  // device table at run time, so either makes every replay exact
  if (b->graphs && n >= 3 && (constant || b->rec_betas)) {
    // graph replays for all but the last iteration, which also writes gp
    rc = b->e->iterate_graph(p, betas[0], n - 1);
    if (!rc) rc = b->e->iterate(p, &betas[n - 1], 1, true);
  } else {
    rc = b->e->iterate(p, betas.data(), n, true);
  }
  if (rc) return rc;
  b->iteration += n;
  return 0;
}

int support_update(SpxBatch* b) {
  for (int i = 0; i < b->sup->size; ++i) {
    SpSupportAlgorithm* a = b->sup->algorithms[i];
    int rc = 0;
    if (a->type == SpSupportArea) {
      SpSupportAreaParameters* p = (SpSupportAreaParameters*)a->params;
      rc = b->e->shrinkwrap_area(spx_smap_interpolate(p->blur_radius_map, (real)b->iteration),
                                 spx_smap_interpolate(p->area, (real)b->iteration));
    } else if (a->type == SpSupportThreshold) {
      SpSupportThresholdParameters* p = (SpSupportThresholdParameters*)a->params;
      rc = b->e->shrinkwrap_threshold(spx_smap_interpolate(p->blur_radius_map, (real)b->iteration),
                                      spx_smap_interpolate(p->threshold, (real)b->iteration));
    } else if (a->type == SpSupportCentreImage) {
      rc = b->e->support_centre_image();
    } else if (a->type == SpSupportClose) {
      rc = b->e->support_close(((SpSupportCloseParameters*)a->params)->size);
    } else if (a->type != SpSupportStatic) {
      return fail(-1, "batched support updates: area, threshold, centre image, close and static only");
    }
    if (rc) return rc;
  }
  return 0;
}

}  // namespace

extern "C" {

SpxBatch* spx_batch_create(int nx, int ny, int nz, int nrec) {
  // batches run the fused passes only: an axis without a Stockham plan
  // (Bluestein transforms) is refused for any number of reconstructions
  const int dims[3] = {nx, ny, nz};
  for (int d : dims)
    if (d > 1 && !spx::find_length(d)) {
      fail(-11, "batches need axis lengths that are products of 2, 3, 5 and 7 up to 8192");
      return nullptr;
    }
  spx::Engine* e = spx::Engine::create(nx, ny, nz, nrec);
  if (!e) return nullptr;
  SpxBatch* b = new SpxBatch();
  b->e = e;
  b->nx = nx;
  b->ny = ny;
  b->nz = nz;
  b->nrec = nrec;
  b->N = (long)nx * ny * nz;
  b->amp.reset((float*)calloc((size_t)b->N * nrec, sizeof(float)));
  if (!b->amp) {
    fail(-20, "host memory for the amplitudes");
    spx_batch_free(b);
    return nullptr;
  }
  return b;
}

void spx_batch_free(SpxBatch* b) {
  if (!b) return;
  delete b->e;
  delete b;
}

int spx_batch_size(const SpxBatch* b) { return b ? b->nrec : 0; }

int spx_batch_set_amplitudes(SpxBatch* b, int rec, const float* amplitudes, const int* mask) {
  if (!check(b, rec) || !amplitudes || !mask) return fail(-1, "bad batch arguments");
  const long n = rec < 0 ? b->N * b->nrec : b->N, off = rec < 0 ? 0 : rec * b->N;
  std::vector<uint8_t> fl((size_t)n);
  if (int rc = b->e->download_flags(fl.data(), rec)) return rc;
  for (long i = 0; i < n; ++i) {
    fl[i] = (uint8_t)((fl[i] & SpPixelInsideSupport) | (mask[i] ? SpPixelMeasuredAmplitude : 0));
    b->amp.get()[off + i] = amplitudes[i];
  }
  return b->e->upload_amplitudes(amplitudes, fl.data(), rec);
}

int spx_batch_set_support(SpxBatch* b, int rec, const float* support) {
  if (!check(b, rec) || !support) return fail(-1, "bad batch arguments");
  const long n = rec < 0 ? b->N * b->nrec : b->N;
  std::vector<uint8_t> fl((size_t)n);
  if (int rc = b->e->download_flags(fl.data(), rec)) return rc;
  for (long i = 0; i < n; ++i)
    fl[i] = (uint8_t)((fl[i] & SpPixelMeasuredAmplitude) | (support[i] != 0.f ? SpPixelInsideSupport : 0));
  return b->e->upload_flags(fl.data(), rec);
}

int spx_batch_set_data(SpxBatch* b, int rec, const float* amplitudes, const int* mask, const float* support) {
  if (!check(b, rec) || !amplitudes || !mask || !support) return fail(-1, "bad batch arguments");
  const long n = rec < 0 ? b->N * b->nrec : b->N, off = rec < 0 ? 0 : rec * b->N;
  std::vector<uint8_t> fl((size_t)n);
  for (long i = 0; i < n; ++i) {
    fl[i] = (uint8_t)((mask[i] ? SpPixelMeasuredAmplitude : 0) | (support[i] != 0.f ? SpPixelInsideSupport : 0));
    b->amp.get()[off + i] = amplitudes[i];
  }
  return b->e->upload_amplitudes(amplitudes, fl.data(), rec);
}

int spx_batch_set_model(SpxBatch* b, int rec, const float* model) {
  if (!check(b, rec) || !model) return fail(-1, "bad batch arguments");
  return b->e->upload_model((const float2*)model, rec);
}

int spx_batch_init_model(SpxBatch* b, int flags, unsigned long long seed) {
  if (!b) return fail(-1, "bad batch arguments");
  int rc = b->e->init_model_from_amplitudes(b->amp.get(), flags, seed);
  if (!rc) rc = b->e->finish_model_init(flags);
  return rc;
}

int spx_batch_init_model_rec(SpxBatch* b, int rec, int flags, unsigned long long seed) {
  if (!b || rec < 0 || rec >= b->nrec) return fail(-1, "bad batch arguments");
  int rc = b->e->init_model_from_amplitudes(b->amp.get() + (size_t)rec * b->N, flags, seed, rec);
  if (!rc) rc = b->e->finish_model_init(flags, rec);
  return rc;
}

int spx_batch_query(SpxBatch* b) {
  if (!b) return fail(-1, "bad batch arguments");
  const hipError_t e = hipStreamQuery(b->e->s);
  if (e == hipSuccess) return 1;
  if (e == hipErrorNotReady) return 0;
  return fail(-20, hipGetErrorString(e));
}

long long spx_batch_estimate_bytes(int nx, int ny, int nz, int nrec) {
  // five complex64 arrays (two iterates, the row-transformed iterate, the
  // model before projection, scratch), two float arrays (encoded amplitudes,
  // |blur|), one flag byte, and the per-reconstruction select scratch
  const long long N = (long long)nx * ny * nz;
  return (long long)nrec * (N * (5 * 8 + 2 * 4 + 1) + (long long)(spx::kHistStride + spx::kCtlStride) * 4 + 4096);
}

int spx_batch_set_algorithm(SpxBatch* b, SpPhasingAlgorithm* alg) {
  if (!b || !alg) return fail(-2, "Algorithm is NULL!");
  if (alg->type != SpHIO && alg->type != SpRAAR && alg->type != SpER && alg->type != SpDiffMap)
    return fail(-7, "unknown phasing algorithm");
  b->alg = alg;
  return 0;
}

int spx_batch_set_rec_algorithms(SpxBatch* b, const int* alg_types, const float* betas) {
  if (!b) return fail(-1, "bad batch arguments");
  if (alg_types)
    for (int r = 0; r < b->nrec; ++r)
      if (alg_types[r] != SpHIO && alg_types[r] != SpRAAR && alg_types[r] != SpER)
        return fail(-7, "per-reconstruction algorithms are ER, HIO or RAAR");
  b->rec_betas = betas != nullptr;
  return b->e->set_rec_params(alg_types, betas);
}

int spx_batch_set_active(SpxBatch* b, const int* active) {
  if (!b) return fail(-1, "bad batch arguments");
  return b->e->set_rec_active(active);
}

int spx_batch_set_support_schedule(SpxBatch* b, SpSupportArray* sup) {
  if (!b) return fail(-1, "bad batch arguments");
  b->sup = sup;
  return 0;
}

int spx_batch_use_graphs(SpxBatch* b, int on) {
  if (!b) return fail(-1, "bad batch arguments");
  b->graphs = on != 0;
  return 0;
}

int spx_batch_iteration(const SpxBatch* b) { return b ? b->iteration : -1; }

int spx_batch_iterate(SpxBatch* b, int iterations) {
  if (!b) return -1;
  if (!b->alg) return -2;
  SpSupportArray* sup = b->sup;
  if (sup && sup->size > 0 && sup->algorithms[0]->function) {
    if (sup->update_period < 1) return fail(-12, "support update_period must be >= 1");
    while (iterations) {
      int to_update = sup->update_period - 1 - b->iteration % sup->update_period;
      const int n = iterations < to_update ? iterations : to_update;
      if (int rc = run_phasing(b, n)) return rc;
      iterations -= n;
      to_update -= n;
      if (to_update == 0 && iterations > 0) {
        if (int rc = support_update(b)) return rc;
        b->iteration++;
        iterations -= 1;
      }
    }
    return 0;
  }
  return run_phasing(b, iterations);
}

static void efourier_of_sums(const SpxBatch* b, const double* s, float* ef, float* fcfo) {
  for (int r = 0; r < b->nrec; ++r) {
    ef[r] = (float)sqrt(s[4 * r] / s[4 * r + 1]);
    if (fcfo) fcfo[r] = (float)(s[4 * r + 2] / s[4 * r + 3]);
  }
}

int spx_batch_efourier(SpxBatch* b, float* ef, float* fcfo) {
  if (!b || !ef) return fail(-1, "bad batch arguments");
  std::vector<double> s((size_t)4 * b->nrec);
  if (int rc = b->e->efourier(s.data(), nullptr)) return rc;
  efourier_of_sums(b, s.data(), ef, fcfo);
  return 0;
}

int spx_batch_efourier_enqueue(SpxBatch* b) {
  if (!b) return fail(-1, "bad batch arguments");
  return b->e->efourier_enqueue();
}

int spx_batch_efourier_fetch(SpxBatch* b, float* ef, float* fcfo) {
  if (!b || !ef) return fail(-1, "bad batch arguments");
  std::vector<double> s((size_t)4 * b->nrec);
  if (int rc = b->e->efourier_fetch(s.data())) return rc;
  efourier_of_sums(b, s.data(), ef, fcfo);
  return 0;
}

int spx_batch_ereal(SpxBatch* b, float* out) {
  if (!b || !out) return fail(-1, "bad batch arguments");
  std::vector<double> s((size_t)2 * b->nrec);
  if (int rc = b->e->ereal(s.data())) return rc;
  for (int r = 0; r < b->nrec; ++r) out[r] = (float)sqrt(s[2 * r] / s[2 * r + 1]);
  return 0;
}

int spx_batch_support_fraction(SpxBatch* b, float* out) {
  if (!b || !out) return fail(-1, "bad batch arguments");
  std::vector<double> s((size_t)b->nrec);
  if (int rc = b->e->support_count(s.data())) return rc;
  for (int r = 0; r < b->nrec; ++r) out[r] = (float)s[r] / (float)b->N;
  return 0;
}

int spx_batch_get_model(SpxBatch* b, int rec, float* out) {
  if (!check(b, rec) || !out) return fail(-1, "bad batch arguments");
  return b->e->download(b->e->g1(), (float2*)out, rec);
}

int spx_batch_get_model_before_projection(SpxBatch* b, int rec, float* out) {
  if (!check(b, rec) || !out) return fail(-1, "bad batch arguments");
  return b->e->download(b->e->gp_buf(), (float2*)out, rec);
}

int spx_batch_get_fmodel(SpxBatch* b, int rec, float* out) {
  if (!check(b, rec) || !out) return fail(-1, "bad batch arguments");
  float2* dev = nullptr;
  if (hipMalloc(&dev, sizeof(float2) * (size_t)b->N * b->nrec) != hipSuccess)
    return fail(-20, "device memory for the Fourier model");
  int rc = b->e->fmodel(dev);
  if (!rc) rc = b->e->download(dev, (float2*)out, rec);
  (void)hipFree(dev);
  return rc;
}

int spx_batch_get_support(SpxBatch* b, int rec, int* out) {
  if (!check(b, rec) || !out) return fail(-1, "bad batch arguments");
  const long n = rec < 0 ? b->N * b->nrec : b->N;
  std::vector<uint8_t> fl((size_t)n);
  if (int rc = b->e->download_flags(fl.data(), rec)) return rc;
  for (long i = 0; i < n; ++i) out[i] = fl[i] & SpPixelInsideSupport;
  return 0;
}

int spx_batch_time_iterations(SpxBatch* b, int iterations, float* ms) {
  if (!b || !b->alg || !ms) return fail(-1, "bad batch arguments");
  std::vector<float> betas((size_t)(iterations > 0 ? iterations : 1));
  for (int i = 0; i < iterations; ++i) betas[i] = beta_of(b, b->iteration + i);
  const int rc = b->e->time_iterations(params_of(b), betas.data(), iterations, ms);
  if (!rc) b->iteration += iterations;
  return rc;
}

int spx_batch_synchronize(SpxBatch* b) {
  if (!b) return fail(-1, "bad batch arguments");
  return hipStreamSynchronize(b->e->s) == hipSuccess ? 0 : fail(-20, "stream synchronisation failed");
}

}  // extern "C"
