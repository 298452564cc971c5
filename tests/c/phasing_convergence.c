/* A C caller of the phasing boundary, compiled with gcc against
 * include/spimage.h and linked to libspimage_hip.so, as the reference's C
 * tests and Hawk link libspimage.
 *
 * 1. Layout: the boundary's structs must match the reference headers'
 *    (tests/golden/reference_layout.json, passed in as -DREF_* macros by
 *    tests/test_c_abi.py), checked at compile time.
 * 2. With argument "run": the reference's convergence case
 *    (tests/PhasingTests.c:10-59): a random complex 4x4 object in an 8x8
 *    frame, its Fourier amplitudes with every pixel measured, the object's
 *    own support, HIO beta = 0.8 from random phases (SpModelRandomPhases),
 *    iterated one step at a time until sum |model change|^2 <= 1e-10 or 300
 *    iterations; up to 10 fresh attempts.  Exit status 0 when it converged. */
#include <stddef.h>
#include <stdio.h>
#include <string.h>

#include "spimage.h"

#if defined(REF_SIZEOF_IMAGE)
_Static_assert(sizeof(Image) == REF_SIZEOF_IMAGE, "Image differs from the reference");
_Static_assert(sizeof(SpSupportArray) == REF_SIZEOF_SPSUPPORTARRAY, "SpSupportArray differs from the reference");
_Static_assert(offsetof(SpSupportArray, update_period) == REF_SPSUPPORTARRAY_UPDATE_PERIOD,
               "SpSupportArray.update_period differs");
_Static_assert(offsetof(SpPhaser, algorithm) == REF_SPPHASER_ALGORITHM, "SpPhaser.algorithm differs");
_Static_assert(offsetof(SpPhaser, sup_algorithm) == REF_SPPHASER_SUP_ALGORITHM, "SpPhaser.sup_algorithm differs");
_Static_assert(offsetof(SpPhaser, iteration) == REF_SPPHASER_ITERATION, "SpPhaser.iteration differs");
_Static_assert(offsetof(SpPhaser, model) == REF_SPPHASER_MODEL, "SpPhaser.model differs");
_Static_assert(offsetof(SpPhaser, engine) == REF_SPPHASER_ENGINE, "SpPhaser.engine differs");
_Static_assert(offsetof(SpPhaser, gp) == REF_SPPHASER_GP, "SpPhaser.gp differs");
#endif

static double integrate2(const Image *a) {
  double s = 0;
  for (long long i = 0; i < sp_image_size(a); i++) s += sp_cabs2(a->image->data[i]);
  return s;
}

static int converge_once(SpPhasingAlgorithm *alg, double tol, int *iters) {
  const int size = 4, over = 2, n = size * over;
  Image *solution = sp_image_alloc(n, n, 1);
  for (long long i = 0; i < sp_image_size(solution); i++) solution->image->data[i] = sp_cinit(0, 0);
  for (int x = 0; x < size; x++)
    for (int y = 0; y < size; y++)
      solution->image->data[y * n + x] = sp_cinit(p_drand48() - 0.5f, p_drand48() - 0.5f);
  Image *f = sp_image_fft(solution);
  if (!f) return -100;
  for (long long i = 0; i < sp_image_size(f); i++) {
    f->image->data[i] = sp_cinit(sp_cabs(f->image->data[i]), 0); /* sp_image_rephase(f, SP_ZERO_PHASE) */
    f->mask->data[i] = 1;
  }
  Image *support = sp_image_duplicate(solution, SP_COPY_ALL);
  for (long long i = 0; i < sp_image_size(support); i++)
    if (sp_cabs(support->image->data[i])) support->image->data[i] = sp_cinit(1, 0);

  SpPhaser *ph = sp_phaser_alloc();
  int rc = sp_phaser_init(ph, alg, NULL, SpEngineAutomatic);
  if (rc == 0) {
    sp_phaser_set_amplitudes(ph, f);
    rc = sp_phaser_init_model(ph, NULL, SpModelRandomPhases);
  }
  if (rc == 0) rc = sp_phaser_init_support(ph, support, 0, 0);
  double change = 1e30;
  int i = 0;
  while (rc == 0) {
    rc = sp_phaser_iterate(ph, 1);
    if (rc) break;
    change = integrate2(sp_phaser_model_change(ph));
    i++;
    if (change <= tol || i >= 300) break;
  }
  if (rc == 0 && ph->iteration != i) rc = -101; /* the public iteration counter */
  sp_phaser_free(ph);
  sp_image_free(f);
  sp_image_free(support);
  sp_image_free(solution);
  *iters = i;
  if (rc) return rc;
  return change <= tol ? 1 : 0;
}

int main(int argc, char **argv) {
  if (argc < 2 || strcmp(argv[1], "run") != 0) {
    printf("layout ok\n");
    return 0;
  }
  if (sp_cuda_get_device_type() != SpCUDAHardwareDevice) {
    fprintf(stderr, "no HIP device\n");
    return 2;
  }
  int grid = 0, block = 0;
  sp_cuda_launch_parameters(8 * 8, &grid, &block);
  if (block != 64 || grid != 1) {
    fprintf(stderr, "sp_cuda_launch_parameters(64): grid %d block %d\n", grid, block);
    return 3;
  }
  sp_srand(7);
  sp_smap *beta = sp_smap_create_from_pair(0, 0.8f);
  SpPhasingAlgorithm *alg = sp_phasing_hio_alloc(beta, SpNoConstraints);
  for (int attempt = 0; attempt <= 10; attempt++) {
    int iters = 0;
    const int r = converge_once(alg, 1e-10, &iters);
    if (r < 0) {
      fprintf(stderr, "attempt %d: error %d (%s)\n", attempt, r, spx_last_error());
      return 4;
    }
    printf("attempt %d: %s after %d iterations\n", attempt, r ? "converged" : "not converged", iters);
    if (r == 1) return 0;
  }
  return 1;
}
