/* Batched phasing: many independent reconstructions of one array shape
 * resident on the GPU and advanced together, one launch per pass for the
 * whole batch.
 *
 * This has no counterpart in the reference, whose Reconstructor runs random
 * restarts one after another (reference src/_spimage_reconstructor.py:677-740,
 * reconstruct_loop).  Every reconstruction of a batch follows the same
 * arithmetic as a single phaser (sp_phaser_iterate, reference
 * src/phasing.c:895-1258): a batch of one gives bit-identical iterates.
 *
 * Reconstruction r of a batch of nrec holds its own amplitudes, measured
 * mask, support and iterates.  The phasing algorithm and the support
 * schedule (an SpSupportArray of area / threshold / centre-image / close /
 * static updates, with the reference's update_period semantics) apply to
 * all reconstructions, each update to each reconstruction's own support;
 * ER, HIO and RAAR may be mixed per reconstruction with per-reconstruction
 * beta.
 *
 * All functions return 0 on success and a negative code on error
 * (spx_last_error() holds the message).  Host arrays are in the reference's
 * layout: (nz, ny, nx), x fastest; complex values are interleaved float pairs.
 */
#ifndef SPX_SP_BATCH_H
#define SPX_SP_BATCH_H 1

#include "sp_phasing.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct SpxBatch SpxBatch;

spimage_EXPORT SpxBatch *spx_batch_create(int nx, int ny, int nz, int nrec);
spimage_EXPORT void spx_batch_free(SpxBatch *b);
spimage_EXPORT int spx_batch_size(const SpxBatch *b);

/* inputs of reconstruction `rec` (rec < 0: all, arrays of nrec * N).  The
 * setters below copy the host arrays before they return and queue the
 * uploads on the batch's stream without waiting for it: iterations queued
 * earlier still see the old values, those queued later the new ones. */
spimage_EXPORT int spx_batch_set_amplitudes(SpxBatch *b, int rec, const float *amplitudes, const int *mask);
spimage_EXPORT int spx_batch_set_support(SpxBatch *b, int rec, const float *support);
/* amplitudes, mask and support together: the same state as
 * spx_batch_set_amplitudes then spx_batch_set_support, without reading the
 * other input's flags back from the device, and queued without waiting */
spimage_EXPORT int spx_batch_set_data(SpxBatch *b, int rec, const float *amplitudes, const int *mask,
                                      const float *support);
spimage_EXPORT int spx_batch_set_model(SpxBatch *b, int rec, const float *model_complex);
/* SpModelRandomPhases / SpModelZeroPhases / SpModelRandomValues start for all
 * reconstructions from their amplitudes (reference src/phasing.c:626-679);
 * reconstruction r draws from the generator seeded with seed + r */
spimage_EXPORT int spx_batch_init_model(SpxBatch *b, int flags, unsigned long long seed);
/* The same start for reconstruction `rec` only, seeded with `seed`: equal to
 * a single phaser's sp_phaser_init_model(ph, NULL, flags) after
 * spx_phaser_set_seed(ph, seed).  The other reconstructions keep their state. */
spimage_EXPORT int spx_batch_init_model_rec(SpxBatch *b, int rec, int flags, unsigned long long seed);

/* algorithm of all reconstructions (the reference's allocators build it) */
spimage_EXPORT int spx_batch_set_algorithm(SpxBatch *b, SpPhasingAlgorithm *alg);
/* per-reconstruction ER/HIO/RAAR and beta (overrides the algorithm's type
 * and beta schedule; NULL restores them) */
spimage_EXPORT int spx_batch_set_rec_algorithms(SpxBatch *b, const int *alg_types, const float *betas);
spimage_EXPORT int spx_batch_set_support_schedule(SpxBatch *b, SpSupportArray *sup);
/* Which reconstructions iterations advance: nrec entries, 1 = active,
 * 0 = idle (its iterate, support and Fourier-space state stay as they are;
 * carrying the iterate over costs a copy of it per iteration), 2 = vacant
 * (nothing is kept: its work-groups exit at once, and the slot must be set
 * again, e.g. with spx_batch_set_data and spx_batch_init_model_rec, before
 * it is made active); NULL = all active.  Support updates and metrics still
 * cover every reconstruction. */
spimage_EXPORT int spx_batch_set_active(SpxBatch *b, const int *active);
/* replay each iteration from a captured HIP graph (1) or launch kernels (0) */
spimage_EXPORT int spx_batch_use_graphs(SpxBatch *b, int on);

spimage_EXPORT int spx_batch_iterate(SpxBatch *b, int iterations);
spimage_EXPORT int spx_batch_iteration(const SpxBatch *b);

/* per-reconstruction metrics, nrec values each (reference src/phasing.c:803-893) */
spimage_EXPORT int spx_batch_efourier(SpxBatch *b, float *efourier, float *fcfo);
spimage_EXPORT int spx_batch_ereal(SpxBatch *b, float *ereal);
spimage_EXPORT int spx_batch_support_fraction(SpxBatch *b, float *fraction);
/* spx_batch_efourier in two halves, for a host loop that feeds several
 * batches: _enqueue queues the error sums behind the batch's queued work and
 * returns at once; _fetch returns them (waiting for the batch if it is still
 * running; without waiting once spx_batch_query reports 1).  One enqueue per
 * fetch; -1 from a fetch without an enqueue. */
spimage_EXPORT int spx_batch_efourier_enqueue(SpxBatch *b);
spimage_EXPORT int spx_batch_efourier_fetch(SpxBatch *b, float *efourier, float *fcfo);

/* results of reconstruction `rec` (rec < 0: all) */
spimage_EXPORT int spx_batch_get_model(SpxBatch *b, int rec, float *model_complex);
spimage_EXPORT int spx_batch_get_support(SpxBatch *b, int rec, int *support);
/* the model before projection of the last iteration (sp_phaser_model_before_projection,
 * reference src/phasing.c:496-513) and the Fourier transform of the current
 * model (sp_phaser_fmodel, src/phasing.c:450-472), complex, row-major */
spimage_EXPORT int spx_batch_get_model_before_projection(SpxBatch *b, int rec, float *model_complex);
spimage_EXPORT int spx_batch_get_fmodel(SpxBatch *b, int rec, float *fmodel_complex);

/* device time of `iterations` iterations (no support updates), HIP events */
spimage_EXPORT int spx_batch_time_iterations(SpxBatch *b, int iterations, float *ms);
/* wait for all queued work of the batch */
spimage_EXPORT int spx_batch_synchronize(SpxBatch *b);
/* 1 when all work queued for the batch has finished, 0 while some is still
 * running (never blocks) */
spimage_EXPORT int spx_batch_query(SpxBatch *b);
/* Device bytes a batch of nrec reconstructions of (nz, ny, nx) allocates for
 * ER / HIO / RAAR with area or threshold support updates (the difference
 * map adds 8 bytes per pixel); used to admit work by free device memory. */
spimage_EXPORT long long spx_batch_estimate_bytes(int nx, int ny, int nz, int nrec);

#ifdef __cplusplus
}
#endif
#endif
