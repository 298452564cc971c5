"""The continuous-batching server (libspimage_amd/serve.py) against requests
run alone through the reference-style phaser API.

A request served in a mix of other requests (other shapes, algorithms,
iteration counts, slots reused by earlier requests, batches created while
others run) must give the same model bit for bit, and the same Fourier
error, as the same request run alone with sp_phaser_iterate."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pattern(shape, seed):
    from libspimage_amd import workloads
    return workloads.make_pattern(shape, n_gauss=5, radius=tuple(s / 6 for s in shape), photons=1e6, beamstop=2,
                                  seed=seed)


def _requests(shapes_algs_its, seed0=500):
    from libspimage_amd import serve
    pats = {}
    out = []
    for k, (shape, alg, its) in enumerate(shapes_algs_its):
        if shape not in pats:
            pats[shape] = _pattern(shape, seed0 + len(pats))
        p = pats[shape]
        out.append(serve.Request(k, p.amplitudes, p.mask, p.support, alg, its, seed=seed0 * 100 + k,
                                 beta=0.9 if alg == "hio" else 0.87, arrival=0.002 * k))
    return out


# 64 x 64 and 48 x 48 run row-major, 128 x 128 in the tiled layout, 16^3 is 3-D
MIX = [((64, 64), "hio", 100), ((48, 48), "er", 50), ((64, 64), "raar", 150), ((16, 16, 16), "hio", 100),
       ((64, 64), "er", 50), ((48, 48), "hio", 150), ((64, 64), "hio", 100), ((16, 16, 16), "raar", 50),
       ((64, 64), "raar", 100), ((48, 48), "raar", 100), ((64, 64), "hio", 50), ((16, 16, 16), "er", 150),
       ((128, 128), "hio", 100), ((128, 128), "raar", 50), ((128, 128), "er", 150), ((128, 128), "hio", 200)]


@pytest.mark.parametrize("realtime", [False, True])
def test_served_equals_alone(sp, realtime):
    from libspimage_amd import serve
    reqs = _requests(MIX)
    # at most 2 slots per 64x64 batch: later requests reuse freed slots and
    # extra batches are created while others run
    srv = serve.Server(quantum=50, max_pixels_per_batch=2 * 128 * 128, reserve_bytes=0, keep_models=True)
    res = srv.run(reqs, realtime=realtime)
    assert sorted(r.rid for r in res) == list(range(len(reqs)))
    if not realtime:
        # all requests queued at once: 64x64 (6 requests, 8 slots per batch)
        # and 128x128 (4 requests, 2 slots) need several batches
        assert srv.batches_created >= 5
    srv.close()
    plain = serve.PlainRunner()
    by_id = {r.rid: r for r in res}
    for req in reqs:
        ef, fc, model = plain.run_one(req, keep_model=True)
        got = by_id[req.rid]
        assert np.array_equal(got.model, model), (req.rid, req.amplitudes.shape, float(np.abs(got.model - model).max()))
        assert abs(got.efourier - ef) <= 1e-6 * abs(ef), req.rid
        assert abs(got.fcfo - fc) <= 1e-6 * abs(fc), req.rid
        assert got.finished >= got.admitted >= (got.arrival if realtime else 0.0)


# lengths with radix-7 stages: 2-D 56 x 98 and 49 x 64, 3-D 14 x 21 x 28
MIX7 = [((56, 98), "hio", 100), ((49, 64), "raar", 50), ((56, 98), "er", 150), ((14, 21, 28), "hio", 100),
        ((49, 64), "hio", 100), ((56, 98), "raar", 50), ((14, 21, 28), "er", 50), ((49, 64), "er", 150),
        ((56, 98), "hio", 200)]


def test_served_radix7_shapes_equal_alone(sp):
    """7-smooth shapes take batches and the server like the 2-3-5 ones; freed
    slots (vacant) are reused by later requests of the class."""
    from libspimage_amd import serve
    reqs = _requests(MIX7, seed0=700)
    srv = serve.Server(quantum=50, max_pixels_per_batch=2 * 56 * 98, reserve_bytes=0, keep_models=True)
    res = srv.run(reqs, realtime=True)
    srv.close()
    assert sorted(r.rid for r in res) == list(range(len(reqs)))
    plain = serve.PlainRunner()
    by_id = {r.rid: r for r in res}
    for req in reqs:
        ef, fc, model = plain.run_one(req, keep_model=True)
        assert np.array_equal(by_id[req.rid].model, model), (req.rid, req.amplitudes.shape)
        assert abs(by_id[req.rid].efourier - ef) <= 1e-6 * abs(ef), req.rid


def test_batch_slot_start_equals_single_seeded_start(sp):
    """spx_batch_init_model_rec(rec, seed) == spx_phaser_set_seed(seed) +
    sp_phaser_init_model(NULL, SpModelRandomPhases)."""
    from libspimage_amd import serve
    reqs = _requests([((40, 40), "hio", 50)] * 3)
    b = sp.Batch((40, 40), 3)
    for r, q in enumerate(reqs):
        b.set_data(r, q.amplitudes, q.mask, q.support)
        b.init_model_rec(r, sp.SpModelRandomPhases, q.seed)
    m = b.model()
    b.close()
    plain = serve.PlainRunner()
    for r, q in enumerate(reqs):
        ph = sp.sp_phaser_alloc()
        assert sp.sp_phaser_init(ph, plain._alg(q), None, sp.SpEngineHIP) == 0
        img = sp.image_from_array(q.amplitudes, q.mask)
        sp.sp_phaser_set_amplitudes(ph, img)
        sp.spx_phaser_set_seed(ph, q.seed)
        assert sp.sp_phaser_init_model(ph, None, sp.SpModelRandomPhases) == 0
        single = sp.image_view(sp.sp_phaser_model(ph)).copy()
        sp.sp_phaser_free(ph)
        sp.sp_image_free(img)
        assert np.array_equal(m[r], single), r
    assert not np.array_equal(m[0], m[1])


def test_admission_respects_free_memory(sp):
    """With a reserve larger than the free device memory nothing can be
    admitted; the capacity rule reports 0 slots instead of allocating."""
    from libspimage_amd import serve
    free, total = sp.device_memory()
    assert 0 < free <= total
    srv = serve.Server(reserve_bytes=total + 1)
    assert srv._capacity((64, 64), 4) == 0
    srv2 = serve.Server(reserve_bytes=0, max_pixels_per_batch=1 << 20)
    assert srv2._capacity((64, 64), 5) == 8
    assert srv2._capacity((64, 64), 1000) == 256


def test_prepare_tunes_ahead_without_changing_results(sp):
    """spx_tune_shape / Server.prepare: tuning ahead for every batch size
    leaves the served results equal to the same requests run alone; an
    unsupported length is refused."""
    from libspimage_amd import serve
    shapes = [(40, 40), (36, 36, 36)]
    reqs = _requests([((40, 40), "hio", 50), ((40, 40), "raar", 100), ((36, 36, 36), "er", 50)], seed0=700)
    srv = serve.Server(quantum=50, max_pixels_per_batch=8 * 40 * 40, reserve_bytes=0, keep_models=True)
    srv.prepare(shapes)
    res = {r.rid: r for r in srv.run(reqs, realtime=False)}
    srv.close()
    plain = serve.PlainRunner()
    for q in reqs:
        ef, fc, model = plain.run_one(q, keep_model=True)
        assert np.array_equal(res[q.rid].model, model), q.rid
    with pytest.raises(ValueError):
        sp.tune_shape((11, 11), 2)


def test_served_volumes_of_the_serving_mix(sp):
    """The served workload's 3-D classes (64^3 and 128^3) mixed with a 2-D
    class: each served result equals the request run alone, bit for bit."""
    from libspimage_amd import serve
    mix = [((64, 64, 64), "hio", 100), ((128, 128, 128), "er", 50), ((64, 64, 64), "raar", 50),
           ((256, 256), "hio", 100), ((128, 128, 128), "hio", 100), ((64, 64, 64), "er", 150)]
    reqs = _requests(mix, seed0=800)
    srv = serve.Server(quantum=50, reserve_bytes=0, keep_models=True)
    res = {r.rid: r for r in srv.run(reqs, realtime=True)}
    srv.close()
    plain = serve.PlainRunner()
    for q in reqs:
        ef, fc, model = plain.run_one(q, keep_model=True)
        assert np.array_equal(res[q.rid].model, model), (q.rid, q.amplitudes.shape)
        assert abs(res[q.rid].efourier - ef) <= 1e-6 * abs(ef), q.rid
