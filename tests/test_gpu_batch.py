"""The batched form (include/sp_batch.h) against single phasers.

A batch runs the same per-line arithmetic as one phaser, so every
reconstruction of a batch must equal the same reconstruction run alone,
bit for bit (models, supports) and to float rounding for the reductions;
graph replay must equal eager launches."""
import numpy as np
import pytest

from _helpers import GpuPhaser

pytestmark = pytest.mark.gpu


def patterns(shape, n, seed=0):
    from libspimage_amd import workloads
    out = []
    for r in range(n):
        p = workloads.make_pattern(shape, n_gauss=6, radius=tuple(s / 6 for s in shape), photons=1e6, beamstop=2,
                                   seed=seed + r)
        out.append(p)
    return out


def run_single(sp, p, alg, n, beta=0.9, sup_array=None):
    with GpuPhaser(sp, alg, p.amplitudes, p.mask.astype(bool), p.support.astype(bool), p.start, beta=beta,
                   sup_array=sup_array) as g:
        assert g.iterate(n) == 0
        return g.get("model"), g.get("support").real.astype(np.int32), g.efourier()[0]


def make_batch(sp, pats, alg_ptr, sup=None, graphs=False):
    b = sp.Batch(pats[0].amplitudes.shape, len(pats))
    for r, p in enumerate(pats):
        b.set_data(r, p.amplitudes, p.mask, p.support)
        b.set_model(r, p.start)
    b.set_algorithm(alg_ptr)
    if sup is not None:
        b.set_support_schedule(sup)
    b.use_graphs(graphs)
    return b


@pytest.mark.parametrize("shape", [(64, 64), (96, 80), (16, 16, 16)])
def test_batch_equals_single_phasers(sp, shape):
    pats = patterns(shape, 3, seed=40)
    hio = sp.sp_phasing_hio_alloc(sp.smap([(0, 0.9)]), 0)
    b = make_batch(sp, pats, hio)
    b.iterate(12)
    got = b.model()
    ef, _ = b.efourier()
    for r, p in enumerate(pats):
        m, s, e = run_single(sp, p, "hio", 12)
        assert np.array_equal(got[r], m), r
        assert abs(ef[r] - e) <= 1e-6 * abs(e)
    b.close()


def test_batch_shrinkwrap_equals_single(sp):
    pats = patterns((128, 128), 4, seed=50)
    blur, area = [(0, 3.0), (30, 1.0)], [(0, 0.1)]
    sup = sp.sp_support_array_init(sp.sp_support_area_alloc(sp.smap(blur), sp.smap(area)), 7)
    raar = sp.sp_phasing_raar_alloc(sp.smap([(0, 0.87)]), 0)
    b = make_batch(sp, pats, raar, sup)
    b.iterate(30)
    assert b.iteration == 30
    got, gsup = b.model(), b.support()
    frac = b.support_fraction()
    for r, p in enumerate(pats):
        sup1 = sp.sp_support_array_init(sp.sp_support_area_alloc(sp.smap(blur), sp.smap(area)), 7)
        m, s, _ = run_single(sp, p, "raar", 30, beta=0.87, sup_array=sup1)
        assert np.array_equal(gsup[r], s), r
        assert np.array_equal(got[r], m), r
        assert abs(frac[r] - s.mean()) < 1e-6
    b.close()


@pytest.mark.parametrize("shape", [(128, 128), (24, 20, 16)])
def test_batch_centre_image_and_close_equal_single(sp, shape):
    """Area shrinkwrap followed by the centre-image and closing updates, each
    applied to every reconstruction's own support on the device, equals the
    same support array on single phasers."""
    pats = patterns(shape, 3, seed=70)

    def array():
        a = sp.sp_support_array_init(sp.sp_support_area_alloc(sp.smap([(0, 2.0)]), sp.smap([(0, 0.12)])), 5)
        sp.sp_support_array_append(a, sp.sp_support_centre_image_alloc())
        sp.sp_support_array_append(a, sp.sp_support_close_alloc(1))
        return a

    hio = sp.sp_phasing_hio_alloc(sp.smap([(0, 0.9)]), 0)
    b = make_batch(sp, pats, hio, array())
    b.iterate(15)
    got, gsup = b.model(), b.support()
    for r, p in enumerate(pats):
        m, s, _ = run_single(sp, p, "hio", 15, sup_array=array())
        assert np.array_equal(gsup[r], s), r
        assert np.array_equal(got[r], m), r
    b.close()


def test_per_reconstruction_algorithms(sp):
    pats = patterns((64, 64), 3, seed=60)
    algs = np.array([4, 1, 2], np.int32)  # ER, HIO, RAAR
    betas = np.array([0.0, 0.8, 0.6], np.float32)
    hio = sp.sp_phasing_hio_alloc(sp.smap([(0, 0.9)]), 0)
    b = make_batch(sp, pats, hio)
    b.set_rec_algorithms(algs, betas)
    b.iterate(8)
    got = b.model()
    for r, (name, beta) in enumerate([("er", 0.0), ("hio", 0.8), ("raar", 0.6)]):
        m, _, _ = run_single(sp, pats[r], name, 8, beta=beta)
        assert np.array_equal(got[r], m), name
    b.close()


def test_graph_replay_equals_eager(sp):
    """Graph replays interleaved with eager work and synchronisations give the
    same iterates as eager launches, on every chunk."""
    pats = patterns((128, 96), 3, seed=70)
    hio = sp.sp_phasing_hio_alloc(sp.smap([(0, 0.9)]), 0)
    b_eager = make_batch(sp, pats, hio, graphs=False)
    b_graph = make_batch(sp, pats, hio, graphs=True)
    for chunk in (5, 9, 4, 11):
        b_eager.iterate(chunk)
        b_graph.iterate(chunk)
        b_graph.synchronize()
        e1, _ = b_graph.efourier()  # eager work between replays
        assert np.array_equal(b_graph.model(), b_eager.model()), chunk
        e0, _ = b_eager.efourier()
        assert np.array_equal(e0, e1)
    b_eager.close()
    b_graph.close()


def test_batch_random_start_streams_differ(sp):
    pats = patterns((64, 64), 2, seed=80)
    hio = sp.sp_phasing_hio_alloc(sp.smap([(0, 0.9)]), 0)
    b = make_batch(sp, pats, hio)
    b.init_model(sp.SpModelRandomPhases, 99)
    m = b.model()
    for r, p in enumerate(pats):
        F = np.fft.fftn(m[r].astype(np.complex128))
        assert np.allclose(np.abs(F), p.amplitudes, atol=2e-5 * p.amplitudes.max())
    # same amplitudes in both slots still give different phases
    b.set_data(1, pats[0].amplitudes, pats[0].mask, pats[0].support)
    b.init_model(sp.SpModelRandomPhases, 99)
    m = b.model()
    assert not np.allclose(m[0], m[1])
    b.close()


@pytest.mark.parametrize("shape", [(64, 64), (128, 128), (16, 16, 16)])
def test_idle_reconstructions_are_left_unchanged(sp, shape):
    """spx_batch_set_active: idle slots keep their iterate while the active
    ones advance exactly as they do alone (row-major, tiled and 3-D)."""
    pats = patterns(shape, 3, seed=90)
    hio = sp.sp_phasing_hio_alloc(sp.smap([(0, 0.9)]), 0)
    b = make_batch(sp, pats, hio)
    before = b.model()
    b.set_active([1, 0, 1])
    b.iterate(7)
    got = b.model()
    assert np.array_equal(got[1], before[1])
    for r in (0, 2):
        m, s, e = run_single(sp, pats[r], "hio", 7)
        assert np.array_equal(got[r], m), r
    # reactivated, the idle one continues from where it stopped
    b.set_active(None)
    b.iterate(3)
    m, s, e = run_single(sp, pats[1], "hio", 3)
    assert np.array_equal(b.model(1), m)
    b.close()


@pytest.mark.parametrize("shape", [(64, 64), (128, 128), (16, 16, 16)])
def test_vacant_slots_do_no_work_and_can_be_refilled(sp, shape):
    """spx_batch_set_active state 2 (vacant): the active slots advance exactly
    as they do alone; a vacant slot set again (model and data) and made
    active runs like a fresh reconstruction (row-major, tiled and 3-D)."""
    pats = patterns(shape, 3, seed=93)
    hio = sp.sp_phasing_hio_alloc(sp.smap([(0, 0.9)]), 0)
    b = make_batch(sp, pats, hio)
    b.set_active([1, 2, 1])
    b.iterate(7)
    got = b.model()
    for r in (0, 2):
        m, s, e = run_single(sp, pats[r], "hio", 7)
        assert np.array_equal(got[r], m), r
    # the vacant slot did no work: after an odd number of iterations the
    # current ping-pong buffer is the one its start was not written to,
    # still as the batch was created (zero), not an iterate
    assert not np.any(got[1])
    # slot 1 takes a reconstruction afresh
    b.set_data(1, pats[1].amplitudes, pats[1].mask, pats[1].support)
    b.set_model(1, pats[1].start)
    b.set_active(None)
    b.iterate(3)
    m, s, e = run_single(sp, pats[1], "hio", 3)
    assert np.array_equal(b.model(1), m)
    m, s, e = run_single(sp, pats[0], "hio", 10)
    assert np.array_equal(b.model(0), m)
    b.close()


@pytest.mark.parametrize("shape", [(64, 64), (16, 16, 16)])
def test_queued_efourier_equals_blocking(sp, shape):
    """spx_batch_efourier_enqueue / _fetch: the errors queued behind the
    iterations equal the blocking spx_batch_efourier's bit for bit, the
    queued call leaves the iterates alone, and a fetch needs an enqueue."""
    pats = patterns(shape, 3, seed=95)
    hio = sp.sp_phasing_hio_alloc(sp.smap([(0, 0.9)]), 0)
    b = make_batch(sp, pats, hio)
    b.iterate(9)
    b.efourier_enqueue()
    while not b.done():
        pass
    ef_q, fc_q = b.efourier_fetch()
    ef, fc = b.efourier()
    assert np.array_equal(ef_q, ef) and np.array_equal(fc_q, fc)
    with pytest.raises(RuntimeError):
        b.efourier_fetch()
    # queued between iterations, the trajectory is unchanged
    b.iterate(4)
    b.efourier_enqueue()
    b.iterate(3)
    b.efourier_fetch()
    c = make_batch(sp, pats, hio)
    c.iterate(16)
    assert np.array_equal(b.model(), c.model())
    b.close()
    c.close()


def test_pass_shapes_do_not_change_results(sp, monkeypatch):
    """The tuner's free choices (wide or narrow projection pass, tiled or
    row-major layout, non-temporal hint or not) compute the same bits: a
    batch forced to the wide row-major pass without the hint equals single
    phasers run with the defaults.  360 has 60 threads per line, so its
    wide pass has 8 columns per work-group against 4."""
    shape = (360, 360)
    pats = patterns(shape, 3, seed=97)
    hio = sp.sp_phasing_hio_alloc(sp.smap([(0, 0.9)]), 0)
    for k, v in (("SPX_COL_WIDE", "1"), ("SPX_NO_T4", "1"), ("SPX_NT", "0")):
        monkeypatch.setenv(k, v)
    b = make_batch(sp, pats, hio)
    for k in ("SPX_COL_WIDE", "SPX_NO_T4", "SPX_NT"):
        monkeypatch.delenv(k)
    b.iterate(6)
    got = b.model()
    for r, p in enumerate(pats):
        m, s, e = run_single(sp, p, "hio", 6)
        assert np.array_equal(got[r], m), r
    b.close()


def test_null_tables_restore_shared_algorithm(sp):
    """spx_batch_set_rec_algorithms(b, NULL, NULL) drops the per-reconstruction
    tables: the batch then follows the shared algorithm and beta schedule,
    exactly like a batch that never had tables."""
    pats = patterns((64, 64), 3, seed=61)
    hio = sp.sp_phasing_hio_alloc(sp.smap([(0, 0.9)]), 0)
    b = make_batch(sp, pats, hio)
    b.set_rec_algorithms(np.array([4, 2, 2], np.int32), np.array([0.0, 0.5, 0.6], np.float32))
    b.iterate(3)
    # restart every slot from its start and drop the tables
    for r, p in enumerate(pats):
        b.set_model(r, p.start)
    b.set_rec_algorithms(None, None)
    b.iterate(6)
    c = make_batch(sp, pats, hio)
    c.iterate(6)
    assert np.array_equal(b.model(), c.model())
    b.close()
    c.close()


def test_graphs_with_algorithm_table_follow_beta_schedule(sp):
    """Graphs on, per-reconstruction algorithm types but no per-reconstruction
    betas, and a beta schedule that changes every iteration: each chunk equals
    eager iteration (a replayed graph would freeze the first beta)."""
    pats = patterns((64, 64), 3, seed=62)
    ramp = sp.sp_phasing_hio_alloc(sp.smap([(0, 0.95), (40, 0.6)]), 0)
    algs = np.array([1, 2, 1], np.int32)
    b_eager = make_batch(sp, pats, ramp, graphs=False)
    b_graph = make_batch(sp, pats, ramp, graphs=True)
    for b in (b_eager, b_graph):
        b.set_rec_algorithms(algs, None)
    for chunk in (6, 9, 5):
        b_eager.iterate(chunk)
        b_graph.iterate(chunk)
        b_graph.synchronize()
        assert np.array_equal(b_graph.model(), b_eager.model()), chunk
    b_eager.close()
    b_graph.close()


@pytest.mark.parametrize("shape", [(256, 256), (384, 384), (512, 512)])
def test_batch_sizes_of_the_batched_workload(sp, o32, shape):
    """The batched workload's other size classes: every reconstruction of a
    batch equals its single phaser bit for bit, and one HIO iteration of each
    agrees with the FP32 oracle (relative L2 <= 1e-5)."""
    from _helpers import rel
    pats = patterns(shape, 2, seed=63)
    hio = sp.sp_phasing_hio_alloc(sp.smap([(0, 0.9)]), 0)
    b = make_batch(sp, pats, hio)
    b.iterate(1)
    one = b.model()
    b.iterate(4)
    got = b.model()
    alg = o32.alg(1, 0, beta=[(0, 0.9)])
    for r, p in enumerate(pats):
        m, _, _ = run_single(sp, p, "hio", 5)
        assert np.array_equal(got[r], m), r
        oph = o32.phaser(shape)
        oph.set_inputs(p.amplitudes, p.mask.astype(bool), p.support.astype(bool), p.start)
        assert oph.iterate(alg, 1) == 0
        assert rel(one[r], oph.g1) <= 1e-5, r
    b.close()


def test_set_data_equals_separate_setters(sp):
    """spx_batch_set_data builds the flags from the mask and support on the
    host; the state must equal spx_batch_set_amplitudes then
    spx_batch_set_support (which read the other input's flags back)."""
    pats = patterns((96, 80), 3, seed=120)
    hio = sp.sp_phasing_hio_alloc(sp.smap([(0, 0.9)]), 0)
    sup = sp.sp_support_array_init(sp.sp_support_area_alloc(sp.smap([(0, 2.0)]), sp.smap([(0, 0.1)])), 5)
    a = make_batch(sp, pats, hio, sup=sup)
    b = sp.Batch((96, 80), 3)
    for r, p in enumerate(pats):
        b.set_amplitudes(r, p.amplitudes, p.mask)
        b.set_support(r, p.support)
        b.set_model(r, p.start)
    b.set_algorithm(hio)
    b.set_support_schedule(sup)
    for x in (a, b):
        x.iterate(12)
    assert np.array_equal(a.model(), b.model())
    assert np.array_equal(a.support(), b.support())
    a.close()
    b.close()


def test_uploads_copy_the_host_arrays_at_once(sp):
    """The setters queue their uploads without waiting for the device: one
    host buffer overwritten right after each call must give the same batch
    as separate arrays (the upload copied it before returning)."""
    shape = (128, 96)
    pats = patterns(shape, 6, seed=130)
    hio = sp.sp_phasing_hio_alloc(sp.smap([(0, 0.9)]), 0)
    ref = make_batch(sp, pats, hio)
    ref.iterate(4)
    amp = np.empty(shape, np.float32)
    msk = np.empty(shape, np.int32)
    sup = np.empty(shape, np.float32)
    b = sp.Batch(shape, len(pats))
    b.set_algorithm(hio)
    for r, p in enumerate(pats):
        amp[...], msk[...], sup[...] = p.amplitudes, p.mask, p.support
        b.set_data(r, amp, msk, sup)
        amp.fill(np.nan)
        msk.fill(0)
        sup.fill(0)
        b.set_model(r, p.start)
    b.set_active([1, 0, 1, 1, 0, 1])
    b.set_active(None)
    b.iterate(4)
    assert np.array_equal(b.model(), ref.model())
    assert np.array_equal(b.support(), ref.support())
    ref.close()
    b.close()


def test_uploads_through_a_wrapped_ring(sp):
    """Twelve 1024^2 slots upload about 108 MB of amplitudes and flags, more
    than the 64 MiB pinned ring the uploads go through, so the ring wraps
    while earlier uploads may still be queued.  The batch must equal one
    filled with a device synchronisation after every slot."""
    shape = (1024, 1024)
    pats = patterns(shape, 3, seed=140)
    hio = sp.sp_phasing_hio_alloc(sp.smap([(0, 0.9)]), 0)
    out = []
    for sync in (False, True):
        b = sp.Batch(shape, 12)
        b.set_algorithm(hio)
        for r in range(12):
            p = pats[r % 3]
            b.set_data(r, p.amplitudes, p.mask, p.support)
            b.init_model_rec(r, sp.SpModelRandomPhases, 1000 + r)
            if sync:
                b.synchronize()
        b.iterate(2)
        out.append(b.model())
        b.close()
    assert np.array_equal(out[0], out[1])
