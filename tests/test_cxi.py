"""CXI files through libspimage_amd.cxi (reference src/_spimage_cxi.py),
checked against h5dump, the HDF5 distribution's own reader, and by round
trips: every supported type, whole datasets and single events, scalars, and
the reader's picking, filtering, sorting and multi-file options."""
import re
import shutil
import subprocess

import numpy as np
import pytest

cxi = pytest.importorskip("libspimage_amd.cxi")


def _h5lib_or_skip():
    try:
        cxi._lib()
    except OSError as e:
        pytest.skip(f"libhdf5 unavailable: {e}")


@pytest.fixture
def sample(tmp_path):
    _h5lib_or_skip()
    rng = np.random.default_rng(0)
    N = 5
    data = {"entry_1": {
        "data_1": {"data": rng.random((N, 6, 7)).astype(np.float32),
                   "mask": rng.random((N, 6, 7)) > 0.5},
        "model": (rng.random((N, 6, 7)) + 1j * rng.random((N, 6, 7))).astype(np.complex64),
        "model64": rng.random((N, 3)) + 1j * rng.random((N, 3)),
        "photons": np.array([30.0, 10.0, 50.0, 20.0, 40.0]),
        "tag": np.array([3, 1, 4, 1, 5], np.int16),
        "u": np.arange(N, dtype=np.uint32),
        "i8": np.arange(N, dtype=np.int8) - 2,
    }}
    path = tmp_path / "a.cxi"
    w = cxi.CXIWriter(str(path), N)
    w.write_stack(data)
    w.write_slice({"result": {"score": 0.25}})
    for i in range(N):
        w.write_slice({"per_event": {"err": float(i) / 10}}, i=i)
    w.close()
    return path, data, N


def _dump_values(path, dset):
    exe = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
    out = subprocess.run([exe, "-m", "%.9g", "-d", dset, "-y", "-w", "0", str(path)], capture_output=True, text=True, check=True).stdout
    body = out.split("DATA {", 1)[1].split("ATTRIBUTE", 1)[0]  # the values, before the "axes" attribute
    return out, [float(v) for v in re.findall(r"-?\d+(?:\.\d+)?(?:e[-+]?\d+)?", body)]


def test_h5dump_reads_what_the_writer_wrote(sample):
    path, data, N = sample
    if not (shutil.which("h5dump") or __import__("os").path.exists("/opt/conda/bin/h5dump")):
        pytest.skip("h5dump not available")
    out, vals = _dump_values(path, "/entry_1/data_1/data")
    assert "H5T_IEEE_F32LE" in out and "( 5, 6, 7 )" in out
    assert np.allclose(vals, data["entry_1"]["data_1"]["data"].ravel(), rtol=1e-6)
    out, vals = _dump_values(path, "/entry_1/model")
    assert '"r"' in out and '"i"' in out  # the compound h5py writes for complex
    m = data["entry_1"]["model"].ravel()
    want = np.stack([m.real, m.imag], axis=1).ravel()
    assert np.allclose(vals, want, rtol=1e-6)
    out, vals = _dump_values(path, "/entry_1/tag")
    assert "H5T_STD_I16LE" in out and vals == [3, 1, 4, 1, 5]


def test_round_trip_every_type(sample):
    path, data, N = sample
    with cxi.H5File(str(path)) as f:
        e = data["entry_1"]
        assert np.array_equal(f.read("/entry_1/data_1/data"), e["data_1"]["data"])
        assert np.array_equal(f.read("/entry_1/data_1/mask").astype(bool), e["data_1"]["mask"])
        assert f.read("/entry_1/data_1/mask").dtype == np.uint8
        assert np.array_equal(f.read("/entry_1/model"), e["model"])
        assert f.read("/entry_1/model64").dtype == np.complex128
        assert np.array_equal(f.read("/entry_1/model64"), e["model64"])
        assert np.array_equal(f.read("/entry_1/tag"), e["tag"]) and f.read("/entry_1/tag").dtype == np.int16
        assert np.array_equal(f.read("/entry_1/u"), e["u"]) and f.read("/entry_1/u").dtype == np.uint32
        assert np.array_equal(f.read("/entry_1/i8"), e["i8"])
        assert np.array_equal(f.read("/entry_1/model", index=3), e["model"][3])
        assert f.read("/result/score").tolist() == [0.25]
        assert np.allclose(f.read("/per_event/err"), np.arange(N) / 10)
        assert "/entry_1/data_1" in f and "/entry_1/nothing" not in f
        with pytest.raises(IndexError):
            f.read("/entry_1/model", index=N)


def test_reader_events(sample):
    path, data, N = sample
    stack = {"I": "/entry_1/data_1/data", "F": "/entry_1/model"}
    r = cxi.CXIReader(str(path), stack, def_stack_ds="/entry_1/photons",
                      dsnames_single={"score": "/result/score"})
    got = []
    while (d := r.get_next()) is not None:
        got.append(d)
    assert [d["i"] for d in got] == list(range(N))
    for k, d in enumerate(got):
        assert np.array_equal(d["I"], data["entry_1"]["data_1"]["data"][k])
        assert np.array_equal(d["F"], data["entry_1"]["model"][k])
        assert d["score"].tolist() == [0.25] and d["filename"] == "a.cxi"
    # first / count
    r = cxi.CXIReader(str(path), stack, def_stack_ds="/entry_1/photons", ifirst=1, nevents=2)
    assert [r.get_next()["I"][0, 0] for _ in range(2)] == list(data["entry_1"]["data_1"]["data"][1:3, 0, 0])
    assert r.get_next() is None
    # filters: a value window and a list of indices
    r = cxi.CXIReader(str(path), stack, def_stack_ds="/entry_1/photons",
                      event_filters={"p": {"dataset_name": "/entry_1/photons", "vmin": 15, "vmax": 45}})
    sel = []
    while (d := r.get_next()) is not None:
        sel.append(float(d["I"][0, 0]))
    want = [float(v) for v, p in zip(data["entry_1"]["data_1"]["data"][:, 0, 0], data["entry_1"]["photons"])
            if 15 <= p <= 45]
    assert sel == want
    r = cxi.CXIReader(str(path), stack, def_stack_ds="/entry_1/photons",
                      event_filters={"t": {"dataset_name": "/entry_1/tag", "indices": [1, 5]}})
    n = 0
    while r.get_next() is not None:
        n += 1
    assert n == 3
    # sorting by a dataset
    r = cxi.CXIReader(str(path), stack, def_stack_ds="/entry_1/photons", sorting_dataset="/entry_1/photons")
    order = [float(r.get_next()["I"][0, 0]) for _ in range(N)]
    idx = np.argsort(data["entry_1"]["photons"])
    assert order == [float(v) for v in data["entry_1"]["data_1"]["data"][idx, 0, 0]]


def test_reader_over_a_directory(tmp_path):
    _h5lib_or_skip()
    for k, n in enumerate((2, 3)):
        w = cxi.CXIWriter(str(tmp_path / f"run{k}.cxi"), n)
        w.write_stack({"i": np.arange(n) + 10 * k, "v": np.full((n, 2, 2), float(k), np.float32)})
        w.close()
    r = cxi.CXIReader(str(tmp_path), {"v": "/v", "i": "/i"})
    seen = []
    while (d := r.get_next()) is not None:
        seen.append((d["filename_i"], int(d["i"])))
    assert seen == [(0, 0), (0, 1), (1, 10), (1, 11), (1, 12)]


def test_writer_rejects_strings(tmp_path):
    _h5lib_or_skip()
    w = cxi.CXIWriter(str(tmp_path / "s.cxi"), 1)
    with pytest.raises(TypeError):
        w.write_slice({"name": np.array(["x"])})
    w.close()


def test_reader_configuration_errors_raise(sample):
    """Bad event counts raise ValueError; the library never ends the process."""
    path, data, N = sample
    with pytest.raises(ValueError):
        cxi.CXIReader(str(path), {}, def_stack_ds="/entry_1/photons", nevents=-1)
    with pytest.raises(ValueError):
        cxi.CXIReader(str(path), {}, def_stack_ds="/entry_1/photons", ifirst=3, nevents=N)
    with pytest.raises(ValueError):
        cxi.CXIReader(str(path), {}, def_stack_ds="/entry_1/photons", pick_events="sideways")


def test_reader_random_pick_is_a_subset(sample):
    path, data, N = sample
    r = cxi.CXIReader(str(path), {"p": "/entry_1/photons"}, def_stack_ds="/entry_1/photons",
                      pick_events="random", nevents=3)
    got = []
    while (d := r.get_next()) is not None:
        got.append(float(d["p"]))
    assert len(got) == 3 and len(set(got)) == 3
    assert set(got) <= set(data["entry_1"]["photons"].tolist())


def test_writer_stack_needs_n_events(tmp_path):
    _h5lib_or_skip()
    w = cxi.CXIWriter(str(tmp_path / "n.cxi"), 4)
    with pytest.raises(ValueError):
        w.write_stack({"x": np.zeros((3, 2))})
    w.close()
