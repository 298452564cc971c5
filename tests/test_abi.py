"""The C ABI without a GPU: the library loads, exports every symbol the
headers in include/ declare, and its host-side logic (schedules, images,
argument checks, error codes) behaves like the reference's."""
import ctypes as C
import re
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


@pytest.fixture(scope="module")
def spm():
    from libspimage_amd import spimage
    return spimage


def declared_symbols():
    names = set()
    for h in sorted((ROOT / "include").glob("*.h")):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"^\s*#.*$", "", text, flags=re.M)
        for m in re.finditer(r"spimage_EXPORT\s+[^;{]*?\b(\w+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_every_declared_symbol_is_exported(spm):
    names = declared_symbols()
    assert len(names) > 80
    missing = [n for n in names if not hasattr(spm.lib, n)]
    assert not missing, missing


def test_smap_schedule_matches_oracle(spm, o64):
    """The engine's Cardano solution of the schedule curve agrees with the
    oracle's bisection (include/spimage/map.h:103-141)."""
    pairs = [(0, 3.0), (300, 0.7)]
    m = spm.smap(pairs)
    for x in (-3, 0, 1, 20, 150, 299, 300, 400):
        assert abs(spm.sp_smap_interpolate(m, x) - o64.bezier(pairs, x)) <= 2e-7 * 3.0
    three = spm.smap([(100, 0.9), (0, 0.5), (400, 0.2)])  # inserted out of order
    keys = three.contents.keys.contents
    assert [keys.data[i] for i in range(keys.used)] == [0.0, 100.0, 400.0]
    assert spm.sp_smap_max(three) == pytest.approx(0.9)
    assert spm.sp_smap_min(three) == pytest.approx(0.2)
    for x in (50, 100, 250):
        assert abs(spm.sp_smap_interpolate(three, x) - o64.bezier([(0, 0.5), (100, 0.9), (400, 0.2)], x)) < 1e-6


def test_dm_gamma_selection_matches_oracle(spm, o32):
    for beta in (0.7, 0.8, 0.9):
        alg = spm.sp_phasing_diff_map_alloc(spm.smap([(0, beta)]), float("inf"), 0.0, 0)
        raw = (C.c_float * 2).from_address(alg.contents.params + 8)
        g1, g2 = o32.dm_gammas(beta)
        assert raw[0] == pytest.approx(g1, rel=1e-6) and raw[1] == pytest.approx(g2, rel=1e-6)
    # finite gamma1 keeps both user values
    alg = spm.sp_phasing_diff_map_alloc(spm.smap([(0, 0.8)]), 0.5, 0.25, 0)
    raw = (C.c_float * 2).from_address(alg.contents.params + 8)
    assert (raw[0], raw[1]) == (0.5, 0.25)


def test_images_and_duplicate(spm):
    a = (np.arange(12).reshape(3, 4) * (1 + 2j)).astype(np.complex64)
    img = spm.image_from_array(a, np.ones((3, 4)))
    assert img.contents.num_dimensions == 2
    assert list(img.contents.detector.contents.image_center)[:2] == [1.5, 1.0]
    d = spm.sp_image_duplicate(img, spm.SP_COPY_DATA)
    assert np.array_equal(spm.image_view(d), a)
    assert not spm.mask_view(d).any()
    d2 = spm.sp_image_duplicate(img, spm.SP_COPY_ALL)
    assert spm.mask_view(d2).all()
    for i in (img, d, d2):
        spm.sp_image_free(i)
    v = spm.image_from_array(np.zeros((2, 3, 4), np.complex64))
    assert v.contents.num_dimensions == 3 and spm.image_view(v).shape == (2, 3, 4)
    spm.sp_image_free(v)


def test_init_error_codes_without_gpu(spm):
    if spm.spx_device_available():
        pytest.skip("a GPU is present; the no-device path is not reachable")
    ph = spm.sp_phaser_alloc()
    er = spm.sp_phasing_er_alloc(0)
    assert spm.sp_phaser_init(None, er, None, spm.SpEngineHIP) == -1
    assert spm.sp_phaser_init(ph, None, None, spm.SpEngineHIP) == -2
    assert spm.sp_phaser_init(ph, er, None, spm.SpEngineCPU) == -9
    assert spm.sp_phaser_init(ph, er, None, spm.SpEngineHIP) == -10
    assert spm.sp_phaser_iterate(ph, 1) == -3
    spm.sp_phaser_free(ph)
    assert spm.spx_fft_device(None, None, 8, 8, 1, 3) < 0
    # tuning ahead needs the device, and refuses rather than touching it
    assert spm.spx_tune_shape(64, 64, 1, 4) == -20
    with pytest.raises(ValueError):
        spm.tune_shape((64, 64), 4)


def test_supported_lengths(spm):
    for L in (2, 8, 128, 384, 1024, 4096, 80, 96, 640, 960, 1000, 3, 5, 15, 5120, 6561, 8192,
              7, 49, 343, 448, 1029, 1400, 2401, 5488, 7168):
        assert spm.spx_supported_length(L) == 1, L
    # 6250 = 5 10 5 5 5 and 7350 = 7 15 10 7 need over 1024 threads per
    # line; 11, 13, 77 run through Bluestein transforms, not the Stockham
    # passes
    for L in (11, 13, 77, 6250, 5250, 7350, 7560, 7840, 16384, 1):
        assert spm.spx_supported_length(L) == 0, L


def test_ifft_of_unphased_image_aborts():
    """sp_image_ifft aborts on an unphased image, like the reference (src/fft.c:58-61)."""
    code = (
        "import sys; sys.path.insert(0, %r)\n"
        "from libspimage_amd import spimage as sp\n"
        "import numpy as np\n"
        "img = sp.image_from_array(np.zeros((4, 4), np.complex64))\n"
        "img.contents.phased = 0\n"
        "img.contents.shifted = 1\n"
        "sp.sp_image_ifft(img)\n"
    ) % str(ROOT)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "unphased" in r.stderr


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU; --gpus N > 1 exits with a message."""
    import subprocess
    import sys
    from pathlib import Path
    root = Path(__file__).resolve().parents[1]
    r = subprocess.run([sys.executable, str(root / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_fixed_array_transform_rejects_before_touching_the_device(spm):
    """spx_fft_noalloc checks its size limit (2^21 elements) and direction
    before any device work, so these run without a GPU."""
    a = np.zeros(4, np.complex64)
    assert spm.spx_fft_noalloc(a.ctypes.data, a.ctypes.data, 2048, 1025, 1, -1) == -12
    assert "fixed device array" in spm.last_error()
    assert spm.spx_fft_noalloc(a.ctypes.data, a.ctypes.data, 2, 2, 1, 0) == -1
    assert spm.spx_fft_noalloc(None, a.ctypes.data, 2, 2, 1, -1) == -1
