"""Shared set-up for the GPU parity tests: the same seeded inputs go into a
HIP phaser (through the C ABI) and an oracle phaser."""
from __future__ import annotations

import numpy as np

ALG_IDS = {"hio": 1, "raar": 2, "dm": 3, "er": 4}


def rel(a, b) -> float:
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def random_case(shape, seed=0, beamstop=True, fill=0.25):
    """Random real block object, its exact amplitudes, measured mask with a
    small corner beamstop, a support slightly larger than the block and a
    random start model."""
    rng = np.random.default_rng(seed)
    obj = np.zeros(shape)
    sl = tuple(slice(0, max(2, int(s * fill))) for s in shape)
    obj[sl] = rng.random(obj[sl].shape)
    amp = np.abs(np.fft.fftn(obj)).astype(np.float32)
    meas = np.ones(shape, bool)
    if beamstop:
        meas.flat[:3] = False
        amp[~meas] = 0
    sup = np.zeros(shape, bool)
    sup[tuple(slice(0, max(2, int(s * fill)) + 2) for s in shape)] = True
    g0 = ((rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) * 0.05).astype(np.complex64)
    return obj, amp, meas, sup, g0


class GpuPhaser:
    """A HIP phaser built through the C ABI; frees everything it allocated."""

    def __init__(self, sp, alg, amp, meas, sup, model, beta=0.9, constraints=0, sup_array=None,
                 gamma1=float("inf"), gamma2=0.0, beta_map=None):
        self.sp = sp
        self.images = []
        self.beta = sp.smap(beta_map or [(0, beta)])
        if alg == "er":
            self.alg = sp.sp_phasing_er_alloc(constraints)
        elif alg == "hio":
            self.alg = sp.sp_phasing_hio_alloc(self.beta, constraints)
        elif alg == "raar":
            self.alg = sp.sp_phasing_raar_alloc(self.beta, constraints)
        else:
            self.alg = sp.sp_phasing_diff_map_alloc(self.beta, gamma1, gamma2, constraints)
        self.ph = sp.sp_phaser_alloc()
        rc = sp.sp_phaser_init(self.ph, self.alg, sup_array, sp.SpEngineHIP)
        assert rc == 0, rc
        a = self._img(amp, meas)
        sp.sp_phaser_set_amplitudes(self.ph, a)
        if model is not None:
            assert sp.sp_phaser_init_model(self.ph, self._img(model), 0) == 0
        if sup is not None:
            assert sp.sp_phaser_init_support(self.ph, self._img(np.asarray(sup, np.float32)), 0, 0.0) == 0

    def _img(self, data, mask=None):
        img = self.sp.image_from_array(data, mask)
        self.images.append(img)
        return img

    def iterate(self, n):
        return self.sp.sp_phaser_iterate(self.ph, n)

    def get(self, name):
        img = getattr(self.sp, f"sp_phaser_{name}")(self.ph)
        assert bool(img), name
        return self.sp.image_view(img).copy()

    def gammas(self):
        import ctypes as C
        raw = (C.c_float * 2).from_address(self.alg.contents.params + 8)
        return float(raw[0]), float(raw[1])

    def efourier(self):
        import ctypes as C
        f = C.c_float()
        e = self.sp.sp_phaser_efourier(self.ph, C.byref(f))
        return float(e), float(f.value)

    def close(self):
        self.sp.sp_phaser_free(self.ph)
        for img in self.images:
            self.sp.sp_image_free(img)
        self.images = []

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
