"""ORACLE -- test infrastructure, not part of the product.

ctypes wrapper of ``spimage_oracle.c`` (built into ``oracle/_build`` by the
Makefile beside it).  Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU
baseline leg of ``bench.py`` import this module, and only as the checker or
as the timed CPU baseline.

``Oracle(precision)`` loads the float32 build (the reference's FP32
arithmetic) or the float64 build (the FP64 oracle).  ``OraclePhaser`` holds a
phaser's arrays as numpy views and runs the reference's iteration and support
scheduler on them.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
BUILD = HERE / "_build"

HIO, RAAR, DIFFMAP, ER = 1, 2, 3, 4
SUP_NONE, SUP_THRESHOLD, SUP_AREA, SUP_CENTRE, SUP_TEMPLATE, SUP_STATIC, SUP_CLOSE = 0, 1, 2, 3, 4, 5, 6


def build(force: bool = False) -> None:
    """Compile both oracle builds with the Makefile beside this file."""
    args = ["make", "-s", "-C", str(HERE)]
    if force:
        args.append("-B")
    subprocess.run(args, check=True)


def _define_structs(R):
    OrAlg_ = type("OrAlg", (C.Structure,), {"_fields_": [
        ("type", C.c_int), ("constraints", C.c_int), ("nbeta", C.c_int),
        ("beta_keys", R * 32), ("beta_vals", R * 32), ("gamma1", R), ("gamma2", R)]})
    OrSup_ = type("OrSup", (C.Structure,), {"_fields_": [
        ("type", C.c_int), ("period", C.c_int), ("nblur", C.c_int),
        ("blur_keys", R * 32), ("blur_vals", R * 32), ("nval", C.c_int),
        ("val_keys", R * 32), ("val_vals", R * 32), ("close_size", C.c_int),
        ("tmpl_abs", C.c_void_p), ("tmpl_sorted", C.c_void_p), ("tmpl_area", R)]})
    return OrAlg_, OrSup_


class Oracle:
    _cache: dict[str, "Oracle"] = {}

    def __new__(cls, precision: str = "f32"):
        if precision in cls._cache:
            return cls._cache[precision]
        self = super().__new__(cls)
        path = BUILD / f"liboracle_{precision}.so"
        if not path.exists():
            build()
        self.lib = C.CDLL(str(path))
        self.precision = precision
        self.R = C.c_float if precision == "f32" else C.c_double
        self.np_real = np.float32 if precision == "f32" else np.float64
        self.np_cplx = np.complex64 if precision == "f32" else np.complex128
        self.OrAlg, self.OrSup = _define_structs(self.R)
        L, R, P = self.lib, self.R, C.c_void_p
        L.or_phaser_create.restype = P
        L.or_phaser_create.argtypes = [C.c_int, C.c_int, C.c_int]
        L.or_phaser_free.argtypes = [P]
        for name in ("or_g0", "or_g1", "or_gp", "or_phased_amp"):
            getattr(L, name).restype = P
            getattr(L, name).argtypes = [P]
        for name in ("or_amp", "or_amp_min", "or_amp_max", "or_flags", "or_iteration", "or_objective",
                     "or_have_margins"):
            getattr(L, name).restype = P
            getattr(L, name).argtypes = [P]
        L.or_iterate.restype = C.c_int
        L.or_iterate.argtypes = [P, C.POINTER(self.OrAlg), C.POINTER(self.OrSup), C.c_int]
        L.or_iterate_multi.restype = C.c_int
        L.or_iterate_multi.argtypes = [P, C.POINTER(self.OrAlg), C.POINTER(self.OrSup), C.c_int, C.c_int]
        L.or_template_prepare.argtypes = [P, C.c_int, C.c_int, C.c_int, R, P, P, C.POINTER(R)]
        L.or_support_template_update.argtypes = [P, P, P, R, R]
        L.or_support_close_update.argtypes = [P, C.c_int]
        L.or_support_centre_image.argtypes = [P]
        L.or_efourier.restype = R
        L.or_efourier.argtypes = [P, C.POINTER(R)]
        L.or_ereal.restype = R
        L.or_ereal.argtypes = [P]
        L.or_support_fraction.restype = R
        L.or_support_fraction.argtypes = [P]
        L.or_last_threshold.restype = R
        L.or_last_threshold.argtypes = [P]
        L.or_fft.argtypes = [P, P, C.c_int, C.c_int, C.c_int, C.c_int]
        L.or_gaussian_blur.argtypes = [P, P, C.c_int, C.c_int, C.c_int, R]
        L.or_support_area_update.argtypes = [P, R, R]
        L.or_support_threshold_update.argtypes = [P, R, R]
        L.or_bezier.restype = R
        L.or_bezier.argtypes = [C.POINTER(R), C.POINTER(R), C.c_int, R]
        L.or_module_projection.restype = C.c_int
        L.or_module_projection.argtypes = [P, P, P, P, P, C.c_long, C.c_int, C.c_int]
        L.or_apply_constraints.argtypes = [P, P, C.c_long, C.c_int]
        L.or_apply_fourier_constraints.argtypes = [P, C.c_long, C.c_int]
        L.or_model_change.argtypes = [P, P]
        L.or_dm_gammas.argtypes = [R, C.POINTER(R), C.POINTER(R)]
        L.or_support_log.restype = C.c_int
        L.or_support_log.argtypes = [P, C.POINTER(C.c_int), C.c_int]
        L.or_set_fft_backend.restype = C.c_int
        L.or_set_fft_backend.argtypes = [C.c_int, C.c_char_p]
        L.or_set_fft_threads.restype = C.c_int
        L.or_set_fft_threads.argtypes = [C.c_int]
        cls._cache[precision] = self
        return self

    # --- stateless helpers -------------------------------------------------
    def fft(self, x: np.ndarray, sign: int) -> np.ndarray:
        a = np.ascontiguousarray(x, dtype=self.np_cplx)
        out = np.empty_like(a)
        nz, ny, nx = _zyx(a.shape)
        self.lib.or_fft(a.ctypes.data, out.ctypes.data, nx, ny, nz, sign)
        return out

    def blur(self, x: np.ndarray, radius: float) -> np.ndarray:
        a = np.ascontiguousarray(x, dtype=self.np_cplx)
        out = np.empty_like(a)
        nz, ny, nx = _zyx(a.shape)
        self.lib.or_gaussian_blur(a.ctypes.data, out.ctypes.data, nx, ny, nz, radius)
        return out

    def bezier(self, pairs, x: float) -> float:
        n = len(pairs)
        keys = (self.R * n)(*[p[0] for p in pairs])
        vals = (self.R * n)(*[p[1] for p in pairs])
        return float(self.lib.or_bezier(keys, vals, n, x))

    def dm_gammas(self, beta0: float) -> tuple[float, float]:
        g1, g2 = self.R(), self.R()
        self.lib.or_dm_gammas(beta0, C.byref(g1), C.byref(g2))
        return float(g1.value), float(g2.value)

    def set_fft_backend(self, backend: int, libpath: str = "") -> int:
        return int(self.lib.or_set_fft_backend(backend, libpath.encode()))

    def set_fft_threads(self, nthreads: int) -> int:
        """Threads of the FFTW3-API library's plans (backend 1 only)."""
        return int(self.lib.or_set_fft_threads(nthreads))

    def alg(self, type_: int, constraints: int = 0, beta=None, gamma1: float = 0.0, gamma2: float = 0.0):
        a = self.OrAlg()
        a.type = type_
        a.constraints = constraints
        pairs = list(beta or [(0, 0.0)])
        a.nbeta = len(pairs)
        for i, (k, v) in enumerate(pairs):
            a.beta_keys[i] = k
            a.beta_vals[i] = v
        a.gamma1 = gamma1
        a.gamma2 = gamma2
        return a

    def template(self, tmpl: np.ndarray, radius: float):
        """sp_support_template_alloc's set-up: (|blurred|, sorted descending, area fraction)."""
        a = np.ascontiguousarray(tmpl, dtype=self.np_cplx)
        nz, ny, nx = _zyx(a.shape)
        absv = np.empty(a.shape, self.np_real)
        srt = np.empty(a.size, self.np_real)
        area = self.R()
        self.lib.or_template_prepare(a.ctypes.data, nx, ny, nz, radius, absv.ctypes.data, srt.ctypes.data,
                                     C.byref(area))
        return absv, srt, float(area.value)

    def sup(self, type_: int = SUP_NONE, period: int = 1, blur=None, values=None, close_size: int = 0,
            template=None):
        """One support algorithm; `template` is the tuple returned by template()."""
        s = self.OrSup()
        s.type = type_
        s.period = period
        s.close_size = close_size
        if template is not None:
            absv, srt, area = (np.ascontiguousarray(template[0], self.np_real),
                               np.ascontiguousarray(template[1], self.np_real), template[2])
            s.tmpl_abs = absv.ctypes.data
            s.tmpl_sorted = srt.ctypes.data
            s.tmpl_area = area
            s._keep = (absv, srt)  # the arrays must outlive the structure
        for name, pairs in (("blur", blur or [(0, 0.0)]), ("val", values or [(0, 0.0)])):
            setattr(s, "nblur" if name == "blur" else "nval", len(pairs))
            keys = getattr(s, f"{name}_keys")
            vals = getattr(s, f"{name}_vals")
            for i, (k, v) in enumerate(pairs):
                keys[i] = k
                vals[i] = v
        return s

    def phaser(self, shape) -> "OraclePhaser":
        return OraclePhaser(self, shape)


def _zyx(shape):
    if len(shape) == 2:
        return 1, shape[0], shape[1]
    return shape[0], shape[1], shape[2]


class OraclePhaser:
    """Arrays of one phaser as numpy views over the oracle's memory.

    Re-read `g1`, `g0` and `gp` after every call: the oracle swaps its g0/g1
    buffers the way the reference does (src/phasing.c:1146-1148)."""

    def __init__(self, o: Oracle, shape):
        self.o = o
        self.shape = tuple(shape)
        nz, ny, nx = _zyx(self.shape)
        self.N = nx * ny * nz
        self.h = o.lib.or_phaser_create(nx, ny, nz)

    def __del__(self):
        try:
            self.o.lib.or_phaser_free(self.h)
        except Exception:
            pass

    def _view(self, fn, ctype, dtype):
        p = getattr(self.o.lib, fn)(self.h)
        buf = (ctype * self.N).from_address(p)
        buf._owner = self  # the view keeps the phaser (and its memory) alive
        return np.frombuffer(buf, dtype=dtype).reshape(self.shape)

    def _cview(self, fn):
        p = getattr(self.o.lib, fn)(self.h)
        buf = (self.o.R * (2 * self.N)).from_address(p)
        buf._owner = self
        return np.frombuffer(buf, dtype=self.o.np_cplx).reshape(self.shape)

    @property
    def g1(self):
        return self._cview("or_g1")

    @property
    def g0(self):
        return self._cview("or_g0")

    @property
    def gp(self):
        return self._cview("or_gp")

    @property
    def phased_amp(self):
        return self._cview("or_phased_amp")

    @property
    def amp(self):
        return self._view("or_amp", self.o.R, self.o.np_real)

    @property
    def amp_min(self):
        return self._view("or_amp_min", self.o.R, self.o.np_real)

    @property
    def amp_max(self):
        return self._view("or_amp_max", self.o.R, self.o.np_real)

    @property
    def flags(self):
        return self._view("or_flags", C.c_int, np.int32)

    def _int(self, fn):
        v = C.c_int.from_address(getattr(self.o.lib, fn)(self.h))
        v._owner = self
        return v

    @property
    def iteration(self) -> int:
        return self._int("or_iteration").value

    @iteration.setter
    def iteration(self, v: int):
        self._int("or_iteration").value = v

    @property
    def objective(self) -> int:
        return self._int("or_objective").value

    @objective.setter
    def objective(self, v: int):
        self._int("or_objective").value = v

    def set_margins(self, lo, hi):
        self.amp_min[...] = lo
        self.amp_max[...] = hi
        self._int("or_have_margins").value = 1

    def set_inputs(self, amp, measured, support, model):
        """Amplitudes, measured mask, support mask and start model (g1)."""
        self.amp[...] = amp
        self.flags[...] = (np.asarray(support, bool) * 1) | (np.asarray(measured, bool) * 2)
        self.g1[...] = model
        self.g0[...] = 0
        self.gp[...] = 0

    def iterate(self, alg, n: int, sup=None) -> int:
        """`sup`: one support algorithm, or a list of them run in order at each update."""
        if isinstance(sup, (list, tuple)):
            arr = (self.o.OrSup * len(sup))(*sup)
            return int(self.o.lib.or_iterate_multi(self.h, C.byref(alg), arr, len(sup), n))
        return int(self.o.lib.or_iterate(self.h, C.byref(alg), C.byref(sup) if sup is not None else None, n))

    def efourier(self):
        f = self.o.R()
        e = self.o.lib.or_efourier(self.h, C.byref(f))
        return float(e), float(f.value)

    def ereal(self) -> float:
        return float(self.o.lib.or_ereal(self.h))

    def support_fraction(self) -> float:
        return float(self.o.lib.or_support_fraction(self.h))

    def support_log(self) -> list[int]:
        """Iterations consumed by support updates so far (first 256)."""
        buf = (C.c_int * 256)()
        n = self.o.lib.or_support_log(self.h, buf, 256)
        return list(buf[: min(n, 256)])

    def support_area_update(self, radius: float, area: float) -> float:
        self.o.lib.or_support_area_update(self.h, radius, area)
        return float(self.o.lib.or_last_threshold(self.h))

    def support_threshold_update(self, radius: float, rel: float) -> float:
        self.o.lib.or_support_threshold_update(self.h, radius, rel)
        return float(self.o.lib.or_last_threshold(self.h))

    def support_template_update(self, template, area: float) -> float:
        """`template`: the tuple of Oracle.template()."""
        absv = np.ascontiguousarray(template[0], self.o.np_real)
        srt = np.ascontiguousarray(template[1], self.o.np_real)
        self.o.lib.or_support_template_update(self.h, absv.ctypes.data, srt.ctypes.data, area, template[2])
        return float(self.o.lib.or_last_threshold(self.h))

    def support_close_update(self, size: int) -> None:
        self.o.lib.or_support_close_update(self.h, size)

    def support_centre_image(self) -> None:
        self.o.lib.or_support_centre_image(self.h)
