"""ctypes mirror of the reference's ``spimage`` Python module for the phasing path.

The reference builds ``spimage_pybackend`` with SWIG from its headers
(reference src/spimage_pybackend.i:163-185) and re-exports it from
``spimage.py`` (reference src/spimage.py:11).  This module binds the same C
names from ``libspimage_hip.so`` so that code written against the reference
reads the same: ``sp_phaser_alloc``, ``sp_phaser_init``,
``sp_phaser_set_amplitudes``, ``sp_phaser_iterate``, ``sp_smap_alloc`` and so
on.  Arrays cross the boundary as numpy views in the reference typemaps'
orientation: ``(y, x)`` for 2-D images and ``(z, y, x)`` for 3-D ones
(reference src/spimage_pybackend.i:79-101).

The library is loaded from ``libspimage_amd/_lib``; importing this module
raises if it has not been built, so nothing silently falls back to a CPU path.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

_LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libspimage_hip.so"
if not _LIB_PATH.exists():
    raise ImportError(
        f"{_LIB_PATH} is missing: build it with `python -m libspimage_amd.build` "
        "(the HIP engine has no CPU fallback)"
    )
lib = C.CDLL(str(_LIB_PATH))

# ---------------------------------------------------------------------------
# enums (reference include/spimage/phasing.h:14-20, support_update.h:9)
# ---------------------------------------------------------------------------
SpModelRandomPhases, SpModelZeroPhases, SpModelRandomValues, SpModelMaskedOutZeroed = 1, 2, 4, 256
SpSupportFromPatterson = 1
SpHIO, SpRAAR, SpDiffMap, SpER = 1, 2, 3, 4
SpNoConstraints, SpRealObject, SpPositiveRealObject, SpPositiveComplexObject = 0, 1, 2, 4
SpPositivityFlipping, SpCentrosymmetricObject, SpAmplitudeErrorMargin = 8, 16, 32
SpEngineAutomatic, SpEngineCPU, SpEngineCUDA = 0, 1, 2
SpEngineHIP = 2
SpPixelInsideSupport, SpPixelMeasuredAmplitude = 1, 2
SpRecoverPhases, SpRecoverAmplitudes = 0, 1
SpSupportThreshold, SpSupportArea, SpSupportCentreImage, SpSupportTemplate, SpSupportStatic, SpSupportClose = 1, 2, 3, 4, 5, 6
SP_COPY_DATA, SP_COPY_DETECTOR, SP_COPY_MASK, SP_COPY_ALL = 1, 0, 2, 3
SpCUDANoDevice, SpCUDAEmulatedDevice, SpCUDAHardwareDevice = 0, 1, 2  # reference cuda_util.h:101


# ---------------------------------------------------------------------------
# structs (layouts of include/sp_core.h, sp_phasing.h, sp_support.h)
# ---------------------------------------------------------------------------
class Complex(C.Structure):
    _fields_ = [("re", C.c_float), ("im", C.c_float)]


class sp_3matrix(C.Structure):
    _fields_ = [("x", C.c_uint), ("y", C.c_uint), ("z", C.c_uint), ("data", C.POINTER(C.c_float))]


class sp_i3matrix(C.Structure):
    _fields_ = [("x", C.c_uint), ("y", C.c_uint), ("z", C.c_uint), ("data", C.POINTER(C.c_int))]


class sp_c3matrix(C.Structure):
    _fields_ = [("x", C.c_uint), ("y", C.c_uint), ("z", C.c_uint), ("data", C.POINTER(Complex))]


class Detector(C.Structure):
    _fields_ = [
        ("image_center", C.c_float * 3),
        ("pixel_size", C.c_float * 3),
        ("detector_distance", C.c_float),
        ("wavelength", C.c_float),
        ("orientation", C.c_void_p),
    ]


class Image(C.Structure):
    _fields_ = [
        ("phased", C.c_int),
        ("scaled", C.c_int),
        ("image", C.POINTER(sp_c3matrix)),
        ("mask", C.POINTER(sp_i3matrix)),
        ("detector", C.POINTER(Detector)),
        ("shifted", C.c_int),
        ("rec_coords", C.POINTER(sp_3matrix)),
        ("num_dimensions", C.c_int),
    ]


class sp_list(C.Structure):
    _fields_ = [("data", C.POINTER(C.c_float)), ("used", C.c_uint), ("size", C.c_uint)]


class sp_smap(C.Structure):
    _fields_ = [("keys", C.POINTER(sp_list)), ("values", C.POINTER(sp_list))]


class SpPhasingAlgorithm(C.Structure):
    _fields_ = [("type", C.c_int), ("params", C.c_void_p)]


class SpSupportAlgorithm(C.Structure):
    _fields_ = [("type", C.c_int), ("params", C.c_void_p), ("function", C.c_void_p)]


class SpSupportArray(C.Structure):
    _fields_ = [("size", C.c_int), ("algorithms", C.POINTER(C.POINTER(SpSupportAlgorithm))), ("update_period", C.c_int)]


class SpPhaser(C.Structure):
    _fields_ = [
        ("amplitudes", C.POINTER(sp_3matrix)),
        ("amplitudes_min", C.POINTER(sp_3matrix)),
        ("amplitudes_max", C.POINTER(sp_3matrix)),
        ("phased_amplitudes", C.POINTER(sp_c3matrix)),
        ("pixel_flags", C.POINTER(sp_i3matrix)),
        ("phasing_objective", C.c_int),
        ("algorithm", C.POINTER(SpPhasingAlgorithm)),
        ("sup_algorithm", C.POINTER(SpSupportArray)),
        ("iteration", C.c_int),
        ("image_size", C.c_int),
        ("nx", C.c_int),
        ("ny", C.c_int),
        ("nz", C.c_int),
        ("model", C.POINTER(Image)),
        ("model_iteration", C.c_int),
        ("old_model", C.POINTER(Image)),
        ("old_model_iteration", C.c_int),
        ("model_change", C.POINTER(Image)),
        ("model_change_iteration", C.c_int),
        ("support", C.POINTER(Image)),
        ("support_iteration", C.c_int),
        ("amplitudes_image", C.POINTER(Image)),
        ("fmodel", C.POINTER(Image)),
        ("fmodel_iteration", C.c_int),
        ("model_before_projection", C.POINTER(Image)),
        ("model_before_projection_iteration", C.c_int),
        ("engine", C.c_int),
        ("g0", C.POINTER(Image)),
        ("g1", C.POINTER(Image)),
        ("gp", C.POINTER(Image)),
        ("hip_engine", C.c_void_p),
    ]


PImage = C.POINTER(Image)
PSmap = C.POINTER(sp_smap)
PPhaser = C.POINTER(SpPhaser)
PAlg = C.POINTER(SpPhasingAlgorithm)
PSupAlg = C.POINTER(SpSupportAlgorithm)
PSupArr = C.POINTER(SpSupportArray)


def _bind(name, restype, *argtypes, alias=None):
    f = getattr(lib, name)
    f.restype = restype
    f.argtypes = list(argtypes)
    globals()[alias or name] = f
    return f


# images
_bind("_sp_image_alloc", PImage, C.c_int, C.c_int, C.c_int, C.c_char_p, C.c_int)
_bind("_sp_image_free", None, PImage, C.c_char_p, C.c_int)
_bind("_sp_image_duplicate", PImage, PImage, C.c_int, C.c_char_p, C.c_int)


def sp_image_alloc(x: int, y: int, z: int) -> PImage:
    return _sp_image_alloc(x, y, z, b"python", 0)


def sp_image_free(img: PImage) -> None:
    _sp_image_free(img, b"python", 0)


def sp_image_duplicate(img: PImage, flags: int) -> PImage:
    return _sp_image_duplicate(img, flags, b"python", 0)


_bind("p_drand48", C.c_float)
_bind("sp_srand", None, C.c_int)

# schedules: the header's inline sp_smap_* map onto these exports
_bind("spx_smap_alloc", PSmap, C.c_uint, alias="sp_smap_alloc")
_bind("spx_smap_free", None, PSmap, alias="sp_smap_free")
_bind("spx_smap_insert", None, PSmap, C.c_float, C.c_float, alias="sp_smap_insert")
_bind("spx_smap_create_from_pair", PSmap, C.c_float, C.c_float, alias="sp_smap_create_from_pair")
_bind("spx_smap_interpolate", C.c_float, PSmap, C.c_float, alias="sp_smap_interpolate")
_bind("spx_smap_max", C.c_float, PSmap, alias="sp_smap_max")
_bind("spx_smap_min", C.c_float, PSmap, alias="sp_smap_min")

# algorithms
_bind("sp_phasing_hio_alloc", PAlg, PSmap, C.c_int)
_bind("sp_phasing_raar_alloc", PAlg, PSmap, C.c_int)
_bind("sp_phasing_diff_map_alloc", PAlg, PSmap, C.c_float, C.c_float, C.c_int)
_bind("sp_phasing_er_alloc", PAlg, C.c_int)

# phaser
_bind("sp_phaser_alloc", PPhaser)
_bind("sp_phaser_free", None, PPhaser)
for _n in ("model", "model_with_support", "model_before_projection", "old_model", "fmodel",
           "fmodel_with_mask", "model_change", "support", "amplitudes"):
    _bind(f"sp_phaser_{_n}", PImage, PPhaser)
_bind("sp_phaser_ereal", C.c_float, PPhaser)
_bind("sp_phaser_efourier", C.c_float, PPhaser, C.POINTER(C.c_float))
_bind("sp_phaser_support_fraction", C.c_float, PPhaser)
for _n in ("model", "support", "phased_amplitudes", "amplitudes"):
    _bind(f"sp_phaser_set_{_n}", None, PPhaser, PImage)
_bind("sp_phaser_set_amplitudes_margins", None, PPhaser, PImage, PImage)
_bind("sp_phaser_set_objective", None, PPhaser, C.c_int)
_bind("sp_phaser_init", C.c_int, PPhaser, PAlg, PSupArr, C.c_int)
_bind("sp_phaser_set_phasing_algorithm", C.c_int, PPhaser, PAlg)
_bind("sp_phaser_set_support_algorithm", C.c_int, PPhaser, PSupArr)
_bind("sp_phaser_init_model", C.c_int, PPhaser, PImage, C.c_int)
_bind("sp_phaser_init_support", C.c_int, PPhaser, PImage, C.c_int, C.c_float)
_bind("sp_phaser_iterate", C.c_int, PPhaser, C.c_int)

# support
_bind("sp_support_threshold_alloc", PSupAlg, PSmap, PSmap)
_bind("sp_support_area_alloc", PSupAlg, PSmap, PSmap)
_bind("sp_support_static_alloc", PSupAlg)
_bind("sp_support_close_alloc", PSupAlg, C.c_int)
_bind("sp_support_centre_image_alloc", PSupAlg)
_bind("sp_support_template_alloc", PSupAlg, PImage, C.c_float, PSmap)
_bind("sp_support_array_alloc", PSupArr, C.c_int, C.c_int)
_bind("sp_support_array_init", PSupArr, PSupAlg, C.c_int)
_bind("sp_support_array_set", None, PSupArr, C.c_int, PSupAlg)
_bind("sp_support_array_append", None, PSupArr, PSupAlg)
_bind("sp_support_array_update", C.c_int, PSupArr, PPhaser)

# FFT and blur
_bind("sp_image_fft", PImage, PImage)
_bind("sp_image_ifft", PImage, PImage)
_bind("sp_image_fft_fast", None, PImage, PImage)
_bind("sp_image_ifft_fast", None, PImage, PImage)
_bind("sp_gaussian_blur", PImage, PImage, C.c_float)

# device queries of the reference's CUDA glue (reference include/spimage/cuda_util.h:101-109)
_bind("sp_cuda_get_device_type", C.c_int)
_bind("sp_cuda_launch_parameters", None, C.c_int64, C.POINTER(C.c_int), C.POINTER(C.c_int))

# plain-pointer engine ABI
_bind("spx_device_available", C.c_int)
_bind("spx_last_error", C.c_char_p)
_bind("spx_supported_length", C.c_int, C.c_int)
_bind("spx_fft_variant", C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int))
_bind("spx_tune_shape", C.c_int, C.c_int, C.c_int, C.c_int, C.c_int)


def tune_shape(shape, nrec: int = 1) -> None:
    """Runs the tuner for arrays of `shape` ((y, x) or (z, y, x)) in batches
    of `nrec` now instead of at the first engine or batch of that shape."""
    dims = (shape[-1], shape[-2], shape[0] if len(shape) == 3 else 1)
    if spx_tune_shape(*dims, nrec) != 0:
        raise ValueError(last_error())


def fft_variant(L: int, tiled: bool = False) -> tuple[int, int]:
    """(row, column) elements-per-thread of the tuned transforms of L x L
    arrays (in either layout)."""
    r, c = C.c_int(), C.c_int()
    if spx_fft_variant(L, 1 if tiled else 0, C.byref(r), C.byref(c)) != 0:
        raise ValueError(last_error())
    return r.value, c.value
_bind("spx_fft_device", C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int)
_bind("spx_fft_noalloc", C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int)
_bind("spx_phaser_time_iterations", C.c_int, PPhaser, C.c_int, C.POINTER(C.c_float))
_bind("spx_phaser_time_passes", C.c_int, PPhaser, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float))
_bind("spx_phaser_tiled", C.c_int, PPhaser)
_bind("spx_phaser_nt_hints", C.c_int, PPhaser)
_bind("spx_phaser_device_arrays", C.c_int, PPhaser, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
      C.POINTER(C.c_void_p))
_bind("spx_phaser_set_seed", None, PPhaser, C.c_ulonglong)
_bind("spx_phaser_last_threshold", C.c_int, PPhaser, C.POINTER(C.c_float))


# batched phasing (include/sp_batch.h)
PBatch = C.c_void_p
PF = C.POINTER(C.c_float)
PI = C.POINTER(C.c_int)
_bind("spx_batch_create", PBatch, C.c_int, C.c_int, C.c_int, C.c_int)
_bind("spx_batch_free", None, PBatch)
_bind("spx_batch_size", C.c_int, PBatch)
_bind("spx_batch_set_amplitudes", C.c_int, PBatch, C.c_int, PF, PI)
_bind("spx_batch_set_support", C.c_int, PBatch, C.c_int, PF)
_bind("spx_batch_set_data", C.c_int, PBatch, C.c_int, PF, PI, PF)
_bind("spx_batch_set_model", C.c_int, PBatch, C.c_int, PF)
_bind("spx_batch_init_model", C.c_int, PBatch, C.c_int, C.c_ulonglong)
_bind("spx_batch_set_algorithm", C.c_int, PBatch, PAlg)
_bind("spx_batch_set_rec_algorithms", C.c_int, PBatch, PI, PF)
_bind("spx_batch_set_support_schedule", C.c_int, PBatch, PSupArr)
_bind("spx_batch_use_graphs", C.c_int, PBatch, C.c_int)
_bind("spx_batch_iterate", C.c_int, PBatch, C.c_int)
_bind("spx_batch_iteration", C.c_int, PBatch)
_bind("spx_batch_efourier", C.c_int, PBatch, PF, PF)
_bind("spx_batch_efourier_enqueue", C.c_int, PBatch)
_bind("spx_batch_efourier_fetch", C.c_int, PBatch, PF, PF)
_bind("spx_batch_ereal", C.c_int, PBatch, PF)
_bind("spx_batch_support_fraction", C.c_int, PBatch, PF)
_bind("spx_batch_get_model", C.c_int, PBatch, C.c_int, PF)
_bind("spx_batch_get_support", C.c_int, PBatch, C.c_int, PI)
_bind("spx_batch_time_iterations", C.c_int, PBatch, C.c_int, PF)
_bind("spx_batch_synchronize", C.c_int, PBatch)
_bind("spx_batch_init_model_rec", C.c_int, PBatch, C.c_int, C.c_int, C.c_ulonglong)
_bind("spx_batch_set_active", C.c_int, PBatch, C.POINTER(C.c_int))
_bind("spx_batch_get_model_before_projection", C.c_int, PBatch, C.c_int, PF)
_bind("spx_batch_get_fmodel", C.c_int, PBatch, C.c_int, PF)
_bind("spx_batch_query", C.c_int, PBatch)
_bind("spx_batch_estimate_bytes", C.c_longlong, C.c_int, C.c_int, C.c_int, C.c_int)
_bind("spx_device_memory", C.c_int, C.POINTER(C.c_ulonglong), C.POINTER(C.c_ulonglong))
_bind("spx_host_alloc", C.c_void_p, C.c_ulonglong)
_bind("spx_host_free", None, C.c_void_p)

# PRTF (include/sp_prtf.h)
_bind("spx_prtf", C.c_int, PF, PI, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, PF, PF, PF, PI, PF)


def device_memory() -> tuple[int, int]:
    """(free, total) bytes of device memory."""
    f, t = C.c_ulonglong(), C.c_ulonglong()
    if spx_device_memory(C.byref(f), C.byref(t)) != 0:
        raise RuntimeError(last_error())
    return f.value, t.value


def _fptr(a):
    return a.ctypes.data_as(PF)


def _iptr(a):
    return a.ctypes.data_as(PI)


class Batch:
    """nrec same-shape reconstructions advanced together on the GPU.

    Arrays are numpy, shaped like one image ((y, x) or (z, y, x)) for a
    single reconstruction or with a leading reconstruction axis for all."""

    def __init__(self, shape, nrec):
        self.shape = tuple(int(s) for s in shape)
        z, y, x = (1,) + self.shape if len(self.shape) == 2 else self.shape
        self.nrec = nrec
        self.h = spx_batch_create(x, y, z, nrec)
        if not self.h:
            raise RuntimeError(f"spx_batch_create failed: {last_error()}")
        self._keep = []
        self._stage = None  # page-locked staging for the downloads of all reconstructions

    def close(self):
        if self.h:
            spx_batch_free(self.h)
            self.h = None
        if self._stage is not None:
            spx_host_free(self._stage[0])
            self._stage = None

    def _staging(self, dtype):
        """A view of the page-locked staging buffer as nrec arrays of `dtype`,
        or None when page-locked memory is not available."""
        nbytes = self.nrec * int(np.prod(self.shape)) * np.dtype(dtype).itemsize
        if self._stage is None or self._stage[1] < nbytes:
            if self._stage is not None:
                spx_host_free(self._stage[0])
            p = spx_host_alloc(nbytes)
            self._stage = (p, nbytes) if p else None
            if p is None:
                return None
        buf = (C.c_char * nbytes).from_address(self._stage[0])
        return np.frombuffer(buf, dtype=dtype).reshape((self.nrec,) + self.shape)

    def __del__(self):
        self.close()

    def _check(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed ({rc}): {last_error()}")

    def set_data(self, rec, amplitudes, mask, support):
        a = np.ascontiguousarray(amplitudes, np.float32)
        m = np.ascontiguousarray(mask, np.int32)
        s = np.ascontiguousarray(support, np.float32)
        self._check(spx_batch_set_data(self.h, rec, _fptr(a), _iptr(m), _fptr(s)), "set_data")

    def set_model(self, rec, model):
        g = np.ascontiguousarray(model, np.complex64)
        self._check(spx_batch_set_model(self.h, rec, g.ctypes.data_as(PF)), "set_model")

    def init_model(self, flags, seed):
        self._check(spx_batch_init_model(self.h, flags, seed), "init_model")

    def set_algorithm(self, alg):
        self._keep.append(alg)
        self._check(spx_batch_set_algorithm(self.h, alg), "set_algorithm")

    def set_rec_algorithms(self, algs=None, betas=None):
        a = None if algs is None else np.ascontiguousarray(algs, np.int32)
        b = None if betas is None else np.ascontiguousarray(betas, np.float32)
        # copied to the device before the call returns: nothing to keep alive
        self._check(spx_batch_set_rec_algorithms(self.h, None if a is None else _iptr(a),
                                                 None if b is None else _fptr(b)), "set_rec_algorithms")

    def set_support_schedule(self, sup):
        self._keep.append(sup)
        self._check(spx_batch_set_support_schedule(self.h, sup), "set_support_schedule")

    def use_graphs(self, on=True):
        self._check(spx_batch_use_graphs(self.h, 1 if on else 0), "use_graphs")

    def iterate(self, n):
        self._check(spx_batch_iterate(self.h, n), "iterate")

    @property
    def iteration(self):
        return spx_batch_iteration(self.h)

    def efourier(self):
        e = np.empty(self.nrec, np.float32)
        f = np.empty(self.nrec, np.float32)
        self._check(spx_batch_efourier(self.h, _fptr(e), _fptr(f)), "efourier")
        return e, f

    def efourier_enqueue(self):
        """Queues the Fourier errors behind the queued work; efourier_fetch
        returns them (without waiting once done() is true)."""
        self._check(spx_batch_efourier_enqueue(self.h), "efourier_enqueue")

    def efourier_fetch(self):
        e = np.empty(self.nrec, np.float32)
        f = np.empty(self.nrec, np.float32)
        self._check(spx_batch_efourier_fetch(self.h, _fptr(e), _fptr(f)), "efourier_fetch")
        return e, f

    def ereal(self):
        e = np.empty(self.nrec, np.float32)
        self._check(spx_batch_ereal(self.h, _fptr(e)), "ereal")
        return e

    def support_fraction(self):
        e = np.empty(self.nrec, np.float32)
        self._check(spx_batch_support_fraction(self.h, _fptr(e)), "support_fraction")
        return e

    def model(self, rec=-1):
        n = self.nrec if rec < 0 else 1
        out = np.empty((n,) + self.shape, np.complex64)
        self._check(spx_batch_get_model(self.h, rec, out.ctypes.data_as(PF)), "get_model")
        return out if rec < 0 else out[0]

    def _complex(self, fn, rec, what, staged=False):
        n = self.nrec if rec < 0 else 1
        out = self._staging(np.complex64) if (staged and rec < 0) else None
        if out is None:
            out = np.empty((n,) + self.shape, np.complex64)
        self._check(fn(self.h, rec, out.ctypes.data_as(PF)), what)
        return out if rec < 0 else out[0]

    # staged=True: the result of all reconstructions lands in page-locked
    # staging memory and is returned as a view of it, valid until the next
    # staged call (copy it out before then)
    def model_before_projection(self, rec=-1, staged=False):
        return self._complex(spx_batch_get_model_before_projection, rec, "get_model_before_projection", staged)

    def fmodel(self, rec=-1, staged=False):
        return self._complex(spx_batch_get_fmodel, rec, "get_fmodel", staged)

    def support(self, rec=-1):
        n = self.nrec if rec < 0 else 1
        out = np.empty((n,) + self.shape, np.int32)
        self._check(spx_batch_get_support(self.h, rec, _iptr(out)), "get_support")
        return out if rec < 0 else out[0]

    def time_iterations(self, n):
        ms = C.c_float()
        self._check(spx_batch_time_iterations(self.h, n, C.byref(ms)), "time_iterations")
        return ms.value

    def synchronize(self):
        self._check(spx_batch_synchronize(self.h), "synchronize")

    def set_active(self, active=None):
        """nrec states: 1 active, 0 idle (left unchanged by iterations), 2
        vacant (nothing kept until the slot is set again); None for all
        active."""
        a = None if active is None else np.ascontiguousarray(active, np.int32)
        self._check(spx_batch_set_active(self.h, None if a is None else _iptr(a)), "set_active")

    def init_model_rec(self, rec, flags, seed):
        self._check(spx_batch_init_model_rec(self.h, rec, flags, seed), "init_model_rec")

    def done(self) -> bool:
        """True when the queued work has finished (never blocks)."""
        r = spx_batch_query(self.h)
        if r < 0:
            raise RuntimeError(f"spx_batch_query failed: {last_error()}")
        return r == 1

    def set_amplitudes(self, rec, amplitudes, mask):
        a = np.ascontiguousarray(amplitudes, np.float32)
        m = np.ascontiguousarray(mask, np.int32)
        self._check(spx_batch_set_amplitudes(self.h, rec, _fptr(a), _iptr(m)), "set_amplitudes")

    def set_support(self, rec, support):
        s = np.ascontiguousarray(support, np.float32)
        self._check(spx_batch_set_support(self.h, rec, _fptr(s)), "set_support")


# ---------------------------------------------------------------------------
# numpy helpers (the reference typemaps' orientation)
# ---------------------------------------------------------------------------
def _shape(img: PImage) -> tuple[int, ...]:
    m = img.contents.image.contents
    return (m.y, m.x) if m.z == 1 else (m.z, m.y, m.x)


def image_view(img: PImage) -> np.ndarray:
    """complex64 view of an Image's data, (y, x) or (z, y, x)."""
    m = img.contents.image.contents
    n = m.x * m.y * m.z
    buf = C.cast(m.data, C.POINTER(C.c_float * (2 * n))).contents
    return np.frombuffer(buf, dtype=np.complex64).reshape(_shape(img))


def mask_view(img: PImage) -> np.ndarray:
    m = img.contents.mask.contents
    n = m.x * m.y * m.z
    buf = C.cast(m.data, C.POINTER(C.c_int * n)).contents
    return np.frombuffer(buf, dtype=np.int32).reshape(_shape(img))


def image_from_array(data: np.ndarray, mask: np.ndarray | None = None) -> PImage:
    """New Image holding `data` ((y, x) or (z, y, x)) and `mask`."""
    a = np.asarray(data)
    if a.ndim == 2:
        z, (y, x) = 1, a.shape
    elif a.ndim == 3:
        z, y, x = a.shape
    else:
        raise ValueError("images are 2-D or 3-D")
    img = sp_image_alloc(x, y, z)
    image_view(img)[...] = a.reshape(_shape(img)).astype(np.complex64)
    if mask is not None:
        mask_view(img)[...] = np.asarray(mask).reshape(_shape(img)).astype(np.int32)
    return img


def smap(pairs) -> PSmap:
    """sp_smap from (key, value) pairs."""
    m = sp_smap_alloc(max(1, len(pairs)))
    for k, v in pairs:
        sp_smap_insert(m, float(k), float(v))
    return m


def last_error() -> str:
    return spx_last_error().decode(errors="replace")
