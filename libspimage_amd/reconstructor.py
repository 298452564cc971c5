"""Reconstructor front-end over the HIP engine.

The public interface is that of the reference's ``spimage.Reconstructor``
(reference src/_spimage_reconstructor.py:9-740): the same setters, the same
keyword arguments, the same output dictionaries.  ``reconstruct()`` drives
one SpPhaser through the phasing API exactly like the reference
(src/_spimage_reconstructor.py:553-675).  ``reconstruct_loop(n)`` does not
repeat that n times (src/_spimage_reconstructor.py:677-740): it advances all
n random starts together in one SpxBatch (include/sp_batch.h), one launch per
pass for all of them, with the same schedule of algorithm and support
changes and the same outputs.  Reconstruction r of the loop equals
``reconstruct(seed=seed + r)`` bit for bit.

Differences from the reference:

* ``seed`` (optional, both calls) fixes the random-phase starts;
* the ``threshold`` support algorithm uses its own ``threshold_init`` /
  ``threshold_final`` schedule (the reference reads undefined attributes
  there, src/_spimage_reconstructor.py:473, so only ``area`` and ``static``
  work in it);
* arrays of ``reconstruct_loop`` are (n, y, x) / (n, z, y, x) also for
  non-square shapes (the reference allocates (n, x, y),
  src/_spimage_reconstructor.py:683-685).

``center_image=True`` appends the centre-image update to the support
algorithm's array, as the reference does (src/_spimage_reconstructor.py:
480-484); it runs on the device after every support update, for every
reconstruction of a loop.
"""
from __future__ import annotations

import ctypes as C
import logging
import os

import numpy as np

from . import spimage as sp

logger = logging.getLogger("RECONSTRUCTOR")
logger.setLevel("WARNING")

_SUPPORT_KWARGS = {"area": ["update_period", "blur_init", "blur_final", "area_init", "area_final"],
                   "threshold": ["update_period", "blur_init", "blur_final", "threshold_init", "threshold_final"],
                   "static": []}
_PHASING_KWARGS = {"raar": ["beta_init", "beta_final"], "hio": ["beta_init", "beta_final"], "er": [],
                   "diffmap": ["beta_init", "beta_final", "gamma1", "gamma2"]}



def _shift_into(dst, src, axes):
    """dst[...] = np.fft.fftshift(src, axes) in one pass: the shift (a roll by
    n // 2 on each axis) as block copies, converting to dst's type on the way."""
    import itertools
    blocks = []
    for ax in axes:
        n = src.shape[ax]
        h = n // 2
        # (destination slice, source slice) of the two halves of the roll
        blocks.append(((slice(h, n), slice(0, n - h)), (slice(0, h), slice(n - h, n))))
    for combo in itertools.product(*blocks):
        d = [slice(None)] * src.ndim
        s_ = [slice(None)] * src.ndim
        for ax, (ds, ss) in zip(axes, combo):
            d[ax], s_[ax] = ds, ss
        dst[tuple(d)] = src[tuple(s_)]

_POOL = None


def _pool():
    global _POOL
    if _POOL is None:
        from concurrent.futures import ThreadPoolExecutor
        _POOL = ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1), thread_name_prefix="spx-outputs")
    return _POOL


def _shift_into_recs(dst, src, axes):
    """_shift_into for arrays of many reconstructions (axis 0), split across
    threads by reconstruction: numpy releases the interpreter lock in its
    copy and conversion loops, so the outputs of a batch are assembled on
    several cores."""
    n = src.shape[0]
    k = min(n, _pool()._max_workers)
    if k < 2 or src.size < (1 << 20):
        _shift_into(dst, src, axes)
        return
    bounds = [n * i // k for i in range(k + 1)]
    futs = [_pool().submit(_shift_into, dst[a:b], src[a:b], axes) for a, b in zip(bounds, bounds[1:])]
    for f in futs:
        f.result()


class Reconstructor:
    def __init__(self, use_gpu=True):
        if not use_gpu:
            raise ValueError("this engine phases on the GPU only (SpEngineCPU is not provided)")
        self._keep = []
        self.clear()

    # -- configuration (reference src/_spimage_reconstructor.py:22-311) -------
    def clear(self):
        self._amplitudes = None
        self._mask = None
        self._initial_support_config = None
        self._number_of_iterations = None
        self._number_of_outputs_images = None
        self._number_of_outputs_scores = None
        self._support_configs = []
        self._phasing_configs = []
        self._i_support = None
        self._i_phasing = None
        self._keep = []

    def set_intensities(self, intensities, shifted=False):
        A = np.array(intensities, dtype="float")
        if (A < 0).sum() > 0:
            logger.info("Intensities found that have negative values. These values are set to zero.")
            A[A < 0] = 0.
        A = np.sqrt(A)
        self._amplitudes = A if shifted else np.fft.fftshift(A)

    def set_mask(self, mask, shifted=False):
        M = np.array(mask, dtype="bool")
        self._mask = M if shifted else np.fft.fftshift(M)

    def set_number_of_iterations(self, number_of_iterations):
        self._number_of_iterations = number_of_iterations

    def set_number_of_outputs_images(self, number_of_outputs_images):
        self._number_of_outputs_images = number_of_outputs_images

    def set_number_of_outputs_scores(self, number_of_outputs_scores):
        self._number_of_outputs_scores = number_of_outputs_scores

    def set_number_of_outputs(self, number_of_outputs):
        self._number_of_outputs_images = number_of_outputs
        self._number_of_outputs_scores = number_of_outputs

    def set_initial_support(self, **kwargs):
        for k in ("radius", "area", "support_mask"):
            if k in kwargs:
                self._initial_support_config = {k: kwargs[k]}
                return
        logger.error("set_initial_support requires one of the following key word arguments: radius, area, support_mask")

    def set_support_algorithm(self, type, number_of_iterations=None, center_image=False, **kwargs):
        self._support_configs = []
        self._i_support = None
        self.append_support_algorithm(type, number_of_iterations=number_of_iterations, center_image=center_image,
                                      **kwargs)

    def append_support_algorithm(self, type, number_of_iterations, center_image=False, **kwargs):
        if number_of_iterations is None and self._support_configs:
            logger.error("Only one support algorithm may leave number_of_iterations unset.")
            return
        if type not in _SUPPORT_KWARGS:
            logger.error("append_support_algorithm accepts algorithms of the following types: %s",
                         list(_SUPPORT_KWARGS))
            return
        conf = {"type": type, "number_of_iterations": number_of_iterations, "center_image": bool(center_image)}
        for k in _SUPPORT_KWARGS[type]:
            if k not in kwargs:
                logger.error("support algorithm %s requires the argument %s", type, k)
                return
            conf[k] = kwargs[k]
        if number_of_iterations is not None:
            self._i_support = (self._i_support or 0) + number_of_iterations
        self._support_configs.append(conf)

    def set_phasing_algorithm(self, type, **kwargs):
        self._phasing_configs = []
        self._i_phasing = None
        kw = dict(kwargs)
        kw.setdefault("number_of_iterations", None)
        self.append_phasing_algorithm(type, **kw)

    def append_phasing_algorithm(self, type, **kwargs):
        if "number_of_iterations" not in kwargs:
            logger.error("append_phasing_algorithm requires the keyword argument 'number_of_iterations'.")
            return
        n = kwargs["number_of_iterations"]
        if n is None and self._phasing_configs:
            logger.error("Only one phasing algorithm may leave number_of_iterations unset.")
            return
        if type not in _PHASING_KWARGS:
            logger.error("append_phasing_algorithm accepts algorithms of the following types: %s",
                         list(_PHASING_KWARGS))
            return
        for k in _PHASING_KWARGS[type]:
            if k not in kwargs:
                logger.error("phasing algorithm %s requires the argument %s", type, k)
                return
        conf = dict(kwargs)
        conf["type"] = type
        conf.setdefault("constraints", "")
        if n is not None:
            self._i_phasing = (self._i_phasing or 0) + n
        self._phasing_configs.append(conf)

    # -- set-up (reference src/_spimage_reconstructor.py:313-551) ------------
    def _ready(self):
        ok = True
        if self._amplitudes is None:
            logger.error("Reconstruction cannot start! You need to set the intensities.")
            ok = False
        if self._initial_support_config is None:
            logger.error("Reconstruction cannot start! You need to set the initial support.")
            ok = False
        if not self._support_configs or not self._phasing_configs:
            logger.error("Reconstruction cannot start! You need to set the support and phasing algorithms.")
            ok = False
        if self._number_of_outputs_images is None or self._number_of_outputs_scores is None:
            logger.error("Cannot prepare reconstruction. Number of outputs need to be set.")
            ok = False
        for i in (self._i_support, self._i_phasing):
            if i is not None and i != self._number_of_iterations:
                logger.error("Cannot prepare reconstruction. Algorithms are not in line with the number of iterations.")
                ok = False
        return ok

    def _prepare(self):
        A = np.ascontiguousarray(self._amplitudes, np.float32)
        M = np.ones(A.shape, bool) if self._mask is None else self._mask
        shape = A.shape
        cfg = self._initial_support_config
        if "support_mask" in cfg:
            S = np.asarray(cfg["support_mask"], np.float32)
        else:
            axes = np.meshgrid(*[np.arange(n) - (n - 1) / 2. for n in shape], indexing="ij")
            R = np.sqrt(sum(a * a for a in axes))
            if "radius" in cfg:
                r = cfg["radius"]
            else:
                r = np.sqrt(cfg["area"] * np.prod(shape) / np.pi)
            S = np.float32(np.fft.fftshift(R) < r)
        self._keep = []
        total = self._number_of_iterations
        supports, i = [], 0
        for conf in self._support_configs:
            n = conf["number_of_iterations"] if conf["number_of_iterations"] is not None else total
            if conf["type"] in ("area", "threshold"):
                blur = sp.smap([(i, conf["blur_init"]), (i + n, conf["blur_final"])])
                key = "area" if conf["type"] == "area" else "threshold"
                val = sp.smap([(i, conf[key + "_init"]), (i + n, conf[key + "_final"])])
                self._keep += [blur, val]
                alloc = sp.sp_support_area_alloc if conf["type"] == "area" else sp.sp_support_threshold_alloc
                arr = sp.sp_support_array_init(alloc(blur, val), conf["update_period"])
            else:
                arr = sp.sp_support_array_init(sp.sp_support_static_alloc(), n)
            if conf.get("center_image"):
                sp.sp_support_array_append(arr, sp.sp_support_centre_image_alloc())
            supports.append((n, arr))
            i += n
        phasings, i = [], 0
        for conf in self._phasing_configs:
            n = conf["number_of_iterations"] if conf["number_of_iterations"] is not None else total
            c = sp.SpNoConstraints
            cons = conf.get("constraints", "")
            if "enforce_positivity" in cons and "enforce_real" in cons:
                c |= sp.SpPositiveRealObject
            elif "enforce_real" in cons:
                c |= sp.SpRealObject
            elif "enforce_positivity" in cons:
                c |= sp.SpPositiveComplexObject
            if "enforce_centrosymmetry" in cons:
                c |= sp.SpCentrosymmetricObject
            t = conf["type"]
            if t in ("raar", "hio", "diffmap"):
                beta = sp.smap([(i, conf["beta_init"]), (i + n, conf["beta_final"])])
                self._keep.append(beta)
            if t == "raar":
                alg = sp.sp_phasing_raar_alloc(beta, c)
            elif t == "hio":
                alg = sp.sp_phasing_hio_alloc(beta, c)
            elif t == "diffmap":
                alg = sp.sp_phasing_diff_map_alloc(beta, conf["gamma1"], conf["gamma2"], c)
            else:
                alg = sp.sp_phasing_er_alloc(c)
            phasings.append((n, alg))
            i += n
        self._keep += [supports, phasings]
        return A, M, S, supports, phasings

    def _schedule(self, supports, phasings):
        """(iteration, change_algorithm, change_support, out_image, out_score)
        at every point where the reference's loop stops
        (src/_spimage_reconstructor.py:578-611)."""
        total = self._number_of_iterations
        out_img = set(np.int64(np.linspace(0, total, self._number_of_outputs_images).round()).tolist())
        out_sco = set(np.int64(np.linspace(0, total, self._number_of_outputs_scores).round()).tolist())
        ia = isup = 0
        a0 = s0 = 0
        for it in range(total + 1):
            ca = it - a0 == phasings[ia][0] and it != total
            cs = it - s0 == supports[isup][0] and it != total
            if ca or cs or it in out_img or it in out_sco:
                yield it, (ia + 1 if ca else None), (isup + 1 if cs else None), it in out_img, it in out_sco
            if ca:
                ia, a0 = ia + 1, it
            if cs:
                isup, s0 = isup + 1, it

    def _index_arrays(self):
        total = self._number_of_iterations
        return (np.int64(np.linspace(0, total, self._number_of_outputs_images).round()),
                np.int64(np.linspace(0, total, self._number_of_outputs_scores).round()))

    # -- one reconstruction (reference src/_spimage_reconstructor.py:553-675) -
    def reconstruct(self, seed=None):
        """One reconstruction from random starting phases; returns the
        reference's output dictionary."""
        if not self._ready():
            return None
        A, M, S, supports, phasings = self._prepare()
        shape = A.shape
        ph = sp.sp_phaser_alloc()
        imgs = [sp.image_from_array(A, M.astype(np.int32)), sp.image_from_array(S)]
        try:
            if sp.sp_phaser_init(ph, phasings[0][1], supports[0][1], sp.SpEngineCUDA) != 0:
                raise RuntimeError(f"sp_phaser_init failed: {sp.last_error()}")
            sp.sp_phaser_set_amplitudes(ph, imgs[0])
            if seed is not None:
                sp.spx_phaser_set_seed(ph, int(seed))
            if sp.sp_phaser_init_model(ph, None, sp.SpModelRandomPhases) != 0:
                raise RuntimeError(sp.last_error())
            if sp.sp_phaser_init_support(ph, imgs[1], 0, 0.0) != 0:
                raise RuntimeError(sp.last_error())
            n_img, n_sco = self._number_of_outputs_images, self._number_of_outputs_scores
            real_space = np.zeros((n_img,) + shape, np.complex128)
            support = np.zeros((n_img,) + shape, bool)
            fourier_space = np.zeros((n_img,) + shape, np.complex128)
            mask = np.zeros((n_img,) + shape, bool)
            ferr, rerr, ssize = np.zeros(n_sco), np.zeros(n_sco), np.zeros(n_sco)
            done = k_img = k_sco = 0
            for it, ca, cs, oi, os_ in self._schedule(supports, phasings):
                if it:
                    if sp.sp_phaser_iterate(ph, it - done) != 0:
                        raise RuntimeError(sp.last_error())
                    done = it
                if ca is not None:
                    ph.contents.algorithm = phasings[ca][1]
                if cs is not None:
                    ph.contents.sup_algorithm = supports[cs][1]
                if oi:
                    if it > 0:
                        img = sp.image_view(sp.sp_phaser_model_before_projection(ph))
                        msk = sp.image_view(sp.sp_phaser_support(ph)).real != 0
                    else:
                        img = sp.image_view(sp.sp_phaser_model(ph))
                        msk = S != 0
                    real_space[k_img] = np.fft.fftshift(img)
                    support[k_img] = np.fft.fftshift(msk)
                    fourier_space[k_img] = np.fft.fftshift(sp.image_view(sp.sp_phaser_fmodel(ph)))
                    mask[k_img] = np.fft.fftshift(sp.mask_view(sp.sp_phaser_amplitudes(ph)) != 0)
                    k_img += 1
                if os_:
                    fc = C.c_float()
                    ferr[k_sco] = sp.sp_phaser_efourier(ph, C.byref(fc))
                    rerr[k_sco] = sp.sp_phaser_ereal(ph)
                    ssize[k_sco] = sp.sp_phaser_support_fraction(ph)
                    k_sco += 1
        finally:
            sp.sp_phaser_free(ph)
            for im in imgs:
                sp.sp_image_free(im)
        ii, isc = self._index_arrays()
        return {"iteration_index_images": ii, "iteration_index_scores": isc, "real_space": real_space,
                "support": support, "fourier_space": fourier_space, "mask": mask, "real_error": rerr,
                "fourier_error": ferr, "support_size": ssize}

    # -- many reconstructions at once (reference src/_spimage_reconstructor.py:677-740)
    def _loop_of_singles(self, n, seed, full_output):
        """reconstruct_loop's output from n single reconstructions, seeds
        seed .. seed + n - 1."""
        singles = [self.reconstruct(seed=seed + r) for r in range(n)]
        out = {"real_space_final": np.stack([o["real_space"][-1] for o in singles]),
               "fourier_space_final": np.stack([o["fourier_space"][-1] for o in singles]),
               "support_final": np.stack([o["support"][-1] for o in singles]),
               "scores_final": {k: np.array([o[k][-1] for o in singles])
                                for k in ("real_error", "fourier_error", "support_size")}}
        if full_output:
            out["single_outputs"] = singles
        return out

    def reconstruct_loop(self, Nrepeats, full_output=False, seed=None):
        """Nrepeats reconstructions from independent random starts, advanced
        together on the GPU in one batch (one after another when an axis
        length has no Stockham plan and runs through Bluestein transforms);
        reconstruction r equals reconstruct(seed=seed + r)."""
        if not self._ready():
            return None
        A, M, S, supports, phasings = self._prepare()
        shape = A.shape
        n = int(Nrepeats)
        if seed is None:
            seed = int(np.random.randint(0, 2**31 - 1))
        if not all(L == 1 or sp.spx_supported_length(int(L)) for L in shape):
            # an axis without a Stockham plan runs through Bluestein transforms,
            # which batches do not take: the starts run one after another
            return self._loop_of_singles(n, int(seed), full_output)
        b = sp.Batch(shape, n)
        try:
            for r in range(n):
                b.set_data(r, A, M.astype(np.int32), S)
            b.init_model(sp.SpModelRandomPhases, int(seed))
            b.set_algorithm(phasings[0][1])
            b.set_support_schedule(supports[0][1])
            n_img, n_sco = self._number_of_outputs_images, self._number_of_outputs_scores
            real_space = np.zeros((n, n_img) + shape, np.complex128)
            support = np.zeros((n, n_img) + shape, bool)
            fourier_space = np.zeros((n, n_img) + shape, np.complex128)
            amp_mask = np.fft.fftshift(M)
            ferr, rerr, ssize = np.zeros((n, n_sco)), np.zeros((n, n_sco)), np.zeros((n, n_sco))
            axes = tuple(range(1, 1 + len(shape)))
            done = k_img = k_sco = 0
            for it, ca, cs, oi, os_ in self._schedule(supports, phasings):
                if it:
                    b.iterate(it - done)
                    done = it
                if ca is not None:
                    b.set_algorithm(phasings[ca][1])
                if cs is not None:
                    b.set_support_schedule(supports[cs][1])
                if oi:
                    img = b.model_before_projection(staged=True) if it > 0 else b.model()
                    _shift_into_recs(real_space[:, k_img], img, axes)
                    msk = b.support() != 0 if it > 0 else np.broadcast_to(S != 0, (n,) + shape)
                    _shift_into_recs(support[:, k_img], msk, axes)
                    _shift_into_recs(fourier_space[:, k_img], b.fmodel(staged=True), axes)
                    k_img += 1
                if os_:
                    ef, _ = b.efourier()
                    ferr[:, k_sco] = ef
                    rerr[:, k_sco] = b.ereal()
                    ssize[:, k_sco] = b.support_fraction()
                    k_sco += 1
        finally:
            b.close()
        out = {"real_space_final": real_space[:, -1], "fourier_space_final": fourier_space[:, -1],
               "support_final": support[:, -1],
               "scores_final": {"real_error": rerr[:, -1], "fourier_error": ferr[:, -1],
                                "support_size": ssize[:, -1]}}
        if full_output:
            ii, isc = self._index_arrays()
            out["single_outputs"] = [
                {"iteration_index_images": ii, "iteration_index_scores": isc, "real_space": real_space[r],
                 "support": support[r], "fourier_space": fourier_space[r],
                 "mask": np.broadcast_to(amp_mask, (n_img,) + shape).copy(), "real_error": rerr[r],
                 "fourier_error": ferr[r], "support_size": ssize[r]} for r in range(n)]
        return out
