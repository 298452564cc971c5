"""Continuous-batching phasing server over the batched C ABI (include/sp_batch.h).

A request is one reconstruction: its measured amplitudes, mask and static
support, an algorithm (ER, HIO or RAAR with its beta), an iteration count
and the seed of its random-phase start (the reference's SpModelRandomPhases
recipe, src/phasing.c:626-646, drawn from the engine's counter-based
generator).  The result is the final model and its Fourier error / FcFo
(src/phasing.c:855-893).

Scheduling:

* size classes -- requests of one array shape share SpxBatch objects (one
  launch per pass advances every reconstruction of a batch); different
  shapes never mix, so no reconstruction is padded or resized and each
  result equals the same request run alone through sp_phaser_iterate;
* continuous batching -- a batch advances in quanta of `quantum`
  iterations.  Between quanta finished reconstructions
  leave and waiting requests of the class take their slots; only when more
  requests wait than the class has free slots (in running batches too,
  which take requests at the end of their quantum) is a further batch
  created, with at least `min_slots` slots, so that requests arriving one by
  one share launches instead of each opening a batch of one;
* admission by free device memory -- a batch is created only if its
  estimated footprint (spx_batch_estimate_bytes) fits into the free device
  memory minus a reserve; otherwise idle batches of other classes are
  released first and the requests wait;
* free slots cost nothing: the batch's activity table (spx_batch_set_active)
  marks them vacant, so the pass kernels skip their work-groups without
  carrying a stale iterate over;
* every batch owns a HIP stream, so quanta of different classes overlap on
  the device; the host loop only polls (spx_batch_query) and never waits
  for one class while another could be fed: the Fourier errors of the
  reconstructions that finish with a quantum are queued behind it
  (spx_batch_efourier_enqueue) and read once it is done;
* batches live until memory is needed: device memory is freed only when an
  admission does not fit (freeing synchronises the whole device, which
  would stall every other class), unless `idle_release` seconds are given.

Request iteration counts must be multiples of `quantum`.
"""
from __future__ import annotations

import ctypes as C
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import spimage as sp

ALGORITHMS = {"er": sp.SpER, "hio": sp.SpHIO, "raar": sp.SpRAAR}


@dataclass
class Request:
    rid: int
    amplitudes: np.ndarray       # (y, x) or (z, y, x) float32
    mask: np.ndarray             # same shape, nonzero = measured
    support: np.ndarray          # same shape, nonzero = inside the support
    algorithm: str               # "er" | "hio" | "raar"
    iterations: int
    seed: int
    beta: float = 0.9
    arrival: float = 0.0         # seconds after the start of the run


@dataclass
class Result:
    rid: int
    efourier: float
    fcfo: float
    arrival: float
    admitted: float
    finished: float
    model: np.ndarray | None = None

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


@dataclass
class _Slot:
    req: Request
    admitted: float
    done: int = 0


@dataclass
class _Batch:
    batch: sp.Batch
    slots: list = field(default_factory=list)  # _Slot | None per reconstruction
    busy: bool = False
    idle_since: float = 0.0
    algs: np.ndarray | None = None
    betas: np.ndarray | None = None


class Server:
    def __init__(self, quantum: int = 50, max_pixels_per_batch: int = 1 << 24, reserve_bytes: int = 8 << 30,
                 keep_models: bool = False, idle_release: float | None = None, min_slots: int = 4):
        self.quantum = quantum
        self.min_slots = max(1, min_slots)
        self.max_pixels = max_pixels_per_batch
        self.reserve = reserve_bytes
        self.keep_models = keep_models
        self.idle_release = idle_release
        self.classes: dict[tuple, list[_Batch]] = {}
        self.queues: dict[tuple, deque] = {}
        self.results: list[Result] = []
        # the batch's base algorithm only fixes the constraints (none); the
        # per-reconstruction tables choose ER/HIO/RAAR and beta
        self._beta_map = sp.smap([(0, 0.9)])
        self._base = sp.sp_phasing_hio_alloc(self._beta_map, 0)
        self.batches_created = 0

    def prepare(self, shapes) -> None:
        """Runs the tuner for every shape at each batch size this server may
        create (and for one reconstruction), so that no timing run of the
        tuner falls into serving."""
        for shape in shapes:
            cap_max = max(1, self.max_pixels // int(np.prod(shape)))
            cap = 1
            while True:
                sp.tune_shape(shape, cap)
                if cap >= cap_max:
                    break
                cap = min(cap * 2, cap_max)

    # -- admission -----------------------------------------------------------
    def _capacity(self, shape, waiting: int) -> int:
        npix = int(np.prod(shape))
        cap_max = max(1, self.max_pixels // npix)
        cap = 1
        while (cap < waiting or cap < self.min_slots) and cap < cap_max:
            cap *= 2
        cap = min(cap, cap_max)
        free, _ = sp.device_memory()
        dims = (shape[-1], shape[-2], shape[0] if len(shape) == 3 else 1)
        while cap >= 1 and sp.spx_batch_estimate_bytes(*dims, cap) > free - self.reserve:
            cap //= 2
        return cap

    def _release_idle(self, keep_shape=None, older_than: float | None = None) -> bool:
        released = False
        now = time.perf_counter()
        for shape, blist in self.classes.items():
            if shape == keep_shape:
                continue
            for bt in list(blist):
                idle = not bt.busy and all(s is None for s in bt.slots)
                if idle and (older_than is None or now - bt.idle_since >= older_than):
                    bt.batch.close()
                    blist.remove(bt)
                    released = True
        return released

    def _new_batch(self, shape, waiting: int) -> _Batch | None:
        cap = self._capacity(shape, waiting)
        if cap < 1 and self._release_idle(keep_shape=shape):
            cap = self._capacity(shape, waiting)
        if cap < 1:
            return None
        b = sp.Batch(shape, cap)
        b.set_algorithm(self._base)
        b.use_graphs(False)
        bt = _Batch(b, [None] * cap, algs=np.full(cap, sp.SpER, np.int32), betas=np.zeros(cap, np.float32))
        self.classes.setdefault(shape, []).append(bt)
        self.batches_created += 1
        return bt

    def _fill(self, bt: _Batch, q: deque, now: float) -> bool:
        changed = False
        for r, s in enumerate(bt.slots):
            if not q:
                break
            if s is not None:
                continue
            req = q.popleft()
            b = bt.batch
            b.set_data(r, req.amplitudes, req.mask, req.support)
            b.init_model_rec(r, sp.SpModelRandomPhases, req.seed)
            bt.algs[r] = ALGORITHMS[req.algorithm]
            bt.betas[r] = req.beta
            bt.slots[r] = _Slot(req, now)
            changed = True
        if changed:
            bt.batch.set_rec_algorithms(bt.algs, bt.betas)
            self._sync_active(bt)
        return changed

    @staticmethod
    def _sync_active(bt: _Batch):
        # free slots are vacant: the pass kernels skip them without carrying
        # their stale iterate over (a request that takes the slot sets it
        # afresh)
        bt.batch.set_active([2 if s is None else 1 for s in bt.slots])

    # -- progress ------------------------------------------------------------
    def _launch(self, bt: _Batch):
        """Queues the batch's next quantum (and, if a reconstruction finishes
        with it, the Fourier errors behind it)."""
        bt.batch.iterate(self.quantum)
        if any(s is not None and s.done + self.quantum >= s.req.iterations for s in bt.slots):
            bt.batch.efourier_enqueue()
        bt.busy = True

    def _finish_quantum(self, bt: _Batch, now_fn):
        for s in bt.slots:
            if s is not None:
                s.done += self.quantum
        fin = [r for r, s in enumerate(bt.slots) if s is not None and s.done >= s.req.iterations]
        if not fin:
            return
        ef, fc = bt.batch.efourier_fetch()
        t = now_fn()
        for r in fin:
            s = bt.slots[r]
            model = bt.batch.model(r) if self.keep_models else None
            self.results.append(Result(s.req.rid, float(ef[r]), float(fc[r]), s.req.arrival, s.admitted, t, model))
            bt.slots[r] = None
        self._sync_active(bt)
        if all(s is None for s in bt.slots):
            bt.idle_since = time.perf_counter()

    def run(self, requests: list[Request], realtime: bool = True) -> list[Result]:
        """Serves `requests` (arrival times relative to the call when
        `realtime`, otherwise all at once) and returns their results in
        completion order."""
        for r in requests:
            if r.iterations % self.quantum:
                raise ValueError(f"request {r.rid}: iterations must be a multiple of {self.quantum}")
            if r.algorithm not in ALGORITHMS:
                raise ValueError(f"request {r.rid}: unknown algorithm {r.algorithm!r}")
            if not all(n == 1 or sp.spx_supported_length(int(n)) for n in r.amplitudes.shape):
                raise ValueError(f"request {r.rid}: shape {r.amplitudes.shape} has an axis without a Stockham plan; "
                                 "batches do not take such shapes (PlainRunner runs them)")
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        t0 = time.perf_counter()

        def now():
            return time.perf_counter() - t0

        n_total = len(requests) + len(self.results)
        while len(self.results) < n_total:
            progressed = False
            t = now()
            while pending and (not realtime or pending[0].arrival <= t):
                req = pending.popleft()
                self.queues.setdefault(tuple(req.amplitudes.shape), deque()).append(req)
            # quanta that have finished
            for blist in self.classes.values():
                for bt in blist:
                    if bt.busy and bt.batch.done():
                        bt.busy = False
                        self._finish_quantum(bt, now)
                        progressed = True
            # admission into free slots, then new batches for what still
            # waits; a filled batch starts at once, so that the device works
            # while the host fills the next one
            for shape, q in self.queues.items():
                if not q:
                    continue
                for bt in self.classes.get(shape, []):
                    if not bt.busy and self._fill(bt, q, now()):
                        self._launch(bt)
                        progressed = True
                while q:
                    # free slots of running batches take requests when their quantum ends
                    free = sum(s is None for bt in self.classes.get(shape, []) for s in bt.slots)
                    if free >= len(q):
                        break
                    bt = self._new_batch(shape, len(q) - free)
                    if bt is None:
                        break
                    self._fill(bt, q, now())
                    self._launch(bt)
                    progressed = True
            # launch the next quantum of every idle batch with work
            for blist in self.classes.values():
                for bt in blist:
                    if not bt.busy and any(s is not None for s in bt.slots):
                        self._launch(bt)
                        progressed = True
            if self.idle_release is not None:
                self._release_idle(older_than=self.idle_release)
            if not progressed:
                busy = any(bt.busy for blist in self.classes.values() for bt in blist)
                stuck = [s for s, q in self.queues.items() if q and not self.classes.get(s)]
                if stuck and not busy and not pending:
                    raise MemoryError(f"requests of shape {stuck[0]} do not fit into free device memory "
                                      f"(reserve {self.reserve} bytes)")
                time.sleep(50e-6)
        return self.results

    def close(self):
        for blist in self.classes.values():
            for bt in blist:
                bt.batch.close()
        self.classes.clear()


class PlainRunner:
    """The reference-style loop: one SpPhaser per request (sp_phaser_alloc /
    init / set_amplitudes / init_model / init_support / iterate / efourier /
    free), nothing shared between requests but the algorithm objects."""

    def __init__(self):
        self._maps = {}
        self._algs = {}

    def _alg(self, req: Request):
        key = (req.algorithm, req.beta)
        if key not in self._algs:
            if req.algorithm == "er":
                self._algs[key] = sp.sp_phasing_er_alloc(0)
            else:
                m = sp.smap([(0, req.beta)])
                self._maps[key] = m
                alloc = sp.sp_phasing_hio_alloc if req.algorithm == "hio" else sp.sp_phasing_raar_alloc
                self._algs[key] = alloc(m, 0)
        return self._algs[key]

    def run_one(self, req: Request, keep_model: bool = False):
        """(efourier, fcfo, model or None) of `req` run alone."""
        ph = sp.sp_phaser_alloc()
        if sp.sp_phaser_init(ph, self._alg(req), None, sp.SpEngineHIP) != 0:
            raise RuntimeError(sp.last_error())
        imgs = [sp.image_from_array(req.amplitudes, req.mask), sp.image_from_array(req.support)]
        try:
            sp.sp_phaser_set_amplitudes(ph, imgs[0])
            sp.spx_phaser_set_seed(ph, req.seed)
            if sp.sp_phaser_init_model(ph, None, sp.SpModelRandomPhases) != 0:
                raise RuntimeError(sp.last_error())
            if sp.sp_phaser_init_support(ph, imgs[1], 0, 0.0) != 0:
                raise RuntimeError(sp.last_error())
            if sp.sp_phaser_iterate(ph, req.iterations) != 0:
                raise RuntimeError(sp.last_error())
            fc = C.c_float()
            ef = sp.sp_phaser_efourier(ph, C.byref(fc))
            model = sp.image_view(sp.sp_phaser_model(ph)).copy() if keep_model else None
        finally:
            sp.sp_phaser_free(ph)
            for im in imgs:
                sp.sp_image_free(im)
        return float(ef), float(fc.value), model

    def run(self, requests: list[Request]) -> tuple[list[Result], float]:
        """Runs every request in arrival order.  Returns the results on a
        simulated FIFO timeline (start = max(arrival, previous finish), with
        the measured service times) and the total service seconds."""
        out, busy_until, total = [], 0.0, 0.0
        for r in sorted(requests, key=lambda r: r.arrival):
            t0 = time.perf_counter()
            ef, fc, _ = self.run_one(r)
            dt = time.perf_counter() - t0
            total += dt
            start = max(r.arrival, busy_until)
            busy_until = start + dt
            out.append(Result(r.rid, ef, fc, r.arrival, start, busy_until))
        return out, total


def latency_summary(results: list[Result]) -> dict:
    lat = np.array([r.latency for r in results])
    return {"p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "max_s": float(lat.max())}
