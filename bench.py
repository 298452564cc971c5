#!/usr/bin/env python3
"""Benchmark of the HIP phasing engine on one MI355X.

Headline: HIO iterations per second on a 1024 x 1024 complex64 pattern (the
seeded blob pattern of libspimage_amd.workloads.blob_1024, static support,
beta = 0.9), inputs resident in HBM, timed with HIP events on the engine's
stream.  A step is one HIO iteration: the fused Fourier-space pass (column
FFT, modulus projection, inverse column FFT) and the fused real-space pass
(inverse row FFT, 1/N, HIO update, support constraint, forward row FFT of the
next iterate).

Prints one JSON line with the roofline of the dominant pass (algorithmic
bytes from the 78 N bytes of a 2-D HIO iteration: 41 N for the real-space
pass, 37 N for the Fourier-space pass), the whole-iteration roofline, the
1024^2 RAAR + shrinkwrap reconstruction (300 iterations, area support every
20), the batched workload (128 patterns x 8 starts), the served workload
(Poisson 200 req/s), the 256^3 DM + ER reconstruction, the Reconstructor
front-end's reconstruct_loop, the PRTF over many reconstructions, the
iteration time of 2-D shapes with mixed-radix lengths, and a CPU baseline:
the oracle's restatement of the reference loop on one host core.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
                    [--no-batched] [--no-serving] [--no-volume]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import subprocess
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

PEAK_HBM = 8.0e12          # MI355X HBM3E, bytes/s
BYTES_HIO = 78             # algorithmic bytes per pixel of one 2-D HIO iteration
BYTES_REAL_PASS = 41       # inverse + forward x passes (32) + g0 (8) + support mask (1)
BYTES_FOURIER_PASS = 37    # forward + inverse y passes (32) + amplitudes (4) + measured mask (1)
SAMPLES = 5               # headline: median of this many timed samples
MKL_CANDIDATES = ["/opt/conda/lib/libmkl_rt.so", "/opt/conda/lib/libmkl_rt.so.2", "/usr/lib/x86_64-linux-gnu/libmkl_rt.so"]


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--warmup", type=int, default=300)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the many-reconstruction workloads (batched, Reconstructor, PRTF, mixed-radix shapes)")
    ap.add_argument("--no-serving", action="store_true", help="skip the served workload")
    ap.add_argument("--no-volume", action="store_true", help="skip the 256^3 reconstruction")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--cpu-baseline-only", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--cpu-spec", default="{}", help=argparse.SUPPRESS)
    ap.add_argument("--only-headline", action="store_true",
                    help="only the 1024^2 HIO timing (a short program for profiler runs)")
    ap.add_argument("--only-batched", action="store_true",
                    help="only the many-reconstruction batched workload (a short program for profiler runs)")
    ap.add_argument("--only-volume", action="store_true",
                    help="only the 256^3 reconstruction (a short program for profiler runs)")
    ap.add_argument("--only-serving-capacity", action="store_true",
                    help="only the all-queued run of the served workload (a short program for kernel traces)")
    ap.add_argument("--trace-summary", metavar="CSV",
                    help="summarise a rocprofv3 kernel trace (busy and idle time, overlap, top kernels); no GPU")
    ap.add_argument("--trace-window-s", type=float, default=None,
                    help="with --trace-summary: only the kernels of the trace's last SECONDS")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only launch 20 plain 64 MiB device-to-device copies (counter calibration)")
    return ap.parse_args()


def _profile_bytes(name):
    """A committed counter summary under profiles/ (HBM bytes rocprof counted
    per pixel or voxel of a workload's passes), or {} when absent."""
    try:
        return json.loads((ROOT / "profiles" / name).read_text())
    except (OSError, ValueError):
        return {}


def make_hio_phaser(sp, pattern, beta=0.9, alg="hio", sup_array=None):
    keep = []
    bmap = sp.smap([(0, beta)])
    keep.append(bmap)
    a = sp.sp_phasing_hio_alloc(bmap, 0) if alg == "hio" else sp.sp_phasing_raar_alloc(bmap, 0)
    ph = sp.sp_phaser_alloc()
    rc = sp.sp_phaser_init(ph, a, sup_array, sp.SpEngineHIP)
    if rc != 0:
        raise RuntimeError(f"sp_phaser_init failed ({rc}): {sp.last_error()}")
    imgs = [sp.image_from_array(pattern.amplitudes, pattern.mask),
            sp.image_from_array(pattern.start),
            sp.image_from_array(pattern.support)]
    sp.sp_phaser_set_amplitudes(ph, imgs[0])
    if sp.sp_phaser_init_model(ph, imgs[1], 0) != 0:
        raise RuntimeError(sp.last_error())
    if sp.sp_phaser_init_support(ph, imgs[2], 0, 0.0) != 0:
        raise RuntimeError(sp.last_error())
    return ph, (a, imgs, keep)


def gpu_one_step(sp, kind, beta, amp, mask, support, model):
    """One GPU iteration of `kind` from `model` with a fresh phaser (no
    support updates): the state after it, for the accuracy probes."""
    import math
    bmap = sp.smap([(0, beta)])
    if kind == "hio":
        a = sp.sp_phasing_hio_alloc(bmap, 0)
    elif kind == "raar":
        a = sp.sp_phasing_raar_alloc(bmap, 0)
    elif kind == "dm":
        a = sp.sp_phasing_diff_map_alloc(bmap, math.inf, math.inf, 0)
    else:
        a = sp.sp_phasing_er_alloc(0)
    ph = sp.sp_phaser_alloc()
    imgs = [sp.image_from_array(amp, mask), sp.image_from_array(model), sp.image_from_array(support)]
    try:
        if sp.sp_phaser_init(ph, a, None, sp.SpEngineHIP) != 0:
            raise RuntimeError(sp.last_error())
        sp.sp_phaser_set_amplitudes(ph, imgs[0])
        if sp.sp_phaser_init_model(ph, imgs[1], 0) != 0 or sp.sp_phaser_init_support(ph, imgs[2], 0, 0.0) != 0:
            raise RuntimeError(sp.last_error())
        if sp.sp_phaser_iterate(ph, 1) != 0:
            raise RuntimeError(sp.last_error())
        return sp.image_view(sp.sp_phaser_model(ph)).copy()
    finally:
        sp.sp_phaser_free(ph)
        for im in imgs:
            sp.sp_image_free(im)


def phaser_state(sp, ph):
    """(current model, current support) of a phaser, as host arrays."""
    import numpy as np
    model = sp.image_view(sp.sp_phaser_model(ph)).copy()
    sup = (sp.image_view(sp.sp_phaser_support(ph)).real > 0).astype(np.float32)
    return model, sup


def save_probe(probe_dir, name, kind, beta, pattern, support, before, after):
    """A state before one GPU iteration and the GPU's state after it; the CPU
    leg repeats the iteration in the FP64 oracle."""
    if not probe_dir:
        return
    import numpy as np
    np.savez(Path(probe_dir) / f"{name}.npz", kind=kind, beta=beta, amp=pattern.amplitudes.astype(np.float32),
             mask=pattern.mask.astype(np.uint8), support=np.asarray(support).astype(np.uint8),
             before=before.astype(np.complex64), after=after.astype(np.complex64))


def gpu_bench(args, probe_dir=None):
    from libspimage_amd import spimage as sp
    from libspimage_amd import workloads
    if not sp.spx_device_available():
        raise SystemExit("bench.py: no HIP device visible")
    p = workloads.blob_1024()
    N = p.amplitudes.size
    ph, keep = make_hio_phaser(sp, p)
    ms = C.c_float()
    if sp.spx_phaser_time_iterations(ph, max(args.warmup, 1), C.byref(ms)) != 0:
        raise RuntimeError(sp.last_error())
    # the median of SAMPLES samples of at least 100 iterations each
    n_sample = max(args.steps, 100)
    samples = []
    for _ in range(SAMPLES):
        if sp.spx_phaser_time_iterations(ph, n_sample, C.byref(ms)) != 0:
            raise RuntimeError(sp.last_error())
        samples.append(ms.value / 1e3 / n_sample)
    t_iter = sorted(samples)[SAMPLES // 2]
    # per-pass durations (events between the two passes of each iteration)
    mf, mr = C.c_float(), C.c_float()
    npass = max(200, args.steps // 4)
    if sp.spx_phaser_time_passes(ph, npass, C.byref(mf), C.byref(mr)) != 0:
        raise RuntimeError(sp.last_error())
    t_f_ev, t_r_ev = mf.value / 1e3 / npass, mr.value / 1e3 / npass
    # each HIP event between passes adds a dispatch gap: the difference
    # between the event-bracketed and the plain iteration time, split evenly
    # over the two passes, is taken off each pass
    gap = max(0.0, (t_f_ev + t_r_ev - t_iter) / 2)
    t_f, t_r = t_f_ev - gap, t_r_ev - gap
    layout = sp.spx_phaser_tiled(ph)
    tiled = layout == 1
    nt_hints = sp.spx_phaser_nt_hints(ph)
    before, sup_now = phaser_state(sp, ph)
    save_probe(probe_dir, "hio_1024", "hio", 0.9, p, sup_now, before,
               gpu_one_step(sp, "hio", 0.9, p.amplitudes, p.mask, sup_now, before))
    sp.sp_phaser_free(ph)

    passes = {
        "real_space_pass": {"bytes_per_pixel": BYTES_REAL_PASS, "seconds": t_r, "event_seconds": t_r_ev},
        "fourier_space_pass": {"bytes_per_pixel": BYTES_FOURIER_PASS, "seconds": t_f, "event_seconds": t_f_ev},
    }
    dom_name = max(passes, key=lambda k: passes[k]["seconds"])
    dom = passes[dom_name]
    achieved = dom["bytes_per_pixel"] * N / dom["seconds"]
    traffic = None
    tfile = ROOT / "profiles" / "hio_1024_hbm_traffic.json"
    if tfile.exists():
        try:
            entry = json.loads(tfile.read_text()).get(dom_name)
            traffic = entry["bytes"] if isinstance(entry, dict) else entry
        except (ValueError, OSError, KeyError, TypeError):
            traffic = None
    roofline = {
        "bound": "hbm", "kernel": dom_name, "achieved": achieved / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
        "frac": achieved / PEAK_HBM, "traffic": traffic,
        "algorithmic_bytes": dom["bytes_per_pixel"] * N, "avg_launch_us": dom["seconds"] * 1e6,
    }
    it_achieved = BYTES_HIO * N / t_iter
    iteration_roofline = {
        "achieved": it_achieved / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s", "frac": it_achieved / PEAK_HBM,
        "algorithmic_bytes": BYTES_HIO * N,
        "passes_us": {k: v["seconds"] * 1e6 for k, v in passes.items()},
        "passes_event_bracketed_us": {k: v["event_seconds"] * 1e6 for k, v in passes.items()},
        "fft_elements_per_thread": dict(zip(("row", "column"), sp.fft_variant(1024, tiled))),
        "timing": f"median of {SAMPLES} samples of {n_sample} iterations (HIP events on the engine stream)",
        "samples_us_per_iteration": [t * 1e6 for t in samples],
        "fourier_space_layout": {1: "4-column tiled", 2: "row-major, transposed Fourier-space pass"}.get(layout,
                                                                                                   "row-major"),
        "nt_hints": nt_hints,
    }

    if args.only_headline:
        return t_iter, roofline, iteration_roofline, None
    # the 1024^2 reconstruction: RAAR beta 0.87, 300 iterations, area
    # shrinkwrap every 20 (blur 3.0 -> 0.7 px, 8 % of the pixels)
    bmap, amap = sp.smap([(0, 3.0), (300, 0.7)]), sp.smap([(0, 0.08)])
    salg = sp.sp_support_area_alloc(bmap, amap)
    sarr = sp.sp_support_array_init(salg, 20)
    ph2, keep2 = make_hio_phaser(sp, p, beta=0.87, alg="raar", sup_array=sarr)
    sp.sp_phaser_iterate(ph2, 20)  # warm the shrinkwrap kernels
    sp.sp_phaser_efourier(ph2, None)
    ph3, keep3 = make_hio_phaser(sp, p, beta=0.87, alg="raar", sup_array=sarr)
    sp.sp_phaser_efourier(ph3, None)  # synchronises the stream
    t0 = time.perf_counter()
    rc = sp.sp_phaser_iterate(ph3, 300)
    fc = C.c_float()
    ef = sp.sp_phaser_efourier(ph3, C.byref(fc))  # synchronising read-back
    t_rec = time.perf_counter() - t0
    recon = {"workload": "RAAR 1024x1024, beta 0.87, 300 iterations, area shrinkwrap every 20",
             "rc": rc, "seconds": t_rec, "iter_per_s": 300 / t_rec, "final_efourier": ef,
             "support_fraction": sp.sp_phaser_support_fraction(ph3)}
    before, sup_now = phaser_state(sp, ph3)
    save_probe(probe_dir, "raar_1024_shrinkwrap", "raar", 0.87, p, sup_now, before,
               gpu_one_step(sp, "raar", 0.87, p.amplitudes, p.mask, sup_now, before))
    sp.sp_phaser_free(ph2)
    sp.sp_phaser_free(ph3)
    return t_iter, roofline, iteration_roofline, recon


def disc_bench(probe_dir=None):
    """The 256^2 disc workload in full on the GPU: HIO beta 0.9 for 200
    iterations, then ER for 20, static support, from the pattern's
    random-phase start; wall time from the first iteration to a
    synchronising error read-back."""
    from libspimage_amd import spimage as sp
    from libspimage_amd import workloads
    p = workloads.disc_256(0)
    er = sp.sp_phasing_er_alloc(0)
    ph, keep = make_hio_phaser(sp, p)
    sp.spx_phaser_time_iterations(ph, 20, C.byref(C.c_float()))  # warm-up (tuning, code objects)
    sp.sp_phaser_free(ph)
    ph, keep = make_hio_phaser(sp, p)
    sp.sp_phaser_efourier(ph, None)
    t0 = time.perf_counter()
    rc = sp.sp_phaser_iterate(ph, 200)
    rc = rc or sp.sp_phaser_set_phasing_algorithm(ph, er) or sp.sp_phaser_iterate(ph, 20)
    fc = C.c_float()
    ef = sp.sp_phaser_efourier(ph, C.byref(fc))
    t = time.perf_counter() - t0
    out = {"workload": "256x256 disc pattern (20 Gaussians, 1e5 photons, beamstop 4 px), static support: HIO beta 0.9 "
                       "200 iterations then ER 20", "rc": rc, "seconds": t, "final_efourier": ef, "fcfo": fc.value,
           "ereal": sp.sp_phaser_ereal(ph)}
    sp.sp_phaser_free(ph)
    if probe_dir:
        # one HIO step from the state at iteration 200 of the same run
        ph, keep = make_hio_phaser(sp, p)
        sp.sp_phaser_iterate(ph, 200)
        before, sup_now = phaser_state(sp, ph)
        sp.sp_phaser_free(ph)
        save_probe(probe_dir, "hio_256_disc", "hio", 0.9, p, sup_now, before,
                   gpu_one_step(sp, "hio", 0.9, p.amplitudes, p.mask, sup_now, before))
    return out


def batched_bench(n_patterns=128, n_starts=8, iters=500, period=25, plain_per_size=None, seed=2000, quantum=25):
    """Many-reconstruction workload: n_patterns seeded disc patterns with sizes
    drawn from {128, 256, 384, 512}^2, n_starts random-phase starts each, HIO
    beta 0.9 for `iters` iterations with an area shrinkwrap every `period`
    (blur 3.0 -> 0.7 px over the run, 8 % of the pixels).

    Batched: one SpxBatch per size class advances all its reconstructions
    together (one launch per pass for the whole class).  The classes run
    concurrently, each batch on its own stream, fed by one host loop in
    quanta of `quantum` iterations; the same classes run one after another
    for comparison.  Plain: the reference-style loop, one SpPhaser per
    reconstruction from the same starts, timed on every reconstruction (or on
    `plain_per_size` per size class, extrapolated to the class's count).  Inputs
    are uploaded before the clock starts in every form; the clock stops after
    synchronising efourier read-backs."""
    import numpy as np
    from libspimage_amd import spimage as sp
    from libspimage_amd import workloads
    rng = np.random.default_rng(seed)
    sizes = rng.choice([128, 256, 384, 512], n_patterns)
    classes = {}
    for i, n in enumerate(sizes):
        classes.setdefault(int(n), []).append(workloads.disc_scaled(int(n), seed=seed + 1 + i))
    keep = []
    hio = sp.sp_phasing_hio_alloc(sp.smap([(0, 0.9)]), 0)
    bmap, amap = sp.smap([(0, 3.0), (iters, 0.7)]), sp.smap([(0, 0.08)])
    keep += [bmap, amap]
    sarr = sp.sp_support_array_init(sp.sp_support_area_alloc(bmap, amap), period)

    def build(n, pats, graphs=False):
        b = sp.Batch((n, n), len(pats) * n_starts)
        for j, p in enumerate(pats):
            for s_ in range(n_starts):
                b.set_data(j * n_starts + s_, p.amplitudes, p.mask, p.support)
        b.init_model(sp.SpModelRandomPhases, seed + n)
        b.set_algorithm(hio)
        b.set_support_schedule(sarr)
        b.use_graphs(graphs)
        return b

    def concurrent(graphs):
        bs = {n: build(n, pats, graphs) for n, pats in sorted(classes.items())}
        for b in bs.values():
            b.synchronize()
        t0 = time.perf_counter()
        for q in range(0, iters, quantum):
            for b in bs.values():
                b.iterate(min(quantum, iters - q))
        ef = {n: b.efourier()[0] for n, b in bs.items()}
        return bs, ef, time.perf_counter() - t0

    # serial classes
    t_serial = 0.0
    ef_serial, starts = {}, {}
    for n, pats in sorted(classes.items()):
        b = build(n, pats)
        n_plain = len(pats) * n_starts if plain_per_size is None else min(plain_per_size, len(pats) * n_starts)
        starts[n] = [b.model(r) for r in range(n_plain)]
        b.synchronize()
        t0 = time.perf_counter()
        b.iterate(iters)
        ef_serial[n], _ = b.efourier()
        t_serial += time.perf_counter() - t0
        b.close()
    # concurrent classes replaying captured HIP graphs (one capture per
    # class, one graph launch per iteration between support updates)
    bg, ef_graph, t_graph = concurrent(True)
    for b in bg.values():
        b.close()
    # concurrent classes, eager launches
    bs, ef_conc, t_conc = concurrent(False)
    # device time of each class's plain HIO iteration (no support update),
    # against the survey's 78 N bytes and against the bytes rocprof counted
    # for that class's two passes
    phys = _profile_bytes("batched_hbm_traffic.json")
    device, survey_bytes = {}, 0.0
    for n, b in bs.items():
        nrec = len(classes[n]) * n_starts
        b.time_iterations(5)
        us = b.time_iterations(50) / 50 * 1e3
        row = {"recons": nrec, "hio_us_per_iteration": us,
               "frac_survey_bytes": 78.0 * n * n * nrec / (us * 1e-6) / PEAK_HBM}
        bpp = phys.get(f"{n}x{n}", {}).get("bytes_per_pixel") if phys else None
        if bpp:
            row["physical_bytes_per_pixel"] = bpp
            row["frac_physical_bytes"] = bpp * n * n * nrec / (us * 1e-6) / PEAK_HBM
        device[f"{n}x{n}"] = row
        survey_bytes += 78.0 * n * n * nrec * iters
    for b in bs.values():
        b.close()
    same = all(np.array_equal(ef_conc[n], ef_serial[n]) for n in classes)
    same_graph = all(np.array_equal(ef_graph[n], ef_conc[n]) for n in classes)
    # plain loop over the first reconstructions of each class
    t_plain, per_size, worst = 0.0, {}, 0.0
    for n, pats in sorted(classes.items()):
        nrec = len(pats) * n_starts
        n_plain = len(starts[n])
        tp, ef_p = 0.0, []
        for r in range(n_plain):
            p = pats[r // n_starts]
            ph = sp.sp_phaser_alloc()
            if sp.sp_phaser_init(ph, hio, sarr, sp.SpEngineHIP) != 0:
                raise RuntimeError(sp.last_error())
            imgs = [sp.image_from_array(p.amplitudes, p.mask), sp.image_from_array(starts[n][r]),
                    sp.image_from_array(p.support)]
            sp.sp_phaser_set_amplitudes(ph, imgs[0])
            if sp.sp_phaser_init_model(ph, imgs[1], 0) != 0 or sp.sp_phaser_init_support(ph, imgs[2], 0, 0.0) != 0:
                raise RuntimeError(sp.last_error())
            sp.sp_phaser_efourier(ph, None)  # synchronises the uploads
            t0 = time.perf_counter()
            sp.sp_phaser_iterate(ph, iters)
            ef = sp.sp_phaser_efourier(ph, None)
            tp += time.perf_counter() - t0
            ef_p.append(ef)
            sp.sp_phaser_free(ph)
            for im in imgs:
                sp.sp_image_free(im)
        tp_class = tp / n_plain * nrec
        diff = max(abs(float(ef_serial[n][r]) - e) / max(abs(e), 1e-30) for r, e in enumerate(ef_p))
        worst = max(worst, diff)
        t_plain += tp_class
        per_size[f"{n}x{n}"] = {"recons": nrec, "plain_recons_timed": n_plain, "plain_s": tp_class,
                                "mean_efourier": float(np.mean(ef_serial[n]))}
    total = n_patterns * n_starts
    return {"workload": f"{n_patterns} patterns x {n_starts} random starts, sizes {{128,256,384,512}}^2, "
                        f"HIO beta 0.9, {iters} iterations, area shrinkwrap every {period}",
            "recons": total, "batched_seconds": t_conc, "batched_recons_per_s": total / t_conc,
            "batched_form": f"one batch per size class, the classes concurrently on their own streams "
                            f"(quanta of {quantum} iterations)",
            "serial_classes_seconds": t_serial, "serial_classes_recons_per_s": total / t_serial,
            "concurrent_equals_serial": same,
            "graph_replay_seconds": t_graph, "graph_replay_recons_per_s": total / t_graph,
            "graph_replay_equals_eager": same_graph,
            "plain_seconds": t_plain, "plain_recons_per_s": total / t_plain,
            "speedup": t_plain / t_conc, "speedup_serial_classes": t_plain / t_serial,
            "plain_sample": ("every reconstruction" if plain_per_size is None else
                             f"{plain_per_size} reconstructions per size class, extrapolated"),
            "efourier_max_rel_diff_batched_vs_plain": worst, "per_size": per_size,
            "hio_iterations_device": device,
            "whole_run_frac_survey_bytes": survey_bytes / t_conc / PEAK_HBM,
            "whole_run_note": "the survey's 78 N bytes of every reconstruction-iteration over the concurrent run's "
                              "wall time, shrinkwrap updates included in the time but not in the bytes"}


def mixed_radix_bench(sizes=(720, 750, 810, 900, 972, 1000, 768, 960, 1024), nrec=16, iters=100, seed=7):
    """HIO iterations (device time, HIP events) of 2-D shapes whose lengths
    are not powers of two, as the served mix has them, and of 1024^2 for
    comparison: one reconstruction and batches of `nrec`, per
    reconstruction-iteration and against the 78*N-byte roofline.  The tuner
    picks layout, variants, column width and cache hints per shape."""
    from libspimage_amd import spimage as sp
    from libspimage_amd import workloads
    hio = sp.sp_phasing_hio_alloc(sp.smap([(0, 0.9)]), 0)
    out = {}
    for n in sizes:
        p = workloads.serving_pattern((n, n), seed)
        row = {}
        for cap in (1, nrec):
            b = sp.Batch((n, n), cap)
            for r in range(cap):
                b.set_data(r, p.amplitudes, p.mask, p.support)
            b.init_model(sp.SpModelRandomPhases, 1)
            b.set_algorithm(hio)
            b.time_iterations(10)
            us = b.time_iterations(iters) / iters * 1e3
            b.close()
            row[f"x{cap}"] = {"us_per_iteration": us, "us_per_reconstruction_iteration": us / cap,
                              "frac": 78.0 * n * n * cap / (us * 1e-6) / 8e12}
        out[f"{n}x{n}"] = row
    return {"workload": f"HIO beta 0.9, serving-mix patterns, 1 and {nrec} reconstructions per launch, "
                        f"{iters} iterations timed", "shapes": out}


def volume_bench(dm_its=200, er_its=100, period=20, probe_dir=None):
    """3-D reconstruction of the 256^3 ellipsoid pattern: difference map
    (beta 0.8, gamma1 = gamma2 = infinity -> Elser's values) for `dm_its`
    iterations, then ER for `er_its`, with an area shrinkwrap every `period`
    (blur 3.0 -> 1.0 voxels over the run, 4 % of the voxels).  Also the
    device time of plain DM and ER iterations (no support updates) against
    their algorithmic bytes (3-D: DM 211 N, ER 102 N)."""
    import math
    from libspimage_amd import spimage as sp
    from libspimage_amd import workloads
    p = workloads.ellipsoid_256()
    N = p.amplitudes.size
    total = dm_its + er_its
    beta = sp.smap([(0, 0.8)])
    dm = sp.sp_phasing_diff_map_alloc(beta, math.inf, math.inf, 0)
    er = sp.sp_phasing_er_alloc(0)
    bmap, amap = sp.smap([(0, 3.0), (total, 1.0)]), sp.smap([(0, 0.04)])
    sarr = sp.sp_support_array_init(sp.sp_support_area_alloc(bmap, amap), period)

    def phaser(alg, sup):
        ph = sp.sp_phaser_alloc()
        if sp.sp_phaser_init(ph, alg, sup, sp.SpEngineHIP) != 0:
            raise RuntimeError(sp.last_error())
        imgs = [sp.image_from_array(p.amplitudes, p.mask), sp.image_from_array(p.start),
                sp.image_from_array(p.support)]
        sp.sp_phaser_set_amplitudes(ph, imgs[0])
        if sp.sp_phaser_init_model(ph, imgs[1], 0) != 0 or sp.sp_phaser_init_support(ph, imgs[2], 0, 0.0) != 0:
            raise RuntimeError(sp.last_error())
        for im in imgs:
            sp.sp_image_free(im)
        return ph

    per_it = {}
    phys = _profile_bytes("volume_256_hbm_traffic.json")
    for name, alg, nbytes in (("dm", dm, 211), ("er", er, 102)):
        ph = phaser(alg, None)
        ms = C.c_float()
        if sp.spx_phaser_time_iterations(ph, 5, C.byref(ms)) != 0 or \
                sp.spx_phaser_time_iterations(ph, 40, C.byref(ms)) != 0:
            raise RuntimeError(sp.last_error())
        t = ms.value / 1e3 / 40
        per_it[name] = {"us_per_iteration": t * 1e6, "algorithmic_bytes": nbytes * N,
                        "achieved_GBps": nbytes * N / t / 1e9, "frac": nbytes * N / t / PEAK_HBM}
        bpv = phys.get(name, {}).get("bytes_per_voxel") if phys else None
        if bpv:
            per_it[name].update({"physical_bytes": bpv * N, "physical_GBps": bpv * N / t / 1e9,
                                 "frac_physical_bytes": bpv * N / t / PEAK_HBM})
        if name == "dm" and probe_dir:
            before, sup_now = phaser_state(sp, ph)
            save_probe(probe_dir, "dm_256cubed", "dm", 0.8, p, sup_now, before,
                       gpu_one_step(sp, "dm", 0.8, p.amplitudes, p.mask, sup_now, before))
        sp.sp_phaser_free(ph)
    ph = phaser(dm, sarr)
    sp.sp_phaser_efourier(ph, None)
    t0 = time.perf_counter()
    rc = sp.sp_phaser_iterate(ph, dm_its)
    if rc == 0:
        rc = sp.sp_phaser_set_phasing_algorithm(ph, er)
    if rc == 0:
        rc = sp.sp_phaser_iterate(ph, er_its)
    fc = C.c_float()
    ef = sp.sp_phaser_efourier(ph, C.byref(fc))
    t_rec = time.perf_counter() - t0
    frac = sp.sp_phaser_support_fraction(ph)
    sp.sp_phaser_free(ph)
    return {"workload": f"256^3 ellipsoid pattern (200 Gaussians, 1e8 photons, beamstop 5): DM beta 0.8 "
                        f"(Elser gammas) {dm_its} iterations then ER {er_its}, area shrinkwrap every {period}",
            "rc": rc, "seconds": t_rec, "iter_per_s": total / t_rec, "final_efourier": ef, "fcfo": fc.value,
            "support_fraction": frac, "plain_iterations": per_it}


def reconstructor_bench(repeats=32, single_sample=None, seed=600):
    """The Reconstructor front-end on a 256^2 disc pattern: RAAR (beta 0.9 ->
    0.8) for 200 iterations then ER for 50, area shrinkwrap every 20, 5
    image / score outputs.  reconstruct_loop(repeats) (all random starts in
    one batch) against the reference's serial loop of reconstruct() calls,
    timed on all of them (or on `single_sample`, extrapolated)."""
    import numpy as np
    from libspimage_amd import workloads
    from libspimage_amd.reconstructor import Reconstructor
    p = workloads.disc_256(seed)
    R = Reconstructor()
    R.set_intensities(np.fft.fftshift(p.amplitudes.astype(np.float64) ** 2))
    R.set_mask(np.fft.fftshift(p.mask.astype(bool)))
    R.set_number_of_iterations(250)
    R.set_number_of_outputs(5)
    R.set_initial_support(radius=45)
    R.set_support_algorithm("area", update_period=20, blur_init=3.0, blur_final=0.7, area_init=0.1, area_final=0.08)
    R.append_phasing_algorithm("raar", number_of_iterations=200, beta_init=0.9, beta_final=0.8)
    R.append_phasing_algorithm("er", number_of_iterations=50)
    # warm-up: the tuner for one reconstruction and for the batch of
    # `repeats` (each batch size class is tuned at its first use), code objects
    R.reconstruct(seed=1)
    R.reconstruct_loop(repeats, seed=seed + 10 * repeats)
    t0 = time.perf_counter()
    loop = R.reconstruct_loop(repeats, seed=seed)
    t_loop = time.perf_counter() - t0
    n_single = repeats if single_sample is None else min(single_sample, repeats)
    t0 = time.perf_counter()
    singles = [R.reconstruct(seed=seed + r) for r in range(n_single)]
    t_single = (time.perf_counter() - t0) / n_single
    same = all(float(loop["scores_final"]["fourier_error"][r]) == float(singles[r]["fourier_error"][-1])
               for r in range(n_single))
    return {"workload": f"Reconstructor, 256x256 disc pattern, {repeats} random starts, RAAR 200 + ER 50, "
                        "area shrinkwrap every 20, 5 outputs",
            "reconstruct_loop_seconds": t_loop, "loop_recons_per_s": repeats / t_loop,
            "serial_reconstruct_seconds_each": t_single, "serial_recons_per_s": 1.0 / t_single,
            "speedup": t_single * repeats / t_loop, "serial_reconstructions_timed": n_single,
            "final_errors_equal": same}


def prtf_bench(shape=(256, 256), counts=(32, 128), seed=700):
    """PRTF over many reconstructions (the reference's prtf(), translate and
    enantio on): noisy copies of one complex object with random shifts,
    reflections and constant phases.  The GPU path against the oracle's
    numpy restatement of the same algorithm on one host core (its
    alignment decisions must agree: the aligned masks are compared)."""
    import numpy as np
    from libspimage_amd import prtf as gp
    from oracle import prtf_oracle as P
    rng = np.random.default_rng(seed)
    obj = np.zeros(shape, complex)
    sl = tuple(slice(n // 4, n // 4 + n // 3) for n in shape)
    blk = obj[sl].shape
    obj[sl] = (rng.random(blk) + 0.3) * np.exp(0.5j * rng.random(blk))
    sup = (np.abs(obj) > 0).astype(np.int32)
    nmax = max(counts)
    imgs = np.empty((nmax,) + shape, np.complex64)
    masks = np.empty((nmax,) + shape, np.int32)
    for k in range(nmax):
        b = np.flip(obj) if k % 3 == 2 else obj
        b = b * np.exp(1j * rng.uniform(-3, 3)) + 0.01 * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))
        s = tuple(int(v) for v in rng.integers(-8, 9, len(shape)))
        imgs[k] = np.roll(b, s, axis=(0, 1))
        masks[k] = np.roll(sup, s, axis=(0, 1))
    gp.prtf(imgs[:4], masks[:4])  # warm-up (tuning, code objects)
    res = {"workload": f"PRTF of noisy shifted / reflected / phase-shifted copies, {shape[0]}x{shape[1]}, "
                       "translate + enantio", "runs": []}
    for n in counts:
        t0 = time.perf_counter()
        g = gp.prtf(imgs[:n], masks[:n], full_out=True)
        t = time.perf_counter() - t0
        res["runs"].append({"reconstructions": n, "seconds": t, "reconstructions_per_s": n / t})
    n0 = counts[0]
    t0 = time.perf_counter()
    o = P.prtf(imgs[:n0], masks[:n0], full_out=True)
    t_cpu = time.perf_counter() - t0
    g = gp.prtf(imgs[:n0], masks[:n0], full_out=True)
    res["cpu_oracle"] = {"reconstructions": n0, "seconds": t_cpu, "cores": 1,
                         "kind": "port (numpy float64 restatement of prtf())"}
    res["speedup_vs_cpu"] = t_cpu / res["runs"][0]["seconds"]
    res["aligned_masks_equal"] = bool(np.array_equal(g["masks"], o["masks"]))
    res["max_prtf_difference"] = float(np.abs(g["prtf"] - o["prtf"]).max())
    return res


def _serving_pattern(args):
    from libspimage_amd import workloads
    shape, seed = args
    return workloads.serving_pattern(shape, seed)


def _serving_setup(seconds, rate, seed, workers):
    """The served workload's trace and patterns, and a warm server: the
    tuner runs for every shape at every batch size the server may create
    (Server.prepare), then every shape goes once through both paths
    (twiddle tables, first launches).  Returns (trace, requests, plain)."""
    from concurrent.futures import ThreadPoolExecutor
    from libspimage_amd import spimage as sp
    from libspimage_amd import serve, workloads
    sizes = [n for n in workloads.smooth_sizes(128, 1024) if sp.spx_supported_length(n)]
    trace = workloads.serving_trace(rate, seconds, seed, sizes_2d=sizes)
    shapes = sorted({s for _, s, _, _, _ in trace})
    # threads (numpy's transforms release the interpreter lock): no child
    # processes to outlive the run
    with ThreadPoolExecutor(workers) as ex:
        pats = dict(zip(shapes, ex.map(_serving_pattern, [(s, seed + i) for i, s in enumerate(shapes)])))

    def requests(realtime=True, keep=None):
        return [serve.Request(k, pats[s].amplitudes, pats[s].mask, pats[s].support, alg, its, sd,
                              beta=0.9 if alg == "hio" else 0.87, arrival=t if realtime else 0.0)
                for k, (t, s, alg, its, sd) in enumerate(trace) if keep is None or keep[k]]

    srv = serve.Server()
    srv.prepare(shapes)
    warm = [serve.Request(-1 - i, pats[s].amplitudes, pats[s].mask, pats[s].support, "hio", 50, 1)
            for i, s in enumerate(shapes)]
    plain = serve.PlainRunner()
    for r in warm:
        plain.run_one(r)
    srv.run(warm, realtime=False)
    srv.close()
    return trace, requests, plain


def serving_capacity(seconds=3.0, rate=200.0, seed=4000, workers=8):
    """Only the all-queued run of the served workload, after the same
    warm-up (a short program for kernel traces: its kernels are the last
    `seconds` of the trace)."""
    from libspimage_amd import serve
    trace, requests, _ = _serving_setup(seconds, rate, seed, workers)
    srv = serve.Server()
    t0 = time.perf_counter()
    srv.run(requests(False), realtime=False)
    t = time.perf_counter() - t0
    srv.close()
    return {"requests": len(trace), "seconds": t, "recons_per_s": len(trace) / t}


def trace_summary(path, window_s=None):
    """Summary of a rocprofv3 kernel trace (CSV): over the kernels that start
    in the last `window_s` seconds before the trace's last kernel ends (all
    of them without a window), the time with no kernel, one, two, ...
    running, the union of their busy time, and the kernels with the most
    time (template arguments kept, call arguments dropped)."""
    import csv
    import re
    rows = []
    with open(path) as f:
        for r in csv.DictReader(f):
            rows.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), re.sub(r"\(.*", "", r["Kernel_Name"])))
    if not rows:
        return {"kernels": 0}
    rows.sort()
    end = max(e for _, e, _ in rows)
    if window_s is not None:
        lo = end - int(window_s * 1e9) - 20_000_000  # 20 ms of slack before the window
        rows = [r for r in rows if r[0] >= lo]
    events = sorted([(s, 1) for s, _, _ in rows] + [(e, -1) for _, e, _ in rows])
    busy_at = {}
    cur, last = 0, events[0][0]
    for t, d in events:
        busy_at[cur] = busy_at.get(cur, 0) + (t - last)
        cur += d
        last = t
    per = {}
    for s, e, n in rows:
        c, tot = per.get(n, (0, 0))
        per[n] = (c + 1, tot + e - s)
    top = sorted(per.items(), key=lambda kv: -kv[1][1])[:15]
    span = max(e for _, e, _ in rows) - rows[0][0]
    return {"kernels": len(rows), "span_s": span / 1e9,
            "busy_union_s": (span - busy_at.get(0, 0)) / 1e9, "idle_s": busy_at.get(0, 0) / 1e9,
            "seconds_with_k_kernels_running": {str(k): v / 1e9 for k, v in sorted(busy_at.items())},
            "sum_of_kernel_durations_s": sum(e - s for s, e, _ in rows) / 1e9,
            "top_kernels": [{"kernel": n, "calls": c, "total_s": t / 1e9, "mean_us": t / c / 1e3}
                            for n, (c, t) in top]}


def serving_bench(seconds=3.0, rate=200.0, seed=4000, workers=8):
    """Served workload: a Poisson trace at `rate` requests/s for `seconds`
    (85 % 2-D n x n with n FFT-smooth in [128, 1024], 15 % 3-D n^3 with n in
    {64, 80, 96, 128}; ER / HIO (beta 0.9) / RAAR (beta 0.87); 100-1000
    iterations) through the continuous-batching server, in real time and
    with every request queued at once (capacity), and through the plain
    reference-style loop (one SpPhaser per request, in arrival order)."""
    import numpy as np
    from libspimage_amd import serve
    trace, requests, plain = _serving_setup(seconds, rate, seed, workers)

    srv = serve.Server()
    t0 = time.perf_counter()
    live = srv.run(requests(True), realtime=True)
    t_live = time.perf_counter() - t0
    created_live = srv.batches_created
    srv.close()
    srv = serve.Server()
    t0 = time.perf_counter()
    offline = srv.run(requests(False), realtime=False)
    t_off = time.perf_counter() - t0
    srv.close()
    # below capacity: the same trace thinned at random to half its rate (a
    # thinned Poisson process is a Poisson process of the lower rate)
    keep = np.random.default_rng(seed + 1).random(len(trace)) < 0.5
    srv = serve.Server()
    t0 = time.perf_counter()
    light = srv.run(requests(True, keep), realtime=True)
    t_light = time.perf_counter() - t0
    srv.close()
    plain_res, t_plain = plain.run(requests(True))
    ef_plain = {r.rid: r.efourier for r in plain_res}
    diff = max(abs(r.efourier - ef_plain[r.rid]) / max(abs(ef_plain[r.rid]), 1e-30) for r in live)
    n = len(trace)
    return {"workload": f"Poisson {rate:g} req/s for {seconds:g} s ({n} requests): 85 % 2-D n^2, n FFT-smooth in "
                        "[128, 1024]; 15 % 3-D n^3, n in {64, 80, 96, 128}; ER / HIO 0.9 / RAAR 0.87 uniform; "
                        "100-1000 iterations; static support; random-phase starts",
            "requests": n, "shapes": len({s for _, s, _, _, _ in trace}),
            "served": {"latency": serve.latency_summary(live), "makespan_s": t_live, "recons_per_s": n / t_live,
                       "batches_created": created_live},
            "served_half_rate": {"offered_req_per_s": rate / 2, "requests": int(keep.sum()),
                                 "latency": serve.latency_summary(light), "makespan_s": t_light,
                                 "recons_per_s": int(keep.sum()) / t_light},
            "capacity_all_queued": {"seconds": t_off, "recons_per_s": n / t_off},
            "plain_loop": {"seconds": t_plain, "recons_per_s": n / t_plain,
                           "latency_fifo": serve.latency_summary(plain_res)},
            "capacity_speedup": t_plain / t_off,
            "efourier_max_rel_diff_served_vs_plain": diff}


# ---------------------------------------------------------------------------
# CPU legs (a child process; the only place the oracle runs)
# ---------------------------------------------------------------------------
def _cpu_threads() -> int:
    """Host threads for the all-cores CPU figure: this process's CPU share
    (the GPU box exports it as OMP_NUM_THREADS) within the affinity mask."""
    avail = len(os.sched_getaffinity(0))
    for var in ("SPX_CPU_THREADS", "OMP_NUM_THREADS"):
        v = os.environ.get(var, "")
        if v.isdigit() and int(v) > 0:
            return min(int(v), avail)
    return avail


def _cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _cpu_oracle(threads: int):
    """The oracle's FP32 build timing the reference loop: MKL's FFTW3 API,
    planned per call like src/fft.c:177-189, with `threads` plan threads
    (the reference's sp_init_fft, src/fft.c:34-49; its phaser never calls
    it, so 1 is the reference default)."""
    from oracle.oracle import Oracle
    o = Oracle("f32")
    backend = "own double-precision FFT"
    for path in MKL_CANDIDATES:
        if os.path.exists(path) and o.set_fft_backend(1, path) == 0:
            backend = f"FFTW3 API of {os.path.basename(path)} (MKL), plan per call"
            if threads > 1:
                backend += f", {threads} FFT threads" if o.set_fft_threads(threads) == 0 else \
                    ", FFT threads unavailable"
            break
    return o, backend


def _cpu_hio_1024(o, seconds):
    """The headline loop: HIO beta 0.9 on the 1024^2 blob pattern, static
    support (src/phasing.c:1142-1172)."""
    from libspimage_amd import workloads
    p = workloads.blob_1024()
    ph = o.phaser(p.amplitudes.shape)
    ph.set_inputs(p.amplitudes, p.mask.astype(bool), p.support.astype(bool), p.start)
    alg = o.alg(1, 0, beta=[(0, 0.9)])
    ph.iterate(alg, 1)
    n = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        ph.iterate(alg, 1)
        n += 1
    dt = time.perf_counter() - t0
    return {"iters": n, "seconds": dt, "iter_per_s": n / dt}


def _cpu_disc_256(o):
    """The 256^2 disc workload in full: HIO beta 0.9 for 200 iterations,
    then ER for 20, static support, from the pattern's random-phase start."""
    from libspimage_amd import workloads
    p = workloads.disc_256(0)
    ph = o.phaser(p.amplitudes.shape)
    ph.set_inputs(p.amplitudes, p.mask.astype(bool), p.support.astype(bool), p.start)
    t0 = time.perf_counter()
    ph.iterate(o.alg(1, 0, beta=[(0, 0.9)]), 200)
    ph.iterate(o.alg(4, 0), 20)
    dt = time.perf_counter() - t0
    ef, fc = ph.efourier()
    return {"seconds": dt, "final_efourier": ef, "fcfo": fc, "ereal": ph.ereal()}


def _cpu_raar_1024(o):
    """The 1024^2 reconstruction in full: RAAR beta 0.87 for 300 iterations
    with an area shrinkwrap every 20 (blur 3.0 -> 0.7 px, 8 % of the pixels;
    the reference's qsort select, src/support_update.c:136-151)."""
    from libspimage_amd import workloads
    from oracle.oracle import SUP_AREA
    p = workloads.blob_1024()
    ph = o.phaser(p.amplitudes.shape)
    ph.set_inputs(p.amplitudes, p.mask.astype(bool), p.support.astype(bool), p.start)
    sup = o.sup(SUP_AREA, period=20, blur=[(0, 3.0), (300, 0.7)], values=[(0, 0.08)])
    t0 = time.perf_counter()
    ph.iterate(o.alg(2, 0, beta=[(0, 0.87)]), 300, sup)
    dt = time.perf_counter() - t0
    ef, fc = ph.efourier()
    return {"seconds": dt, "iter_per_s": 300 / dt, "final_efourier": ef, "support_fraction": ph.support_fraction()}


def _cpu_volume_256(o, dm_sample=5, er_sample=3):
    """The 256^3 reconstruction, timed on a sample and extrapolated: DM
    (beta 0.8, Elser's gammas) and ER iterations and one area shrinkwrap
    (blur 3.0, 4 %), scaled to the run's 190 DM + 95 ER iterations and 15
    updates (200 DM + 100 ER with an update every 20, each consuming one)."""
    from libspimage_amd import workloads
    p = workloads.ellipsoid_256()
    ph = o.phaser(p.amplitudes.shape)
    ph.set_inputs(p.amplitudes, p.mask.astype(bool), p.support.astype(bool), p.start)
    g1, g2 = o.dm_gammas(0.8)
    t0 = time.perf_counter()
    ph.iterate(o.alg(3, 0, beta=[(0, 0.8)], gamma1=g1, gamma2=g2), dm_sample)
    t_dm = (time.perf_counter() - t0) / dm_sample
    t0 = time.perf_counter()
    ph.iterate(o.alg(4, 0), er_sample)
    t_er = (time.perf_counter() - t0) / er_sample
    t0 = time.perf_counter()
    ph.support_area_update(3.0, 0.04)
    t_area = time.perf_counter() - t0
    total = 190 * t_dm + 95 * t_er + 15 * t_area
    return {"seconds_extrapolated": total, "dm_seconds_per_iteration": t_dm, "er_seconds_per_iteration": t_er,
            "area_update_seconds": t_area,
            "sample": f"{dm_sample} DM + {er_sample} ER iterations and 1 area update timed, extrapolated to "
                      "190 DM + 95 ER + 15 updates"}


def _accuracy(probe_dir):
    """Each probe holds a state before one GPU iteration and the GPU's state
    after it.  The same iteration in the FP64 oracle (its own FFT in double)
    and in the FP32 oracle (the reference's CPU arithmetic with MKL's FFTW3
    API) gives the GPU's relative L2 error against FP64, next to the
    reference CPU path's own."""
    import numpy as np
    from oracle.oracle import Oracle
    out = {}
    for f in sorted(Path(probe_dir).glob("*.npz")):
        d = np.load(f)
        kind = str(d["kind"])
        beta = float(d["beta"])
        ids = {"hio": 1, "raar": 2, "dm": 3, "er": 4}
        res = {}
        for prec in ("f64", "f32"):
            o = Oracle(prec)
            kw = {}
            if kind == "dm":
                kw = dict(zip(("gamma1", "gamma2"), o.dm_gammas(beta)))
            alg = o.alg(ids[kind], 0, beta=[(0, beta)], **kw)
            ph = o.phaser(d["amp"].shape)
            ph.set_inputs(d["amp"], d["mask"].astype(bool), d["support"].astype(bool), d["before"])
            ph.iterate(alg, 1)
            res[prec] = np.array(ph.g1, np.complex128)
            del ph
        ref = res["f64"]
        nrm = float(np.linalg.norm(ref))
        out[f.stem] = {"iteration": kind, "shape": list(d["amp"].shape),
                       "gpu_vs_fp64_rel_l2": float(np.linalg.norm(d["after"].astype(np.complex128) - ref) / nrm),
                       "reference_cpu_fp32_vs_fp64_rel_l2": float(np.linalg.norm(res["f32"] - ref) / nrm)}
    return out


def cpu_baseline_worker(spec_json: str):
    """Runs in a child process: the CPU tasks named in the spec."""
    spec = json.loads(spec_json)
    threads = int(spec.get("threads", 1))
    o, backend = _cpu_oracle(threads)
    res = {"threads": threads, "backend": backend}
    for task in spec["tasks"]:
        if task == "hio_1024":
            res[task] = _cpu_hio_1024(o, float(spec.get("seconds", 12.0)))
        elif task == "disc_256":
            res[task] = _cpu_disc_256(o)
        elif task == "raar_1024":
            res[task] = _cpu_raar_1024(o)
        elif task == "volume_256":
            res[task] = _cpu_volume_256(o)
        elif task == "accuracy":
            res[task] = _accuracy(spec["probe_dir"])
    print(json.dumps(res))


def _run_cpu_child(spec: dict, threads: int):
    env = dict(os.environ, MKL_NUM_THREADS=str(threads))
    if threads == 1:
        env.update(MKL_THREADING_LAYER="SEQUENTIAL", OMP_NUM_THREADS="1")
    else:
        env.update(MKL_THREADING_LAYER="GNU", OMP_NUM_THREADS=str(threads))
    spec = dict(spec, threads=threads)
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--cpu-baseline-only", "--cpu-spec", json.dumps(spec)],
                       capture_output=True, text=True, env=env, timeout=900)
    if r.returncode != 0:
        return {"error": r.stderr[-400:]}
    return json.loads(r.stdout.strip().splitlines()[-1])


def cpu_baseline(seconds: float, probe_dir: str | None, volume: bool):
    """The reference loop on the host cores: one thread (the reference's
    default) and all threads of this process's CPU share, plus the accuracy
    of each GPU result against the FP64 oracle."""
    tasks1 = ["hio_1024", "disc_256", "raar_1024"] + (["volume_256"] if volume else [])
    if probe_dir:
        tasks1.append("accuracy")
    one = _run_cpu_child({"tasks": tasks1, "seconds": seconds, "probe_dir": probe_dir}, 1)
    nthr = _cpu_threads()
    many = _run_cpu_child({"tasks": ["hio_1024", "raar_1024"] + (["volume_256"] if volume else []),
                           "seconds": seconds}, nthr) if nthr > 1 else None
    hio = one.get("hio_1024")
    base = {"value": hio["iter_per_s"] if hio else None, "unit": "iter/s", "cores": 1, "kind": "port",
            "sample": (f"{hio['iters']} HIO iterations of the 1024x1024 blob pattern in {hio['seconds']:.1f} s "
                       f"(oracle restatement of src/phasing.c:1142-1172, FFT: {one.get('backend')})")
            if hio else f"failed: {one.get('error')}",
            "cpu_model": _cpu_model(), "nproc": os.cpu_count(), "threads_all_cores": nthr}
    if many and many.get("hio_1024"):
        base["all_cores"] = {"value": many["hio_1024"]["iter_per_s"], "unit": "iter/s", "threads": nthr,
                             "backend": many.get("backend"),
                             "note": "FFTs threaded like the reference's sp_init_fft; the element-wise loops of "
                                     "src/phasing.c stay serial as in the reference"}
    workloads_cpu = {}
    for key in ("disc_256", "raar_1024", "volume_256"):
        entry = {}
        if key in one:
            entry["threads_1"] = one[key]
        if many and key in many:
            entry[f"threads_{nthr}"] = many[key]
        if entry:
            workloads_cpu[key] = entry
    return base, workloads_cpu, one.get("accuracy"), {k: v for k, v in (("one", one), ("many", many))
                                                     if v and "error" in v}


def copy_calibration():
    """20 launches of torch's copy kernel over 64 MiB: the known byte count
    against which HBM read/write counters are scaled."""
    import torch
    n = 64 << 20
    src = torch.ones(n // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(20):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n, "copy_bytes_written": n, "launches": 20}))


def main():
    args = parse()
    if args.trace_summary:
        print(json.dumps(trace_summary(args.trace_summary, args.trace_window_s)))
        return
    if args.copy_calibration:
        copy_calibration()
        return
    if args.cpu_baseline_only:
        cpu_baseline_worker(args.cpu_spec)
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(independent reconstructions shard by id across GPUs with no exchange)", file=sys.stderr)
        raise SystemExit(2)
    import shutil
    import tempfile
    if args.only_volume:
        print(json.dumps({"volume_256": volume_bench()}))
        return
    if args.only_batched:
        print(json.dumps({"batched": batched_bench()}))
        return
    if args.only_serving_capacity:
        print(json.dumps({"serving_capacity": serving_capacity()}))
        return
    if args.only_headline:
        args.no_batched = args.no_serving = args.no_volume = args.no_cpu_baseline = True
    probe_dir = None if args.no_cpu_baseline else tempfile.mkdtemp(prefix="spx_accuracy_")
    try:
        t_iter, roofline, it_roof, recon = gpu_bench(args, probe_dir)
        disc = None if args.only_headline else disc_bench(probe_dir)
        batched = None if args.no_batched else batched_bench()
        serving = None if args.no_serving else serving_bench()
        volume = None if args.no_volume else volume_bench(probe_dir=probe_dir)
        reconstructor = None if args.no_batched else reconstructor_bench()
        prtf = None if args.no_batched else prtf_bench()
        mixed = None if args.no_batched else mixed_radix_bench()
        cpu = cpu_workloads = accuracy = cpu_errors = None
        if not args.no_cpu_baseline:
            cpu, cpu_workloads, accuracy, cpu_errors = cpu_baseline(args.cpu_seconds, probe_dir, volume is not None)
    finally:
        if probe_dir:
            shutil.rmtree(probe_dir, ignore_errors=True)
    # GPU against CPU on the measured workloads (same inputs, same schedule)
    if cpu_workloads:
        gpu_side = {"disc_256": (disc, "seconds"), "raar_1024": (recon, "seconds"),
                    "volume_256": (volume, "seconds")}
        for key, entry in cpu_workloads.items():
            g, field = gpu_side[key]
            if not g:
                continue
            entry["gpu_seconds"] = g[field]
            for tk, tv in list(entry.items()):
                if not tk.startswith("threads_"):
                    continue
                cs = tv.get("seconds", tv.get("seconds_extrapolated"))
                entry[f"speedup_vs_{tk}"] = cs / g[field] if cs else None
                if "final_efourier" in tv and "final_efourier" in g:
                    entry[f"final_efourier_rel_diff_vs_{tk}"] = abs(g["final_efourier"] - tv["final_efourier"]) / \
                        max(abs(tv["final_efourier"]), 1e-30)
    out = {
        "metric": "HIO iter/s @1024^2 fp32 (%HBM roofline)",
        "value": 1.0 / t_iter,
        "unit": "iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_iter * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "HIO 1024x1024 complex64, seeded blob pattern (1e7 photons, beamstop 8 px), "
                               "static support, beta 0.9", "shape": [1024, 1024], "algorithm": "HIO"},
        "roofline": roofline,
        "iteration_roofline": it_roof,
        "reconstruction_1024": recon,
        "disc_256": disc,
        "batched": batched,
        "serving": serving,
        "volume_256": volume,
        "reconstructor": reconstructor,
        "prtf": prtf,
        "mixed_radix_2d": mixed,
        "cpu_baseline": cpu,
        "cpu_workloads": cpu_workloads,
        "accuracy_vs_fp64": accuracy,
    }
    if cpu_errors:
        out["cpu_leg_errors"] = cpu_errors
    print(json.dumps(out))


if __name__ == "__main__":
    main()
