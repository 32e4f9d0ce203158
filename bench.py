"""Benchmark of the template-matching search on one MI355X.

Workload (the production configuration of the headline metric): one 5760 x 4092 float32
micrograph with 400 inserted particles at SNR 0.05, 1000 templates of 180 x 180 from a uniform
projection-direction grid, in-plane step 5 deg (72 angles), norm_type=1 on the whole image,
window 320 (234 blocks), defocus 2.0 um, astigmatism 50 A (dfdiff 0.005 um) at 30 deg.

A step is one full search (template and image preprocessing, 72 x 1000 x 234 correlation planes,
per-pixel statistics, records and NMS) with the image and templates already resident in HBM.
value = template.angle.pixel correlations per second = 1000 * 72 * 5760 * 4092 / s.

The roofline entry is for the dominant correlation kernel (the column pass): its algorithmic
bytes per launch (the inter-pass buffer written once, every template and block spectrum of the
launch read once) over its average launch time, from HIP events recorded around every launch on
the engine's stream during the timed steps; `kernels` gives the same for the row pass (buffer
read once plus one read-modify-write of the per-pixel records), `traffic` and
`hbm_frac_counters` the copy-calibrated HBM bytes rocprofv3 counted for the same kernels
(profiles/corr_c1_traffic.json), and `limiter` their shader counters (profiles/corr_c1_valu.json).
The CPU baseline is the FP64 oracle on a bounded sample of the same workload (see
cpu_baseline.sample).

The second half of the metric ("refined peaks/s at tol") is the `refine` object: the 2000 best
picks at z >= 5 after NMS (64 px) of the same search, perturbed, refined by 20 ascent steps on
the GPU, with a sample refined by the FP64 oracle from the same starts to check the tolerance.
The `batch_refine` object times a batch of 4 uneven micrographs searched and refined by
gs_search_batch_refine against the per-micrograph form; the `c0` object the norm_type=0
configuration (1440 x 1440, 8 templates, 36 angles, window 720) with its oracle comparison.
`--batch K` measures a batch of K uneven micrographs instead, `--c0` the norm_type=0
configuration (each its own JSON line).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gisspa_v2.0_amd"))

import numpy as np  # noqa: E402

METRIC = "template·angle·pixel correlations/s + % HBM peak; refined peaks/s at tol"
HBM_PEAK = 8.0e12


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--angles", type=int, default=0, help="limit angles (exploration only)")
    ap.add_argument("--templates", type=int, default=1000)
    ap.add_argument("--chunk", type=int, default=0)
    ap.add_argument("--row-variant", type=int, default=-1, help="-1: tune on one angle")
    ap.add_argument("--col-variant", type=int, default=-1, help="-1: tune on one angle")
    ap.add_argument("--overlap", type=int, default=0,
                    help="1: row pass on a second stream (blurs per-kernel timing)")
    ap.add_argument("--batch", type=int, default=0,
                    help="K > 0: uneven-micrograph batch workload instead (its own JSON line)")
    ap.add_argument("--c0", action="store_true",
                    help="norm_type=0 configuration (1440x1440, 8 templates, 36 angles, window 720) "
                         "instead (its own JSON line)")
    ap.add_argument("--reference-records", action="store_true",
                    help="the production search with reference_records = 1 (the reference's own "
                         "per-chunk records from a second pass) instead (its own JSON line)")
    ap.add_argument("--refine-only", action="store_true",
                    help="search restricted to --angles (default 8) then only the refinement leg "
                         "(for profiling the refinement kernels)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch-refine", type=int, default=4,
                    help="K uneven micrographs for the batched + refined leg of the default line (0: skip)")
    ap.add_argument("--c0-leg", type=int, default=1,
                    help="1: add the norm_type=0 configuration as the c0 object of the default line")
    ap.add_argument("--cpu-templates", type=int, default=4)
    ap.add_argument("--cpu-angles", type=int, default=4)
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(cfg, ds, nt, na):
    """FP64 oracle, marginal rate between 1 and 1+na angles on the first nt templates."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    oracle.load()
    cores = oracle.num_threads()
    tmpl = ds.templates[:nt]
    times = []
    for a in (1, 1 + na):
        t0 = time.perf_counter()
        oracle.search_norm(cfg, ds.image, tmpl, ds.ctf, n_angles=a)
        times.append(time.perf_counter() - t0)
        log(f"cpu oracle: {nt} templates x {a} angles: {times[-1]:.2f} s")
    H, W = ds.image.shape
    rate = nt * na * W * H / max(times[1] - times[0], 1e-9)
    return {
        "value": rate,
        "unit": "correlations/s",
        "cores": cores,
        "kind": "port",
        "sample": (f"FP64 oracle on the same micrograph, first {nt} templates, marginal time of "
                   f"{na} extra angles over all blocks (runs with 1 and {1 + na} angles: "
                   f"{times[0]:.1f} s and {times[1]:.1f} s)"),
    }


REFINE_PEAKS = 2000
# the top REFINE_PEAKS picks are NMS-separated peaks (64 px, the search's own radius) at z >= 5
REFINE_SELECT = (5.0, 64.0)
REFINE_TOP = 300  # picks refined per micrograph in the batch workload
REFINE_TOL = (0.05, 0.05, 0.1, 5.0)  # px, px, deg, Angstrom


def refine_leg(eng, cfg, ds, no_cpu, sample=16):
    """Adjoint refinement of the top REFINE_PEAKS picks of the search just run (re-selected at
    z >= 5 with the search's 64 px NMS, so that every pick is a distinct peak), each perturbed
    by U(+-0.5 px) in x and y, U(+-2.5 deg)
    in theta and U(+-300 A) in defocus, 20 ascent steps of 0.1 (scale 2 px, 2 px, 3 deg,
    500 A).  'at tol': a sample of the picks is refined from the same start by the FP64 oracle
    with the same frozen mu/sigma; within_tol is the fraction whose four parameters agree
    within (0.05 px, 0.05 px, 0.1 deg, 5 A)."""
    recs = eng.select(REFINE_SELECT[0], REFINE_SELECT[1])[:REFINE_PEAKS]
    pk = eng.peaks_from_records(recs)
    rng = np.random.default_rng(3000)
    k = len(pk)
    pk["dx"] = rng.uniform(-0.5, 0.5, k)
    pk["dy"] = rng.uniform(-0.5, 0.5, k)
    pk["theta"] += rng.uniform(-2.5, 2.5, k)
    pk["ddf"] = rng.uniform(-300.0, 300.0, k)
    eng.refine(pk[:16])  # warm (whitened template spectra, buffers)
    t0 = time.perf_counter()
    out = eng.refine(pk)
    dt = time.perf_counter() - t0
    dev_ms = eng.refine_timing()
    log(f"refinement: {k} picks in {dt:.3f} s (device {dev_ms:.1f} ms)")
    res = {
        "picks": k, "selection": f"z >= {REFINE_SELECT[0]}, NMS radius {REFINE_SELECT[1]} px",
        "min_z_selected": float(recs["score"].min()) if k else None, "iterations": 20, "peaks_per_s": k / dt, "wall_s": dt, "device_ms": dev_ms,
        "mean_z_start": float(out["score0"].mean()), "mean_z_refined": float(out["score"].mean()),
        "tolerance": list(REFINE_TOL),
    }
    if no_cpu:
        return res
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    oracle.load()
    idx = np.linspace(0, k - 1, sample).astype(int)
    mu_all = eng.stage(4, np.float64)
    sg_all = eng.stage(5, np.float64)
    l = cfg.window_size
    q = pk["block"][idx].astype(np.int64) * l * l + pk["pixel"][idx]
    opk = np.zeros(sample, dtype=oracle.PEAK_DTYPE)
    for n in pk.dtype.names:
        opk[n] = pk[n][idx]
    R = oracle.Refiner(cfg, ds.image, ds.templates, ds.ctf)
    R.float_geometry(True)
    t0 = time.perf_counter()
    ref, zo, _, _ = R.run(opk, mu_all[q], sg_all[q], iters=20, step=0.1, scale=(2.0, 2.0, 3.0, 500.0))
    cpu_dt = time.perf_counter() - t0
    R.close()
    dev = np.stack([np.abs(out["peak"][n][idx] - ref[n]) for n in ("dx", "dy", "theta", "ddf")], 1)
    ok = np.all(dev <= np.array(REFINE_TOL), axis=1)
    res.update({
        "oracle_sample": sample, "within_tol": float(ok.mean()),
        "max_deviation": [float(v) for v in dev.max(0)],
        "max_score_rel_dev": float(np.max(np.abs(out["score"][idx] - zo) / np.abs(zo))),
        "cpu_peaks_per_s": sample / cpu_dt, "cpu_cores": oracle.num_threads(),
    })
    log(f"refinement oracle sample: {ok.mean() * 100:.0f}% within tolerance, {cpu_dt:.1f} s on CPU")
    return res


def batch_specs(k, seed=2000):
    """K uneven micrographs as the batching configuration draws them: size from {3710x3838,
    4096x4096, 5760x4092}, N_t ~ U{200..1000} templates of a size drawn from {128, 180, 256},
    defocus U(0.8, 3.0) um and U{100..500} particles at SNR 0.05, each with its own template set."""
    rng = np.random.default_rng(seed)
    sizes = [(3710, 3838), (4096, 4096), (5760, 4092)]
    specs = []
    for i in range(k):
        W, H = sizes[int(rng.integers(3))]
        nt = int(rng.integers(200, 1001))
        ts = int(rng.choice([128, 180, 256]))
        df = float(rng.uniform(0.8, 3.0))
        npart = int(rng.integers(100, 501))
        specs.append((1000 + i, W, H, nt, ts, df, npart))
    return specs


def make_item(spec):
    """One micrograph of batch_specs: (item for the search API, dataset, description)."""
    import synth

    seed, W, H, nt, ts, df, npart = spec
    ds = synth.make_dataset(seed, W, H, n_templates=nt, tsize=ts, n_particles=npart, snr=0.05, defocus=df)
    item = {"image": ds.image, "ctf": ds.ctf, "templates": ds.templates}
    return item, ds, f"{W}x{H}/{nt}x{ts}/df{df:.2f}/{npart}p"


def _one_thread():
    # pool workers: one BLAS/FFT thread each, the pool supplies the parallelism
    from threadpoolctl import threadpool_limits

    threadpool_limits(1)


def batch_items(k, seed=2000, workers=1):
    """Build the K micrographs of batch_specs, in a process pool when workers > 1 (only before
    anything has used the GPU: the pool forks this process)."""
    specs = batch_specs(k, seed)
    if workers > 1:
        import multiprocessing as mp

        with mp.get_context("fork").Pool(min(workers, k), initializer=_one_thread) as pool:
            made = []
            for i, m in enumerate(pool.imap(make_item, specs)):
                made.append(m)
                log(f"micrograph {i}: {m[2]}")
    else:
        made = []
        for i, spec in enumerate(specs):
            made.append(make_item(spec))
            log(f"micrograph {i}: {made[-1][2]}")
    items = [m[0] for m in made]
    dss = [m[1] for m in made]
    desc = [m[2] for m in made]
    work = sum(nt * 72 * W * H for (_, W, H, nt, _, _, _) in specs)
    return items, work, desc, dss


def batch_refine_leg(k, no_cpu, sample=8):
    """Batched + refined form against the plain per-micrograph form on the same GPU: K uneven
    micrographs searched by gs_search_batch_refine (one engine, staging thread, graph replay,
    top REFINE_TOP picks of each micrograph refined right after its search) against the
    reference's driver loop followed by refinement (an engine per micrograph, gs_search, then
    gs_refine of its top REFINE_TOP picks).  A sample of refined picks of the last micrograph is
    checked against the FP64 oracle refiner from the same start."""
    import gisspa

    t0 = time.perf_counter()
    items, work, desc, dss = batch_items(k)
    log(f"batch data ready in {time.perf_counter() - t0:.1f} s")
    cfg = c1_config(gisspa)
    with gisspa.Engine(cfg) as e:  # warm the library on the first item
        e.load_templates(items[0]["templates"])
        e.set_limits(1, 0)
        e.search(items[0]["image"], items[0]["ctf"])
    with gisspa.Engine(cfg) as e:
        e.set_graph(True)
        t1 = time.perf_counter()
        recs_b, refs_b = e.search_batch_refine(items, REFINE_TOP)
        t_batch = time.perf_counter() - t1
    log(f"batch + refinement: {t_batch:.2f} s")
    t1 = time.perf_counter()
    refs_s, last = [], None
    for it in items:
        with gisspa.Engine(cfg) as e:
            e.load_templates(it["templates"])
            r = e.search(it["image"], it["ctf"])[:REFINE_TOP]
            pk = e.peaks_from_records(r) if len(r) else None
            refs_s.append(e.refine(pk) if len(r) else None)
            if it is items[-1] and len(r):
                last = (pk, e.stage(4, np.float64), e.stage(5, np.float64))
    t_seq = time.perf_counter() - t1
    log(f"per-micrograph search + refinement: {t_seq:.2f} s")
    same = all((x is None and len(y) == 0) or
               (x is not None and np.array_equal(x["score"], y["score"]) and np.array_equal(x["peak"], y["peak"]))
               for x, y in zip(refs_s, refs_b))
    n_ref = int(sum(len(r) for r in refs_b))
    res = {
        "micrographs": desc, "top_k": REFINE_TOP, "refined_picks": n_ref,
        "seconds_batched_refined": t_batch, "seconds_per_micrograph": t_seq,
        "micrographs_per_s": k / t_batch, "micrographs_per_s_per_micrograph_form": k / t_seq,
        "gain": t_seq / t_batch, "correlations_per_s_incl_refinement": work / t_batch,
        "refined_identical_to_per_micrograph_form": bool(same),
    }
    if not no_cpu and last is not None:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle

        oracle.load()
        pk, mu_all, sg_all = last
        idx = np.linspace(0, len(pk) - 1, min(sample, len(pk))).astype(int)
        l = cfg.window_size
        q = pk["block"][idx].astype(np.int64) * l * l + pk["pixel"][idx]
        opk = np.zeros(len(idx), dtype=oracle.PEAK_DTYPE)
        for n in pk.dtype.names:
            opk[n] = pk[n][idx]
        ds = dss[-1]
        R = oracle.Refiner(cfg, ds.image, ds.templates, ds.ctf)
        R.float_geometry(True)
        ref, zo, _, _ = R.run(opk, mu_all[q], sg_all[q], iters=20, step=0.1, scale=(2.0, 2.0, 3.0, 500.0))
        R.close()
        got = refs_b[-1]
        dev = np.stack([np.abs(got["peak"][n][idx] - ref[n]) for n in ("dx", "dy", "theta", "ddf")], 1)
        ok = np.all(dev <= np.array(REFINE_TOL), axis=1)
        res.update({"oracle_sample": int(len(idx)), "within_tol": float(ok.mean()),
                    "max_deviation": [float(v) for v in dev.max(0)],
                    "max_score_rel_dev": float(np.max(np.abs(got["score"][idx] - zo) / np.abs(zo)))})
        log(f"batch refinement oracle sample: {ok.mean() * 100:.0f}% within tolerance")
    return res


def c1_config(gisspa):
    return gisspa.SearchConfig(angpix=1.0, phistep=5.0, kk=3.0, energy=300.0, cs=2.7, highres=8.0,
                               lowres=100.0, diameter=128.0, threshold=7.0, window_size=320,
                               overlap=0, norm_type=1, invert=0, phase_flip=0, nms_radius=64.0)


def batch_main(a):
    """Uneven-micrograph batch: K micrographs with sizes drawn from {3710x3838, 4096x4096,
    5760x4092}, defocus U(0.8, 3.0) um, N_t ~ U{200..1000} templates of a size drawn from
    {128, 180, 256}, U{100..500} particles, each micrograph with its own template set.  Three ways
    to run them (the second only for K <= 8):
    `per_engine` = the reference's driver loop (an engine created, loaded, run and destroyed per
    micrograph), `one_engine` = one engine and gs_search per micrograph, `batch` =
    gs_search_batch (staging thread + graph replay).  value = correlations/s of `batch`.  The
    batch + refinement form (gs_search_batch_refine, top REFINE_TOP picks per micrograph) is
    compared with the same work run per micrograph (engine, gs_search, gs_refine)."""
    t0 = time.perf_counter()
    # the micrographs are synthesised in a process pool before this process touches the GPU
    items, work, desc, _ = batch_items(a.batch, workers=min(16, os.cpu_count() or 1))
    import gisspa

    log(f"batch data ready in {time.perf_counter() - t0:.1f} s")
    cfg = c1_config(gisspa)
    times = {}
    # warm the library (code objects, twiddle tables) on the first item
    with gisspa.Engine(cfg) as e:
        e.load_templates(items[0]["templates"])
        e.set_limits(1, 0)
        e.search(items[0]["image"], items[0]["ctf"])
    recs = {}
    t1 = time.perf_counter()
    recs["per_engine"] = []
    for i, it in enumerate(items):
        with gisspa.Engine(cfg) as e:
            e.load_templates(it["templates"])
            recs["per_engine"].append(e.search(it["image"], it["ctf"]))
        if i % 16 == 15:
            log(f"per-engine loop: {i + 1} micrographs in {time.perf_counter() - t1:.1f} s")
    times["per_engine"] = time.perf_counter() - t1
    log(f"per-engine loop: {times['per_engine']:.2f} s")
    if a.batch <= 8:  # a third form for small batches only (it adds a full pass over the work)
        t1 = time.perf_counter()
        recs["one_engine"] = []
        with gisspa.Engine(cfg) as e:
            for it in items:
                e.load_templates(it["templates"])
                recs["one_engine"].append(e.search(it["image"], it["ctf"]))
        times["one_engine"] = time.perf_counter() - t1
        log(f"one-engine loop: {times['one_engine']:.2f} s")
    with gisspa.Engine(cfg) as e:
        e.set_graph(True)
        t1 = time.perf_counter()
        recs["batch"] = e.search_batch(items)
        times["batch"] = time.perf_counter() - t1
        graphs = e.graph_stats()
    log(f"batch: {times['batch']:.2f} s, graph captures/launches {graphs}")
    # C4: the same batch with the top REFINE_TOP picks of every micrograph refined in place
    with gisspa.Engine(cfg) as e:
        e.set_graph(True)
        t1 = time.perf_counter()
        recs_c4, refs_c4 = e.search_batch_refine(items, REFINE_TOP)
        times["batch_refine"] = time.perf_counter() - t1
    n_ref = int(sum(len(r) for r in refs_c4))
    log(f"batch + refinement of {n_ref} picks: {times['batch_refine']:.2f} s")
    # the same work run sequentially per micrograph, as the plain search followed by the plain
    # refinement would be: an engine per micrograph, gs_search, then gs_refine of its top picks
    t1 = time.perf_counter()
    refs_seq = []
    for i, it in enumerate(items):
        with gisspa.Engine(cfg) as e:
            e.load_templates(it["templates"])
            r = e.search(it["image"], it["ctf"])[:REFINE_TOP]
            refs_seq.append(e.refine(e.peaks_from_records(r)) if len(r) else None)
        if i % 16 == 15:
            log(f"per-engine search + refinement: {i + 1} micrographs in {time.perf_counter() - t1:.1f} s")
    times["per_engine_refine"] = time.perf_counter() - t1
    log(f"per-engine search + refinement: {times['per_engine_refine']:.2f} s")
    seq_same = all((x is None and len(y) == 0) or
                   (x is not None and np.array_equal(x["score"], y["score"]) and
                    np.array_equal(x["peak"], y["peak"]))
                   for x, y in zip(refs_seq, refs_c4))
    same = all(
        len(x) == len(y) and all(np.array_equal(x[n], y[n]) for n in x.dtype.names)
        for k in ("one_engine", "batch") if k in recs for x, y in zip(recs["per_engine"], recs[k]))
    out = {
        "metric": "template·angle·pixel correlations/s over an uneven micrograph batch",
        "value": work / times["batch"],
        "unit": "correlations/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 1,
        "ms_per_step": times["batch"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded blob-density projections, CTF, N(0,1) noise)",
        "config": {"workload": f"batch of {a.batch} uneven micrographs, norm_type=1, window 320, "
                               "72 angles", "micrographs": desc},
        "seconds": times,
        "speedup_vs_per_engine": times["per_engine"] / times["batch"],
        "speedup_vs_one_engine": times["one_engine"] / times["batch"] if "one_engine" in times else None,
        "records_identical": bool(same),
        "graph_captures_launches": list(graphs),
        "batch_refine": {
            "top_k": REFINE_TOP, "refined_picks": n_ref, "seconds": times["batch_refine"],
            "refined_picks_per_s_marginal": n_ref / max(times["batch_refine"] - times["batch"], 1e-9),
            "records_identical": bool(all(
                len(x) == len(y) and all(np.array_equal(x[n], y[n]) for n in x.dtype.names)
                for x, y in zip(recs["batch"], recs_c4))),
            "mean_z_start": float(np.mean(np.concatenate([r["score0"] for r in refs_c4]))) if n_ref else None,
            "mean_z_refined": float(np.mean(np.concatenate([r["score"] for r in refs_c4]))) if n_ref else None,
            "micrographs_per_s": a.batch / times["batch_refine"],
            "micrographs_per_s_sequential": a.batch / times["per_engine_refine"],
            "speedup_vs_sequential": times["per_engine_refine"] / times["batch_refine"],
            "refined_identical_to_sequential": bool(seq_same),
        },
    }
    print(json.dumps(out), flush=True)


def c0_leg(steps, warmup, no_cpu):
    """norm_type=0 configuration: 1440x1440 micrograph with 40 particles at SNR 0.1, 8 templates
    of 128x128, 36 angles (phistep 10), window 720 (4 tiles of one block), phase_flip=1,
    defocus 1.5 um.  Rate = correlations/s = 8 * 36 * 1440^2 / s per search; without no_cpu the
    FP64 oracle runs the same search in full (CPU baseline and pick comparison)."""
    import gisspa
    import synth

    cfg = gisspa.SearchConfig(angpix=1.0, phistep=10.0, kk=3.0, energy=300.0, cs=2.7, highres=8.0,
                              lowres=100.0, diameter=128.0, threshold=7.0, window_size=720, overlap=0,
                              norm_type=0, invert=0, phase_flip=1, nms_radius=64.0)
    ds = synth.make_dataset(1000, 1440, 1440, n_templates=8, tsize=128, n_particles=40, snr=0.1, defocus=1.5)
    eng = gisspa.Engine(cfg)
    eng.load_templates(ds.templates)
    eng.set_image(ds.image, ds.ctf)
    for _ in range(warmup):
        eng.run()
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.run()
    dt = (time.perf_counter() - t0) / steps
    total_ms, corr_ms = eng.timing()
    recs = eng.records()
    eng.close()
    work = 8 * 36 * 1440 * 1440
    out = {"workload": "norm_type=0: 1440x1440, 8 templates 128x128, 36 angles, window 720, phase_flip=1",
           "correlations_per_s": work / dt, "ms_per_search": dt * 1e3, "steps": steps, "warmup": warmup,
           "picks": int(len(recs)), "pipeline_ms": {"device_total": total_ms, "tile_angle_loop": corr_ms}}
    if not no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle

        oracle.load()
        t0 = time.perf_counter()
        opicks, _ = oracle.search_nonorm(cfg, ds.image, ds.templates, ds.ctf)
        cdt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": work / cdt, "unit": "correlations/s", "cores": oracle.num_threads(),
                               "kind": "port", "sample": f"the whole search on the FP64 oracle ({cdt:.1f} s)"}
        out["picks_identical_to_oracle"] = bool(
            [(int(p["template_id"]), int(p["angle_index"]), int(p["block"])) for p in recs] ==
            [(int(p["template_id"]), int(p["angle_index"]), int(p["block"])) for p in opicks])
    return out


def c0_main(a):
    """The norm_type=0 configuration as its own JSON line (see c0_leg)."""
    leg = c0_leg(a.steps, a.warmup, a.no_cpu)
    out = {
        "metric": "template·angle·pixel correlations/s (norm_type=0 configuration)",
        "value": leg["correlations_per_s"], "unit": "correlations/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": leg["ms_per_search"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic (seeded blob-density projections, CTF, N(0,1) noise)",
        "config": {"workload": leg["workload"], "picks": leg["picks"]},
        "pipeline_ms": leg["pipeline_ms"],
    }
    if "cpu_baseline" in leg:
        out["cpu_baseline"] = leg["cpu_baseline"]
        out["config"]["picks_identical_to_oracle"] = leg["picks_identical_to_oracle"]
    print(json.dumps(out), flush=True)


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported "
              "(the search shards by micrograph or block with no exchange between GPUs)",
              file=sys.stderr)
        sys.exit(2)
    if a.batch:
        batch_main(a)
        return
    if a.c0:
        c0_main(a)
        return
    import gisspa
    import synth

    t0 = time.perf_counter()
    ds = synth.make_dataset(1001, 5760, 4092, n_templates=a.templates, tsize=180, n_particles=400,
                            snr=0.05, defocus=2.0, dfdiff=0.005, dfang=30.0, grid=True)
    log(f"dataset ready in {time.perf_counter() - t0:.1f} s")
    cfg = c1_config(gisspa)
    if a.reference_records:
        cfg.reference_records = 1
    eng = gisspa.Engine(cfg, device=0)
    eng.load_templates(ds.templates, ds.eulers[:, :2])
    eng.set_image(ds.image, ds.ctf)
    if a.reference_records:
        # the same search emitting the reference's records: both passes, default kernel options
        if a.angles:
            eng.set_limits(a.angles, 0)
        for _ in range(a.warmup):
            eng.run()
        t1 = time.perf_counter()
        for _ in range(a.steps):
            eng.run()
        dt = (time.perf_counter() - t1) / a.steps
        n_ang = a.angles if a.angles else int(round(360 / cfg.phistep))
        H, W = ds.image.shape
        print(json.dumps({
            "metric": "template·angle·pixel correlations/s (reference-records mode)",
            "value": a.templates * n_ang * W * H / dt, "unit": "correlations/s", "n_gpus": 1,
            "steps": a.steps, "warmup": a.warmup, "ms_per_step": dt * 1e3, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
            "data": "synthetic (seeded blob-density projections, CTF, N(0,1) noise)",
            "config": {"workload": "production-size search with reference_records = 1: per-pixel statistics "
                                   "pass, then the reference's per-(angle, block, template, 256-chunk) records",
                       "angles": n_ang, "records": int(len(eng.records()))},
        }), flush=True)
        eng.close()
        return
    if a.refine_only:
        eng.set_limits(a.angles or 8, 0)
        eng.run()
        log(f"search restricted to {a.angles or 8} angles done")
        print(json.dumps({"refine": refine_leg(eng, cfg, ds, True)}), flush=True)
        return
    if a.angles:
        eng.set_limits(a.angles, 0)
    # result-invariant kernel options: the library's autotuner on a one-angle sample, unless given
    if a.col_variant < 0 and a.row_variant < 0 and not a.chunk:
        t1 = time.perf_counter()
        col, row, chunk = eng.autotune()
        log(f"autotune: column variant {col}, row variant {row}, {chunk} templates per launch "
            f"({time.perf_counter() - t1:.1f} s)")
    else:
        # an option not given keeps the library default (column variant 12 with row variant 8,
        # the row-packed layout); row variants 7 and 8 read only that layout (written by column
        # variants 8 to 13), so another column variant defaults to row variant 6
        col = a.col_variant if a.col_variant >= 0 else 12
        row = a.row_variant if a.row_variant >= 0 else (8 if 8 <= col <= 13 else 6)
        chunk = a.chunk
        eng.set_variant(0, col)
        eng.set_variant(1, row)
        if chunk:
            eng.set_chunk(chunk)
    eng.set_variant(2, max(a.overlap, 0))
    variant = {0: col, 1: row, 2: max(a.overlap, 0), "chunk": chunk}
    for i in range(a.warmup):
        t1 = time.perf_counter()
        eng.run()
        log(f"warmup {i}: {time.perf_counter() - t1:.2f} s")
    eng.set_profiling(True)
    t1 = time.perf_counter()
    for i in range(a.steps):
        eng.run()
        log(f"step {i} done")
    dt = time.perf_counter() - t1
    recs = eng.records()
    col_ms, col_n = eng.kernel_timing(0)
    row_ms, row_n = eng.kernel_timing(1)
    total_ms, corr_ms = eng.timing()
    H, W = ds.image.shape
    n_ang = a.angles if a.angles else int(round(360 / cfg.phistep))
    work = a.templates * n_ang * W * H
    value = work * a.steps / dt
    l = cfg.window_size
    nb = 18 * 13
    kx = l // 2 + 1
    # planes per launch = templates per launch x blocks; launches per step = angles x chunks
    launches_per_step = col_n / a.steps
    planes_per_launch = a.templates * n_ang * nb / launches_per_step
    tc = planes_per_launch / nb
    # Algorithmic HBM bytes per launch, i.e. what the two-pass transform must move at least:
    # column pass = inter-pass buffer written once per plane + each template spectrum and each
    # block spectrum read once per launch (all blocks share a template spectrum, so its reads
    # per plane are cache traffic, not HBM); row pass = inter-pass buffer read once per plane +
    # one read-modify-write of the per-pixel records (sum, sum of squares: double; best: float;
    # argmax: int32) per launch.
    zbytes = 8 * l * (l // 2)
    col_bytes = planes_per_launch * zbytes + (tc + nb) * 8 * l * kx
    row_bytes = planes_per_launch * zbytes + nb * l * l * 24 * 2
    # measured HBM bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated), from
    # the committed counter summary of the same kernel variant, scaled to this launch's planes
    tpath = os.path.join(ROOT, "profiles", "corr_c1_traffic.json")
    traffic = {"column_pass": None, "row_pass": None}
    if os.path.exists(tpath) and l == 320:
        # demangled names, every template argument spelled out as rocprofv3 prints them
        col_name = {0: "k_colpass<320, 0, 0, 8, 16, false>", 3: "k_colpass<320, 1, 0, 8, 16, false>",
                    4: "k_colpass<320, 1, 1, 8, 16, false>", 5: "k_colpass<320, 0, 1, 8, 16, false>",
                    6: "k_colpass<320, 1, 2, 8, 16, false>", 7: "k_colpass<320, 1, 3, 8, 16, false>",
                    8: "k_colpass<320, 1, 4, 8, 16, false>", 12: "k_colpass<320, 1, 4, 4, 4, true>",
                    13: "k_colpass<320, 1, 4, 2, 8, true>"}.get(variant[0], "")
        row_name = "k_rowpass_pix_wave<320, %s, false>" % {
            0: "false, 4, false, false, false", 1: "true, 4, false, false, false",
            2: "true, 8, false, false, false", 3: "false, 8, false, false, false",
            4: "true, 4, true, false, false", 5: "false, 4, true, false, false",
            6: "true, 8, true, false, false", 7: "true, 8, true, true, false",
            8: "true, 8, true, true, true", 9: "true, 8, true, false, true"}.get(variant[1], "")
        recs_t = json.load(open(tpath))["kernels"]
        for key, kname in (("column_pass", col_name), ("row_pass", row_name)):
            rec = recs_t.get(kname)
            if rec:
                per_plane = (rec["fetch_bytes_per_launch"] + rec["write_bytes_per_launch"]) / (
                    rec["templates_per_launch"] * rec["blocks"])
                traffic[key] = per_plane * planes_per_launch
    kernels = {}
    for key, ms, nbytes in (("column_pass", col_ms, col_bytes), ("row_pass", row_ms, row_bytes)):
        bw = nbytes / (ms * 1e-3)
        kernels[key] = {
            "avg_launch_ms": ms, "algorithmic_bytes_per_launch": nbytes, "achieved_GBps": bw / 1e9,
            "frac": bw / HBM_PEAK, "traffic": traffic[key],
            "hbm_frac_counters": None if traffic[key] is None else traffic[key] / (ms * 1e-3) / HBM_PEAK,
        }
    dominant = "column_pass" if col_ms >= row_ms else "row_pass"
    dom = kernels[dominant]
    # work ratio of the whole correlation stage: 24 l (l/2 + 1) bytes per plane (template
    # spectrum read + inter-pass write + read) over the angle loop's wall time; the template
    # spectrum reads it counts per plane are mostly served by L2, so this is not an HBM fraction
    stage_bytes = a.templates * n_ang * nb * 24 * l * kx
    stage_bw = stage_bytes / (corr_ms * 1e-3)
    valu = None
    vpath = os.path.join(ROOT, "profiles", "corr_c1_valu.json")
    if os.path.exists(vpath):
        valu = json.load(open(vpath)).get("kernels")
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "correlations/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": dt * 1e3 / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded blob-density projections, CTF, N(0,1) noise)",
        "config": {
            "workload": "plain search, production size: 5760x4092 fp32 micrograph, 1000 templates "
                        "180x180, 72 in-plane angles, norm_type=1, window 320 (234 blocks)",
            "templates": a.templates, "angles": n_ang, "image": [W, H], "window_size": l,
            "blocks": nb, "picks": int(len(recs)), "column_pass_variant": variant[0],
            "row_pass_variant": variant[1], "pass_overlap": variant[2],
            "templates_per_launch": variant["chunk"] or None,
        },
        "roofline": {
            "bound": "hbm", "achieved": dom["achieved_GBps"], "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": dom["frac"], "traffic": dom["traffic"],
            "traffic_source": "profiles/corr_c1_traffic.json: rocprofv3 FETCH_SIZE and WRITE_SIZE of the "
                              "same kernel variant, copy-calibrated, scaled to this launch's planes",
            "kernel": dominant, "avg_launch_ms": dom["avg_launch_ms"],
            "algorithmic_bytes_per_launch": dom["algorithmic_bytes_per_launch"],
            "hbm_frac_counters": dom["hbm_frac_counters"],
            "kernels": kernels,
            "limiter": valu,
            "stage_work_ratio": stage_bw / HBM_PEAK,
        },
        "pipeline_ms": {"device_total": total_ms, "angle_loop": corr_ms},
        "copy_ceiling_GBps": gisspa.copy_bandwidth(1 << 30, 20),
    }
    if not a.angles and a.templates == 1000:
        out["refine"] = refine_leg(eng, cfg, ds, a.no_cpu)
    eng.close()
    if not a.angles and a.templates == 1000 and a.batch_refine > 0:
        out["batch_refine"] = batch_refine_leg(a.batch_refine, a.no_cpu)
    if not a.angles and a.templates == 1000 and a.c0_leg:
        out["c0"] = c0_leg(20, 3, a.no_cpu)
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(cfg, ds, a.cpu_templates, a.cpu_angles)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
