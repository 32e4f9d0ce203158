"""Batching: gs_search_batch, graph replay of the angle loop, and the autotuner.

Batching is pinned by record-for-record equality with the plain per-micrograph form (one
gs_search per micrograph on a fresh engine, graphs off): the batch runs the same kernels in the
same order, so every field must match exactly.  The micrographs are uneven on purpose
(different sizes, block counts, CTFs, template counts and template sizes).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cfg(gs, threshold=5.0):
    return gs.SearchConfig(angpix=1.0, phistep=10.0, kk=3.0, energy=300.0, cs=2.7, highres=8.0,
                           lowres=100.0, diameter=128.0, threshold=threshold, window_size=320,
                           overlap=0, norm_type=1, invert=0, phase_flip=0, nms_radius=64.0)


def _micrographs():
    import synth

    a = synth.make_dataset(2001, 1280, 960, n_templates=6, tsize=128, n_particles=6, snr=0.5,
                           defocus=1.2)
    b = synth.make_dataset(2002, 960, 704, n_templates=6, tsize=128, n_particles=4, snr=0.5,
                           defocus=2.4, dfdiff=0.01, dfang=45.0)
    c = synth.make_dataset(2003, 1088, 1024, n_templates=9, tsize=180, n_particles=5, snr=0.5,
                           defocus=0.9)
    return a, b, c


def _same(x, y):
    assert x.dtype == y.dtype and len(x) == len(y)
    for name in x.dtype.names:
        assert np.array_equal(x[name], y[name]), name


def test_batch_equals_per_micrograph_search(gs):
    a, b, c = _micrographs()
    cfg = _cfg(gs)
    # plain form: one engine, one gs_search per micrograph, template set reloaded when it changes
    plain = []
    eng = gs.Engine(cfg)
    eng.load_templates(a.templates)
    plain.append(eng.search(a.image, a.ctf))
    plain.append(eng.search(b.image, b.ctf))
    eng.load_templates(c.templates)
    plain.append(eng.search(c.image, c.ctf))
    eng.close()
    assert sum(len(p) for p in plain) > 0

    plain.insert(1, plain[0])  # micrograph a twice
    for graph in (False, True):
        eb = gs.Engine(cfg)
        eb.set_graph(graph)
        got = eb.search_batch([
            {"image": a.image, "ctf": a.ctf, "templates": a.templates},
            {"image": a.image, "ctf": a.ctf},
            {"image": b.image, "ctf": b.ctf},  # keeps the template set of the first item
            {"image": c.image, "ctf": c.ctf, "templates": c.templates},
        ])
        assert len(got) == 4
        for g, p in zip(got, plain):
            _same(g, p)
        # only the shape that occurs twice (a) is captured: once, then replayed; b and c, which
        # occur once, run eagerly
        caps, launches = eb.graph_stats()
        assert (caps, launches) == ((1, 2) if graph else (0, 0))
        eb.close()


def test_batch_refine_equals_search_then_refine(gs):
    """gs_search_batch_refine = per micrograph: gs_search, then gs_refine of its top picks."""
    a, b, _ = _micrographs()
    cfg = _cfg(gs)
    eng = gs.Engine(cfg)
    eng.load_templates(a.templates)
    plain, prefs = [], []
    for ds in (a, b):
        r = eng.search(ds.image, ds.ctf)
        plain.append(r)
        prefs.append(eng.refine(eng.peaks_from_records(r[:3]), iterations=5))
    eng.close()
    eb = gs.Engine(cfg)
    recs, refs = eb.search_batch_refine([{"image": a.image, "ctf": a.ctf, "templates": a.templates},
                                         {"image": b.image, "ctf": b.ctf}], 3, iterations=5)
    eb.close()
    assert len(recs) == 2 and len(refs) == 2
    for x, y in zip(recs, plain):
        _same(x, y)
    for x, y in zip(refs, prefs):
        assert len(x) == len(y) > 0
        for n in ("score", "score0", "grad"):
            assert np.array_equal(x[n], y[n]), n
        for n in x["peak"].dtype.names:
            assert np.array_equal(x["peak"][n], y["peak"][n]), n


def test_graph_replays_match_eager_runs(gs):
    """The captured angle loop, replayed several times with other work in between (a search of
    another shape, a standalone transform, host synchronisation), gives the eager result every
    time, including the per-pixel statistics."""
    a, b, _ = _micrographs()
    cfg = _cfg(gs)
    eng = gs.Engine(cfg)
    eng.load_templates(a.templates)
    ref_a = eng.search(a.image, a.ctf)
    ref_a_best = eng.stage(6, np.float32)
    ref_a_sigma = eng.stage(5, np.float64)
    ref_b = eng.search(b.image, b.ctf)
    eng.set_graph(True)
    rng = np.random.default_rng(5)
    x = (rng.standard_normal((4, 64, 64)) + 1j * rng.standard_normal((4, 64, 64))).astype(np.complex64)
    for rep in range(3):
        g = eng.search(a.image, a.ctf)
        _same(g, ref_a)
        assert np.array_equal(eng.stage(6, np.float32), ref_a_best)
        assert np.array_equal(eng.stage(5, np.float64), ref_a_sigma)
        # other work between replays: a search of a different (smaller) shape and a transform
        _same(eng.search(b.image, b.ctf), ref_b)
        y = gs.fft2_c2c(x, -1)
        np.testing.assert_allclose(y, np.fft.fft2(x), rtol=0, atol=1e-3 * np.abs(np.fft.fft2(x)).max())
    caps, launches = eng.graph_stats()
    assert caps == 2 and launches == 6
    # turning graphs off drops them; the eager path still gives the same records
    eng.set_graph(False)
    _same(eng.search(a.image, a.ctf), ref_a)
    eng.close()


def test_autotune_is_result_invariant(gs):
    a, _, _ = _micrographs()
    cfg = _cfg(gs)
    eng = gs.Engine(cfg)
    eng.load_templates(a.templates)
    ref = eng.search(a.image, a.ctf)
    col, row, chunk = eng.autotune()
    assert (col, row) in ((12, 8), (7, 6), (3, 1)) and chunk in (64, 128, 256, 512)
    got = eng.search(a.image, a.ctf)
    # the options change only the kernel decomposition (FFT radices, rows per wave) and the
    # grouping of the per-pixel sums, i.e. rounding: scores within the parity bar, same picks
    key = lambda r: [(int(p["x"]), int(p["y"]), int(p["template_id"]), int(p["angle_index"])) for p in r]
    assert key(got) == key(ref)
    np.testing.assert_allclose(got["score"], ref["score"], rtol=1e-4)
    eng.close()
