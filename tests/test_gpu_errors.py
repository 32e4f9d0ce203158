"""Error paths of the C ABI on a device: each bad call returns a status with a message (never
exits or throws across the ABI, unlike the reference's printf + exit(-1), src/utils/utils.h:24-38),
leaves the engine usable, and a following valid search gives the records a fresh engine gives.
Empty results (no pick above threshold) are an empty record list, not an error."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def case(gs):
    import synth

    cfg = gs.SearchConfig(angpix=1.0, phistep=30.0, kk=3.0, energy=300.0, cs=2.7, highres=8.0,
                          lowres=100.0, diameter=64.0, threshold=4.0, window_size=320, overlap=0,
                          norm_type=1, invert=0, phase_flip=0, nms_radius=32.0)
    ds = synth.make_dataset(3301, 640, 640, n_templates=4, tsize=64, n_particles=4, snr=0.8,
                            defocus=1.5, diameter=64.0)
    with gs.Engine(cfg) as e:
        e.load_templates(ds.templates)
        ref = e.search(ds.image, ds.ctf)
    assert len(ref) > 0
    return cfg, ds, ref


def _same(a, b):
    return len(a) == len(b) and all(np.array_equal(a[n], b[n]) for n in a.dtype.names)


def _still_works(eng, ds, ref):
    eng.load_templates(ds.templates)
    assert _same(eng.search(ds.image, ds.ctf), ref)


def test_bad_calls_leave_the_engine_usable(gs, case):
    cfg, ds, ref = case
    eng = gs.Engine(cfg)
    print("run before any templates", flush=True)
    with pytest.raises(gs.GisspaError, match="no templates") as ei:
        eng.run()
    assert ei.value.status != 0
    print("templates larger than the window", flush=True)
    with pytest.raises(gs.GisspaError, match="larger than window_size"):
        eng.load_templates(np.zeros((2, 330, 330), np.float32))
    eng.load_templates(ds.templates)
    print("run before any image", flush=True)
    with pytest.raises(gs.GisspaError, match="no image"):
        eng.run()
    print("image smaller than the window", flush=True)
    with pytest.raises(gs.GisspaError, match="at least window_size"):
        eng.set_image(np.zeros((300, 640), np.float32), ds.ctf)
    print("non-finite pixels", flush=True)
    for v in (np.nan, np.inf, -np.inf, 3e19):
        img = ds.image.copy()
        img[17, 301] = v
        with pytest.raises(gs.GisspaError, match="non-finite") as ei:
            eng.set_image(img, ds.ctf)
        assert ei.value.status == 1
    print("unknown kernel variants and chunk", flush=True)
    with pytest.raises(gs.GisspaError, match="unknown kernel variant"):
        eng.set_variant(1, 10)
    with pytest.raises(gs.GisspaError, match="unknown kernel variant"):
        eng.set_variant(0, -1)
    with pytest.raises(gs.GisspaError):
        eng.set_chunk(-3)
    print("valid search after the errors", flush=True)
    _still_works(eng, ds, ref)
    print("refinement of a pick outside the blocks", flush=True)
    pk = eng.peaks_from_records(ref[:1])
    bad = pk.copy()
    bad["block"] = 10_000
    with pytest.raises(gs.GisspaError):
        eng.refine(bad)
    bad = pk.copy()
    bad["dx"] = np.nan
    with pytest.raises(gs.GisspaError):
        eng.refine(bad)
    out = eng.refine(pk)
    assert np.isfinite(out["score"]).all()
    _still_works(eng, ds, ref)
    eng.close()


def test_batch_is_validated_before_any_work(gs, case):
    cfg, ds, ref = case
    eng = gs.Engine(cfg)
    good = {"image": ds.image, "ctf": ds.ctf, "templates": ds.templates}
    bad = {"image": ds.image, "ctf": ds.ctf, "templates": np.zeros((2, 400, 400), np.float32)}
    print("batch with an oversized template set in its second item", flush=True)
    with pytest.raises(gs.GisspaError, match="window_size"):
        eng.search_batch([good, bad])
    print("valid batch after the error", flush=True)
    out = eng.search_batch([good, good])
    assert _same(out[0], ref) and _same(out[1], ref)
    eng.close()


def test_no_pick_above_threshold_is_an_empty_list(gs, case):
    cfg, ds, ref = case
    hi = gs.SearchConfig(**{**cfg.__dict__, "threshold": 1e6})
    with gs.Engine(hi) as e:
        e.load_templates(ds.templates)
        recs = e.search(ds.image, ds.ctf)
        assert len(recs) == 0 and recs.dtype == gs.RECORD_DTYPE
        # the statistics of the run are there all the same: a lower threshold re-selects the picks
        again = e.select(cfg.threshold, cfg.nms_radius)
    assert _same(again, ref)
