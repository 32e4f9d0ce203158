"""Engines on separate host threads (the C ABI's contract: an engine belongs to one thread and
separate engines may run concurrently, like the reference's one engine per thread per device,
src/main.cu:58-76): concurrent searches on one GPU give the records of sequential runs, and the
last-error text is per thread."""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _same(a, b):
    return len(a) == len(b) and all(np.array_equal(a[n], b[n]) for n in a.dtype.names)


def test_concurrent_engines_match_sequential_runs(gs):
    import synth

    cfg = gs.SearchConfig(phistep=15.0, window_size=320, norm_type=1, nms_radius=64.0, threshold=6.0)
    data = [synth.make_dataset(60 + k, 640, 640, n_templates=3, tsize=128, n_particles=3, snr=0.6,
                               defocus=1.4 + 0.2 * k) for k in range(3)]

    def run(ds):
        with gs.Engine(cfg) as e:
            e.load_templates(ds.templates)
            return e.search(ds.image, ds.ctf)

    seq = [run(ds) for ds in data]
    assert all(len(r) > 0 for r in seq)
    out = [None] * len(data)
    errs = []

    def worker(k):
        try:
            for _ in range(2):
                out[k] = run(data[k])
        except Exception as ex:  # reported below, on the main thread
            errs.append(repr(ex))

    threads = [threading.Thread(target=worker, args=(k,)) for k in range(len(data))]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=240)
    assert not any(t.is_alive() for t in threads), "a search thread did not finish"
    assert not errs, errs
    for k in range(len(data)):
        assert _same(out[k], seq[k]), k


def test_last_error_is_per_thread(gs):
    """Two threads fail at the same time with different errors; afterwards each still reads its
    own message from gs_last_error."""
    cfg = gs.SearchConfig(window_size=320, norm_type=1)
    lib = gs.load()
    seen = {}

    def fail(tag, barrier):
        with gs.Engine(cfg) as e:
            barrier.wait()
            try:
                if tag == "a":
                    e.search(np.zeros((640, 640), np.float32), (1.0, 0.0, 0.0))  # no templates loaded
                else:
                    e.set_chunk(-5)  # invalid argument
            except gs.GisspaError:
                seen[tag] = lib.gs_last_error().decode()
            barrier.wait()
            seen[tag + "_after"] = lib.gs_last_error().decode()

    b = threading.Barrier(2)
    ts = [threading.Thread(target=fail, args=(t, b)) for t in ("a", "b")]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=120)
    assert "a" in seen and "b" in seen and seen["a"] != seen["b"], seen
    assert seen["a_after"] == seen["a"] and seen["b_after"] == seen["b"], seen
