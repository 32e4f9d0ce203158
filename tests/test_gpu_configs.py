"""GPU parity at the benchmark configurations' real shapes.

* C0 (norm_type=0): 1440x1440 micrograph, 40 particles at SNR 0.1, 8 templates of 128^2, 36
  angles, window 720 (4 tiles, 1 block each), phase_flip=1 -- run in full on both sides.
* C1 frame (norm_type=1): the full 5760x4092 micrograph (whole-frame FFT of 4092 = 2^2*3*11*31
  rows, 18 x 13 = 234 blocks with the clamped last block row at 3772), restricted to the first 8
  templates and 8 angles so the FP64 oracle finishes in seconds; mu and sigma then run over the
  same restricted hypothesis set on both sides.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _key(a):
    return [(int(p["x"]), int(p["y"]), int(p["template_id"]), int(p["angle_index"])) for p in a]


def test_c0_nonorm_full_config(gs, orc):
    import synth

    cfg = gs.SearchConfig(angpix=1.0, phistep=10.0, kk=3.0, energy=300.0, cs=2.7, highres=8.0,
                          lowres=100.0, diameter=128.0, threshold=7.0, window_size=720, overlap=0,
                          norm_type=0, invert=0, phase_flip=1, nms_radius=64.0)
    ds = synth.make_dataset(1000, 1440, 1440, n_templates=8, tsize=128, n_particles=40, snr=0.1,
                            defocus=1.5)
    opicks, orecs = orc.search_nonorm(cfg, ds.image, ds.templates, ds.ctf)
    eng = gs.Engine(cfg)
    eng.load_templates(ds.templates)
    g = eng.search(ds.image, ds.ctf)
    assert len(orecs) > 0
    # every (tile, angle, template) record above threshold, then the NMS pick list
    assert len(g) == len(opicks)
    assert [(int(p["template_id"]), int(p["angle_index"]), int(p["block"])) for p in g] == \
        [(int(p["template_id"]), int(p["angle_index"]), int(p["block"])) for p in opicks]
    np.testing.assert_allclose(g["score"], opicks["score"], rtol=1e-4)
    np.testing.assert_allclose(g["x"], opicks["x"], atol=1e-3)
    np.testing.assert_allclose(g["y"], opicks["y"], atol=1e-3)
    assert np.array_equal(g["pixel"], opicks["pixel"])


def test_c1_frame_norm_restricted(gs, orc):
    import synth

    cfg = gs.SearchConfig(angpix=1.0, phistep=5.0, kk=3.0, energy=300.0, cs=2.7, highres=8.0,
                          lowres=100.0, diameter=128.0, threshold=4.5, window_size=320, overlap=0,
                          norm_type=1, invert=0, phase_flip=0, nms_radius=64.0)
    # 8 templates x 8 angles: z over 64 hypotheses can exceed the threshold (max z < sqrt(63))
    ds = synth.make_dataset(1001, 5760, 4092, n_templates=8, tsize=180, n_particles=60, snr=0.3,
                            defocus=2.0, dfdiff=0.005, dfang=30.0, grid=True)
    na = 8
    opicks, ost = orc.search_norm(cfg, ds.image, ds.templates, ds.ctf, want_stats=True, n_angles=na)
    eng = gs.Engine(cfg)
    eng.load_templates(ds.templates)
    eng.set_limits(na, 0)
    g = eng.search(ds.image, ds.ctf)
    sd = eng.stage(5, np.float64)
    best = eng.stage(6, np.float32).astype(np.float64)
    mu = eng.stage(4, np.float64)
    assert sd.size == 234 * 320 * 320
    z_o = (ost["best"] - ost["mu"]) / ost["sigma"]
    z_g = (best - mu) / sd
    assert np.max(np.abs(z_g - z_o) / np.maximum(1.0, np.abs(z_o))) < 1e-4
    assert len(opicks) > 0
    assert _key(g) == _key(opicks)
    np.testing.assert_allclose(g["score"], opicks["score"], rtol=1e-4)
