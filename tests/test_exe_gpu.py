"""End-to-end test of the config-driven executable on the GPU: MRC micrographs and template
stack, LST and Euler files, and a reference-style config go in; the record lines that come out
must be the oracle's picks in the reference's line format (src/norm.cu:610-615)."""
import os
import re
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "gisspa_v2.0_amd", "lib", "gisspa_main")
LINE = re.compile(r"^(\d+)\t(\S+)\tdefocus=([-\d.]+)\tdfdiff=([-\d.]+)\tdfang=([-\d.]+)\t"
                  r"euler=([-\d.]+),([-\d.]+),([-\d.]+)\tcenter=([-\d.]+),([-\d.]+)\tscore=([-\d.]+)$")


def test_executable_matches_oracle(tmp_path, orc, gs):
    import synth

    cfg = gs.SearchConfig(angpix=1.0, phistep=10.0, kk=3.0, energy=300.0, cs=2.7, highres=8.0,
                          lowres=100.0, diameter=128.0, threshold=7.0, window_size=320, norm_type=1,
                          nms_radius=64.0)
    mics = []
    for k, seed in enumerate((21, 22)):
        ds = synth.make_dataset(seed, 640, 640, n_templates=4, tsize=128, n_particles=3, snr=0.6,
                                defocus=1.2 + 0.3 * k, dfdiff=0.01, dfang=20.0 * k)
        path = tmp_path / f"mic_{k}.mrc"
        synth.write_mrc(str(path), ds.image)
        mics.append((path, ds))
    templates = mics[0][1].templates  # same seed-independent density and template orientations?
    eul = mics[0][1].eulers
    synth.write_mrc(str(tmp_path / "templates.mrc"), templates)
    (tmp_path / "euler.txt").write_text("".join(f"{j} {a:.4f} {b:.4f} 0.0\n" for j, (a, b, _) in enumerate(eul)))
    (tmp_path / "list.lst").write_text("".join(
        f"{k} {p} defocus={d.ctf[0]:.6f} dfdiff={d.ctf[1]:.6f} dfang={d.ctf[2]:.6f}\n"
        for k, (p, d) in enumerate(mics)))
    (tmp_path / "config.txt").write_text(
        "input = list.lst\ntemplate = templates.mrc\neulerfile = euler.txt\nangpix = 1.0\n"
        "phistep = 10\nkk = 3\nenergy = 300\ncs = 2.7\nHighres = 8\nLowres = 100\n"
        "diameter = 128\nthreshold = 7.0\noutput = out.txt\nfirst = 0\nlast = 2\nGPU_ID = 0\n"
        "window_size = 320\nnorm_type = 1\nnms_radius = 64\n")
    r = subprocess.run([EXE, "config.txt"], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    lines = (tmp_path / "out.txt").read_text().splitlines()
    got = [LINE.match(l).groups() for l in lines]
    expect = []
    for k, (p, d) in enumerate(mics):
        picks, _ = orc.search_norm(cfg, d.image, templates, d.ctf)
        for q in picks:
            j = int(q["template_id"])
            expect.append((k, str(p), j, float(q["psi"]), float(q["x"]), float(q["y"]), float(q["score"])))
    assert len(got) == len(expect) and len(expect) > 0
    for g, e in zip(got, expect):
        k, path, j, psi, x, y, s = e
        assert int(g[0]) == k and g[1] == path
        assert abs(float(g[5]) - eul[j][0]) < 1e-3 and abs(float(g[6]) - eul[j][1]) < 1e-3
        assert float(g[7]) == pytest.approx(psi, abs=1e-6)
        assert float(g[8]) == x and float(g[9]) == y
        assert abs(float(g[10]) - s) <= 1e-4 * abs(s) + 1e-6


def test_executable_reads_hdf5_templates_and_stack_sections(tmp_path, gs):
    """The same search from other file layouts: templates in an EMAN2 HDF5 stack and both
    micrographs in one MRC stack, each read at its LST index (src/utils/image.cpp:9).  The
    record lines must equal those of one-file-per-micrograph MRC inputs."""
    import hdf5_fixture
    import synth

    ds = synth.make_dataset(31, 640, 640, n_templates=3, tsize=128, n_particles=3, snr=0.6,
                            defocus=1.3, dfdiff=0.01, dfang=10.0)
    templates, eul = ds.templates, ds.eulers
    # two micrographs of the same particles: the image and its 180-degree rotation
    images = [ds.image, np.ascontiguousarray(np.rot90(ds.image, 2))]
    (tmp_path / "euler.txt").write_text("".join(f"{j} {a:.4f} {b:.4f} 0.0\n" for j, (a, b, _) in enumerate(eul)))
    common = ("eulerfile = euler.txt\nangpix = 1.0\nphistep = 10\nkk = 3\nenergy = 300\ncs = 2.7\n"
              "Highres = 8\nLowres = 100\ndiameter = 128\nthreshold = 7.0\nfirst = 0\nlast = 2\n"
              "GPU_ID = 0\nwindow_size = 320\nnorm_type = 1\nnms_radius = 64\n")
    ctf = "defocus=1.300000 dfdiff=0.010000 dfang=10.000000"
    # reference layout: one MRC per micrograph, MRC template stack
    synth.write_mrc(str(tmp_path / "templates.mrc"), templates)
    for k, im in enumerate(images):
        synth.write_mrc(str(tmp_path / f"m{k}.mrc"), im)
    (tmp_path / "a.lst").write_text("".join(f"0 m{k}.mrc {ctf}\n" for k in range(2)))
    (tmp_path / "a.txt").write_text("input = a.lst\ntemplate = templates.mrc\noutput = a.out\n" + common)
    # other layout: EMAN2 HDF5 templates, one two-section MRC read at LST indices 0 and 1
    hdf5_fixture.write(tmp_path / "templates.hdf", templates, eman1=False)
    synth.write_mrc(str(tmp_path / "stack.mrc"), np.stack(images))
    (tmp_path / "b.lst").write_text("".join(f"{k} stack.mrc {ctf}\n" for k in range(2)))
    (tmp_path / "b.txt").write_text("input = b.lst\ntemplate = templates.hdf\noutput = b.out\n" + common)
    for c in ("a.txt", "b.txt"):
        r = subprocess.run([EXE, c], cwd=tmp_path, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr
    a = [LINE.match(l).groups() for l in (tmp_path / "a.out").read_text().splitlines()]
    b = [LINE.match(l).groups() for l in (tmp_path / "b.out").read_text().splitlines()]
    assert len(a) > 0 and len(a) == len(b)
    for ga, gb in zip(a, b):  # identical except the micrograph index and path columns
        assert ga[2:] == gb[2:]
    assert [g[0] for g in b] == sorted(g[0] for g in b) and {g[0] for g in b} == {"0", "1"}


def test_executable_resumes_from_its_manifest(tmp_path, gs):
    """A run cut short and resumed with resume = 1 writes the same output file as one whole run:
    finished micrographs are skipped, and lines past the last finished one are cut away."""
    import synth

    mics = []
    for k, seed in enumerate((31, 32, 33)):
        ds = synth.make_dataset(seed, 640, 640, n_templates=4, tsize=128, n_particles=3, snr=0.6,
                                defocus=1.2 + 0.2 * k)
        path = tmp_path / f"mic_{k}.mrc"
        synth.write_mrc(str(path), ds.image)
        mics.append((path, ds))
    synth.write_mrc(str(tmp_path / "templates.mrc"), mics[0][1].templates)
    (tmp_path / "euler.txt").write_text("".join(f"{j} {a:.4f} {b:.4f} 0.0\n"
                                               for j, (a, b, _) in enumerate(mics[0][1].eulers)))
    (tmp_path / "list.lst").write_text("".join(
        f"{k} {p} defocus={d.ctf[0]:.6f} dfdiff={d.ctf[1]:.6f} dfang={d.ctf[2]:.6f}\n"
        for k, (p, d) in enumerate(mics)))

    def run(last, resume, out):
        (tmp_path / "config.txt").write_text(
            "input = list.lst\ntemplate = templates.mrc\neulerfile = euler.txt\nangpix = 1.0\n"
            "phistep = 10\nkk = 3\nenergy = 300\ncs = 2.7\nHighres = 8\nLowres = 100\n"
            f"diameter = 128\nthreshold = 6.0\noutput = {out}\nfirst = 0\nlast = {last}\nGPU_ID = 0\n"
            f"window_size = 320\nnorm_type = 1\nnms_radius = 64\nresume = {resume}\n")
        r = subprocess.run([EXE, "config.txt"], cwd=tmp_path, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr
        return r.stdout

    run(3, 0, "whole.txt")
    whole = (tmp_path / "whole.txt").read_text()
    assert len((tmp_path / "whole.txt.done").read_text().splitlines()) == 3
    assert len(whole.splitlines()) > 0

    run(1, 0, "part.txt")
    first = (tmp_path / "part.txt").read_text()
    assert whole.startswith(first)
    with open(tmp_path / "part.txt", "a") as f:  # lines of a micrograph that was being written
        f.write("1\tmic_1.mrc\tdefocus=1.4")
    out = run(3, 1, "part.txt")
    assert "mic_0.mrc: done in an earlier run, skipped" in out
    assert (tmp_path / "part.txt").read_text() == whole
    out = run(3, 1, "part.txt")  # nothing left to do
    assert out.count("skipped") == 3 and (tmp_path / "part.txt").read_text() == whole
