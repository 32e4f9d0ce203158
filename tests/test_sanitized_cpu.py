"""The host readers and the FP64 oracle built with AddressSanitizer and UndefinedBehaviorSanitizer
(g++ -fsanitize=address,undefined -fno-sanitize-recover=all; CPU only), run on the format
fixtures, on damaged inputs that must end in a clean error, and on small oracle calls.  Any
sanitizer finding aborts the process with a report on stderr, which every check here looks for.

Damaged inputs are the failures the reference meets with exit(-1) or a silent substitute image
(src/EMReader/emdata.cpp:193-226, src/EMReader/DataReader2.cpp:13-53): a missing file, an MRC
header cut short, data cut short, dimensions far beyond the file, a negative dimension, an
unsupported mode, a stack shorter than asked for, a config value that is not a number and a
config line without '='.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

import hdf5_fixture
from test_formats_and_synth import _write_mrc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN_EXE = os.path.join(ROOT, "gisspa_v2.0_amd", "lib", "san", "gisspa_inspect")
EXE = os.path.join(ROOT, "gisspa_v2.0_amd", "lib", "gisspa_main")
SAN_ORACLE = os.path.join(ROOT, "oracle", "lib", "san", "libgisspa_oracle.so")
GOLDEN = os.path.join(ROOT, "tests", "golden")
REPORT_MARKS = ("AddressSanitizer", "runtime error:", "LeakSanitizer", "UndefinedBehaviorSanitizer")


def _run(args, env=None, cwd=None):
    r = subprocess.run([SAN_EXE] + [str(a) for a in args], capture_output=True, text=True, env=env,
                       cwd=cwd, timeout=120)
    for m in REPORT_MARKS:
        assert m not in r.stderr, r.stderr[-4000:]
    return r


@pytest.fixture(scope="module")
def san_exe():
    if not os.path.exists(SAN_EXE):
        pytest.skip("sanitizer build of the host readers absent (make -C gisspa_v2.0_amd san)")
    return SAN_EXE


def test_parse_fixtures_under_sanitizers(san_exe):
    r = _run(["--parse-dump", "parse_config.txt"], cwd=GOLDEN)
    assert r.returncode == 0, r.stderr
    assert "config angpix" in r.stdout and r.stdout.count("\nlst ") + r.stdout.startswith("lst ") >= 1
    if os.path.exists(EXE):  # the same dump as the executable's own build
        ref = subprocess.run([EXE, "--parse-dump", "parse_config.txt"], capture_output=True, text=True,
                             cwd=GOLDEN, check=True)
        assert r.stdout == ref.stdout


@pytest.mark.parametrize("mode,big", [(0, False), (1, False), (2, False), (6, False), (2, True), (1, True)])
def test_mrc_modes_under_sanitizers(san_exe, tmp_path, mode, big):
    rng = np.random.default_rng(100 + mode + big)
    a = (rng.standard_normal((3, 11, 19)) * 100).astype(np.float32)
    if mode != 2:
        a = np.clip(a, {0: -128, 1: -32768, 6: 0}[mode], {0: 127, 1: 32767, 6: 65535}[mode]).astype(np.int64)
    p = tmp_path / "m.mrc"
    _write_mrc(p, a, mode, big)
    r = _run(["--mrc-info", p])
    assert r.returncode == 0 and r.stdout.split()[:4] == ["mrc", "19", "11", "3"], r.stderr
    for idx in (0, 2, 5):
        r = _run(["--image-info", p, idx])
        assert r.returncode == 0 and r.stdout.split()[:3] == ["image", "19", "11"], r.stderr
    r = _run(["--stack-info", p, 3])
    assert r.returncode == 0, r.stderr


@pytest.mark.parametrize("eman1", [False, True])
def test_hdf5_stacks_under_sanitizers(san_exe, tmp_path, eman1):
    stack = np.random.default_rng(7).standard_normal((3, 9, 13)).astype(np.float32)
    p = tmp_path / "s.hdf"
    hdf5_fixture.write(p, stack, eman1)
    for i in range(3):
        r = _run(["--image-info", p, i])
        assert r.returncode == 0 and r.stdout.startswith("image "), r.stderr
    r = _run(["--stack-info", p, 4])
    assert r.returncode != 0 and "no image 3" in r.stderr


def _header(nx, ny, nz, mode, ext=0):
    h = np.zeros(256, dtype="<i4")
    h[0:4] = (nx, ny, nz, mode)
    h[23] = ext
    return h.tobytes()


DAMAGED = {
    "missing": (None, ["--mrc-info"], "cannot open"),
    "short header": (b"\x00" * 300, ["--mrc-info"], "truncated MRC header"),
    "short data": (_header(19, 11, 3, 2) + b"\x00" * (19 * 11 * 4 * 2), ["--mrc-info"], "truncated MRC data"),
    "huge dims": (_header(1 << 30, 1 << 30, 1000, 2) + b"\x00" * 64, ["--mrc-info"], "truncated MRC data"),
    "huge section": (_header(2147483647, 2147483647, 1, 2) + b"\x00" * 64, ["--image-info"], "truncated MRC data"),
    "ext beyond file": (_header(4, 4, 1, 2, ext=1 << 30) + b"\x00" * 64, ["--mrc-info"], "truncated MRC data"),
    "negative dim": (_header(19, -11, 1, 2) + b"\x00" * 4096, ["--mrc-info"], "bad MRC dimensions"),
    "mode 3": (_header(4, 4, 1, 3) + b"\x00" * 256, ["--mrc-info"], "unsupported MRC mode"),
    "stack too short": (_header(4, 4, 2, 2) + b"\x00" * 128, ["--stack-info", "@", 3], "sections"),
}


@pytest.mark.parametrize("case", sorted(DAMAGED))
def test_damaged_images_end_in_clean_errors(san_exe, tmp_path, case):
    data, args, msg = DAMAGED[case]
    p = tmp_path / "d.mrc"
    if data is not None:
        p.write_bytes(data)
    args = [a if a != "@" else p for a in args]
    if "@" not in DAMAGED[case][1]:
        args = args[:1] + [p] + args[1:]
    r = _run(args)
    assert r.returncode == 1 and msg in r.stderr, (r.returncode, r.stderr)


@pytest.mark.parametrize("text,msg", [
    ("angpix = abc\n", "bad value for angpix"),
    ("first = 99999999999\n", "bad value for first"),
    ("angpix : 1.0\n", "malformed line"),
])
def test_damaged_configs_end_in_clean_errors(san_exe, tmp_path, text, msg):
    p = tmp_path / "c.txt"
    p.write_text(text)
    r = _run(["--parse-dump", p])
    assert r.returncode == 1 and msg in r.stderr, r.stderr


ORACLE_SCRIPT = r"""
import sys
import numpy as np
sys.path[:0] = [sys.argv[1] + "/gisspa_v2.0_amd", sys.argv[1] + "/oracle"]
import gisspa, oracle, synth
rng = np.random.default_rng(3)
x = rng.standard_normal((12, 20)) + 1j * rng.standard_normal((12, 20))
assert np.allclose(oracle.fft2d(x, -1), np.fft.fft2(x))
cfg = gisspa.SearchConfig(phistep=90.0, window_size=320, norm_type=1, nms_radius=32.0, threshold=1.0)
ds = synth.make_dataset(5, 400, 352, n_templates=2, tsize=64, n_particles=2, snr=0.8, diameter=48.0)
picks, st = oracle.search_norm(cfg, ds.image, ds.templates, ds.ctf, want_stats=True)
recs = oracle.search_norm_records(cfg, ds.image, ds.templates, ds.ctf)
assert len(picks) > 0 and len(recs) > 0
R = oracle.Refiner(cfg, ds.image, ds.templates, ds.ctf)
pk = np.zeros(2, dtype=oracle.PEAK_DTYPE)
pk["block"] = picks["block"][:1].repeat(2)
pk["pixel"] = picks["pixel"][:1].repeat(2)
pk["template_id"] = picks["template_id"][:1].repeat(2)
pk["theta"] = picks["psi"][:1].repeat(2)
q = pk["block"].astype(np.int64) * 320 * 320 + pk["pixel"]
R.run(pk, st["mu"][q], st["sigma"][q], iters=2, step=0.1, scale=(2.0, 2.0, 3.0, 500.0))
R.close()
cfg0 = gisspa.SearchConfig(phistep=90.0, window_size=720, norm_type=0, threshold=1.0)
d0 = synth.make_dataset(6, 720, 720, n_templates=1, tsize=64, n_particles=2, snr=0.8, diameter=48.0)
oracle.search_nonorm(cfg0, d0.image, d0.templates, d0.ctf)
print("oracle under sanitizers ok")
"""


def test_oracle_under_sanitizers(tmp_path):
    if not os.path.exists(SAN_ORACLE):
        pytest.skip("sanitizer build of the oracle absent (make -C oracle san)")
    asan = subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    if not os.path.isabs(asan) or not os.path.exists(asan):
        pytest.skip("no libasan runtime to preload")
    env = dict(os.environ, LD_PRELOAD=asan, GISSPA_ORACLE_LIB=SAN_ORACLE, OMP_NUM_THREADS="2",
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([sys.executable, "-c", ORACLE_SCRIPT, ROOT], capture_output=True, text=True, env=env,
                       timeout=600)
    for m in REPORT_MARKS:
        assert m not in r.stderr, r.stderr[-4000:]
    assert r.returncode == 0 and "oracle under sanitizers ok" in r.stdout, r.stderr[-4000:]
