"""CPU checks of the drop-in boundary: the C ABI library loads and exports every entry point of
include/gisspa.h, rejects bad parameters before touching a device, and the executable parses
config/LST/Euler files exactly like the reference's own parser (built from
src/EMReader/DataReader2.cpp into oracle/_ref when the reference is present)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
EXE = os.path.join(ROOT, "gisspa_v2.0_amd", "lib", "gisspa_main")
REF_PARSE = os.path.join(ROOT, "oracle", "_ref", "ref_parse")


def header_functions():
    text = open(os.path.join(ROOT, "include", "gisspa.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gs_[a-z0-9_]+)\s*\(", text)))


def test_every_header_symbol_is_exported(gs):
    lib = gs.load()
    names = header_functions()
    assert "gs_create" in names and "gs_run" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) <= set(gs.SIGNATURES), set(names) - set(gs.SIGNATURES)


@pytest.mark.parametrize("field,value,status", [
    ("window_size", 321, 1),   # odd
    ("window_size", 0, 1),     # not positive (every even size is built: 320 and 720 run the
                               # tiled correlation kernels, other sizes the generic path)
    ("phistep", 0.0, 1),
    ("angpix", -1.0, 1),
    ("overlap", 400, 1),
])
def test_create_rejects_bad_parameters_without_a_device(gs, field, value, status):
    cfg = gs.SearchConfig()
    setattr(cfg, field, value)
    p = cfg.to_c()
    h = ctypes.c_void_p()
    st = gs.load().gs_create(ctypes.byref(p), 0, ctypes.byref(h))
    assert st == status
    assert gs.load().gs_last_error().decode()


def test_null_arguments(gs):
    lib = gs.load()
    assert lib.gs_create(None, 0, None) == 1
    assert lib.gs_run(None) == 1
    lib.gs_destroy(None)
    lib.gs_free(None)
    counts = (ctypes.c_size_t * 1)()
    p = ctypes.c_void_p()
    assert lib.gs_search_batch(None, None, 1, ctypes.byref(p), counts) == 1
    assert lib.gs_set_graph(None, 1) == 1
    assert lib.gs_autotune(None, None, None, None) == 1


def test_noalloc_transform_checks_arguments_before_any_device_work(gs):
    lib = gs.load()
    assert lib.gs_fft2_c2c_noalloc(0, None, 8, 8, 1, -1, None) == 1
    # beyond the static work buffer: rejected by the size check alone
    assert lib.gs_fft2_c2c_noalloc(0, ctypes.c_void_p(256), 4096, 4096, 1, -1, None) == 1
    assert b"exceed the work buffer" in lib.gs_last_error()


def test_struct_layouts_match_the_header(gs, tmp_path):
    """The ctypes mirrors have the C layout (sizes and field offsets from the compiled header)."""
    fields = {
        "gs_params": (gs.Params, [f[0] for f in gs.Params._fields_]),
        "gs_record": (gs.Record, [f[0] for f in gs.Record._fields_]),
        "gs_batch_item": (gs.BatchItem, [f[0] for f in gs.BatchItem._fields_]),
        "gs_ctf": (gs.Ctf, [f[0] for f in gs.Ctf._fields_]),
        "gs_peak": (gs.Peak, [f[0] for f in gs.Peak._fields_]),
        "gs_refine_opts": (gs.RefineOpts, [f[0] for f in gs.RefineOpts._fields_]),
        "gs_refined": (gs.Refined, [f[0] for f in gs.Refined._fields_]),
    }
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "gisspa.h"', "int main(void) {"]
    for cname, (_, names) in fields.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for n in names:
            lines.append(f'printf("{cname} {n} %zu\\n", offsetof({cname}, {n}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l}
    for cname, (cls, names) in fields.items():
        assert got[(cname, "size")] == ctypes.sizeof(cls), cname
        for n in names:
            assert got[(cname, n)] == getattr(cls, n).offset, (cname, n)
    # the numpy mirrors of gs_peak / gs_refined used for arrays of picks
    assert gs.PEAK_DTYPE.itemsize == got[("gs_peak", "size")]
    assert [gs.PEAK_DTYPE.fields[n][1] for n in fields["gs_peak"][1]] == \
        [got[("gs_peak", n)] for n in fields["gs_peak"][1]]
    assert gs.REFINED_DTYPE.itemsize == got[("gs_refined", "size")]
    assert [gs.REFINED_DTYPE.fields[n][1] for n in fields["gs_refined"][1]] == \
        [got[("gs_refined", n)] for n in fields["gs_refined"][1]]


def test_refine_defaults_and_argument_checks(gs):
    lib = gs.load()
    o = lib.gs_refine_defaults()
    assert (o.iterations, o.step, list(o.scale)) == (20, 0.1, [2.0, 2.0, 3.0, 500.0])
    z = (ctypes.c_double * 1)()
    assert lib.gs_score_and_grad(None, None, 1, z, z) == 1
    assert lib.gs_refine(None, None, 1, None, None) == 1
    p = ctypes.c_void_p()
    n = ctypes.c_size_t()
    assert lib.gs_select(None, 7.0, 64.0, ctypes.byref(p), ctypes.byref(n)) == 1


def _dump(cmd):
    out = subprocess.run(cmd, cwd=GOLDEN, capture_output=True, text=True, check=True).stdout
    return [l for l in out.splitlines() if re.match(r"^(config|lst|euler) ", l)]


@pytest.mark.skipif(not os.path.exists(EXE), reason="executable not built")
def test_executable_parser_matches_reference_parser():
    if not os.path.exists(REF_PARSE):
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    ours = _dump([EXE, "--parse-dump", "parse_config.txt"])
    ref = _dump([REF_PARSE, "parse_config.txt"])
    ours_cfg = {l for l in ours if l.startswith("config")}
    ref_cfg = {l for l in ref if l.startswith("config")}
    assert ref_cfg <= ours_cfg  # ours adds nms_radius and tile_local_centers
    assert [l for l in ours if not l.startswith("config")] == [l for l in ref if not l.startswith("config")]


@pytest.mark.skipif(not os.path.exists(EXE), reason="executable not built")
def test_executable_parse_dump_fixture():
    lines = _dump([EXE, "--parse-dump", "parse_config.txt"])
    assert "config angpix 1.06" in lines and "config window_size 720" in lines
    assert [l for l in lines if l.startswith("lst")] == [
        "lst 0 mic_000.mrc 1.500000 0.020000 30.000000",
        "lst 1 mic_001.mrc 2.250000 0.000000 0.000000",
        "lst 2 /abs/path/mic_002.mrc 0.800000 -0.015000 -45.500000",
    ]
    assert [l for l in lines if l.startswith("euler")] == [
        "euler 12.5 33.25 0", "euler 90 180 10", "euler 45 270 5", "euler 60 120 0"]


@pytest.mark.skipif(not os.path.exists(EXE), reason="executable not built")
def test_executable_reports_missing_keys(tmp_path):
    cfg = tmp_path / "c.txt"
    cfg.write_text("input = a.lst\ntemplate = t.mrc\neulerfile = e.txt\n")
    r = subprocess.run([EXE, str(cfg)], capture_output=True, text=True)
    assert r.returncode != 0 and "angpix" in r.stderr


@pytest.mark.parametrize("changes", [dict(norm_type=0, window_size=720), dict(window_size=330)])
def test_reference_records_limits(gs, changes):
    """reference_records needs norm_type=1 and window_size^2 divisible by 256 (the reference's
    256-element chunks); both are rejected at gs_create, before any device work."""
    cfg = gs.SearchConfig(reference_records=1)
    for k, v in changes.items():
        setattr(cfg, k, v)
    p = cfg.to_c()
    h = ctypes.c_void_p()
    assert gs.load().gs_create(ctypes.byref(p), 0, ctypes.byref(h)) == 3
    assert b"reference_records" in gs.load().gs_last_error()
