"""The C-ABI library loads and exports every symbol include/tt.h declares
(no compute calls: this runs without a GPU)."""
import os
import re

import pytest

from turbulence_tracing_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    src = open(os.path.join(ROOT, "include", "tt.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(tt_\w+)\s*\(", src)))


def test_header_matches_binding():
    assert header_symbols() == sorted(_lib.EXPORTS)


def test_library_exports_every_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libtt.so not built: run python -m turbulence_tracing_amd.build")
    lib = _lib.load()
    for name in header_symbols():
        assert hasattr(lib.dll, name), name
    assert lib.version() == 1


def test_argument_errors_raise_without_gpu():
    """Argument validation happens before any device work."""
    lib = _lib.load()
    g = _lib.Grid(1, 4, 4, 0, 0, 0, 1, 1, 1)
    with pytest.raises(_lib.TTError, match="nodes"):
        lib.tt_grad_stencil(_lib.C.c_void_p(16), 0, 1, 1, 1, _lib.C.byref(g), 0, 1, 1, 1, None, 0,
                            _lib.C.c_void_p(16), None)
    opts = _lib.TraceOpts(0.3, 1.0)
    g = _lib.Grid(4, 4, 4, 0, 0, 0, 1, 1, 1)
    with pytest.raises(_lib.TTError, match="1/m"):
        lib.tt_trace_rk4(_lib.C.c_void_p(16), _lib.C.byref(g), _lib.C.byref(opts), _lib.C.c_void_p(16), 4, 4,
                         None, _lib.C.c_void_p(16), 4, None, None, 4, None, None)
    t = _lib.Train()
    t.n_ops = 1
    t.ops[0] = _lib.Optic(_lib.OP_KNIFE, 0, 0.0, 0.0)
    with pytest.raises(_lib.TTError, match="direction"):
        lib.tt_optics_apply(_lib.C.byref(t), _lib.C.c_void_p(16), 4, 4, _lib.C.c_void_p(16), 4, None)
    with pytest.raises(_lib.TTError, match="sizes"):
        lib.tt_fft_plan_create(1, 4, 4, _lib.C.byref(_lib.P()))


def test_optics_op_tables_match_header():
    src = open(os.path.join(ROOT, "include", "tt.h")).read()
    for name, val in [("DISTANCE", 1), ("LENS", 2), ("APERTURE", 3), ("STOP", 4), ("KNIFE", 5), ("RECT", 6),
                      ("ANNULUS", 7)]:
        assert re.search(rf"TT_OP_{name}\s*=\s*{val}\b", src)
        assert getattr(_lib, "OP_" + name) == val
    assert re.search(rf"#define TT_MAX_OPS {_lib.TT_MAX_OPS}\b", src)


# ctypes mirror of every struct in include/tt.h: sizes and field offsets must
# equal what the C compiler lays out (a mismatch corrupts every call)
STRUCTS = {
    "tt_grid": "Grid", "tt_spectrum": "Spectrum", "tt_trace_opts": "TraceOpts", "tt_batch_item": "BatchItem",
    "tt_batch_info": "BatchInfo", "tt_batch_tuning": "BatchTuning", "tt_optic": "Optic", "tt_train": "Train",
    "tt_hist": "Hist", "tt_detector": "Detector", "tt_gemm_args": "GemmArgs",
}


def test_struct_layouts_match_c(tmp_path):
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "tt.h"', 'int main(void) {']
    for cname, pyname in STRUCTS.items():
        st = getattr(_lib, pyname)
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f in st._fields_:
            lines.append(f'printf("{cname} {f[0]} %zu\\n", offsetof({cname}, {f[0]}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run([cc, "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for cname, pyname in STRUCTS.items():
        st = getattr(_lib, pyname)
        assert got[(cname, "size")] == _lib.C.sizeof(st), cname
        for f in st._fields_:
            assert got[(cname, f[0])] == getattr(st, f[0]).offset, (cname, f[0])


def test_empty_ray_sets_need_no_device_memory():
    """n = 0 on every ray-parallel entry point returns success with NULL
    device pointers and launches nothing (so it runs without a GPU)."""
    lib = _lib.load()
    g = _lib.Grid(4, 4, 4, 0, 0, 0, 1, 1, 1)
    opts = _lib.TraceOpts(0.5, 1.0)
    lib.tt_trace_rk4(None, _lib.C.byref(g), _lib.C.byref(opts), None, 0, 0, None, None, 0, None, None, 0, None, None)
    lib.tt_sample_field(None, _lib.C.byref(g), None, 0, 0, None, 0, None)
    lib.tt_beam_sample(0, 1e-3, 1e-5, 1.0, 2, 1, 0, None, 0, None)
    lib.tt_ray_sort(None, 0, 0, _lib.C.byref(g), 0, None, None, None)
    lib.tt_ray_sort_rays(None, 0, 0, _lib.C.byref(g), 0, None, None, None, None, 0, None)
    lib.tt_path_features(None, None, 4, _lib.C.byref(g), None, 0, 0, None, None, 0, None, 0, None)
    lib.tt_mlp_infer(None, 48, 1, 0, 48, None, None, None, None, None, 8, 8, None)
    lib.tt_surrogate_apply(None, 0, None, 8, 0, None, None, None, 0, None, None, None, None, 0, None)
    t = _lib.Train()
    lib.tt_optics_apply(_lib.C.byref(t), None, 0, 0, None, 0, None)
