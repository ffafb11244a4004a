"""The C-ABI library loads and exports every function include/lyg.h declares; the
synthetic generators are deterministic.  No compute calls (no GPU needed)."""
import ctypes as C
import os

import numpy as np
import pytest

from lyon_amd import _lib, workloads


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(_lib.LIB_PATH)
    declared = _lib.declared_symbols()
    assert len(declared) >= 14
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, missing
    # the ctypes table covers the header exactly
    assert sorted(_lib.SIGNATURES) == declared


def test_generators_are_deterministic():
    a = workloads.stroke_batch(50, 8, 256, 2048.0, seed=0)
    b = workloads.stroke_batch(50, 8, 256, 2048.0, seed=0)
    assert np.array_equal(a.verbs, b.verbs) and np.array_equal(a.points, b.points)
    c = workloads.stroke_batch(50, 8, 256, 2048.0, seed=1)
    assert not np.array_equal(a.points[:10], c.points[:10])
    segs = np.array([a.path(i).num_segments() for i in range(len(a))])
    assert segs.min() >= 8 and segs.max() <= 256
    assert (a.points >= 0).all() and (a.points <= 2048).all()
    o = workloads.stroke_options(100)
    assert set(o["line_join"]) <= {0, 2, 3} and (o["line_width"] >= 0.5).all()


def test_plain_path_is_closed_64_segments():
    p = workloads.plain_path(0)
    assert p.num_segments() == 64 and p.verbs[-1] == 4
    assert np.array_equal(p.points[-1], p.points[0])


def test_header_is_c_and_links():
    """include/lyg.h compiles as C99 and a C program links against liblyg.so and calls a
    host-only entry point (the FFI a C or Rust caller uses)."""
    import shutil
    import subprocess
    import tempfile
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    lib_dir = os.path.join(root, "lyon_amd")
    if not os.path.exists(os.path.join(lib_dir, "liblyg.so")) or not shutil.which("gcc"):
        pytest.skip("liblyg.so not built")
    src = ('#include "lyg.h"\n#include <stdio.h>\n#include <string.h>\n'
           'int main(void) {\n'
           '  lyg_path_view v; lyg_output o; lyg_fill_options f; lyg_stroke_options s;\n'
           '  memset(&v, 0, sizeof v); memset(&o, 0, sizeof o); (void)f; (void)s;\n'
           '  const char* m = lyg_status_string(LYG_ERR_CAPACITY);\n'
           '  printf("%s\\n", m ? m : "");\n'
           '  return lyg_fill_rectangle(0, 0, &o) == LYG_ERR_INVALID_ARGUMENT ? 0 : 3;\n}\n')
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        exe = os.path.join(d, "t")
        with open(c, "w") as fh:
            fh.write(src)
        subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(root, "include"),
                        c, "-o", exe, "-L", lib_dir, "-llyg", "-Wl,-rpath," + lib_dir],
                       check=True, capture_output=True)
        r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
        assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
        assert r.stdout.strip()
