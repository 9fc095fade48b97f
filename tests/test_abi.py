"""The C-ABI library loads and exports every function include/*.h declares
(no GPU calls), and its pure-host parts (parameter files) behave as main.c
expects. Runs without a GPU."""
import ctypes
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")


def declared_functions():
    names = set()
    for fn in os.listdir(INCLUDE):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(INCLUDE, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"#define[^\n]*(\\\n[^\n]*)*", "", text)
        for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(([^;{}]*)\)\s*;", text):
            name = m.group(1)
            if name in ("if", "while", "for", "return", "sizeof"):
                continue
            names.add(name)
    return names


def test_headers_declare_the_drop_in():
    names = declared_functions()
    for fn in ("solve_hydro_2d", "set_timestep", "print_conservation",
               "allocate_data", "copy_buffer", "handle_boundary_2d",
               "initialise_mesh_2d", "initialise_shared_data_2d",
               "get_key_value_parameter", "flowhip_run"):
        assert fn in names


def test_library_exports_every_declared_function():
    import flow_amd
    path = flow_amd.library_path()
    assert os.path.exists(path), "libflowhip.so not built"
    lib = ctypes.CDLL(path)
    missing = [n for n in sorted(declared_functions()) if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_params_reader(tmp_path):
    from flow_amd import native
    L = native.lib()
    p = tmp_path / "t.params"
    p.write_text("# comment\nnx 64\ndt   1.0e-3\n"
                 "problem_0 density=0.125 energy=2.0 xpos=0.0 ypos=0.0 "
                 "width=1.0 height=1.0\n")
    f = str(p).encode()
    assert L.get_int_parameter(b"nx", f) == 64
    assert L.get_double_parameter(b"dt", f) == 1.0e-3
    keys = ctypes.create_string_buffer(256 * 1024)
    vals = (ctypes.c_double * 256)()
    n = ctypes.c_int()
    assert L.get_key_value_parameter(b"problem_0", f, keys, vals, ctypes.byref(n)) == 1
    assert n.value == 6
    assert keys.raw[:7] == b"density"
    assert keys.raw[1024 * 5:1024 * 5 + 6] == b"height"
    assert list(vals[:6]) == [0.125, 2.0, 0.0, 0.0, 1.0, 1.0]
    assert L.get_key_value_parameter(b"problem_1", f, keys, vals, ctypes.byref(n)) == 0


def test_missing_parameter_is_fatal(tmp_path):
    """arch's TERMINATE: a missing key prints and exits non-zero."""
    p = tmp_path / "t.params"
    p.write_text("nx 64\n")
    code = ("import ctypes,sys; sys.path.insert(0, %r); from flow_amd import native;"
            "native.lib().get_int_parameter(b'ny', %r)" % (ROOT, str(p).encode()))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True)
    assert r.returncode != 0
    assert "Could not find the parameter ny" in r.stderr


def test_library_refuses_to_run_without_build(tmp_path, monkeypatch):
    from flow_amd import native
    monkeypatch.setattr(native, "_LIB", str(tmp_path / "nope.so"))
    monkeypatch.setattr(native, "_lib", None)
    with pytest.raises(RuntimeError, match="not built"):
        native.lib()


def test_mesh_sizes_outside_32_bit_indexing_are_refused():
    """Fields are indexed with 32-bit element offsets: (n+5)^2 <= 2^31-1 up to
    n = 46335. Larger meshes (and interiors below 2x2) are refused before any
    device call, not left to overflow an index."""
    import ctypes
    from flow_amd import native
    L = native.lib()
    L.flowhip_mesh_size_ok.restype = ctypes.c_int
    L.flowhip_mesh_size_ok.argtypes = [ctypes.c_int, ctypes.c_int]
    assert (46335 + 5) ** 2 <= 2 ** 31 - 1 < (46336 + 5) ** 2
    assert L.flowhip_mesh_size_ok(46335, 46335) == 1
    assert L.flowhip_mesh_size_ok(2, 2) == 1
    for nx, ny in ((46336, 46336), (46336, 8), (8, 46336), (1, 8), (8, 1), (0, 0)):
        assert L.flowhip_mesh_size_ok(nx, ny) == 0, (nx, ny)
    import pytest
    import flow_amd
    with pytest.raises(RuntimeError):
        flow_amd.Mesh(46336, 46336)
