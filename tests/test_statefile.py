"""The state-file format (flow_amd/statefile.py, flow_amd/csrc/hip/state.c)
and the driver's command line, on the CPU: the byte layout is pinned at fixed
offsets, a write/read round trip is exact, malformed files are refused, and
flow.hip rejects a bad command line before it touches a GPU."""
import os
import struct
import subprocess

import numpy as np
import pytest

from flow_amd import problems
from flow_amd import statefile as sf

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "flow_amd", "lib", "flow.hip")


def _state(n, seed=1):
    rng = np.random.default_rng(seed)
    L = n + 4
    return {name: rng.standard_normal(sf.layout_shape(lay, L)) for name, lay in sf.ARRAYS}


def test_header_layout_is_pinned(tmp_path):
    p = tmp_path / "s.bin"
    sf.write(p, _state(3), 3, step=7, width=2.0, height=2.0, dt=1.25e-4,
             dt_h=1.5e-4, sim_time=0.75)
    b = p.read_bytes()
    assert b[:8] == b"FLOWSTAT"
    assert struct.unpack_from("<II", b, 8) == (1, 6)
    assert struct.unpack_from("<4i", b, 16) == (3, 3, 2, 7)
    assert struct.unpack_from("<5d", b, 32) == (2.0, 2.0, 1.25e-4, 1.5e-4, 0.75)
    # first record: rho, cell layout, L*L values of L = n + 4
    name, lay, _, count = struct.unpack_from("<16sc7sQ", b, 72)
    assert name.rstrip(b"\0") == b"rho" and lay == b"C" and count == 49
    L = 7
    cells, faces = L * L, L * (L + 1)
    assert len(b) == 72 + 6 * 32 + 8 * (2 * cells + 4 * faces)


def test_round_trip_is_exact(tmp_path):
    st = _state(5, seed=3)
    st["rho"][0, 0] = np.nextafter(1.0, 2.0)
    st["e"][1, 1] = -0.0
    p = tmp_path / "s.bin"
    sf.write(p, st, 5, step=2, dt=3e-5, dt_h=2e-5, sim_time=1e-4)
    d = sf.read(p)
    assert (d["n"], d["step"], d["dt"], d["dt_h"]) == (5, 2, 3e-5, 2e-5)
    for name, _ in sf.ARRAYS:
        assert d["arrays"][name].tobytes() == st[name].tobytes(), name


def test_initial_state_from_problem_buffers():
    n = 16
    rho, e = problems.sod_x(n)
    st = sf.initial_state(n, rho, e)
    L = n + 4
    np.testing.assert_array_equal(st["rho"], rho[: L * L].reshape(L, L))
    assert st["u"].shape == (L, L + 1) and not st["u"].any()
    assert st["mom_y"].shape == (L + 1, L) and not st["mom_y"].any()


def test_rectangular_round_trip_and_shapes(tmp_path):
    """A 64 x 40 mesh: Lx = 68 columns, Ly = 44 rows; x-face arrays 44 x 69,
    y-face arrays 45 x 68; the header holds nx then ny."""
    rng = np.random.default_rng(11)
    Lx, Ly = 68, 44
    st = {name: rng.standard_normal(sf.layout_shape(lay, Lx, Ly)) for name, lay in sf.ARRAYS}
    assert st["u"].shape == (44, 69) and st["v"].shape == (45, 68)
    p = tmp_path / "r.bin"
    sf.write(p, st, (64, 40), step=3)
    b = p.read_bytes()
    assert struct.unpack_from("<4i", b, 16) == (64, 40, 2, 3)
    d = sf.read(p)
    assert (d["nx"], d["ny"]) == (64, 40)
    for name, _ in sf.ARRAYS:
        assert d["arrays"][name].tobytes() == st[name].tobytes(), name
    cells, xf, yf = Lx * Ly, Ly * (Lx + 1), (Ly + 1) * Lx
    assert len(b) == 72 + 6 * 32 + 8 * (2 * cells + 2 * xf + 2 * yf)


def test_malformed_files_are_refused(tmp_path):
    p = tmp_path / "s.bin"
    sf.write(p, _state(2), 2)
    good = p.read_bytes()
    (tmp_path / "bad_magic.bin").write_bytes(b"XLOWSTAT" + good[8:])
    (tmp_path / "short.bin").write_bytes(good[:-8])
    huge = bytearray(good)
    struct.pack_into("<2i", huge, 16, 1 << 30, 2)
    (tmp_path / "huge.bin").write_bytes(bytes(huge))
    for bad in ("bad_magic.bin", "short.bin", "huge.bin"):
        with pytest.raises(ValueError):
            sf.read(tmp_path / bad)


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="flow.hip not built")
@pytest.mark.parametrize("args", [[], ["a.params", "b.params"], ["a.params", "--dump"],
                                  ["a.params", "--frobnicate"]])
def test_driver_rejects_bad_command_lines(args):
    r = subprocess.run([DRIVER] + args, capture_output=True, text=True, timeout=60)
    assert r.returncode == 2
    assert "usage:" in r.stderr
