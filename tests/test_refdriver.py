"""flow's own driver (main.c, flow_data.c: compiled unmodified from the
reference checkout and linked against libflowhip.so) running the hip kernel
set on flow.params' known-answer case (flow.tests:1).

The link test needs the reference checkout and runs where it is present; the
GPU test runs the binary that the link step left in tests/refdriver/_ref/.
"""
import json
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REFDIR = os.path.join(HERE, "refdriver")
BINARY = os.path.join(REFDIR, "_ref", "flow_ref_driver")
BINARY_PROF = os.path.join(REFDIR, "_ref", "flow_ref_driver_prof")
REFERENCE = "/root/reference"
LIB = os.path.join(os.path.dirname(HERE), "flow_amd", "lib", "libflowhip.so")


@pytest.mark.skipif(not os.path.exists(os.path.join(REFERENCE, "main.c")),
                    reason="reference checkout not present")
@pytest.mark.skipif(not os.path.exists(LIB), reason="libflowhip.so not built")
def test_reference_driver_links():
    r = subprocess.run(["make", "-B", "-C", REFDIR], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "warning" not in (r.stdout + r.stderr).lower()
    # the three interface entry points are undefined in the driver's objects
    # and resolved by the kernel set's library
    und = subprocess.run(["nm", "-u", os.path.join(REFDIR, "_ref", "main.o")],
                         capture_output=True, text=True, check=True).stdout
    for sym in ("solve_hydro_2d", "set_timestep", "print_conservation",
                "initialise_shared_data_2d", "handle_boundary_2d"):
        assert sym in und
    defs = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True,
                          text=True, check=True).stdout
    for sym in ("solve_hydro_2d", "set_timestep", "print_conservation"):
        assert f" T {sym}" in defs


def _write_case(tmp_path):
    kat = json.load(open(os.path.join(HERE, "golden", "flow_tests_kat.json")))
    (tmp_path / "arch.params").write_text(
        f"width {kat['width']!r}\nheight {kat['height']!r}\n"
        f"max_dt {kat['max_dt']!r}\nsim_end {kat['sim_end']!r}\n")
    run = tmp_path / "flow"
    run.mkdir()
    lines = [f"problem_{k} density={p[0]!r} energy={p[1]!r} xpos={p[2]!r} "
             f"ypos={p[3]!r} width={p[4]!r} height={p[5]!r}"
             for k, p in enumerate(kat["problems"])]
    lines += [f"dt {kat['dt']!r}", f"nx {kat['nx']}", f"ny {kat['nx']}",
              f"iterations {kat['iterations']}", "visit_dump 0"]
    (run / "flow.params").write_text("\n".join(lines) + "\n")
    (run / "flow.tests").write_text(
        f"flow.params density={kat['density']:.12e} energy={kat['energy']:.12e}\n")
    return run, kat


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(BINARY),
                    reason="flow's driver was not linked (needs the reference checkout)")
@pytest.mark.parametrize("schedule", ["march", "stagewise"])
def test_reference_driver_known_answer(tmp_path, schedule):
    run, kat = _write_case(tmp_path)
    env = dict(os.environ, FLOWHIP_SCHEDULE=schedule)
    print(f"case flow driver {schedule} n={kat['nx']}", flush=True)
    r = subprocess.run([BINARY, "flow.params"], cwd=run, env=env,
                       capture_output=True, text=True, timeout=110)
    out = r.stdout
    assert r.returncode == 0, out[-2000:] + r.stderr[-2000:]
    assert f"Expected density {kat['density']:.12e}, result was {kat['density']:.12e}." in out
    assert f"Expected energy  {kat['energy']:.12e}, result was {kat['energy']:.12e}." in out
    assert "PASSED validation." in out
    assert out.count("\nIteration ") == kat["iterations"]


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(BINARY_PROF),
                    reason="flow's driver was not linked (needs the reference checkout)")
def test_reference_driver_prints_kernel_times(tmp_path):
    """main.c built with -DENABLE_PROFILING (arch's switch) and run with
    FLOWHIP_PROFILE=1: PRINT_PROFILING_RESULTS (main.c:123-124) lists the
    kernel set's hipEvent times, one launch of each march kernel per
    iteration, and the known answer still passes."""
    run, kat = _write_case(tmp_path)
    env = dict(os.environ, FLOWHIP_PROFILE="1")
    print(f"case flow driver profiling n={kat['nx']}", flush=True)
    r = subprocess.run([BINARY_PROF, "flow.params"], cwd=run, env=env,
                       capture_output=True, text=True, timeout=110)
    out = r.stdout
    assert r.returncode == 0, out[-2000:] + r.stderr[-2000:]
    assert "PASSED validation." in out
    assert "Kernel times (hipEvents" in out
    rows = {line.split()[0]: line.split() for line in out.splitlines()
            if line.split() and line.split()[0] in ("lagrangian", "advect_mass_energy",
                                                   "advect_momentum", "timestep")}
    for k in ("lagrangian", "advect_mass_energy", "advect_momentum"):
        assert int(rows[k][1]) == kat["iterations"], rows[k]
        assert float(rows[k][2]) > 0.0
    assert int(rows["timestep"][1]) == 1   # the pre-loop set_timestep
    assert "Communication: none" in out
