"""Rectangular meshes (nx != ny) against the oracle, bit for bit: flow's
driver reads nx and ny separately (main.c:20-21), and the kernel set maps
them onto different things (x: the strips a wave covers across its lanes,
y: the chunks it marches down), so a square-only test would not see a
swapped stride or bound. The shapes cover power-of-two spacings on both
axes (the reciprocal-multiply path), a non-power-of-two axis (IEEE
quotients), a mesh one chunk row high, a mesh one strip wide, and odd sides.
Every step of both sweep orders is covered (12 steps), for each schedule."""
import numpy as np
import pytest

from flow_amd import problems
from tests import harness

pytestmark = pytest.mark.gpu

SHAPES = [(64, 40), (40, 96), (37, 91), (128, 32), (512, 24), (24, 512)]
STEPS = 12
SUM_TOL = 1e-13


def rect_problem(n):
    """A blast box off centre and a Sod-like split along x on an nx x ny
    mesh of unit width and height."""
    nx, ny = n
    return problems.boxes(nx, ny, [(1.0, 0.25, 0.0, 0.0, 1.0, 1.0),
                                   (1.0, 250.0, 0.4, 0.3, 0.2, 0.3),
                                   (0.125, 2.0, 0.7, 0.0, 0.3, 1.0)])


def _check(got, ref):
    assert got["steps"] == ref["steps"] == STEPS
    assert np.array_equal(got["dt"], ref["dt"])
    for k, (rel, ndiff) in harness.compare_state(got["arrays"], ref["arrays"],
                                                 ref["L"]).items():
        assert ndiff == 0, (k, rel, ndiff)
    np.testing.assert_allclose(got["mass"], ref["mass"], rtol=SUM_TOL)
    np.testing.assert_allclose(got["energy"], ref["energy"], rtol=SUM_TOL)


@pytest.mark.parametrize("shape", SHAPES, ids=[f"{a}x{b}" for a, b in SHAPES])
def test_rectangular_mesh_bitwise(shape):
    import flow_amd
    ref = harness.oracle_run(rect_problem, shape, STEPS)
    for sched in (flow_amd.MARCH, flow_amd.FUSED, flow_amd.STAGEWISE):
        print(f"case rect {shape[0]}x{shape[1]} schedule={sched} steps={STEPS}",
              flush=True)
        _check(harness.gpu_run(rect_problem, shape, STEPS, sched), ref)


def test_rectangular_mesh_graph_replay():
    import flow_amd
    shape = (96, 40)
    print(f"case rect {shape[0]}x{shape[1]} march graphs steps={STEPS}", flush=True)
    ref = harness.oracle_run(rect_problem, shape, STEPS)
    _check(harness.gpu_run(rect_problem, shape, STEPS, flow_amd.MARCH, graphs=True), ref)
