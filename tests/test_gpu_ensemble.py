"""The batching form: every instance of an ensemble, advanced together
with the others, computes bit for bit what it computes alone (the
single-mesh march schedule, itself bitwise equal to the oracle), in FP64
and in the FP32-state form, with and without graph replay, under every
work-list schedule; the autotuner leaves the state as it found it."""
import numpy as np
import pytest

from flow_amd import problems
from tests import harness

SIZES = [2, 17, 40, 64, 90, 131]
KINDS = ["sod_x", "blast", "sod_y", "blast", "sod_x", "blast"]
STEPS = 7
DT = 1.0e-4
STATE = ("rho", "e", "u", "v", "mom_x", "mom_y", "F_x", "F_y", "rho_old")


def _init(k):
    return problems.ensemble_instance(SIZES[k], KINDS[k], 0.3 + 0.07 * k, 10.0 * (k + 1))


def _ensemble(fp32=False, graphs=False, schedule=None, autotune=False):
    import flow_amd
    ens = flow_amd.Ensemble(SIZES, dt=DT, fp32=fp32)
    for k in range(len(SIZES)):
        ens.set_state(k, *_init(k))
    if schedule:
        ens.set_schedule(*schedule)
    if autotune:
        ens.autotune(trial_steps=2)
    ens.set_graphs(graphs)
    dts, mass, energy, _ = ens.run(STEPS)
    out = [dict(arrays=ens.get_all(k), dt=dts[k], mass=mass[k], energy=energy[k])
           for k in range(len(SIZES))]
    ens.close()
    return out


def _single(k, fp32):
    import flow_amd
    if not fp32:
        return harness.gpu_run(None, SIZES[k], STEPS, flow_amd.MARCH, dt=DT, init=_init(k))
    from tests.test_gpu_fp32 import gpu_run_f32
    return gpu_run_f32(None, SIZES[k], STEPS, DT, init=_init(k))


def _compare(batch, fp32):
    for k in range(len(SIZES)):
        one = _single(k, fp32)
        L = SIZES[k] + 4
        print(f"  instance {k} n={SIZES[k]}", flush=True)
        assert np.array_equal(batch[k]["dt"], one["dt"]), k
        for name in STATE:
            a, b = batch[k]["arrays"][name], one["arrays"][name]
            ndiff = int(np.count_nonzero(a != b))
            assert ndiff == 0, (k, name, ndiff)
        # totals: the same compensated sums, merged in another order
        np.testing.assert_allclose(batch[k]["mass"], one["mass"], rtol=1e-14, atol=0)
        np.testing.assert_allclose(batch[k]["energy"], one["energy"], rtol=1e-14, atol=0)
        assert L > 0


@pytest.mark.gpu
@pytest.mark.parametrize("graphs", [False, True])
def test_ensemble_instances_equal_single_runs_fp64(graphs):
    print(f"case ensemble fp64 graphs={graphs} sizes={SIZES}", flush=True)
    _compare(_ensemble(graphs=graphs), fp32=False)


@pytest.mark.gpu
def test_ensemble_instances_equal_single_runs_fp32():
    print(f"case ensemble fp32 sizes={SIZES}", flush=True)
    _compare(_ensemble(fp32=True, graphs=True), fp32=True)


@pytest.mark.gpu
@pytest.mark.parametrize("schedule", [(16, 0, True), (64, 2, False), (120, 1, True),
                                      (48, 1, False)])
def test_ensemble_schedule_does_not_change_results(schedule):
    print(f"case ensemble schedule={schedule}", flush=True)
    ref = _ensemble()
    got = _ensemble(schedule=schedule, graphs=True)
    for k in range(len(SIZES)):
        for name in STATE:
            assert np.array_equal(got[k]["arrays"][name], ref[k]["arrays"][name]), (k, name)
        assert np.array_equal(got[k]["dt"], ref[k]["dt"])


@pytest.mark.gpu
def test_ensemble_autotune_restores_the_state():
    print("case ensemble autotune", flush=True)
    ref = _ensemble()
    got = _ensemble(autotune=True, graphs=True)
    for k in range(len(SIZES)):
        for name in STATE:
            assert np.array_equal(got[k]["arrays"][name], ref[k]["arrays"][name]), (k, name)
        assert np.array_equal(got[k]["dt"], ref[k]["dt"])


@pytest.mark.gpu
def test_ensemble_instance_matches_oracle():
    """One instance of the ensemble against the oracle directly."""
    print("case ensemble vs oracle", flush=True)
    batch = _ensemble()
    k = 4
    ref = harness.oracle_run(None, SIZES[k], STEPS, dt=DT, init=_init(k))
    cmp = harness.compare_state(batch[k]["arrays"], ref["arrays"], ref["L"])
    for name, (rel, ndiff) in cmp.items():
        assert ndiff == 0, (name, rel, ndiff)
    assert np.array_equal(batch[k]["dt"], ref["dt"])


POW2_SIZES = [8, 32, 64, 128]


def _ensemble_of(sizes, fp32):
    import flow_amd
    ens = flow_amd.Ensemble(sizes, dt=DT, fp32=fp32)
    for k, n in enumerate(sizes):
        ens.set_state(k, *problems.ensemble_instance(n, KINDS[k], 0.3 + 0.07 * k,
                                                     10.0 * (k + 1)))
    dts, mass, energy, _ = ens.run(STEPS)
    out = [dict(arrays=ens.get_all(k), dt=dts[k], mass=mass[k], energy=energy[k])
           for k in range(len(sizes))]
    ens.close()
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("fp32", [False, True])
@pytest.mark.parametrize("sizes", [SIZES, POW2_SIZES], ids=["uneven", "pow2"])
def test_ensemble_geometry_forms_agree(monkeypatch, sizes, fp32):
    """The ensemble's geometry forms give the same bits: row geometry read
    once per item (the default; every instance has one spacing per axis),
    per row (FLOWHIP_UNIFORM=0), and, on power-of-two spacings, IEEE
    quotients in place of exact reciprocal products (FLOWHIP_POW2=0)."""
    envs = [{}, {"FLOWHIP_UNIFORM": "0"}]
    if sizes is POW2_SIZES:
        envs += [{"FLOWHIP_POW2": "0"}, {"FLOWHIP_POW2": "0", "FLOWHIP_UNIFORM": "0"}]
    runs = []
    for env in envs:
        print(f"case ensemble geometry sizes={sizes} fp32={fp32} env={env}", flush=True)
        for k in ("FLOWHIP_UNIFORM", "FLOWHIP_POW2"):
            monkeypatch.delenv(k, raising=False)
        for k, v in env.items():
            monkeypatch.setenv(k, v)
        runs.append(_ensemble_of(sizes, fp32))
    ref = runs[0]
    for got in runs[1:]:
        for k in range(len(sizes)):
            for name in STATE:
                assert np.array_equal(got[k]["arrays"][name], ref[k]["arrays"][name]), (k, name)
            assert np.array_equal(got[k]["dt"], ref[k]["dt"]), k
            np.testing.assert_allclose(got[k]["mass"], ref[k]["mass"], rtol=1e-14, atol=0)
            np.testing.assert_allclose(got[k]["energy"], ref[k]["energy"], rtol=1e-14, atol=0)
