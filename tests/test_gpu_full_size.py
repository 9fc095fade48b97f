"""The benchmark workloads at their stated sizes, asserted on the GPU.

- Multi-blast 8192^2 FP64 (4x4 seeded blasts, p log-U[10, 1000], rho noise
  5%, .params dt 1e-6; the roofline workload): two steps, one of each sweep
  parity, bit for bit against the oracle. At this size the march kernels
  run ~48k work items and index up to (8197)^2 elements per field.
- The FP32-state form on the same workload for 200 steps against the FP64
  GPU state after the same 200 steps (itself bitwise equal to the oracle
  wherever the oracle has been run): per-field relative L-inf <= 1e-4, with
  the relative L1 and the mass/energy drift reported.
- The 256-instance seeded ensemble (sides log-U[64, 1024], Sod-x/Sod-y/
  blast, pressure ratios log-U[10, 1000]) for 12 steps in FP64 and FP32:
  the smallest and the largest instance, a blast and a Sod-y instance bit
  for bit against their single-mesh runs, and one instance against the
  oracle.
"""
import numpy as np
import pytest

import flow_amd
from flow_amd import problems
from tests import harness
from tests.test_gpu_fp32 import field_errors, gpu_run_f32

pytestmark = pytest.mark.gpu

N_BIG = 8192
DT_BIG = 1.0e-6


def test_multiblast_8192_two_steps_bitwise():
    print(f"case multi_blast n={N_BIG} steps=2 march vs oracle", flush=True)
    init = problems.multi_blast(N_BIG)
    ref = harness.oracle_run(None, N_BIG, 2, dt=DT_BIG, init=init)
    got = harness.gpu_run(None, N_BIG, 2, flow_amd.MARCH, dt=DT_BIG, init=init)
    for k, (rel, ndiff) in harness.compare_state(got["arrays"], ref["arrays"],
                                                 ref["L"]).items():
        print(f"  {k:6s} rel_linf={rel:.3e} bitwise_diffs={ndiff}", flush=True)
        assert ndiff == 0, (k, rel, ndiff)
    for k in ("F_x", "F_y", "rho_old"):
        ndiff = int(np.count_nonzero(got["arrays"][k] != ref["arrays"][k]))
        print(f"  {k:6s} bitwise_diffs={ndiff}", flush=True)
        assert ndiff == 0, k
    np.testing.assert_array_equal(got["dt"], ref["dt"])
    np.testing.assert_array_equal(got["undefined"], ref["undefined"])
    # The kernel set's totals are exact fixed-point sums of per-item
    # compensated sums; the oracle's are plain FP64 OpenMP sums, whose
    # rounding over 67M cells reaches ~1e-12. Both are checked against a
    # pairwise sum of the final state (error ~1e-15) of the same cells.
    L = ref["L"]
    v = harness.views(ref["arrays"], L)
    rho_i, e_i = v["rho"][2:L - 2, 2:L - 2], v["e"][2:L - 2, 2:L - 2]
    mass = float(np.sum(rho_i))
    energy = float(np.sum(rho_i * e_i))
    print(f"  mass gpu {got['mass'][-1]!r} oracle {ref['mass'][-1]!r} pairwise {mass!r}",
          flush=True)
    print(f"  energy gpu {got['energy'][-1]!r} oracle {ref['energy'][-1]!r} "
          f"pairwise {energy!r}", flush=True)
    assert abs(got["mass"][-1] - mass) <= 1e-13 * mass
    assert abs(got["energy"][-1] - energy) <= 1e-13 * energy
    np.testing.assert_allclose(got["mass"], ref["mass"], rtol=1e-11)
    np.testing.assert_allclose(got["energy"], ref["energy"], rtol=1e-11)


def _gpu_fields_fp64(init, n, steps, dt):
    m = flow_amd.Mesh(n, n, dt=dt)
    st = flow_amd.DeviceState(m)
    st.upload("rho", init[0])
    st.upload("e", init[1])
    st.reflect_initial()
    steps_done, dts, mass, energy, _ = st.run(steps)
    out = dict(arrays={k: st.download(k) for k in ("rho", "e", "u", "v", "mom_x", "mom_y")},
               dt=dts, mass=mass, energy=energy, steps=steps_done, L=m.nx)
    st.close()
    m.close()
    return out


def test_fp32_state_8192_200_steps_within_bound():
    steps = 200
    print(f"case fp32 multi_blast n={N_BIG} steps={steps} vs FP64 GPU", flush=True)
    init = problems.multi_blast(N_BIG)
    ref = _gpu_fields_fp64(init, N_BIG, steps, DT_BIG)
    got = gpu_run_f32(None, N_BIG, steps, DT_BIG, init=init)
    assert got["steps"] == ref["steps"] == steps
    errs = field_errors(got["arrays"], ref["arrays"], ref["L"])
    for k, (linf, l1) in errs.items():
        print(f"  {k:4s} rel_linf={linf:.3e} rel_l1={l1:.3e}", flush=True)
    m0, e0 = ref["mass"][0], ref["energy"][0]
    print(f"  mass drift FP64 {(ref['mass'][-1] - m0) / m0:.3e} "
          f"FP32 {(got['mass'][-1] - got['mass'][0]) / got['mass'][0]:.3e}; "
          f"energy drift FP64 {(ref['energy'][-1] - e0) / e0:.3e} "
          f"FP32 {(got['energy'][-1] - got['energy'][0]) / got['energy'][0]:.3e}",
          flush=True)
    dt_rel = np.max(np.abs(got["dt"] - ref["dt"]) / np.abs(ref["dt"]))
    print(f"  dt_rel_max={dt_rel:.3e}", flush=True)
    for k, (linf, _) in errs.items():
        assert linf <= 1.0e-4, (k, linf)
    assert dt_rel <= 1.0e-5
    np.testing.assert_allclose(got["mass"], ref["mass"], rtol=1e-6)
    np.testing.assert_allclose(got["energy"], ref["energy"], rtol=1e-6)


ENS_STEPS = 12
ENS_DT = 1.0e-4
STATE = ("rho", "e", "u", "v", "mom_x", "mom_y", "F_x", "F_y", "rho_old")


def _pick(specs):
    sizes = [s[0] for s in specs]
    picks = {"smallest": int(np.argmin(sizes)), "largest": int(np.argmax(sizes))}
    picks["blast"] = next(k for k, s in enumerate(specs) if s[1] == "blast")
    picks["sod_y"] = next(k for k, s in enumerate(specs) if s[1] == "sod_y")
    return picks


@pytest.mark.parametrize("fp32", [False, True])
def test_ensemble_256_instances_equal_single_runs(fp32):
    specs = problems.ensemble_specs(256)
    sizes = [s[0] for s in specs]
    picks = _pick(specs)
    print(f"case ensemble 256 fp32={fp32} cells={sum(n * n for n in sizes)} "
          f"picks={picks}", flush=True)
    ens = flow_amd.Ensemble(sizes, dt=ENS_DT, fp32=fp32)
    for k, (n, kind, split, ratio) in enumerate(specs):
        ens.set_state(k, *problems.ensemble_instance(n, kind, split, ratio))
    ens.set_graphs(True)
    dts, mass, energy, _ = ens.run(ENS_STEPS)
    assert np.all(np.isfinite(dts)) and np.all(dts > 0)
    got = {name: (k, ens.get_all(k), dts[k]) for name, k in picks.items()}
    undefined = ens.undefined
    ens.close()
    print(f"  instances with undefined CFL candidates: "
          f"{int((undefined.sum(axis=1) > 0).sum())}", flush=True)
    for name, (k, arrays, dt) in got.items():
        init = problems.ensemble_instance(*specs[k])
        if fp32:
            one = gpu_run_f32(None, sizes[k], ENS_STEPS, ENS_DT, init=init)
        else:
            one = harness.gpu_run(None, sizes[k], ENS_STEPS, flow_amd.MARCH, dt=ENS_DT,
                                  init=init)
        print(f"  {name}: instance {k} n={sizes[k]} kind={specs[k][1]}", flush=True)
        assert np.array_equal(dt, one["dt"]), name
        for a in STATE:
            ndiff = int(np.count_nonzero(arrays[a] != one["arrays"][a]))
            assert ndiff == 0, (name, a, ndiff)
    if not fp32:
        k, arrays, dt = got["blast"]
        ref = harness.oracle_run(None, sizes[k], ENS_STEPS, dt=ENS_DT,
                                 init=problems.ensemble_instance(*specs[k]))
        for a, (rel, ndiff) in harness.compare_state(arrays, ref["arrays"],
                                                     ref["L"]).items():
            assert ndiff == 0, (a, rel, ndiff)
        assert np.array_equal(dt, ref["dt"])
