#!/usr/bin/env python3
"""Throughput of the hip kernel set on the multi-blast workload.

One step = one solve_hydro_2d + print_conservation pass (main.c:82-91) over an
8192^2 FP64 mesh (4x4 seeded blasts, log-uniform pressures in [10, 1000] on a
rho=1 p=0.1 background, +/-5% seeded density noise). Inputs are resident in
HBM before the timed region; the driver loop runs on the host through the C
ABI (flowhip_run), with one device-to-host transfer of the scalars per
interface call, exactly as main.c would drive it.

Prints ONE JSON line: whole-job cell-updates/s, the roofline of the dominant
kernel (HIP-event durations on the kernel set's stream x its algorithmic
bytes), the full-step fraction of HBM peak at B_alg = 376 B per cell-update,
and the oracle (CPU restatement of omp3) timed on this host's cores on a
bounded sample of the same workload.
"""
import argparse
import json
import os
import re
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("cell-updates/s and % HBM peak at 8192^2, FP64 vs FP32 state; "
          "ensemble solves/s")
HBM_PEAK = 8.0e12            # MI355X HBM3E, B/s
B_ALG_FP64 = 376             # bytes per cell-update: 47 compulsory field sweeps
B_ALG_FP32 = 188
# The march schedule's own compulsory traffic: its three kernels read and
# write 10 + 8 + 11 = 29 field sweeps per cell-update.
B_MARCH_FP64 = 232
B_MARCH_FP32 = 116

# Algorithmic bytes per launch, in field sweeps of 8 B over the interior
# (each distinct array a kernel reads counted once, each it writes once).
KERNEL_SWEEPS = {
    # stagewise schedule (one kernel per omp3 loop)
    "eos": 3, "pressure_accel": 8, "av_stress": 5, "av_accel": 9,
    "shock_heating": 8, "timestep": 4, "copy_rho_old": 2,
    "x_flux": 5, "y_flux": 5, "x_update": 7, "y_update": 7,
    "mom_flux_xx": 3, "mom_flux_xy": 4, "mom_flux_yx": 4, "mom_flux_yy": 3,
    "mom_x_update_u": 5, "mom_y_update_v": 5, "mom_x_update": 3,
    "mom_y_update": 3, "conservation": 2,
    # fused and march schedules: the Lagrangian kernel reads rho, e, mom_x,
    # mom_y and writes mx1, my1, u0, v0, e1, rho_old
    "lagrangian": 10, "advect_mass_energy": 8, "advect_momentum_x": 8,
    "advect_momentum_y": 8, "advect_momentum": 11,
}


def parse_args():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=200)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--n", type=int, default=8192)
    p.add_argument("--dt", type=float, default=1.0e-6,
                   help=".params dt (the first step uses C_T * dt)")
    p.add_argument("--seed", type=int, default=20250101)
    p.add_argument("--schedule", choices=("fused", "march", "stagewise"), default=None)
    p.add_argument("--march-rows", type=int, default=None)
    p.add_argument("--graphs", action="store_true")
    p.add_argument("--timing-steps", type=int, default=20)
    p.add_argument("--cpu-seconds", type=float, default=12.0,
                   help="approximate CPU time budget of the oracle sample")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-other-configs", action="store_true",
                   help="skip the Sod 512^2 and 4096^2 blast lines")
    p.add_argument("--no-fp32", action="store_true",
                   help="skip the FP32-state precision form")
    p.add_argument("--fp32-rows", type=int, default=None,
                   help="chunk height of the FP32-state run (default: --march-rows)")
    p.add_argument("--no-ensemble", action="store_true",
                   help="skip the batching form")
    p.add_argument("--ens-count", type=int, default=256)
    p.add_argument("--ens-steps", type=int, default=300)
    p.add_argument("--ens-dt", type=float, default=1.0e-4,
                   help=".params dt of every ensemble instance")
    return p.parse_args()


def reset_state(st, rho0, e0, mesh, dt_param):
    import flow_amd
    for k in flow_amd.ARRAY_NAMES:
        st.zero(k)
    st.upload("rho", rho0)
    st.upload("e", e0)
    st.reflect_initial()
    mesh.set_dt(0.3 * dt_param, 0.3 * dt_param)


def traffic_from_profiles(kernel, n):
    """HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE /
    WRITE_SIZE (scaled by a copy-kernel calibration), if committed; kernel
    is "<schedule>:<timer name>"."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.md")
    if not os.path.exists(path):
        return None
    pat = re.compile(r"^traffic_bytes_per_launch\s+(\S+)\s+(\d+)\s+([0-9.eE+]+)")
    for line in open(path):
        m = pat.match(line.strip())
        if m and m.group(1) == kernel and int(m.group(2)) == n:
            return float(m.group(3))
    return None


def step_roofline(cups, s_per_step, n, sched_name, prof_run, b_alg, b_sched, suffix=""):
    """The whole step against HBM peak three ways: at the reference
    schedule's fixed B_ALG (kept for comparability), at this schedule's own
    compulsory bytes, and at the bytes rocprofv3's counters measured per
    step (the three march kernels' committed traffic lines), when known."""
    out = {"b_alg_per_cell_update": b_alg,
           "achieved_GBps": cups * b_alg / 1e9,
           "frac": cups * b_alg / HBM_PEAK}
    if sched_name == "march":
        out["schedule_b_per_cell_update"] = b_sched
        out["schedule_frac"] = cups * b_sched / HBM_PEAK
        parts = [traffic_from_profiles(f"{prof_run}:{k}{suffix}", n)
                 for k in ("lagrangian", "advect_mass_energy", "advect_momentum")]
        if all(p is not None for p in parts):
            out["pmc_bytes_per_step"] = sum(parts)
            out["pmc_frac"] = sum(parts) / s_per_step / HBM_PEAK
    return out


def field_errors(got, ref, L):
    """Per-field relative L-inf and L1 error of got against ref over the
    regions omp3 defines (rho, e, u, v)."""
    from tests import harness
    g, r = harness.views(got, L), harness.views(ref, L)
    reg = harness.state_regions(L)
    out = {}
    for k in ("rho", "e", "u", "v"):
        a, b = g[k][reg[k]], r[k][reg[k]]
        d = np.abs(a - b)
        out[k] = {"rel_linf": float(d.max() / np.abs(b).max()),
                  "rel_l1": float(d.sum() / np.abs(b).sum())}
    return out


def fp32_leg(args, n, rho0, e0, ref_state, ref_dts, ref_mass, ref_energy,
             fp64_cups):
    """The FP32-state precision form on the same workload: timed like the
    FP64 run, then its state after the same steps against the FP64 state."""
    import flow_amd
    from flow_amd import native
    mesh = flow_amd.Mesh(n, n, dt=args.dt)
    rows = args.fp32_rows or args.march_rows
    if rows:
        mesh.set_march_rows(rows)
    st = flow_amd.DeviceState32(mesh)

    def reset():
        for k in flow_amd.ARRAY_NAMES:
            st.zero(k)
        st.upload("rho", rho0)
        st.upload("e", e0)
        st.reflect_initial()
        mesh.set_dt(0.3 * args.dt, 0.3 * args.dt)

    reset()
    if args.warmup:
        st.run(args.warmup)
    reset()
    st.run(0)  # the pre-loop set_timestep
    native.sync()
    t0 = time.perf_counter()
    steps, dts, mass, energy, _ = st.run(args.steps, pre_timestep=False)
    native.sync()
    elapsed = time.perf_counter() - t0
    if steps != args.steps:
        raise RuntimeError(f"FP32 run stopped after {steps} steps")
    cups = n * n * steps / elapsed
    got = {k: st.download(k) for k in ("rho", "e", "u", "v", "mom_x", "mom_y")}
    errs = field_errors(got, ref_state, n + 4)
    mesh.reset_timing()
    mesh.set_timing(True)
    reset()
    st.run(args.timing_steps)
    mesh.set_timing(False)
    timers = mesh.timing()
    st.close()
    mesh.close()
    return {
        "value": cups, "unit": "cell-updates/s", "ms_per_step": elapsed / steps * 1e3,
        "speedup_vs_fp64": cups / fp64_cups,
        "step_roofline": step_roofline(cups, elapsed / steps, n, "march", "fp32",
                                       B_ALG_FP32, B_MARCH_FP32, suffix="_f32"),
        "error_vs_fp64": errs,
        "max_field_rel_linf": max(v["rel_linf"] for v in errs.values()),
        "mass_rel_diff": float(abs(mass[-1] - ref_mass[-1]) / abs(ref_mass[-1])),
        "energy_rel_diff": float(abs(energy[-1] - ref_energy[-1]) / abs(ref_energy[-1])),
        "mass_drift": float((mass[-1] - mass[0]) / mass[0]),
        "energy_drift": float((energy[-1] - energy[0]) / energy[0]),
        "dt_rel_diff_max": float(np.max(np.abs(dts[:, 0] - ref_dts[:, 0]) / ref_dts[:, 0])),
        "kernels": {k: v[0] / v[1] for k, v in timers.items()},
        "comparator": "the FP64 GPU state after the same steps (bitwise equal "
                      "to the oracle: parity_vs_oracle)",
    }


def ensemble_leg(args):
    """The batching form on the seeded ensemble workload: args.ens_count
    independent Sod/blast instances of sides log-uniform in [64, 1024],
    every instance its own dt, args.ens_steps steps, FP64 and FP32 state.
    The autotuner picks the chunk height and work order first (from the
    initial state, which it restores); steps are replayed from hipGraphs."""
    import flow_amd
    from flow_amd import problems
    specs = problems.ensemble_specs(count=args.ens_count)
    sizes = [sp[0] for sp in specs]
    cells = float(sum(n * n for n in sizes))
    out = {"instances": len(sizes), "steps": args.ens_steps, "cells": cells,
           "sizes": {"min": min(sizes), "max": max(sizes)},
           "workload": f"{len(sizes)} seeded Sod/blast instances, sides log-U[64,1024], "
                       f"split U[0.3,0.7], pressure ratio log-U[10,1000], own dt each"}
    for label, fp32 in (("fp64", False), ("fp32", True)):
        ens = flow_amd.Ensemble(sizes, dt=args.ens_dt, fp32=fp32)
        for k, sp in enumerate(specs):
            ens.set_state(k, *problems.ensemble_instance(*sp))
        (rows, order, fused), tuned_ms = ens.autotune(trial_steps=3)
        ens.set_graphs(True)
        dts, mass, energy, wall = ens.run(args.ens_steps)
        ok = bool(np.all(np.isfinite(dts)) and np.all(dts[:, :, 0] > 0))
        out[label] = {
            "solves_per_s": len(sizes) / wall,
            "unit": "ensemble solves/s (one solve = one instance advanced "
                    f"{args.ens_steps} steps)",
            "cell_updates_per_s": cells * args.ens_steps / wall,
            "wall_s": wall, "ms_per_step": wall / args.ens_steps * 1e3,
            "autotuned": {"chunk_rows": rows, "work_order": order,
                          "fused_momentum": fused, "trial_ms_per_step": tuned_ms},
            "finite": ok,
            "dt_spread": [float(dts[:, -1, 0].min()), float(dts[:, -1, 0].max())],
        }
        ens.close()
    return out


# omp3's dataflow moves 95 field sweeps per cell-update (760 B in FP64): the
# CPU baseline's effective bandwidth is stated at that figure.
B_OMP3_FP64 = 760


# The affinity mask as the process started: once the oracle's OpenMP runtime
# loads with OMP_PROC_BIND set, it binds the calling thread to one place.
START_AFFINITY = sorted(os.sched_getaffinity(0))


def host_topology():
    """This process's CPUs: the affinity mask it started with, and the
    physical cores, sockets and NUMA nodes behind it (lscpu -p)."""
    allowed = START_AFFINITY
    info = {"affinity_cpus": len(allowed), "logical_cpus": os.cpu_count()}
    try:
        out = subprocess.run(["lscpu", "-p=CPU,CORE,SOCKET,NODE"], capture_output=True,
                             text=True, timeout=30).stdout
        rows = [tuple(int(x) if x else -1 for x in line.split(","))
                for line in out.splitlines() if line and not line.startswith("#")]
        mine = [r for r in rows if r[0] in set(allowed)]
        info["sockets"] = len({r[2] for r in rows})
        info["physical_cores"] = len({(r[2], r[1]) for r in rows})
        info["numa_nodes"] = len({r[3] for r in rows})
        info["affinity_physical_cores"] = len({(r[2], r[1]) for r in mine})
        info["affinity_numa_nodes"] = len({r[3] for r in mine})
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    return info


def oracle_threads():
    """OpenMP threads and binding for the oracle: OMP_NUM_THREADS if set,
    else one per CPU of the affinity mask; cores placed and spread, unless
    the environment says otherwise. Set before the oracle's OpenMP runtime
    starts (it reads them once, when the library loads)."""
    os.environ.setdefault("OMP_PLACES", "cores")
    os.environ.setdefault("OMP_PROC_BIND", "spread")
    threads = int(os.environ.get("OMP_NUM_THREADS", 0) or len(START_AFFINITY))
    os.environ["OMP_NUM_THREADS"] = str(threads)
    return threads


def oracle_sample(n, dt_param, rho0, e0, budget_s, max_steps=None):
    """The oracle from an initial state through main.c's loop (the pre-loop
    set_timestep untimed, then solve_hydro_2d + print_conservation per step,
    timed as main.c:80-93 times them) for as many steps as fit in about
    budget_s seconds (at least one, at most max_steps). Returns
    (cell-updates/s, steps, wall seconds, arrays)."""
    from oracle import pyoracle as po
    m = po.Mesh(n, n, dt=dt_param)
    a = po.new_arrays(m)
    a["rho"][:] = rho0
    a["e"][:] = e0
    po.reflect_initial(m, a)
    po.set_timestep(m, a, 0)
    wall, steps = 0.0, 0
    while True:
        t0 = time.perf_counter()
        po.solve_step(m, a, steps)
        po.conservation(m, a)
        wall += time.perf_counter() - t0
        steps += 1
        if max_steps and steps >= max_steps:
            break
        if wall >= budget_s or wall / steps * (steps + 1) > 2.0 * budget_s:
            break
    return n * n * steps / wall, steps, wall, a


def cpu_baseline(n, dt_param, budget_s, rho0, e0):
    """The oracle on this host's cores, from the multi-blast initial state,
    for as many steps as fit in about budget_s seconds."""
    threads = oracle_threads()
    cups, steps, wall, a = oracle_sample(n, dt_param, rho0, e0, budget_s)
    return {
        "value": cups, "unit": "cell-updates/s",
        "cores": threads, "kind": "port",
        "sample": f"oracle (C+OpenMP restatement of omp3/flow.c) on the "
                  f"{n}x{n} multi-blast, first {steps} steps "
                  f"({wall:.1f} s, {threads} OpenMP threads, "
                  f"OMP_PLACES={os.environ['OMP_PLACES']} "
                  f"OMP_PROC_BIND={os.environ['OMP_PROC_BIND']})",
        "steps": steps, "wall_s": wall,
        "effective_GBps_at_760B": cups * B_OMP3_FP64 / 1e9,
    }, a, steps


def other_configs(args):
    """The reference's CPU configuration (Sod 512^2, 100 steps, .params dt
    1e-4) and the 4096^2 blast (500 steps, .params dt 1e-4): the GPU over
    the whole run (interface-driven, as main.c drives it) and the oracle on
    the same host over config 1 whole and a bounded sample of config 2."""
    import flow_amd
    from flow_amd import native, problems
    oracle_threads()
    out = {}
    for label, fn, n, steps, budget in (("sod_512_100_steps", problems.sod_x, 512, 100, 30.0),
                                        ("blast_4096_500_steps", problems.blast_4096, 4096,
                                         500, 3.0)):
        rho0, e0 = fn(n)
        mesh = flow_amd.Mesh(n, n, dt=1.0e-4)
        st = flow_amd.DeviceState(mesh)
        reset_state(st, rho0, e0, mesh, 1.0e-4)
        st.run(min(steps, 20))                      # warm-up
        reset_state(st, rho0, e0, mesh, 1.0e-4)
        st.set_timestep(0)
        native.sync()
        t0 = time.perf_counter()
        done = st.run(steps, pre_timestep=False)[0]
        native.sync()
        gpu_s = time.perf_counter() - t0
        und = st.undefined
        st.close()
        mesh.close()
        cups, csteps, cwall, _ = oracle_sample(n, 1.0e-4, rho0, e0, budget,
                                               max_steps=steps)
        out[label] = {
            "cells": n * n, "steps": done,
            "gpu_cell_updates_per_s": n * n * done / gpu_s,
            "gpu_ms_per_step": gpu_s / done * 1e3,
            "steps_with_undefined_cfl_candidates": int((und[1:] > 0).sum()),
            "cpu_cell_updates_per_s": cups,
            "cpu_sample": f"{csteps} steps in {cwall:.2f} s",
            "cpu_effective_GBps_at_760B": cups * B_OMP3_FP64 / 1e9,
        }
    return out


def main():
    args = parse_args()
    if args.gpus != 1:
        print(f"bench.py: the hip kernel set runs on one GPU; --gpus {args.gpus} "
              "is not supported (the multi-GPU row-slab design is documented, "
              "not built)", file=sys.stderr)
        sys.exit(2)
    import flow_amd
    from flow_amd import native, problems

    n = args.n
    rho0, e0 = problems.multi_blast(n, seed=args.seed)
    mesh = flow_amd.Mesh(n, n, dt=args.dt)
    fused_ok = hasattr(native.lib(), "fk_lagrangian")
    sched_name = args.schedule or ("march" if fused_ok else "stagewise")
    mesh.set_schedule({"fused": flow_amd.FUSED, "march": flow_amd.MARCH,
                       "stagewise": flow_amd.STAGEWISE}[sched_name])
    if args.march_rows:
        mesh.set_march_rows(args.march_rows)
    mesh.set_graphs(args.graphs)
    st = flow_amd.DeviceState(mesh)

    # warm-up, then the timed run from the initial state
    reset_state(st, rho0, e0, mesh, args.dt)
    if args.warmup:
        st.run(args.warmup)
    reset_state(st, rho0, e0, mesh, args.dt)
    st.set_timestep(0)
    native.sync()
    t0 = time.perf_counter()
    steps, dts, mass, energy, _ = st.run(args.steps, pre_timestep=False)
    native.sync()
    t1 = time.perf_counter()
    if steps != args.steps or not np.all(np.isfinite(dts)):
        raise RuntimeError(f"run stopped after {steps} steps")
    elapsed = t1 - t0
    cups = n * n * steps / elapsed
    fp32 = None
    if not args.no_fp32:
        ref_state = {k: st.download(k) for k in ("rho", "e", "u", "v", "mom_x", "mom_y")}
        fp32 = fp32_leg(args, n, rho0, e0, ref_state, dts, mass, energy, cups)
        del ref_state

    # per-kernel HIP-event timing over the same workload
    reset_state(st, rho0, e0, mesh, args.dt)
    st.set_timestep(0)
    mesh.reset_timing()
    mesh.set_timing(True)
    st.run(args.timing_steps, pre_timestep=False)
    mesh.set_timing(False)
    timers = mesh.timing()
    dom = max((k for k in timers if k in KERNEL_SWEEPS),
              key=lambda k: timers[k][0])
    dom_ms, dom_launches = timers[dom]
    avg_s = dom_ms / dom_launches * 1e-3
    alg_bytes = KERNEL_SWEEPS[dom] * 8.0 * n * n
    achieved = alg_bytes / avg_s
    roofline = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                "unit": "GB/s", "frac": achieved / HBM_PEAK,
                "traffic": traffic_from_profiles(f"{sched_name}:{dom}", n),
                "kernel": dom, "avg_launch_ms": dom_ms / dom_launches,
                "alg_bytes_per_launch": alg_bytes}
    kernels = {k: {"ms_per_launch": v[0] / v[1], "launches_per_step": v[1] / args.timing_steps}
               for k, v in sorted(timers.items(), key=lambda kv: -kv[1][0])}

    # the other way of timing the step: the same run replayed from captured
    # hipGraphs (value stays the interface-driven figure above)
    mesh.set_graphs(not args.graphs)
    reset_state(st, rho0, e0, mesh, args.dt)
    st.set_timestep(0)
    st.run(2, pre_timestep=False)  # capture both sweep parities
    reset_state(st, rho0, e0, mesh, args.dt)
    st.set_timestep(0)
    native.sync()
    g0 = time.perf_counter()
    g_steps = st.run(args.steps, pre_timestep=False)[0]
    native.sync()
    g_elapsed = time.perf_counter() - g0
    mesh.set_graphs(args.graphs)
    other_timing = {"graphs": not args.graphs, "steps": g_steps,
                    "ms_per_step": g_elapsed / g_steps * 1e3,
                    "value": n * n * g_steps / g_elapsed}

    result = {
        "metric": METRIC, "value": cups, "unit": "cell-updates/s", "n_gpus": 1,
        "steps": steps, "warmup": args.warmup,
        "ms_per_step": elapsed / steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"multi-blast {n}x{n} FP64 (4x4 seeded blasts, "
                               f"p log-U[10,1000], rho noise 5%, seed {args.seed}), "
                               f"{steps} steps",
                   "cells": n * n, "schedule": sched_name, "graphs": args.graphs,
                   "dt_param": args.dt},
        "roofline": roofline,
        "step_roofline": step_roofline(cups, elapsed / steps, n, sched_name, "march",
                                       B_ALG_FP64, B_MARCH_FP64),
        "kernels": kernels,
        "graph_replay" if not args.graphs else "direct_launch": other_timing,
        "fp32_state": fp32,
        "device": native.device_name(),
    }

    if not args.no_cpu_baseline:
        base, ora, k = cpu_baseline(n, args.dt, args.cpu_seconds, rho0, e0)
        result["cpu_baseline"] = {kk: base[kk] for kk in
                                  ("value", "unit", "cores", "kind", "sample",
                                   "effective_GBps_at_760B")}
        result["cpu_baseline"]["host"] = host_topology()
        # parity at full size on the same sample: k GPU steps vs k oracle steps
        reset_state(st, rho0, e0, mesh, args.dt)
        got_steps = st.run(k)[0]
        got = st.download_all()
        from tests import harness
        cmp = harness.compare_state(got, ora, n + 4)
        result["parity_vs_oracle"] = {
            "steps": got_steps,
            "max_rel_linf": max(v[0] for v in cmp.values()),
            "bitwise_diffs": sum(v[1] for v in cmp.values())}
    else:
        result["cpu_baseline"] = None

    st.close()
    mesh.close()
    result["ensemble"] = None if args.no_ensemble else ensemble_leg(args)
    if not (args.no_other_configs or args.no_cpu_baseline):
        result["other_configs"] = other_configs(args)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
