#!/usr/bin/env python
"""Throughput of FP64 micromechanical MD on one MI355X: cell-steps per second at 128^3 cells.

Workload (``npt128``): periodic 128 x 128 x 128 grid, type B cells with probability 0.2
(two metastable states), 5% empty cells, NPT at 300 K and 1 bar with a Nose-Hoover chain
(length 3, 100 fs) and the MTK barostat (full 3x3 cell fluctuation, 1000 fs), dt = 1 fs, FP64,
eight-corner nanocell model. Each step evaluates the force three times (MTK pre, main, MTK post).
Cell-steps count grid cells n^3 (empty cells included) times steps. Inputs are synthetic and
resident on the GPU before the timed region starts.

Prints one JSON line. ``roofline`` refers to the dominant kernel (the per-cell force pass),
timed with HIP events on the stream it is launched on; ``cpu_baseline`` times the reference's
algorithm restated as a literal per-cell loop on one host core, on a bounded sample of the same
workload (an 8^3 grid), with the C restatement on all host cores beside it. Further legs
report the other workloads: ``batched_replicas`` (512 uneven replicas as one batch and one at
a time), ``nvt_nhc32`` (32^3 NVT with a Nose-Hoover chain, 2000 steps in one call),
``cg_relaxation`` (CG at 128^3) and ``differentiation`` (the 64 x 16^3 Adam fit).
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FP64 MD cell-steps/s @128³ cells; batched-replica vs plain speedup same GPU"
PEAK_FP64 = 78.6e12      # MI355X FP64 (vector and matrix) peak, FLOP/s
PEAK_HBM = 8.0e12        # MI355X HBM3E, B/s
# algorithmic FLOPs of the eight-corner model: per cell and state, per cell (edges, virial,
# log-sum-exp), per node (gather and Verlet update)
FLOP_CELL_STATE = 2290
FLOP_CELL = 200
FLOP_NODE = 54


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--grid", type=int, default=128)
    ap.add_argument("--model", default="corner8", choices=["corner8", "original"])
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--cpu-baseline", dest="cpu", action="store_true", default=True)
    ap.add_argument("--no-cpu-baseline", dest="cpu", action="store_false")
    ap.add_argument("--cpu-grid", type=int, default=8)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--kernel-launches", type=int, default=20)
    ap.add_argument("--batched", dest="batched", action="store_true", default=True)
    ap.add_argument("--no-batched", dest="batched", action="store_false")
    ap.add_argument("--replicas", type=int, default=512)
    ap.add_argument("--batched-steps", type=int, default=100)
    ap.add_argument("--sequential-steps", type=int, default=20)
    ap.add_argument("--adjoint", dest="adjoint", action="store_true", default=True)
    ap.add_argument("--no-adjoint", dest="adjoint", action="store_false")
    ap.add_argument("--adjoint-steps", type=int, default=500)
    ap.add_argument("--cg", dest="cg", action="store_true", default=True)
    ap.add_argument("--no-cg", dest="cg", action="store_false")
    ap.add_argument("--cg-steps", type=int, default=200)
    ap.add_argument("--nvt", dest="nvt", action="store_true", default=True)
    ap.add_argument("--no-nvt", dest="nvt", action="store_false")
    ap.add_argument("--nvt-steps", type=int, default=2000)
    ap.add_argument("--pmc-probe", type=int, default=0,
                    help="launch only the cell pass this many times, then copy 1 GiB with a "
                         "plain copy kernel (the counter calibration), and exit")
    return ap.parse_args()


def build_workload(n, model):
    import numpy as np
    from micmec_amd import units
    from micmec_amd.pes import ForcePartMechanical, MicMecForceField
    from micmec_amd.sampling import MTKBarostat, NHCThermostat, VerletIntegrator
    from micmec_amd.workloads import make_system
    system = make_system((n, n, n), p_b=0.2, p_hole=0.05, seed=2)
    fpm = ForcePartMechanical(system, model=model)
    mmf = MicMecForceField(system, [fpm])
    np.random.seed(2)
    thermo = NHCThermostat(temp=300.0, timecon=100 * units.femtosecond, chainlength=3)
    baro = MTKBarostat(mmf, temp=300.0, press=1.0 * units.bar,
                       timecon=1000 * units.femtosecond)
    verlet = VerletIntegrator(mmf, timestep=1.0 * units.femtosecond, hooks=[thermo, baro],
                              temp0=300.0)
    return system, fpm, verlet


def kernel_timing(fpm, verlet, launches):
    """Average duration of the cell pass, HIP events on the stream it is launched on."""
    import ctypes
    import torch
    from micmec_amd import _native
    lib = _native.lib()
    eng = verlet._engine
    stream = torch.cuda.current_stream()
    sp = _native.stream_ptr(stream)
    rv = eng.S[49:58]
    for _ in range(3):
        _native.check(lib.mm_field_launch_cells(fpm.field.handle, _native.ptr(eng.pos),
                                                ctypes.c_void_p(rv.data_ptr()), sp))
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(launches):
        _native.check(lib.mm_field_launch_cells(fpm.field.handle, _native.ptr(eng.pos),
                                                ctypes.c_void_p(rv.data_ptr()), sp))
    stop.record(stream)
    torch.cuda.synchronize()
    return start.elapsed_time(stop) / 1e3 / launches


def algorithmic_counts(system, model):
    """FLOPs and compulsory bytes of one cell pass (one force evaluation)."""
    import numpy as np
    from micmec_amd.pes import parse_type_params
    type_ids, ns, *_ = parse_type_params(system.params, system.types)
    per_type = dict(zip(type_ids, ns))
    types = np.rint(np.asarray(system.types, dtype=float)).astype(int)
    states = np.array([per_type[t] for t in types]) if len(types) else np.zeros(0)
    scale = 1.0 if model == "corner8" else 1.0 / 7.0
    flops = float(np.sum(FLOP_CELL + FLOP_CELL_STATE * scale * states))
    # per cell: 8 corner indices + wrap + type (40 B); per node: position gathered once (24 B)
    nbytes = 40.0 * len(types) + 24.0 * system.nnodes
    return flops, nbytes, float(states.mean()) if len(states) else 0.0


def read_traffic():
    """HBM bytes per cell-pass launch from the committed rocprofv3 counter summary, if any."""
    fn = os.path.join(ROOT, "profiles", "cell_kernel_npt128_traffic.txt")
    if not os.path.exists(fn):
        return None
    with open(fn) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    return None


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _oracle_md(s, force):
    import numpy as np
    from micmec_amd import units
    from oracle.integrators import OracleMD
    np.random.seed(2)
    fs = units.femtosecond
    return OracleMD(force, s.pos, np.array(s.domain.rvecs), s.masses, fs,
                    thermo=dict(temp=300.0, timecon=100 * fs),
                    baro=dict(temp=300.0, press=units.bar, timecon=1000 * fs))


def _timed_steps(md, seconds):
    steps = 0
    t0 = time.perf_counter()
    while True:
        md.propagate()
        steps += 1
        if time.perf_counter() - t0 >= seconds:
            break
    return steps, time.perf_counter() - t0


def cpu_baseline(n, seconds, model):
    """The reference's CPU path, restated: the literal per-cell loop of micmec/pes/mmff.py:
    326-403 (one numpy evaluation per cell and state, image shifts from the N x N table of
    mmff.py:258-286, built once and not timed) driving the Verlet/NHC/MTK sequence, on one host
    core, on an n^3 sample of the NPT workload for about ``seconds``; beside it the cost per
    cell evaluation at 3^3..12^3 and the C restatement on all host cores."""
    import numpy as np
    from micmec_amd.workloads import make_system
    from oracle import cforce
    from oracle import mmff as omm
    from oracle.topology import get_mic

    def loop_force(s):
        mic = get_mic(np.rint(s.grid).astype(int), True)
        d = {"pos": s.pos, "rvecs": s.domain.rvecs, "grid": s.grid, "types": s.types,
             "surrounding_nodes": s.surrounding_nodes,
             "surrounding_cells": s.surrounding_cells, "params": s.params}

        def force(p, r):
            d["pos"], d["rvecs"] = p, r
            return omm.compute_loop(d, model, mic=mic)
        return force

    per_cell = {}
    for m in (3, 4, 6, 8, 12):
        s = make_system((m, m, m), p_b=0.2, p_hole=0.05, seed=2)
        f = loop_force(s)
        t0 = time.perf_counter()
        f(s.pos, np.array(s.domain.rvecs))
        per_cell["%d^3" % m] = (time.perf_counter() - t0) / m ** 3 * 1e6
    s = make_system((n, n, n), p_b=0.2, p_hole=0.05, seed=2)
    steps, dt = _timed_steps(_oracle_md(s, loop_force(s)), seconds)
    out = {"value": n ** 3 * steps / dt, "unit": "cell-steps/s", "cores": 1, "kind": "port",
           "sample": "literal per-cell loop (oracle/mmff.py compute_loop, the reference's "
                     "mmff.py:258-403 restated) + Verlet/NHC/MTK (oracle/integrators.py), "
                     "%d^3 grid of the NPT workload, %d steps (3 force evaluations each) in "
                     "%.1f s" % (n, steps, dt),
           "us_per_cell_eval": per_cell,
           "cpu_model": _cpu_model(), "host_cpus": os.cpu_count()}
    if cforce.available():
        m = 32
        s = make_system((m, m, m), p_b=0.2, p_hole=0.05, seed=2)
        steps, dt = _timed_steps(_oracle_md(s, cforce.make_force(s, model)), 0.5 * seconds)
        out["all_cores"] = {
            "value": m ** 3 * steps / dt, "unit": "cell-steps/s", "cores": cforce.num_threads(),
            "sample": "C restatement of the force field (oracle/c, OpenMP over cells) + numpy "
                      "Verlet/NHC/MTK, %d^3 grid of the NPT workload, %d steps in %.1f s"
                      % (m, steps, dt)}
    return out


def batched_vs_sequential(count, steps, seq_steps):
    """Aggregate cell-steps/s of the replica set run as one batch and one replica at a time."""
    import gc
    import numpy as np
    import torch
    from micmec_amd import units
    from micmec_amd.batch import BatchVerlet
    from micmec_amd.pes import ForcePartMechanical, MicMecForceField
    from micmec_amd.sampling import NHCThermostat, VerletIntegrator
    from micmec_amd.workloads import batch_replicas
    reps = batch_replicas(count)
    cells = np.array([int(np.prod(np.shape(s.grid))) for s, _, _ in reps])
    fs = units.femtosecond
    batch = BatchVerlet([r[0] for r in reps], 1.0 * fs, [r[1] for r in reps],
                        [r[2] for r in reps], thermostat=True)
    batch.run(5)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    batch.run(steps)
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    del batch
    gc.collect()
    t_seq = 0.0
    for s, temp, seed in reps:
        mmf = MicMecForceField(s, [ForcePartMechanical(s)])
        np.random.seed(seed)
        v = VerletIntegrator(mmf, timestep=1.0 * fs, temp0=temp,
                             hooks=[NHCThermostat(temp=temp, timecon=100 * fs)])
        v.run(2)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        v.run(seq_steps)
        torch.cuda.synchronize()
        t_seq += time.perf_counter() - t0
        del v, mmf
    value_b = cells.sum() * steps / t_batch
    value_s = cells.sum() * seq_steps / t_seq
    return {"workload": "batch%d" % count, "replicas": count,
            "grid_cells": int(cells.sum()), "ensemble": "NVT NHC(3), dt 1 fs",
            "batched": {"value": value_b, "unit": "cell-steps/s", "steps": steps,
                        "ms_per_step": t_batch / steps * 1e3},
            "one_at_a_time": {"value": value_s, "unit": "cell-steps/s",
                              "steps_per_replica": seq_steps,
                              "ms_per_step_all_replicas": t_seq / seq_steps * 1e3},
            "speedup": value_b / value_s}


def pmc_probe(n, model, launches):
    """Counter probe: the cell pass alone, then a 1 GiB device copy (bytes known exactly)."""
    import torch
    system, fpm, verlet = build_workload(n, model)
    torch.cuda.synchronize()
    t = kernel_timing(fpm, verlet, launches)
    a = torch.empty(1 << 27, dtype=torch.float64, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"pmc_probe": {"cell_pass_launches": launches + 3, "cell_pass_ms": t * 1e3,
                                    "copy_bytes_each_way": a.numel() * 8, "copies": 3}}))


def nvt_nhc32(steps):
    """The mid-size plain workload: 32^3 cells, 20% type B, NVT with a Nose-Hoover chain of 3
    at 300 K (timecon 100 fs), dt 1 fs, the steps of one verlet.run call."""
    import numpy as np
    import torch
    from micmec_amd import units
    from micmec_amd.pes import ForcePartMechanical, MicMecForceField
    from micmec_amd.sampling import NHCThermostat, VerletIntegrator
    from micmec_amd.workloads import make_system, workload
    w = workload("nhc32", steps=steps)
    system = make_system(w["shape"], p_b=w["p_b"], p_hole=w["p_hole"], seed=w["seed"])
    mmf = MicMecForceField(system, [ForcePartMechanical(system)])
    np.random.seed(w["seed"])
    fs = units.femtosecond
    thermo = NHCThermostat(temp=w["temp"], timecon=w["timecon_fs"] * fs, chainlength=w["chain"])
    verlet = VerletIntegrator(mmf, timestep=w["timestep_fs"] * fs, hooks=[thermo],
                              temp0=w["temp"])
    verlet.run(50)
    e0, k0 = verlet.econs, verlet.ekin
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    verlet.run(w["steps"])
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return {"workload": "nhc32", "grid": list(w["shape"]), "cells": system.ncells,
            "steps": w["steps"], "ms_per_step": dt / w["steps"] * 1e3,
            "cell_steps_per_s": system.ncells * w["steps"] / dt,
            "econs_drift_over_ekin": abs(verlet.econs - e0) / k0}


def cg_relaxation(n, steps, model):
    """CG geometry optimisation of the 128^3 system from a 2% random strain (positions only,
    domain fixed): `steps` iterations, thresholds tight enough that all of them run."""
    import torch
    from micmec_amd.pes import ForcePartMechanical, MicMecForceField
    from micmec_amd.sampling import CartesianDOF, CGOptimizer
    from micmec_amd.workloads import make_system
    system = make_system((n, n, n), p_b=0.2, p_hole=0.05, seed=2, strain=0.02)
    mmf = MicMecForceField(system, [ForcePartMechanical(system, model=model)])
    dof = CartesianDOF(mmf, gpos_rms=1e-30, dpos_rms=1e-30)
    opt = CGOptimizer(dof)
    e0 = opt.epot
    torch.cuda.synchronize()
    calls0 = dof._device.calls
    t0 = time.perf_counter()
    opt.run(steps)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    evals = dof._device.calls - calls0
    return {"workload": "cg128", "grid": [n, n, n], "iterations": opt.counter,
            "force_evals": evals, "ms_per_iteration": dt / max(opt.counter, 1) * 1e3,
            "ms_per_force_eval": dt / max(evals, 1) * 1e3,
            "cell_evals_per_s": n ** 3 * evals / dt, "energy_first": e0,
            "energy_last": opt.epot, "gpos_rms_last": float(dof.gpos_rms)}


def differentiation(steps):
    """Force-and-stress fit of the elasticity matrices: 64 configurations on 16^3 grids, each
    Adam step one forward force pass and one dual-number adjoint pass over all of them."""
    import torch
    from micmec_amd.adjoint import fit, fit_workload
    problem, init, target = fit_workload(count=64, n=16)
    field = problem.field
    fit(problem, init, 3)                      # warm-up (library plans, allocator)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    losses, final = fit(problem, init, steps)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    # the adjoint launch alone, HIP events on the current stream
    pos, rv = problem.pos, problem.rvecs
    dpos = torch.randn_like(pos)
    drv = torch.randn_like(rv)
    field.set_params(final)
    field.adjoint(pos, rv, dpos, drv)
    stream = torch.cuda.current_stream()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 20
    ev0.record(stream)
    for _ in range(reps):
        field.adjoint(pos, rv, dpos, drv)
    ev1.record(stream)
    torch.cuda.synchronize()
    t_adj = ev0.elapsed_time(ev1) / reps * 1e-3
    err = float(((final.cv - target.cv).abs().max() / target.cv.abs().max()))
    return {"workload": "fit64x16", "configurations": field.nsys, "cells": field.ncells,
            "adam_steps": steps, "ms_per_adam_step": dt / steps * 1e3,
            "cell_evals_per_s": 2 * field.ncells * steps / dt,
            "adjoint_call_ms": t_adj * 1e3,
            "loss_first": losses[0], "loss_last": losses[-1],
            "cv_max_rel_error_vs_target": err}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on one GPU only; --gpus %d is not supported (this build measures "
              "one GPU's share; multi-GPU sharding is a design note)." % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(2)
    n = args.grid
    if args.pmc_probe > 0:
        pmc_probe(n, args.model, args.pmc_probe)
        return
    system, fpm, verlet = build_workload(n, args.model)
    verlet.use_graph = not args.no_graph
    print("workload built: %d cells (%d non-empty), %d nodes" % (n ** 3, system.ncells,
                                                                 system.nnodes),
          file=sys.stderr, flush=True)
    verlet.run(args.warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    verlet.run(args.steps)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    value = n ** 3 * args.steps / elapsed
    print("timed %d steps in %.3f s" % (args.steps, elapsed), file=sys.stderr, flush=True)

    t_cell = kernel_timing(fpm, verlet, args.kernel_launches)
    flops, nbytes, mean_states = algorithmic_counts(system, args.model)
    achieved = flops / t_cell
    roofline = {"bound": "fp64_valu", "achieved": achieved / 1e12, "peak": PEAK_FP64 / 1e12,
                "unit": "TFLOP/s", "frac": achieved / PEAK_FP64, "traffic": read_traffic(),
                "kernel": "cell_kernel<corner8>" if args.model == "corner8" else
                          "cell_kernel<original>",
                "avg_launch_ms": t_cell * 1e3, "flop_per_launch": flops,
                "compute_unit": "FP64 VALU (no MFMA; 78.6 TFLOP/s is the MI355X FP64 peak)"}
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "npt128", "grid": [n, n, n], "cells_nonempty": system.ncells,
                   "nodes": system.nnodes, "mean_states": mean_states,
                   "ensemble": "NPT NHC(3)+MTK, 300 K, 1 bar", "timestep_fs": 1.0,
                   "force_evals_per_step": 3, "nanocell": args.model,
                   "parallelism": "single GPU", "graph": not args.no_graph},
        "roofline": roofline,
    }
    if args.batched:
        print("batched replicas ...", file=sys.stderr, flush=True)
        out["batched_replicas"] = batched_vs_sequential(args.replicas, args.batched_steps,
                                                        args.sequential_steps)
    if args.nvt:
        print("NVT nhc32 ...", file=sys.stderr, flush=True)
        out["nvt_nhc32"] = nvt_nhc32(args.nvt_steps)
    if args.cg:
        print("CG relaxation ...", file=sys.stderr, flush=True)
        out["cg_relaxation"] = cg_relaxation(n, args.cg_steps, args.model)
    if args.adjoint:
        print("differentiation fit ...", file=sys.stderr, flush=True)
        out["differentiation"] = differentiation(args.adjoint_steps)
    if args.cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_grid, args.cpu_seconds, args.model)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
