"""Benchmark of the Ni-Cr phase-field time step on one MI355X.

Workload (default): 256^3 cells, 32 eta fields evolving (64 Voronoi seeds, each
assigned one of 32 eta fields by colouring, so neighbouring grains differ),
Cr = 0.20 +- U(0.005) noise, O = 0 with the inlet layer at x = 0, Ksp
insertion on, --bounds legacy, repo-default mobilities and dt. One step = one
accepted time step over the whole mesh (mu, mobility/flux/phi trial, eta trial,
acceptance). Inputs are resident in HBM when the timed region starts.

Prints ONE JSON line: value = mixed-precision Gcell-steps/s; the FP64 step is
measured on the same state and reported beside it.
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
PKG_NAME = "ni-cr_oxidation_phase-field_simulation_amd"
HBM_PEAK_GBS = 8000.0


def _pkg():
    if REPO not in sys.path:
        sys.path.insert(0, REPO)
    return importlib.import_module(PKG_NAME)


def _wl():
    return importlib.import_module(PKG_NAME + ".workloads")


def build_state(nicr, shape, n_seeds, n_eta, seed):
    """The 3D evolving-grain workload's state (host arrays)."""
    t0 = time.time()
    state = _wl().evolving_grain_state(shape, n_seeds, n_eta, seed)
    print(f"[bench] state {shape} with {n_seeds} seeds -> {n_eta} eta built in {time.time() - t0:.1f}s",
          file=sys.stderr, flush=True)
    return state


def make_model(nicr, shape, n_eta, precision, state):
    return _wl().evolving_grain_model(precision, state, n_eta)


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter runs
    (profiles/traffic_<kernel>_256cube_32eta.json: FETCH_SIZE and WRITE_SIZE per
    dispatch, each scaled by the same counter's reading for a copy of known size),
    or None when no such measurement is committed."""
    path = os.path.join(REPO, "profiles", f"traffic_{kernel}_256cube_32eta.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return {"bytes_per_launch": d["hbm_bytes_per_dispatch"], "read": d["read_bytes_per_dispatch"],
            "write": d["write_bytes_per_dispatch"], "source": os.path.relpath(path, REPO)}


def time_steps(m, steps, warmup, label):
    """Wall time of `steps` steps after `warmup` untimed ones, with no event
    markers in the stream; then the same number of steps again with a HIP
    event pair around every kernel launch on the context's stream, for the
    per-kernel durations (the markers add a few microseconds per launch, so
    they stay out of the wall-clock window)."""
    r = m.try_step(warmup)
    if r.status != 0:
        raise RuntimeError(f"{label}: warmup step rejected (status {r.status})")
    m.sync()
    t0 = time.perf_counter()
    r = m.try_step(steps)
    m.sync()
    t1 = time.perf_counter()
    if r.status != 0 or r.steps_done != steps:
        raise RuntimeError(f"{label}: timed step rejected (status {r.status})")
    m.reset_kernel_times()
    m.set_timing(True)
    t2 = time.perf_counter()
    r = m.try_step(steps)
    m.sync()
    t3 = time.perf_counter()
    m.set_timing(False)
    if r.status != 0 or r.steps_done != steps:
        raise RuntimeError(f"{label}: timed step rejected (status {r.status})")
    return (t1 - t0), m.kernel_times(), (t3 - t2)


def cpu_baseline_reference(n_eta):
    """The reference's own 3D program (oracle/_ref/nicr3d_emulated: mpi3d-slab
    under its thread-emulated MPI driver, built from the reference sources) on
    the host cores: a 64 x 128 x 128 slab with n_eta evolving grains (one eta
    per grain, the reference's own grain mode), bounds legacy, repo defaults
    otherwise. Per-step time = (T(3N) - T(N)) / 2N, which cancels the
    initialisation and the two field outputs every run writes. None when the
    program was not built."""
    exe = os.path.join(REPO, "oracle", "_ref", "nicr3d_emulated")
    if not os.path.exists(exe) or n_eta < 1:
        return None
    import shutil
    import subprocess
    import tempfile
    shape = (64, 128, 128)
    cells = shape[0] * shape[1] * shape[2]
    # grain diameter of the reference's count rule round(6 V / (pi d^3)) = n_eta
    d = (6.0 * cells / (np.pi * n_eta)) ** (1.0 / 3.0)
    try:
        ranks = min(16, len(os.sched_getaffinity(0)))
    except AttributeError:
        ranks = min(16, os.cpu_count() or 1)
    ranks = max(1, min(ranks, shape[0] // 2))
    n = 10
    times = {}
    for steps in (n, 3 * n):
        out = tempfile.mkdtemp(prefix="nicr_ref_")
        os.rmdir(out)   # the program wants a new or empty directory
        cmd = [exe, str(ranks), "--nx", str(shape[0]), "--ny", str(shape[1]), "--nz", str(shape[2]),
               "--grain-evolution", "on", "--grain-diameter-grid", f"{d:.4f}", "--bounds", "legacy",
               "--steps", str(steps), "--out", out]
        t0 = time.perf_counter()
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=600)
        times[steps] = time.perf_counter() - t0
        shutil.rmtree(out, ignore_errors=True)
        if r.returncode != 0:
            raise RuntimeError(f"reference CPU run failed ({r.returncode}): {r.stdout[-500:]}")
    dt = (times[3 * n] - times[n]) / (2 * n)
    print(f"[bench] reference CPU: {ranks} ranks, {times[n]:.1f} s / {times[3 * n]:.1f} s for {n} / {3 * n} steps",
          file=sys.stderr, flush=True)
    return {"value": cells / dt / 1e9, "unit": "Gcell-steps/s", "cores": ranks, "kind": "reference",
            "sample": f"mpi3d-slab (oracle/_ref/nicr3d_emulated, thread-emulated MPI, {ranks} ranks on {ranks} "
                      f"host cores), {shape[0]}x{shape[1]}x{shape[2]} cells, {n_eta} evolving grains "
                      f"(--grain-diameter-grid {d:.4f}), bounds legacy, FP64; (T({3 * n}) - T({n})) / {2 * n} = "
                      f"{dt * 1e3:.1f} ms per step"}


C1_STEPS = 1000
C1_SHAPE = (128, 128)


def c1_config(nicr):
    """Config C1: the 2D program (mpi2d, run loop 2D:1216-1236 with
    nucleation=center) on 128 x 128 cells, grains off, Cr 0.20 +- U(0.005) from
    std::mt19937_64(20260909), O = 0 in the bulk, the concentration inlet, the
    centre seed of radius 5 cells, Ck legacy, bounds legacy, default dt."""
    return nicr.make_config(2, nx=C1_SHAPE[0], ny=C1_SHAPE[1], grains=False, bc="concentration", ck="legacy",
                            bounds="legacy", nucleation="center", radius_grid=5.0)


def c1_gpu(nicr):
    """C1's 1000 binary64 steps on the GPU from the reference's initial state
    (host noise, centre seed on the device). A first model of the same config
    takes 10 untimed steps so that first-launch costs stay out of the timing.
    Returns the timing, the initial o, cr, p (after the seed) and the final ones."""
    syn = importlib.import_module(PKG_NAME + ".synthetic")
    cfg = c1_config(nicr)
    o, cr, p = syn.noisy_alloy(C1_SHAPE, 0.20, 0.005, seed=20260909)
    models = []
    for _ in range(2):
        m = nicr.Model(cfg)
        for name, a in (("o", o), ("cr", cr), ("p", p)):
            m.set(name, a)
        seeded = m.initialize_center_oxide()
        models.append(m)
    models[0].step(10)
    models[0].sync()
    m = models[1]
    initial = {f: m.get(f) for f in ("o", "cr", "p")}
    m.sync()
    t0 = time.perf_counter()
    r = m.try_step(C1_STEPS)
    m.sync()
    t1 = time.perf_counter()
    if r.status != 0 or r.steps_done != C1_STEPS:
        raise RuntimeError(f"C1: step rejected (status {r.status})")
    fields = {f: m.get(f) for f in ("o", "cr", "p")}
    for mm in models:
        mm.close()
    cells = C1_SHAPE[0] * C1_SHAPE[1]
    return {"seconds": round(t1 - t0, 4), "ms_per_step": round((t1 - t0) / C1_STEPS * 1e3, 4),
            "value": round(cells * C1_STEPS / (t1 - t0) / 1e9, 4), "unit": "Gcell-steps/s",
            "centre_seed_cells": seeded}, initial, fields


def cpu_baseline_c1(nicr, initial, gpu_fields):
    """The reference's own 2D Model (oracle/_ref/libref2d.so, compiled from
    mpi2d/NiCr2D_grains.cpp) on config C1 on one host thread, as mpirun -np 1
    runs it: constructed with C1's config (inlet layer, centre seed), its o, cr,
    p replaced by the GPU run's initial state (its constructor has no Cr noise:
    bulk Cr 0.1986), then 1000 Model::step calls are timed. Also whether its
    final o, cr, p equal the GPU's bit for bit. None when it was not built."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import binding
    if not binding.ref_available(2):
        return None
    ref = binding.RefModel(c1_config(nicr))
    for f, a in initial.items():
        ref.set(f, a)
    t0 = time.perf_counter()
    done = ref.run(C1_STEPS)
    t1 = time.perf_counter()
    if done != C1_STEPS:
        raise RuntimeError(f"C1 reference run stopped after {done} steps")
    equal = all(np.array_equal(ref.get(f), gpu_fields[f]) for f in ("o", "cr", "p"))
    cells = C1_SHAPE[0] * C1_SHAPE[1]
    return {"value": round(cells * C1_STEPS / (t1 - t0) / 1e9, 6), "unit": "Gcell-steps/s", "cores": 1,
            "kind": "reference", "seconds": round(t1 - t0, 3), "ms_per_step": round((t1 - t0) / C1_STEPS * 1e3, 4),
            "sample": f"mpi2d Model (oracle/_ref/libref2d.so) on one host thread: the whole C1 run, "
                      f"{C1_SHAPE[0]}x{C1_SHAPE[1]} cells, {C1_STEPS} steps",
            "fields_bitwise_equal_gpu": bool(equal)}


def cpu_baseline(nicr, n_eta, seed):
    """The CPU oracle (one thread) on a 64 x 256 x 256 slab of the same workload, 12 steps."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import binding
    shape = (64, 256, 256)
    state = build_state(nicr, shape, 16, n_eta, seed)
    o, cr, p, eta = state
    cfg = nicr.make_config(3, nx=shape[0], ny=shape[1], nz=shape[2], grain_evolution=True, n_eta=n_eta,
                           bc="concentration", bounds="legacy", ck="legacy")
    ora = binding.OracleModel(cfg)
    ora.set("o", o)
    ora.set("cr", cr)
    ora.set("p", p)
    for g in range(n_eta):
        ora.set(g, eta[g])
    del state, eta
    nsteps = 12
    t0 = time.perf_counter()
    done = ora.run(nsteps)
    t1 = time.perf_counter()
    cells = shape[0] * shape[1] * shape[2]
    return {"value": cells * done / (t1 - t0) / 1e9, "unit": "Gcell-steps/s", "cores": 1, "kind": "port",
            "sample": f"oracle/liboracle.so, {shape[0]}x{shape[1]}x{shape[2]} cells, {n_eta} eta, "
                      f"{done} FP64 steps, {t1 - t0:.1f} s on one host thread"}


def ensemble_bench(nicr, n_inst, steps, seed):
    """Ensemble of n_inst 2D instances, sides from {64, 128, 256, 512}, Cr in
    U(0.15, 0.30), inlet O supply U(0.5, 1.5) x 0.0023: autotuned graph replay vs
    the same instances stepped one after another as separate contexts."""
    insts, states = _wl().ensemble_instances(n_inst, seed)
    sides = [nx for nx, _, _ in insts]
    cfg = _wl().ensemble_config()
    ens = nicr.Ensemble(cfg, insts)
    for i, st in enumerate(states):
        for name, a in zip(("o", "cr", "p"), st):
            ens.set(i, name, a)
    cells = ens.cells
    tx, grid, tune_ms = ens.autotune(5)
    ens.step(10)
    t0 = time.perf_counter()
    reps = ens.step(steps)
    t1 = time.perf_counter()
    accepted = sum(r.steps_done for r in reps)
    ens.set_tiling(tx, grid, graph=False)
    t2 = time.perf_counter()
    ens.step(steps)
    t3 = time.perf_counter()
    ens.close()
    # the same instances as separate single-instance contexts, one after another
    models = []
    for (nx, ny, lo), st in zip(insts, states):
        c1 = nicr.make_config(2, nx=nx, ny=ny, grains=False, bc="concentration", bounds="legacy", ck="legacy",
                              nucleation="off", layer_o=lo)
        m = nicr.Model(c1)
        for name, a in zip(("o", "cr", "p"), st):
            m.set(name, a)
        m.step(2)
        models.append(m)
    t4 = time.perf_counter()
    for m in models:
        m.step(steps)
    t5 = time.perf_counter()
    for m in models:
        m.close()
    return {"workload": f"{n_inst} 2D instances, sides {sorted(set(int(s) for s in sides))}, Cr U(0.15,0.30), "
                        "inlet O U(0.5,1.5)x0.0023, centre seed r=5, bounds legacy, binary64",
            "cells": cells, "steps": steps, "all_accepted": accepted == steps * n_inst,
            "autotuned": {"tx": tx, "grid": grid, "ms_per_step_dry": round(tune_ms, 4)},
            "graph": {"value": round(cells * steps / (t1 - t0) / 1e9, 4), "ms_per_step": round((t1 - t0) / steps * 1e3, 4)},
            "eager": {"value": round(cells * steps / (t3 - t2) / 1e9, 4), "ms_per_step": round((t3 - t2) / steps * 1e3, 4)},
            "sequential_contexts": {"value": round(cells * steps / (t5 - t4) / 1e9, 4),
                                    "ms_per_step": round((t5 - t4) / steps * 1e3, 4)},
            "unit": "Gcell-steps/s",
            "speedup_vs_sequential": round((t5 - t4) / (t1 - t0), 3)}


def unet_flops_per_cell():
    """Multiply-add flops of one U-Net forward pass per cell of the volume (the
    3x3x3 convolutions; (cin, cout) per level from include/nicr_nn.h)."""
    levels = {0: [(32, 32), (32, 32), (96, 32), (32, 32)], 1: [(32, 64), (64, 64), (192, 64), (64, 64)],
              2: [(64, 128), (128, 128), (384, 128), (128, 128)], 3: [(128, 256), (256, 256)]}
    return sum(2 * 27 * ci * co / 8 ** lv for lv, convs in levels.items() for ci, co in convs)


def surrogate_bench(nicr, n_train, n_run, n_traj, iters, nsteps, tol, seed, settle, noise, runs, fallback,
                    net_fields=("o", "cr", "phi"), min_jump=0):
    """Learned surrogate on the inlet workload: a training set from n_traj plain
    binary64 trajectories of the static-grain inlet generator at n_train^3
    (reference Voronoi grains of --grain-diameter-grid 63, Cr 0.20 +- 0.005,
    O inlet on x = 0, legacy Ck, Ksp scans on the reference schedule), Adam on
    the GPU; then at n_run^3, for each workload in `runs`, nsteps of the
    reference loop with its Ksp schedule by the plain binary64 step and with
    50-step network jumps, both timed; the surrogate's final fields, totals,
    budgets and budget closure compared with the plain run's.
    runs: "evolving" (32 evolving eta from 64 Voronoi seeds, the headline mesh's
    state) and/or "static" (the generator's static grains at n_run^3).
    settle: extra plain steps between a jump's check step and the next proposal;
    fallback: plain steps after a rejected jump;
    noise: largest relative perturbation of the recorded samples;
    net_fields: the fields whose jump takes the network's correction (the
    others jump by the linear extrapolation alone: output gain 0);
    min_jump: shortest last jump before a scan (0: plain steps there)."""
    sur = importlib.import_module(PKG_NAME + ".surrogate")
    wl = _wl()

    def log(msg):
        print(f"[bench] surrogate: {msg}", file=sys.stderr, flush=True)

    J = 50
    # a proposal may undershoot O slightly ahead of the inlet front, where the
    # plain step itself clips (legacy bounds); the check step clips those
    # values and records them, so only values far outside reject a jump
    bounds = (-1e-2, 1.0)
    net = sur.UNet((n_train,) * 3, max_batch=2, train=True, max_samples=n_traj, seed=seed)
    rec = sur.record_inlet_trajectories(net, nicr, n_traj, jump=J, seed=seed, log=log, noise=noise)
    net.fit_norm()
    net.set_augment(True)   # samples through random Y/Z shifts and mirrors (symmetries of the step)
    tr = sur.train_schedule(net, iters, batch=2, lr=1e-3, seed=seed, log=log)
    params, norm = net.params(), net.norm()
    net.close()
    shape = (n_run,) * 3
    cells = n_run ** 3
    inet = sur.UNet(shape, seed=1)
    inet.set_params(params)
    for k, f in enumerate(("o", "cr", "phi")):
        if f not in net_fields:
            norm.gain[k] = 0.0
    inet.set_norm(norm)
    inet.set_timing(True)
    znet = sur.UNet(shape, seed=1)   # an untrained network's head is zero
    znet.set_norm(norm)
    out = {"training": {"generator": f"{n_train}^3 static Voronoi grains (--grain-diameter-grid 63), Cr 0.20+-0.005, "
                                     "O inlet on x=0 (concentration BC), Ck legacy, bounds legacy, Ksp scans",
                        "volume": f"{n_train}^3", "trajectories": rec["trajectories"], "noise": noise,
                        "augment": "random Y/Z shifts and mirrors",
                        "plain_steps": rec["plain_steps"], "record_s": round(rec["seconds"], 1),
                        "iters": tr["iters"], "batch": tr["batch"], "train_s": round(tr["seconds"], 1),
                        "loss_first": round(tr["loss_first"], 4), "loss_last": round(tr["loss_last"], 4)},
           "jump": J, "residual_tol": tol, "settle_steps": settle, "fallback_steps": fallback,
           "network_fields": list(net_fields), "min_jump": min_jump,
           "bounds": list(bounds), "runs": {}}
    for kind in runs:
        t0 = time.perf_counter()
        if kind == "evolving":
            state = wl.evolving_grain_state(n_run)
            make = lambda: wl.evolving_grain_model("fp64", state)
            n_eta = len(state[3])
            desc = f"3D {n_run}^3, {n_eta} evolving eta from 64 Voronoi seeds"
        else:
            gb = wl.static_grain_gb(shape, 63.0, wl.SEED)
            fields = importlib.import_module(PKG_NAME + ".synthetic").noisy_alloy(shape, 0.20, 0.005, seed=wl.SEED)
            make = lambda: wl.inlet_model(nicr, "fp64", fields, gb=gb)
            n_eta = 0
            desc = f"3D {n_run}^3, static Voronoi grains (--grain-diameter-grid 63)"
        plain, sm = make(), make()
        log(f"{kind}: {n_run}^3 models built in {time.perf_counter() - t0:.1f} s")
        t_start, b_start = wl.totals(sm), sm.b
        plain.sync()
        t0 = time.perf_counter()
        rp = plain.run_loop(nsteps)
        plain.sync()
        t_plain = time.perf_counter() - t0
        log(f"{kind}: plain binary64: {nsteps} steps in {t_plain:.2f} s, {rp['scans']} scans")
        sm.sync()
        t0 = time.perf_counter()
        rs = inet.run_loop(sm, nsteps, jump=J, residual_tol=tol, settle_steps=settle, bounds=bounds,
                           fallback_steps=fallback, min_jump=min_jump)
        sm.sync()
        t_sur = time.perf_counter() - t0
        log(f"{kind}: surrogate: {nsteps} steps in {t_sur:.2f} s: {rs['jumps_accepted']} jumps accepted, "
            f"{rs['jumps_rejected']} rejected, {rs['plain_steps']} plain steps, {rs['scans']} scans")
        if rp["steps"] != nsteps or rs["steps"] != nsteps or rs["status"] != 0:
            raise RuntimeError(f"surrogate workload {kind}: a plain step was rejected")
        # the same checked jumps without the network: a zero head proposes x0 + J u0
        zm = make()
        zm.sync()
        t0 = time.perf_counter()
        rz = znet.run_loop(zm, nsteps, jump=J, residual_tol=tol, settle_steps=settle, bounds=bounds,
                           fallback_steps=fallback, min_jump=min_jump)
        zm.sync()
        t_zero = time.perf_counter() - t0
        log(f"{kind}: extrapolation only: {nsteps} steps in {t_zero:.2f} s: {rz['jumps_accepted']} jumps accepted, "
            f"{rz['jumps_rejected']} rejected")
        err_z = wl.field_errors(zm, plain)
        zero_line = {"speedup": round(t_plain / t_zero, 3), "ms": round(t_zero * 1e3, 1),
                     "jumps_accepted": rz["jumps_accepted"], "jumps_rejected": rz["jumps_rejected"],
                     "max_residual": [float(f"{v:.3e}") for v in rz["max_residual"][:3 if n_eta else 2]],
                     "rel_l2_vs_plain": {k: float(f"{v['rel_l2']:.3e}") for k, v in err_z.items()}}
        zm.close()
        err = wl.field_errors(sm, plain)
        tsur, tpl = wl.totals(sm), wl.totals(plain)
        bs, bp = sm.b, plain.b
        srcs = {sp: {k: (bs[k + sp] - b_start[k + sp], bp[k + sp] - b_start[k + sp]) for k in ("bc", "ck", "clip", "nuc")}
                for sp in ("O", "Cr")}
        dt = plain.c.dt
        fwd_ms = inet.forward_ms()
        flops = unet_flops_per_cell() * cells
        out["runs"][kind] = {
            "workload": desc + ", Cr 0.20+-0.005, O inlet on x=0, Ck legacy, bounds legacy, Ksp scans on the "
                               f"reference schedule, binary64 steps; {nsteps} steps from the initial state",
            "speedup": round(t_plain / t_sur, 3),
            "sim_time_per_s_plain_fp64": nsteps * dt / t_plain,
            "sim_time_per_s_surrogate": nsteps * dt / t_sur,
            "ms_plain": round(t_plain * 1e3, 1), "ms_surrogate": round(t_sur * 1e3, 1),
            "scans": rs["scans"], "inserted_cells": [rs["inserted"], rp["inserted"]],
            "jumps_accepted": rs["jumps_accepted"], "jumps_rejected": rs["jumps_rejected"],
            "plain_steps": rs["plain_steps"],
            "max_residual": [float(f"{v:.3e}") for v in rs["max_residual"][:3 if n_eta else 2]],
            "min_rejected_residual": rs["min_rejected_residual"],
            "rel_l2_vs_plain": {k: float(f"{v['rel_l2']:.3e}") for k, v in err.items()},
            "rel_linf_vs_plain": {k: float(f"{v['rel_linf']:.3e}") for k, v in err.items()},
            "totals_rel_vs_plain": {f: float(f"{abs(tsur[f] - tpl[f]) / abs(tpl[f]):.3e}") for f in tsur},
            "closure_surrogate": {k: float(f"{v:.3e}") for k, v in wl.closure(sm, t_start).items()},
            "closure_plain": {k: float(f"{v:.3e}") for k, v in wl.closure(plain, t_start).items()},
            "sources_surrogate_vs_plain": {sp: {k: [float(f"{a:.6e}"), float(f"{b:.6e}")] for k, (a, b) in d.items()}
                                           for sp, d in srcs.items()},
            "extrapolation_only": zero_line,
            "forward_ms": round(fwd_ms, 3),
            "forward_roofline": {"bound": "mfma", "achieved": round(flops / (fwd_ms * 1e-3) / 1e12, 1),
                                 "peak": 2500.0, "unit": "TFLOP/s",
                                 "frac": round(flops / (fwd_ms * 1e-3) / 1e12 / 2500.0, 4), "flops_per_forward": flops},
        }
        plain.close()
        sm.close()
    inet.close()
    znet.close()
    first = out["runs"][runs[0]]
    out["speedup"] = first["speedup"]
    out["workload"] = runs[0]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=256, help="mesh side")
    ap.add_argument("--eta", type=int, default=32)
    ap.add_argument("--seeds", type=int, default=64)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--run-steps", type=int, default=5000,
                    help="steps of the whole-configuration run with the Ksp schedule, mixed and FP64 (0: skip)")
    ap.add_argument("--no-static", action="store_true", help="skip the 128^3 static-grain error run")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the 2D ensemble measurement")
    ap.add_argument("--ensemble-only", action="store_true", help="only the 2D ensemble measurement")
    ap.add_argument("--no-surrogate", action="store_true", help="skip the learned-surrogate measurement")
    ap.add_argument("--surrogate-only", action="store_true", help="only the learned-surrogate measurement")
    ap.add_argument("--sur-traj", type=int, default=200, help="training trajectories of the surrogate")
    ap.add_argument("--sur-iters", type=int, default=3000, help="Adam iterations of the surrogate")
    ap.add_argument("--sur-train-n", type=int, default=128, help="side of the training volumes")
    ap.add_argument("--sur-run-n", type=int, default=256, help="side of the surrogate run")
    ap.add_argument("--sur-steps", type=int, default=5000, help="steps of the timed run (5000: the 256^3 run length)")
    ap.add_argument("--sur-settle", type=int, default=2,
                    help="extra plain steps between a jump's check step and the next proposal")
    ap.add_argument("--sur-fallback", type=int, default=50, help="plain steps after a rejected jump")
    ap.add_argument("--sur-noise", type=float, default=0.2,
                    help="largest perturbation of the recorded training states' Cr (x jump x RMS update)")
    ap.add_argument("--sur-tol", type=float, default=0.15,
                    help="trapezoid residual tolerance of a jump (final fields stay within 1e-3 relative L2)")
    ap.add_argument("--sur-min-jump", type=int, default=12,
                    help="shortest last jump before a Ksp scan (0: the steps left there are plain steps)")
    ap.add_argument("--sur-net-fields", default="cr,phi",
                    help="fields whose jump takes the network's correction (others: linear extrapolation)")
    ap.add_argument("--sur-runs", default="evolving,static",
                    help="surrogate workloads at --sur-run-n: evolving (32 eta) and/or static grains, comma-separated")
    ap.add_argument("--copy-probe-only", action="store_true",
                    help="run only the 1 GiB device copy reference (rocprof counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {args.gpus} is not supported "
              "(several GPUs would each hold an X slab of the mesh, which is not built).", file=sys.stderr)
        sys.exit(2)
    nicr = _pkg()
    if args.copy_probe_only:
        cfg = nicr.make_config(3, nx=8, ny=8, nz=8)
        m = nicr.Model(cfg)
        ms = m.copy_probe(1 << 30, 20)
        print(json.dumps({"copy_probe_bytes": 1 << 30, "ms": ms, "GBs_read_plus_write": 2 * (1 << 30) / ms / 1e6}))
        return
    def run_surrogate():
        return surrogate_bench(nicr, args.sur_train_n, args.sur_run_n, args.sur_traj, args.sur_iters,
                               args.sur_steps, args.sur_tol, 20260909, args.sur_settle, args.sur_noise,
                               tuple(args.sur_runs.split(",")), args.sur_fallback,
                               tuple(args.sur_net_fields.split(",")), args.sur_min_jump)

    if args.surrogate_only:
        print(json.dumps({"surrogate": run_surrogate()}), flush=True)
        return
    if args.ensemble_only:
        print(json.dumps({"ensemble": ensemble_bench(nicr, 64, 200, 20260909)}), flush=True)
        return
    shape = (args.n, args.n, args.n)
    cells = shape[0] * shape[1] * shape[2]
    state = build_state(nicr, shape, args.seeds, args.eta, 20260909)

    m = make_model(nicr, shape, args.eta, "mixed", state)
    print(f"[bench] mixed model resident: {m.device_bytes() / 2**30:.2f} GiB", file=sys.stderr, flush=True)
    copy_ms = m.copy_probe(1 << 30, 10)
    copy_gbs = 2 * (1 << 30) / copy_ms / 1e6
    wall, kt, wall_marked = time_steps(m, args.steps, args.warmup, "mixed")
    m.close()
    ms_per_step = wall / args.steps * 1e3
    value = cells * args.steps / wall / 1e9

    fp64 = None
    if not args.no_fp64:
        m64 = make_model(nicr, shape, args.eta, "fp64", state)
        wall64, kt64, _ = time_steps(m64, max(10, args.steps // 4), 2, "fp64")
        n64 = max(10, args.steps // 4)
        m64.close()
        ms64 = wall64 / n64 * 1e3
        alg64 = (3 + args.eta) * 8 * 2 * cells   # o, cr, phi and eta read and written once, binary64
        fp64 = {"value": cells * n64 / wall64 / 1e9, "ms_per_step": ms64,
                "kernel_ms": {k: v[0] for k, v in kt64.items()},
                "algorithmic_bytes_per_step": alg64, "algorithmic_GBs": round(alg64 / ms64 / 1e6, 1),
                "hbm_fraction": round(alg64 / ms64 / 1e6 / HBM_PEAK_GBS, 4)}
        path64 = os.path.join(REPO, "profiles", "traffic_fp64_step_256cube_32eta.json")
        if shape == (256, 256, 256) and args.eta == 32 and os.path.exists(path64):
            with open(path64) as f:
                t64 = json.load(f)
            fp64["dram"] = {"bytes_per_step": t64["hbm_bytes_per_step"],
                            "GBs": round(t64["hbm_bytes_per_step"] / ms64 / 1e6, 1),
                            "source": os.path.relpath(path64, REPO)}
    # The whole configuration: args.run_steps steps of the reference's run loop
    # with its Ksp schedule (scans before steps 500, 1000, ...), mixed and
    # binary64 from the same state, timed end to end (scans and budget
    # reductions included), and the mixed run's error against the binary64 run.
    config_run = mixed_error = None
    if args.run_steps > 0:
        def log(msg):
            print(f"[bench] config run: {msg}", file=sys.stderr, flush=True)
        res = _wl().mixed_vs_fp64(lambda prec: make_model(nicr, shape, args.eta, prec, state), args.run_steps, log)
        step_b = cells * (48 + 8 * args.eta)
        config_run = {"steps": args.run_steps, "scans": res["mixed"]["scans"],
                      "inserted_cells": res["mixed"]["inserted_cells"]}
        for prec in ("mixed", "fp64"):
            sec = res[prec]["seconds"]
            config_run[prec] = {"seconds": sec, "value": round(cells * args.run_steps / sec / 1e9, 4),
                                "ms_per_step": round(sec / args.run_steps * 1e3, 4)}
        config_run["mixed"]["hbm_fraction"] = round(step_b * args.run_steps / res["mixed"]["seconds"] / 1e9
                                                    / HBM_PEAK_GBS, 4)
        mixed_error = {f"evolving_{args.n}cube_{args.eta}eta_{args.run_steps}steps":
                       {k: res[k] for k in ("error", "totals_rel_vs_fp64")} |
                       {"closure_mixed": res["mixed"]["closure"], "closure_fp64": res["fp64"]["closure"]}}
        if not args.no_static:
            rs = _wl().mixed_vs_fp64(lambda prec: _wl().static_grain_model(prec), 2000, log)
            mixed_error["static_128cube_16grains_2000steps"] = (
                {k: rs[k] for k in ("error", "totals_rel_vs_fp64")} |
                {"closure_mixed": rs["mixed"]["closure"], "closure_fp64": rs["fp64"]["closure"],
                 "seconds": {"mixed": rs["mixed"]["seconds"], "fp64": rs["fp64"]["seconds"]},
                 "value": {p: round(128 ** 3 * 2000 / rs[p]["seconds"] / 1e9, 4) for p in ("mixed", "fp64")}})
    del state

    # Roofline of the dominant kernel, k_eta_rows (one launch per pass of 16
    # eta fields). Algorithmic bytes per launch: 8 B per eta field per cell
    # (binary32 read once + written once) x the fields of one pass x the cells;
    # duration: mean per launch from HIP events on the context stream over the
    # second window of the same steps.
    eta_kernel = "k_eta_rows" if "k_eta_rows" in kt else "k_eta_march"
    eta_ms, eta_count = kt.get(eta_kernel, (float("nan"), 0))
    launches_per_step = eta_count / args.steps if eta_count else float("nan")
    eta_bytes = cells * 8 * args.eta / launches_per_step
    achieved = eta_bytes / (eta_ms * 1e-3) / 1e9
    pde_ms = kt.get("k_pde_march", (float("nan"), 0))[0]
    pde_bytes = cells * 52   # o, cr, phi binary64 read + write, gb binary32 read
    step_bytes = cells * (48 + 8 * args.eta)   # FP64 o/cr/phi + FP32 eta, read + write
    traffic = measured_traffic(eta_kernel)
    out = {
        "metric": "Gcell-steps/s @256^3 mixed vs FP64 GPU, % HBM peak; surrogate speedup @1e-3 L2",
        "value": round(value, 4),
        "unit": "Gcell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64 (mixed); eta fp32",
        "data": "synthetic",
        "config": {"workload": f"3D {shape[0]}^3, {args.eta} evolving eta from {args.seeds} Voronoi seeds, "
                               "Cr 0.20+-0.005, O inlet layer, Ksp on, bounds legacy, mixed precision",
                   "cells": cells, "n_eta": args.eta},
        "step_hbm_fraction": round(step_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "kernel_ms": {k: round(v[0], 5) for k, v in kt.items()},
        "kernel_launches_per_step": {k: round(v[1] / args.steps, 3) for k, v in kt.items()},
        "kernel_timing": f"HIP events around every launch on the context stream over a second window of "
                         f"{args.steps} steps ({wall_marked / args.steps * 1e3:.4f} ms per step with the markers); "
                         "value and ms_per_step come from the window without markers",
        "roofline": {"bound": "hbm", "kernel": eta_kernel, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "algorithmic_bytes_per_launch": int(eta_bytes), "ms_per_launch": round(eta_ms, 5),
                     "traffic": traffic["bytes_per_launch"] if traffic else None,
                     "traffic_detail": traffic},
        "roofline_pde": {"bound": "hbm", "kernel": "k_pde_march",
                         "achieved": round(pde_bytes / (pde_ms * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(pde_bytes / (pde_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                         "algorithmic_bytes_per_launch": int(pde_bytes)},
        "fp64": fp64,
        "config_run": config_run,
        "mixed_error": mixed_error,
        "copy_probe_GBs": round(copy_gbs, 1),
    }
    c1, c1_initial, c1_fields = c1_gpu(nicr)
    out["c1"] = {"workload": f"2D {C1_SHAPE[0]}x{C1_SHAPE[1]} single crystal, centre seed r=5 (80 cells), Cr "
                             "0.20+-0.005 (mt19937_64), O inlet on x=0 (concentration BC), Ck legacy, bounds legacy, "
                             f"{C1_STEPS} binary64 steps", "gpu": c1}
    if not args.no_ensemble:
        out["ensemble"] = ensemble_bench(nicr, 64, 200, 20260909)
    if not args.no_surrogate:
        out["surrogate"] = run_surrogate()
    if not args.no_cpu:
        ref = cpu_baseline_reference(args.eta)
        out["cpu_baseline"] = ref if ref is not None else cpu_baseline(nicr, args.eta, 20260909)
        c1_ref = cpu_baseline_c1(nicr, c1_initial, c1_fields)
        if c1_ref is not None:
            out["c1"]["cpu_reference"] = c1_ref
            out["c1"]["speedup_vs_reference"] = round(c1["value"] / c1_ref["value"], 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
