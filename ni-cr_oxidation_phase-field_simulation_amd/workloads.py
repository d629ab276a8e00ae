"""The benchmark workloads as device models, and their comparison measures.

- 3D static grains (128^3): the reference's own Voronoi construction with
  --grain-diameter-grid 63 (16 grains by mpi3d-slab/NiCr3D_grains.cpp:418-421),
  the GB mask of the fixed-grain mode, Cr 0.20 +- U(0.005) from
  std::mt19937_64 in the reference's loop order, O = 0 with the concentration-BC
  inlet layer, Ksp insertion, Ck legacy, bounds legacy; 2000 steps.
- 3D evolving grains (256^3): 64 Voronoi seeds coloured onto 32 evolving eta
  fields (neighbouring grains differ), the same composition; 5000 steps.

Both run the reference's loop with its Ksp schedule (Model.run_loop). The
error measures compare a mixed-precision run with the binary64 run of the same
state, which is itself bitwise equal to the CPU oracle.
"""
from __future__ import annotations

import os
import time

import numpy as np

from . import make_config, Model
from . import synthetic as syn

SEED = 20260909


def static_grain_gb(shape, diameter: float = 63.0, seed: int = SEED) -> np.ndarray:
    """The GB mask of the reference's static Voronoi grains (--grain-diameter-grid
    `diameter`, --random-seed `seed`) on a 3D mesh `shape` (host array)."""
    from .host import HostState
    nx, ny, nz = shape
    cfg = make_config(3, nx=nx, ny=ny, nz=nz, grain_diameter_grid=diameter, random_seed=seed, grains=True,
                      grain_evolution=False, bc="concentration", ck="legacy", bounds="legacy", nucleation="ksp")
    old = os.environ.get("NICR_HOST_INIT")
    os.environ["NICR_HOST_INIT"] = "device"   # the grain distances on the GPU (bitwise the host loops)
    try:
        hs = HostState(cfg, "concentration", "off")
    finally:
        if old is None:
            os.environ.pop("NICR_HOST_INIT", None)
        else:
            os.environ["NICR_HOST_INIT"] = old
    gb = hs.get("gb")
    hs.close()
    return gb


def inlet_model(nicr_pkg, precision: str, fields, gb=None, eta=None) -> Model:
    """A 3D model of the inlet workloads holding fields = (o, cr, p): the
    concentration-BC O inlet on x = 0, legacy Ck, legacy bounds, Ksp scans;
    static grains with mask gb, or len(eta) evolving eta fields."""
    o, cr, p = fields
    nx, ny, nz = o.shape
    if eta is not None:
        cfg = make_config(3, nx=nx, ny=ny, nz=nz, grain_evolution=True, n_eta=len(eta), bc="concentration",
                          bounds="legacy", ck="legacy", nucleation="ksp", precision=precision)
    else:
        cfg = make_config(3, nx=nx, ny=ny, nz=nz, grains=True, grain_evolution=False, bc="concentration",
                          ck="legacy", bounds="legacy", nucleation="ksp", precision=precision)
    m = Model(cfg)
    m.set("o", o)
    m.set("cr", cr)
    m.set("p", p)
    if eta is not None:
        for g, e in enumerate(eta):
            m.set(g, e)
    else:
        m.set("gb", gb)
    return m


def static_grain_model(precision: str, n: int = 128, diameter: float = 63.0, seed: int = SEED) -> Model:
    """The 3D static-grain workload on the device."""
    gb = static_grain_gb((n, n, n), diameter, seed)
    return inlet_model(None, precision, syn.noisy_alloy((n, n, n), 0.20, 0.005, seed=seed), gb=gb)


def evolving_grain_state(n=256, n_seeds: int = 64, n_eta: int = 32, seed: int = SEED):
    """(o, cr, p, eta) of the 3D evolving-grain workload (host arrays); n is the
    cube side or a shape (nx, ny, nz)."""
    shape = (n, n, n) if isinstance(n, int) else tuple(n)
    labels, dist, _ = syn.voronoi_labels(shape, n_seeds, seed=seed)
    colour = syn.colour_seeds(labels, n_seeds, n_eta)
    o, cr, p = syn.noisy_alloy(shape, 0.20, 0.005, seed=seed)
    eta = syn.eta_from_labels(labels, dist, colour, n_eta)
    return o, cr, p, eta


def evolving_grain_model(precision: str, state, n_eta: int = 32) -> Model:
    o, cr, p, eta = state
    return inlet_model(None, precision, (o, cr, p), eta=[eta[g] for g in range(n_eta)])


def totals(m: Model) -> dict:
    """Device sums of O and Cr (double-double, as Model::sum)."""
    return {f: sum(m.stats(f)[0]) for f in ("o", "cr")}


def closure(m: Model, t0: dict) -> dict:
    """The residual of the reference's diagnostics (2D:697-710): the change of
    each total minus the recorded sources, relative to the initial total."""
    b = m.b
    t = totals(m)
    out = {}
    for f, k in (("o", "O"), ("cr", "Cr")):
        res = (t[f] - t0[f]) - (b["bc" + k] + b["nuc" + k] + b["ck" + k] + b["clip" + k])
        out[f] = abs(res) / abs(t0["cr"])   # O starts at the inlet layer only: scale by the Cr inventory
    return out


def field_errors(a: Model, b: Model) -> dict:
    """Per field, a against the reference run b: rel_linf = max|a - b| / max|b|,
    rel_l2 = ||a - b|| / ||b|| (all eta fields together as one)."""
    out = {}
    for f in ("o", "cr", "p"):
        x, y = a.get(f), b.get(f)
        d = np.abs(x - y)
        out[f] = {"rel_linf": float(d.max() / max(np.abs(y).max(), 1e-300)),
                  "rel_l2": float(np.linalg.norm(x - y) / max(np.linalg.norm(y), 1e-300))}
    if b.n_eta:
        num = den = 0.0
        dmax = ymax = 0.0
        for g in range(b.n_eta):
            x, y = a.get(g), b.get(g)
            num += float(np.sum((x - y) ** 2))
            den += float(np.sum(y * y))
            dmax = max(dmax, float(np.abs(x - y).max()))
            ymax = max(ymax, float(np.abs(y).max()))
        out["eta"] = {"rel_linf": dmax / ymax, "rel_l2": (num / den) ** 0.5}
    return out


def mixed_vs_fp64(make, steps: int, log=None) -> dict:
    """Run the mixed and the binary64 model from the same state for `steps`
    steps of the reference loop (Ksp schedule included) and compare them."""
    out = {"steps": steps}
    runs = {}
    for prec in ("fp64", "mixed"):
        m = make(prec)
        t0 = totals(m)
        m.sync()
        w0 = time.perf_counter()
        r = m.run_loop(steps)
        m.sync()
        wall = time.perf_counter() - w0
        if r["steps"] != steps:
            raise RuntimeError(f"{prec} run rejected a step: {r['report'].status}")
        runs[prec] = m
        out[prec] = {"seconds": round(wall, 3), "scans": r["scans"], "inserted_cells": r["inserted"],
                     "nuclei": m.b["nuclei"], "clipped": m.b["clipped"],
                     "closure": {k: float(f"{v:.3e}") for k, v in closure(m, t0).items()}}
        if log:
            log(f"{prec}: {steps} steps in {wall:.2f} s, {r['scans']} scans, {r['inserted']} cells inserted")
    err = field_errors(runs["mixed"], runs["fp64"])
    out["error"] = {f: {k: float(f"{v:.3e}") for k, v in d.items()} for f, d in err.items()}
    tm, tf = totals(runs["mixed"]), totals(runs["fp64"])
    out["totals_rel_vs_fp64"] = {f: float(f"{abs(tm[f] - tf[f]) / abs(tf[f]):.3e}") for f in tm}
    for m in runs.values():
        m.close()
    return out


def ensemble_state(shape, cr0, layer_o, rng):
    """One 2D instance of the ensemble workload: O = 0 with the inlet row at
    layer_o (Cr 0.182), Cr = cr0 + U(-0.005, 0.005), the centre oxide disk of
    radius 5 cells (the reference's initialize_center_oxide geometry)."""
    nx, ny = shape
    o = np.zeros(shape)
    cr = cr0 + rng.uniform(-0.005, 0.005, size=shape)
    p = np.zeros(shape)
    o[0, :] = layer_o
    cr[0, :] = 0.182
    i, j = np.meshgrid(np.arange(1, nx + 1), np.arange(1, ny + 1), indexing="ij")
    disk = (i - (nx + 1) / 2.0) ** 2 + (j - (ny + 1) / 2.0) ** 2 <= 25.0
    o[disk], cr[disk], p[disk] = syn.OXIDE_O, syn.OXIDE_CR, 1.0
    return o, cr, p


def ensemble_instances(n_inst: int = 64, seed: int = SEED):
    """The 2D ensemble workload: n_inst instances with sides drawn from
    {64, 128, 256, 512}, Cr in U(0.15, 0.30), inlet O U(0.5, 1.5) x 0.0023.
    Returns ([(nx, ny, layer_o)], [(o, cr, p)])."""
    rng = np.random.default_rng(seed)
    sides = rng.choice([64, 128, 256, 512], size=n_inst)
    crs = rng.uniform(0.15, 0.30, size=n_inst)
    supply = rng.uniform(0.5, 1.5, size=n_inst) * syn.LAYER_O
    insts = [(int(s), int(s), float(lo)) for s, lo in zip(sides, supply)]
    states = [ensemble_state((s, s), c, lo, rng) for (s, _, lo), c in zip(insts, crs)]
    return insts, states


def ensemble_config(**kw):
    """The ensemble's shared 2D configuration (binary64, no grains)."""
    base = dict(grains=False, bc="concentration", bounds="legacy", ck="legacy", nucleation="off")
    base.update(kw)
    return make_config(2, nx=8, ny=8, **base)
