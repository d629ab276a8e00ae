"""Mixed precision against the binary64 step over the workloads' own horizons.

The binary64 GPU step is bitwise the CPU oracle (tests/test_gpu_parity.py,
tests/test_gpu_fullsize.py), so it stands in for the oracle at these sizes.
Both runs follow the reference's loop with its Ksp schedule (scans before
steps 500, 1000, ...; mpi3d-slab/NiCr3D_grains.cpp:1365-1372).

Bounds (relative to the binary64 run; about ten times the values measured on
MI355X, recorded in the comments):
- rel_linf = max|mixed - fp64| / max|fp64| and rel_l2 = ||mixed - fp64|| / ||fp64||
  per field (all eta fields together);
- closure: |sum(X) - sum(X0) - (inlet + insertion + Ck + clipping)| / sum(Cr0),
  the residual of the reference's diagnostics, for both runs;
- totals: |sum(X_mixed) - sum(X_fp64)| / sum(X_fp64).
"""
import importlib

import pytest

pytestmark = pytest.mark.gpu

WL = "ni-cr_oxidation_phase-field_simulation_amd.workloads"


def _check(res, bounds):
    for f, (linf, l2) in bounds["fields"].items():
        e = res["error"][f]
        assert e["rel_linf"] <= linf, f"{f}: rel_linf {e['rel_linf']:.3e} > {linf:.1e}"
        assert e["rel_l2"] <= l2, f"{f}: rel_l2 {e['rel_l2']:.3e} > {l2:.1e}"
    for prec in ("fp64", "mixed"):
        for f, v in res[prec]["closure"].items():
            assert v <= bounds["closure"], f"{prec} {f} closure {v:.3e}"
    for f, v in res["totals_rel_vs_fp64"].items():
        assert v <= bounds["totals"][f], f"total {f}: {v:.3e}"
    assert res["mixed"]["scans"] == res["fp64"]["scans"] == res["steps"] // 500


def test_static_grains_128cube_2000_steps(nicr):
    """128^3, 16 static Voronoi grains (the reference's construction with
    --grain-diameter-grid 63), Cr 0.20 +- 0.005, O inlet layer, Ksp, 2000 steps.
    Measured: O 4.4e-8 / 9.7e-9, Cr 1.3e-9 / 9.8e-11, phi 0 (no insertion in this
    run); closure 8e-17 (O), 1.3e-14 (Cr); totals O 6e-9, Cr 1.3e-14."""
    wl = importlib.import_module(WL)
    print("[mixed horizon] 128^3 static grains, 2000 steps", flush=True)
    res = wl.mixed_vs_fp64(lambda prec: wl.static_grain_model(prec), 2000)
    _check(res, {"fields": {"o": (5e-7, 1e-7), "cr": (2e-8, 1e-9), "p": (1e-10, 1e-10)},
                 "closure": 1e-12, "totals": {"o": 1e-7, "cr": 1e-12}})


def test_evolving_grains_256cube_5000_steps(nicr):
    """256^3, 32 evolving eta from 64 seeds, Cr 0.20 +- 0.005, O inlet layer, Ksp,
    5000 steps (10 scans). Measured: O 4.9e-5 / 4.5e-6, Cr 4.1e-6 / 2.6e-8,
    eta 1.5e-4 / 1.7e-5, phi 0; closure 3e-17 (O), 1.0e-14 (Cr); totals O 3.4e-8,
    Cr 1.0e-14."""
    wl = importlib.import_module(WL)
    print("[mixed horizon] 256^3 with 32 evolving eta, 5000 steps", flush=True)
    state = wl.evolving_grain_state()
    res = wl.mixed_vs_fp64(lambda prec: wl.evolving_grain_model(prec, state), 5000)
    _check(res, {"fields": {"o": (5e-4, 5e-5), "cr": (5e-5, 5e-7), "p": (1e-10, 1e-10), "eta": (2e-3, 2e-4)},
                 "closure": 1e-12, "totals": {"o": 5e-7, "cr": 1e-12}})
