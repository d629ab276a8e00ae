"""The initial 3D grain geometry with its distances from the GPU
(nicr_grain_distances) equals the host loops bit for bit: seeds, grain ids, GB
mask, every eta field and the inventories (mpi3d-slab/NiCr3D_grains.cpp:395-494).
The host loops are pinned to the reference by tests/test_host_init.py.
The 256^3 case with 64 evolving grains spans five device chunks of X planes."""
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [
    ["--nx", "31", "--ny", "17", "--nz", "15", "--grain-diameter-grid", "18", "--radius-grid", "3.5",
     "--seed", "2001", "--grain-evolution", "on", "--nucleation", "center", "--bc", "closed"],
    ["--nx", "23", "--ny", "13", "--nz", "11", "--grain-diameter-grid", "6", "--left-face-sites", "6",
     "--bc", "mu", "--nucleation", "off"],
    ["--nx", "64", "--ny", "48", "--nz", "40", "--grain-diameter-grid", "12", "--grain-evolution", "on",
     "--left-face-sites", "5", "--seed", "7"],
    ["--nx", "12", "--ny", "12", "--nz", "12", "--grain-diameter-grid", "40", "--grain-evolution", "on"],
    ["--nx", "256", "--ny", "256", "--nz", "256", "--grain-diameter-grid", "79.4", "--grain-evolution", "on",
     "--bounds", "legacy"],
]


def _state(host, args, mode, monkeypatch):
    monkeypatch.setenv("NICR_HOST_INIT", mode)
    cfg, opt = host.parse_args(3, args)
    return host.HostState(cfg, opt["bc"], opt["nucleation"])


@pytest.mark.parametrize("case", range(len(CASES)))
def test_device_grain_geometry_bitwise(nicr, case, monkeypatch):
    host = importlib.import_module("ni-cr_oxidation_phase-field_simulation_amd.host")
    args = CASES[case]
    print(f"case {case}: {' '.join(args)}", flush=True)
    dev = _state(host, args, "device", monkeypatch)
    ref = _state(host, args, "host", monkeypatch)
    assert dev.n_eta == ref.n_eta
    assert np.array_equal(dev.seeds(), ref.seeds())
    for f in ["o", "cr", "p", "gb", "grain"] + list(range(ref.n_eta)):
        assert np.array_equal(dev.get(f), ref.get(f)), f
    bd, io_d, ic_d = dev.budget()
    br, io_r, ic_r = ref.budget()
    assert io_d == io_r and ic_d == ic_r and bd.nuclei == br.nuclei
    assert dev.center_seed_cells() == ref.center_seed_cells()
    if case == 3:
        assert dev.lib.nicr_host_n_seeds(dev.h) == 1
    dev.close()
    ref.close()
