"""The whole-row eta kernel (k_eta_rows: blocks of whole Z rows, the next X
plane loaded a plane ahead, binary32 cell coefficients through buffer
resources) against the 64-wide marching kernel (k_eta_march, NICR_ETA_ROWS=0 at
model creation). Both evaluate each cell with the same arithmetic in the same
field order, so every eta field, the mask and O/Cr/phi must be bitwise equal,
in binary32 (mixed) and binary64, over several steps. The shapes cover every
row width the row kernel takes (nz = 64, 128, 256), partial row tiles (ny not a
multiple of the block's rows), X segments that end inside the mesh and at its
last plane (the default segments, and three per column: odd and even segment
lengths for the loop unrolled by two), and several passes of fields (first,
middle and last pass)."""
import importlib
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

WL = "ni-cr_oxidation_phase-field_simulation_amd.workloads"


def _model(nicr, precision, state, n_eta, rows, nseg=None):
    """A model created under NICR_ETA_ROWS (row kernel on or off) and, when
    given, NICR_ETA_NSEG (X segments per column of the eta kernels)."""
    wl = importlib.import_module(WL)
    env = {"NICR_ETA_ROWS": "1" if rows else "0", "NICR_ETA_NSEG": None if nseg is None else str(nseg)}
    old = {k: os.environ.get(k) for k in env}
    for k, v in env.items():
        if v is None:
            os.environ.pop(k, None)
        else:
            os.environ[k] = v
    try:
        return wl.evolving_grain_model(precision, state, n_eta)
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


@pytest.mark.parametrize("nseg", [None, 3])
@pytest.mark.parametrize("precision", ["mixed", "fp64"])
@pytest.mark.parametrize("shape,n_seeds,n_eta", [
    ((20, 13, 256), 12, 16),     # two thread rows per block, ny odd: a partial tile
    ((37, 10, 128), 10, 32),     # four rows per block, ny not a multiple of 4; two passes (mixed)
    ((9, 24, 64), 8, 48),        # eight rows per block; three passes (mixed), six (binary64)
])
def test_row_kernel_equals_march_kernel(nicr, precision, shape, n_seeds, n_eta, nseg):
    wl = importlib.import_module(WL)
    state = wl.evolving_grain_state(shape, n_seeds, n_eta)
    a = _model(nicr, precision, state, n_eta, rows=True, nseg=nseg)
    b = _model(nicr, precision, state, n_eta, rows=False)
    try:
        ra, rb = a.try_step(6), b.try_step(6)
        assert ra.status == rb.status == 0 and ra.steps_done == rb.steps_done == 6
        for f in ["o", "cr", "p", "eta_s2", "gb"] + list(range(n_eta)):
            x, y = a.get(f), b.get(f)
            assert np.array_equal(x, y), f"{precision} {shape}: field {f} differs (max {np.max(np.abs(x - y)):.3e})"
        assert a.b == b.b
    finally:
        a.close()
        b.close()
