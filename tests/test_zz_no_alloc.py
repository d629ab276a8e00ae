"""nicr_nucleate_noalloc: the Ksp scan with its scratch in the library's static
device buffer (no device allocation on the call) must equal nicr_nucleate bit for
bit, on two calls in a row (the second on a smaller mesh, reusing the buffer),
and refuse a mesh whose scratch exceeds the buffer."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(nicr, m):
    lib = nicr.load_library()
    fn = lib.nicr_nucleate_noalloc
    fn.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
    fn.restype = ctypes.c_int
    n = ctypes.c_int(-1)
    rc = fn(m._h.h, ctypes.byref(n))
    if rc != 0:
        raise RuntimeError(lib.nicr_last_error().decode())
    return n.value


def _pair(nicr, dim, shape, n_eta, radius):
    cfg = nicr.make_config(dim, nx=shape[0], ny=shape[1], nz=shape[2] if dim == 3 else 1, grains=True,
                           grain_evolution=n_eta > 0, n_eta=n_eta, bc="closed", bounds="legacy", ck="off",
                           nucleation="ksp", radius_grid=radius)
    rng = np.random.default_rng(5)
    o = 0.001 * (1.0 + 0.5 * rng.random(shape))
    models = []
    for _ in range(2):
        m = nicr.Model(cfg)
        m.set("o", o)
        m.set("cr", np.full(shape, 0.2))
        m.set("p", np.zeros(shape))
        for g in range(n_eta):
            m.set(g, np.full(shape, 1.0 / n_eta))
        models.append(m)
    return models


def _same(a, b):
    names = ["o", "cr", "p"] + list(range(a.n_eta))
    for f in names:
        assert np.array_equal(a.get(f), b.get(f)), f
    assert a.b == b.b


@pytest.mark.parametrize("first,second", [((3, (64, 40, 36), 3, 2.5), (2, (31, 29), 0, 3.0)),
                                          ((2, (96, 80), 2, 3.0), (3, (31, 13, 11), 0, 2.5))])
def test_noalloc_scan_equals_scan_twice(nicr, first, second):
    for dim, shape, n_eta, radius in (first, second):
        a, b = _pair(nicr, dim, shape, n_eta, radius)
        na = a.nucleate()
        nb = _noalloc(nicr, b)
        assert na == nb > 0, (dim, shape, na, nb)
        _same(a, b)
        # the scan's effect on later steps is the same too
        ra, rb = a.try_step(3), b.try_step(3)
        assert ra.status == rb.status == 0
        _same(a, b)
        a.close()
        b.close()


def test_noalloc_refuses_a_mesh_beyond_the_buffer(nicr):
    cfg = nicr.make_config(3, nx=256, ny=256, nz=160, grains=False, nucleation="ksp", bounds="legacy")
    m = nicr.Model(cfg)
    with pytest.raises(RuntimeError, match="static buffer holds"):
        _noalloc(nicr, m)
    m.close()
