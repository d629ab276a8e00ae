"""The symmetries that surrogate training applies to its samples
(nicr_sur_set_augment): with periodic Y/Z seams and mirror-symmetric stencils,
the reference step commutes with periodic shifts in Y and Z and with mirrors in
Y and Z, while X (the inlet side, Neumann ghosts) is kept. Checked on the CPU
oracle with the inlet, legacy Ck and legacy bounds: shifts are exact (the same
operations on relabelled cells); a mirror swaps the order of the two faces of
each axis in the flux sums, so it holds to rounding."""
import numpy as np
import pytest

SHAPE = (6, 8, 10)


def _model(nicr, oracle):
    cfg = nicr.make_config(3, nx=SHAPE[0], ny=SHAPE[1], nz=SHAPE[2], grains=True, bc="concentration",
                           ck="legacy", bounds="legacy", nucleation="off")
    return oracle.OracleModel(cfg)


def _state(seed):
    r = np.random.default_rng(seed)
    o = np.zeros(SHAPE)
    o[:2] = 0.0023 * r.uniform(0.5, 1.0, size=(2,) + SHAPE[1:])
    cr = 0.2 + r.uniform(-0.005, 0.005, size=SHAPE)
    p = np.zeros(SHAPE)
    gb = r.uniform(0.0, 1.0, size=SHAPE)
    return {"o": o, "cr": cr, "p": p, "gb": gb}


def _run(nicr, oracle, st, steps):
    m = _model(nicr, oracle)
    for k, v in st.items():
        m.set(k, v)
    r = m.step(steps)
    assert r.steps_done == steps
    return {k: m.get(k) for k in ("o", "cr", "p")}


@pytest.mark.parametrize("sy,sz,fy,fz", [(3, 0, 0, 0), (0, 7, 0, 0), (5, 2, 0, 0), (0, 0, 1, 0), (0, 0, 0, 1),
                                         (2, 9, 1, 1)])
def test_step_commutes_with_yz_shifts_and_mirrors(nicr, oracle, sy, sz, fy, fz):
    def T(a):
        if fy:
            a = a[:, ::-1, :]
        if fz:
            a = a[:, :, ::-1]
        return np.roll(a, (-sy, -sz), axis=(1, 2))

    st = _state(11)
    ref = _run(nicr, oracle, st, 20)
    got = _run(nicr, oracle, {k: np.ascontiguousarray(T(v)) for k, v in st.items()}, 20)
    for k in ("o", "cr"):
        want = T(ref[k])
        if not (fy or fz):
            assert np.array_equal(got[k], want), k
        else:
            assert np.allclose(got[k], want, rtol=1e-12, atol=1e-18), (k, np.abs(got[k] - want).max())
    # a mirror of X is no symmetry: the inlet face is at x = 0
    bad = _run(nicr, oracle, {k: np.ascontiguousarray(v[::-1]) for k, v in st.items()}, 20)
    assert not np.allclose(bad["o"], ref["o"][::-1], rtol=1e-6)
