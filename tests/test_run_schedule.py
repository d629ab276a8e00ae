"""Model.run_loop (the reference run loop without output, mpi2d/NiCr2D_grains.cpp:1216-1223)
queues the steps between Ksp events as single calls; it must insert and step
exactly like the per-step loop of the reference, on the CPU oracle."""
import numpy as np


def _loop(m, nsteps):
    """The reference's per-step loop (2D:1218-1223), one step at a time."""
    c = m.c
    nxt, scans = c.nucleation_period, 0
    for n in range(1, nsteps + 1):
        if n * c.dt + 1.e-12 * c.nucleation_period >= nxt:
            m.nucleate()
            scans += 1
            nxt += c.nucleation_period
        r = m.try_step(1)
        assert r.status == 0
    return scans


def test_run_matches_per_step_loop(nicr, oracle):
    # dt 1e-12, period 5e-11: events before steps 50, 100, 150 (and the
    # accumulated event time drifts by the period's rounding, as in the reference)
    shape = (24, 20)
    cfg = nicr.make_config(2, nx=shape[0], ny=shape[1], grains=False, bc="closed", ck="off", bounds="legacy",
                           nucleation="ksp", dt=1e-12, nucleation_period=5e-11, radius_grid=2.5)
    a, b = oracle.OracleModel(cfg), oracle.OracleModel(cfg)
    rng = np.random.default_rng(3)
    o = 0.004 + 0.002 * rng.random(shape)
    for m in (a, b):
        m.set("o", o)
        m.set("cr", np.full(shape, 0.2))
        m.set("p", np.zeros(shape))
    out = a.run_loop(170)
    scans = _loop(b, 170)
    assert out["steps"] == 170 and out["scans"] == scans == 3
    assert out["inserted"] > 0
    for f in ("o", "cr", "p"):
        assert np.array_equal(a.get(f), b.get(f))
    assert a.b == b.b
    # continuing a run carries the next event time
    out2 = a.run_loop(40, first=171, next_event=out["next_event"])
    scans2 = 0
    nxt = out["next_event"]
    for n in range(171, 211):
        if n * cfg.dt + 1.e-12 * cfg.nucleation_period >= nxt:
            b.nucleate()
            scans2 += 1
            nxt += cfg.nucleation_period
        b.try_step(1)
    assert out2["scans"] == scans2 == 1
    for f in ("o", "cr", "p"):
        assert np.array_equal(a.get(f), b.get(f))
