"""Learned surrogate: BF16 MFMA convolution kernels, U-Net forward and
gradients against a CPU float32 torch reference, training on the GPU, and the
accept/reject logic of surrogate jumps against the plain binary64 step.

Bars:
- one convolution (forward, data gradient): each output within 2^-8 relative
  (+1e-5 of the largest) of torch's float32 result on the same bfloat16 inputs
  -- the device rounds its float32 sums to bfloat16 once;
- weight gradient: float32 sums of exact bfloat16 products, within 1e-4 of the
  largest entry;
- whole network (tests/unet_ref.py rounds to bfloat16 where the device stores
  bfloat16): outputs within 1e-2 relative L2; every parameter block's gradient
  within 5e-2 relative L2 and cosine > 0.999 (different float32 summation orders
  flip single bfloat16 roundings; their effect grows along the backward chain to
  about 2.5e-2 at the 2^3-cell bottleneck, while a wiring mistake -- a wrong skip
  channel range, a missing ReLU mask -- costs O(1));
- a surrogate whose every jump is rejected leaves the model bitwise equal to
  plain stepping (fields and budgets): rejection restores the state exactly.
"""
import importlib

import numpy as np
import pytest
import torch

import unet_ref as R

pytestmark = pytest.mark.gpu

PKG = "ni-cr_oxidation_phase-field_simulation_amd"
OUT_RTOL = 2 ** -8
GRAD_REL_L2 = 5e-2
GRAD_COS = 0.999


def _sur():
    return importlib.import_module(PKG + ".surrogate")


def _syn():
    return importlib.import_module(PKG + ".synthetic")


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32))


CONV_CASES = [
    (2, 4, 6, 5, 32, 32),     # partial block, odd sides: zero pad in X, wrap in Y and Z
    (1, 6, 8, 8, 64, 64),     # 64-wide output tiles
    (1, 4, 4, 8, 96, 32),     # concatenated input
    (1, 2, 4, 4, 384, 128),
    (2, 2, 2, 2, 256, 256),   # bottleneck of a 16^3 volume
    (1, 3, 1, 1, 32, 64),     # a single Y/Z cell: every wrapped tap is the cell itself
    # sides multiples of (4, 16, 16): the halo-tile kernel
    (1, 4, 16, 16, 32, 32),   # one tile: every halo row wraps in Y and Z, X zero padded
    (2, 8, 16, 16, 96, 64),   # several input chunks, two output-channel units
    (1, 4, 16, 32, 64, 32),
    (1, 8, 32, 16, 384, 128),
]


def _conv_case(case, seed):
    nb, s0, s1, s2, cin, cout = case
    rng = np.random.default_rng(seed)
    x = R.bf16(rng.normal(size=(nb, s0, s1, s2, cin)))
    w = R.bf16(rng.normal(0, 1.0 / np.sqrt(27 * cin), size=(27, cout, cin)))
    b = rng.normal(0, 0.1, size=cout).astype(np.float32)
    dy = R.bf16(rng.normal(size=(nb, s0, s1, s2, cout)))
    return x, w, b, dy


def _close(a, ref, rtol=OUT_RTOL):
    tol = rtol * np.abs(ref) + 1e-5 * np.abs(ref).max()
    bad = np.abs(a - ref) > tol
    assert not bad.any(), f"{bad.sum()} of {bad.size} entries off; max |d| {np.abs(a - ref).max():.3e}"


@pytest.mark.parametrize("case", CONV_CASES)
def test_conv_forward_vs_torch(case):
    x, w, b, _ = _conv_case(case, 1)
    y = _sur().conv_check(0, x, w, b, None, relu=True)
    ref = torch.relu(R.conv3(_t(x).permute(0, 4, 1, 2, 3), _t(w), _t(b))).permute(0, 2, 3, 4, 1).numpy()
    _close(y, ref)


@pytest.mark.parametrize("case", CONV_CASES)
def test_conv_data_gradient_vs_torch(case):
    x, w, _, dy = _conv_case(case, 2)
    dx = _sur().conv_check(1, None, w, None, dy)
    xt = _t(x).permute(0, 4, 1, 2, 3).contiguous().requires_grad_(True)
    y = R.conv3(xt, _t(w), None)
    (g,) = torch.autograd.grad(y, xt, _t(dy).permute(0, 4, 1, 2, 3))
    _close(dx, g.permute(0, 2, 3, 4, 1).numpy())


@pytest.mark.parametrize("case", CONV_CASES)
def test_conv_weight_gradient_vs_torch(case):
    x, w, _, dy = _conv_case(case, 3)
    dw = _sur().conv_check(2, x, w, None, dy)
    wt = _t(w).requires_grad_(True)
    y = R.conv3(_t(x).permute(0, 4, 1, 2, 3), wt, None)
    (g,) = torch.autograd.grad(y, wt, _t(dy).permute(0, 4, 1, 2, 3))
    g = g.numpy()
    assert np.abs(dw - g).max() <= 1e-4 * np.abs(g).max()


def _random_net_params(net, seed):
    rng = np.random.default_rng(seed)
    p = net.params()
    convs, _ = R.split_params(p)
    off = 0
    for cin, cout in R.CONVS:
        off += 27 * cin * cout
        p[off:off + cout] = rng.normal(0, 0.05, size=cout)
        off += cout
    p[off:off + 99] = rng.normal(0, 0.3, size=99)
    return p


@pytest.mark.parametrize("shape,nb", [((16, 16, 16), 2), ((8, 16, 24), 1)])
def test_unet_forward_and_gradients_vs_torch(shape, nb):
    sur = _sur()
    net = sur.UNet(shape, max_batch=nb, train=True, seed=7)
    assert net.nparam == R.param_count()
    p = _random_net_params(net, 3)
    net.set_params(p)
    rng = np.random.default_rng(5)
    sz = (nb,) + shape
    fields = np.stack([rng.uniform(0, 2e-3, size=sz), 0.2 + 0.01 * rng.normal(size=sz), rng.uniform(0, 1, size=sz),
                       rng.uniform(0, 1, size=sz), 1e-7 * rng.normal(size=sz), 1e-6 * rng.normal(size=sz),
                       1e-4 * rng.normal(size=sz)], axis=1)
    norm = dict(phi_mean=0.5, phi_scale=0.3, gb_mean=0.5, gb_scale=0.3, c_mean=[-3.5, 0.2, -7.0, -6.0],
                c_scale=[0.5, 0.02, 0.5, 0.5], gain=[1.0, 1.0, 1.0], jump=50)
    net.set_norm(**norm)
    target = rng.normal(size=(nb, 3) + shape).astype(np.float32)
    out = net.forward(fields)
    ref_out, ref_loss, ref_g = R.loss_and_grads(net.params(), fields, target, norm)
    err = np.linalg.norm(out - ref_out) / np.linalg.norm(ref_out)
    assert err < 1e-2, f"forward rel L2 {err:.3e}"
    loss = net.loss_backward(fields, target)
    assert abs(loss - ref_loss / np.prod(shape) / nb) <= 2e-2 * abs(loss)
    g = net.grads()
    off = 0
    names = [f"conv{i}" for i in range(len(R.CONVS))]
    blocks = []
    for (cin, cout), nm in zip(R.CONVS, names):
        blocks += [(nm + ".W", off, off + 27 * cin * cout), (nm + ".b", off + 27 * cin * cout,
                                                             off + 27 * cin * cout + cout)]
        off += 27 * cin * cout + cout
    blocks += [("head.W", off, off + 96), ("head.b", off + 96, off + 99)]
    for nm, a, b in blocks:
        ga, gr = g[a:b], ref_g[a:b]
        if np.linalg.norm(gr) == 0:
            assert np.linalg.norm(ga) == 0, nm
            continue
        rel = np.linalg.norm(ga - gr) / np.linalg.norm(gr)
        cos = float(np.dot(ga, gr) / (np.linalg.norm(ga) * np.linalg.norm(gr)))
        assert rel < GRAD_REL_L2 and cos > GRAD_COS, f"{nm}: gradient rel L2 {rel:.3e}, cosine {cos:.6f}"
    net.close()


def test_inference_network_equals_training_network():
    """Without training buffers the decoder convolutions read the skip
    concatenation in place (coarse tensor upsampled + skip tensor) instead of
    a stored concatenation: the outputs must be bitwise those of the network
    that stores it. 16 x 32 x 32 has halo-kernel levels 0 and 1 and a per-tap
    level 2."""
    sur = _sur()
    shape = (16, 32, 32)
    a = sur.UNet(shape, max_batch=1, train=True, seed=7)
    b = sur.UNet(shape, max_batch=1, train=False, seed=7)
    p = _random_net_params(a, 5)
    a.set_params(p)
    b.set_params(p)
    norm = dict(phi_mean=0.5, phi_scale=0.3, gb_mean=0.5, gb_scale=0.3, c_mean=[-3.5, 0.2, -7.0, -6.0],
                c_scale=[0.5, 0.02, 0.5, 0.5], gain=[1.0, 1.0, 1.0], jump=50)
    a.set_norm(**norm)
    b.set_norm(**norm)
    rng = np.random.default_rng(9)
    sz = (1,) + shape
    fields = np.stack([rng.uniform(0, 2e-3, size=sz), 0.2 + 0.01 * rng.normal(size=sz), rng.uniform(0, 1, size=sz),
                       rng.uniform(0, 1, size=sz), 1e-7 * rng.normal(size=sz), 1e-6 * rng.normal(size=sz),
                       1e-4 * rng.normal(size=sz)], axis=1)
    assert np.array_equal(a.forward(fields), b.forward(fields))
    a.close()
    b.close()


def _closed_model(nicr, shape, seed, n_eta=0):
    state = _syn().closed_alloy(shape, n_seeds=8, seed=seed, n_eta=n_eta)
    return _sur().closed_model(nicr, shape, state)


def _plain_norm(net, J):
    net.set_norm(gb_mean=0.5, gb_scale=0.3, c_mean=[-3.5, 0.2, -7.0, -6.0], c_scale=[0.5, 0.02, 0.5, 0.5],
                 gain=[0.05, 0.05, 0.0], jump=J)



def _open_model(nicr, shape, seed, **change):
    state = _syn().closed_alloy(shape, n_seeds=8, seed=seed)
    kw = dict(grains=True, bc="concentration", ck="legacy", bounds="legacy", nucleation="ksp")
    kw.update(change)
    cfg = nicr.make_config(3, nx=shape[0], ny=shape[1], nz=shape[2], **kw)
    m = nicr.Model(cfg)
    o, cr, p, gb = state
    m.set("o", o); m.set("cr", cr); m.set("p", p); m.set("gb", gb)
    return m


def _totals(m):
    return {k: sum(m.stats(f)[0]) for k, f in (("O", "o"), ("Cr", "cr"))}


def _sources(b, sp):
    return b["bc" + sp] + b["nuc" + sp] + b["ck" + sp] + b["clip" + sp]


def test_jump_sources_are_trapezoid_of_step_sources(nicr):
    """One accepted jump on the inlet model (O inlet, legacy Ck, Ksp) without
    mass rescaling and a zero network head: the fields equal the extrapolated
    state x0 + J (x0 - x_-1) uploaded to a twin and stepped once, bitwise, and
    the budgets equal the twin's plus J/2 (s_pre + s_check) per slot, s the
    per-step sources of the step that gave u0 and of the check step. O on the
    inlet plane x = 0 is not extrapolated (every step resets it to the inlet
    value before its update)."""
    J = 20
    shape = (16, 16, 24)
    a = _open_model(nicr, shape, 5)
    b = _open_model(nicr, shape, 5)
    a.step(200)
    b.step(200)
    names = ("o", "cr", "p")
    xm = {f: b.get(f) for f in names}
    bm = b.b
    b.step(1)
    x0 = {f: b.get(f) for f in names}
    b0 = b.b
    for f in names:
        d = x0[f] - xm[f]
        if f == "o":
            d[0] = 0.0
        b.set(f, x0[f] + J * d)
    b.step(1)
    b1 = b.b
    net = _sur().UNet(shape, seed=1)
    _plain_norm(net, J)
    rep = net.advance(a, J + 2, jump=J, conserve_mass=False, residual_tol=1e30, bounds=(-1.0, 2.0))
    assert rep.status == 0 and rep.steps_done == J + 2 and rep.jumps_accepted == 1 and rep.plain_steps == 2
    for f in names:
        assert np.array_equal(a.get(f), b.get(f)), f
    ba = a.b
    for k in ("bcO", "ckO", "clipO", "nucO", "bcCr", "ckCr", "clipCr", "nucCr"):
        s_pre, s_check = b0[k] - bm[k], b1[k] - b0[k]
        want = b1[k] + 0.5 * J * (s_pre + s_check)
        assert ba[k] == pytest.approx(want, rel=1e-14, abs=1e-300), k
    assert b1["bcO"] - b0["bcO"] > 0          # the inlet feeds O every step
    net.close()


@pytest.mark.parametrize("jc,min_jump", [(12, 5), (12, 13)], ids=["short-jump", "below-min-jump"])
def test_last_short_jump(nicr, jc, min_jump):
    """A call whose steps do not fit a full jump (J = 20): with min_jump <= Jc
    the last cycle jumps Jc = steps - 2 (one plain step for u0, the check
    step): bitwise the extrapolation x0 + Jc (x0 - x_-1) (O held on the inlet
    plane) uploaded to a twin and stepped once, budgets the twin's plus Jc/2
    (s_pre + s_check); with min_jump > Jc every step is a plain step."""
    J = 20
    shape = (16, 16, 24)
    a = _open_model(nicr, shape, 7)
    b = _open_model(nicr, shape, 7)
    a.step(150)
    b.step(150)
    names = ("o", "cr", "p")
    net = _sur().UNet(shape, seed=1)
    _plain_norm(net, J)
    rep = net.advance(a, jc + 2, jump=J, conserve_mass=False, residual_tol=1e30, bounds=(-1.0, 2.0),
                      min_jump=min_jump)
    assert rep.status == 0 and rep.steps_done == jc + 2
    if min_jump > jc:
        assert rep.jumps_accepted == 0 and rep.plain_steps == jc + 2
        b.step(jc + 2)
        for f in names:
            assert np.array_equal(a.get(f), b.get(f)), f
        assert a.b == b.b
        net.close()
        return
    assert rep.jumps_accepted == 1 and rep.plain_steps == 2
    xm = {f: b.get(f) for f in names}
    bm = b.b
    b.step(1)
    x0 = {f: b.get(f) for f in names}
    b0 = b.b
    for f in names:
        d = x0[f] - xm[f]
        if f == "o":
            d[0] = 0.0
        b.set(f, x0[f] + jc * d)
    b.step(1)
    b1 = b.b
    for f in names:
        assert np.array_equal(a.get(f), b.get(f)), f
    ba = a.b
    for k in ("bcO", "ckO", "clipO", "bcCr", "ckCr", "clipCr"):
        want = b1[k] + 0.5 * jc * ((b0[k] - bm[k]) + (b1[k] - b0[k]))
        assert ba[k] == pytest.approx(want, rel=1e-14, abs=1e-300), k
    net.close()


@pytest.mark.parametrize("change", [dict(), dict(bc="mu"), dict(bc="closed")],
                         ids=["inlet-ck-ksp", "mu-inlet", "closed-ck-ksp"])
def test_open_model_jumps_carry_sources(nicr, change):
    """Checked jumps on models whose plain step takes O from an inlet
    (concentration or mu boundary) with legacy Ck and Ksp scans on: with
    conserve_mass the totals move with the recorded sources, so the reference's
    diagnostics (change of a total minus the recorded sources, 2D:697-710)
    close to rounding, and the inlet budget follows plain stepping's within 5 %
    (extrapolation-only jumps: a zero network head)."""
    J = 20
    shape = (16, 16, 24)
    a = _open_model(nicr, shape, 4, **change)
    b = _open_model(nicr, shape, 4, **change)
    a.step(300)
    b.step(300)
    t0, b0 = _totals(a), a.b
    net = _sur().UNet(shape, seed=1)
    _plain_norm(net, J)
    rep = net.advance(a, 200, jump=J, residual_tol=0.3)
    print(f"jumps accepted {rep.jumps_accepted} rejected {rep.jumps_rejected}")
    assert rep.status == 0 and rep.steps_done == 200
    assert rep.jumps_accepted >= 1
    b.step(200)
    t1, b1 = _totals(a), a.b
    tb, bb = _totals(b), b.b
    for sp in ("O", "Cr"):
        res = (t1[sp] - t0[sp]) - (_sources(b1, sp) - _sources(b0, sp))
        print(f"{sp}: closure {res / t0['Cr']:.2e}, sources surrogate {_sources(b1, sp) - _sources(b0, sp):.6e} "
              f"plain {_sources(bb, sp) - _sources(b0, sp):.6e}, total rel vs plain {abs(t1[sp] - tb[sp]) / tb[sp]:.2e}")
        assert abs(res) <= 1e-12 * t0["Cr"]
    if change.get("bc", "concentration") != "closed":
        ds, dp = b1["bcO"] - b0["bcO"], bb["bcO"] - b0["bcO"]
        assert dp > 0 and abs(ds - dp) <= 0.05 * dp
    else:
        assert t1["O"] == pytest.approx(t0["O"], rel=1e-13) and t1["Cr"] == pytest.approx(t0["Cr"], rel=1e-13)
    net.close()


def test_run_loop_scans_on_schedule(nicr):
    """The surrogate run loop calls Model::nucleate where the reference loop
    does (2D:1222) and never jumps across a scan: with every jump rejected it
    equals Model.run_loop bitwise, scans and budgets included."""
    shape = (16, 16, 24)
    a = _open_model(nicr, shape, 6)
    b = _open_model(nicr, shape, 6)
    net = _sur().UNet(shape, seed=1)
    _plain_norm(net, 20)
    ra = net.run_loop(a, 1100, jump=20, residual_tol=1e-12)
    rb = b.run_loop(1100)
    assert ra["status"] == 0 and ra["steps"] == 1100 and ra["jumps_accepted"] == 0 and ra["jumps_rejected"] > 0
    assert ra["scans"] == rb["scans"] == 2 and ra["inserted"] == rb["inserted"]
    for f in ("o", "cr", "p"):
        assert np.array_equal(a.get(f), b.get(f)), f
    assert a.b == b.b
    net.close()


@pytest.mark.parametrize("n_eta,settle,fallback", [(0, 0, 0), (8, 0, 0), (8, 2, 7)],
                         ids=["static", "evolving", "evolving-settle2-fallback7"])
def test_rejected_jumps_equal_plain_steps(nicr, n_eta, settle, fallback):
    """With a residual tolerance no proposal meets, every jump is rejected and
    the model equals plain stepping bitwise (fields, eta, mask and budgets),
    whatever plain steps the loop inserts (settle and fallback steps)."""
    shape = (16, 16, 24)
    a = _closed_model(nicr, shape, 4, n_eta)
    b = _closed_model(nicr, shape, 4, n_eta)
    net = _sur().UNet(shape, seed=1)
    _plain_norm(net, 20)
    rep = net.advance(a, 130, jump=20, residual_tol=1e-12, settle_steps=settle, fallback_steps=fallback)
    assert rep.status == 0 and rep.steps_done == 130
    assert rep.jumps_accepted == 0
    # attempts: the first after 1 + settle plain steps, then one per (fallback
    # + settle) steps while at least jump + 1 + settle steps are left
    fb = fallback or 20
    left, attempts, pre = 130, 0, 1 + settle
    while left >= 20 + 1 + pre:
        left -= pre + fb
        attempts += 1
        pre = settle
    assert rep.jumps_rejected == attempts
    assert rep.plain_steps == 130
    rb = b.step(130)
    assert rb.steps_done == 130
    for f in ("o", "cr", "p", "gb") + tuple(range(n_eta)):
        assert np.array_equal(a.get(f), b.get(f)), f
    assert a.b == b.b
    # the next steps read the restored mask
    a.step(3)
    b.step(3)
    for f in ("o", "cr", "p") + tuple(range(n_eta)):
        assert np.array_equal(a.get(f), b.get(f)), f
    net.close()


def test_evolving_eta_jump_is_checked_extrapolation(nicr):
    """Evolving grains: an accepted jump moves every eta field to the linear
    extrapolation x0 + J (x0 - x_-1) with its mask rebuilt, then takes the check
    step. With a zero network head the O, Cr, phi proposal is the same
    extrapolation, so the result equals the extrapolated state uploaded to a
    twin model and stepped once, bitwise; the eta residual is reported."""
    J = 20
    shape = (16, 16, 24)
    n_eta = 8
    a = _closed_model(nicr, shape, 5, n_eta)
    b = _closed_model(nicr, shape, 5, n_eta)
    a.step(30)
    b.step(30)
    names = ("o", "cr", "p") + tuple(range(n_eta))
    xm = {f: b.get(f) for f in names}
    b.step(1)
    x0 = {f: b.get(f) for f in names}
    for f in names:
        b.set(f, x0[f] + J * (x0[f] - xm[f]))
    rb = b.step(1)
    assert rb.steps_done == 1
    net = _sur().UNet(shape, seed=1)
    _plain_norm(net, J)
    rep = net.advance(a, J + 2, jump=J, conserve_mass=False, residual_tol=1e30, bounds=(-1.0, 2.0))
    assert rep.status == 0 and rep.steps_done == J + 2
    assert rep.jumps_accepted == 1 and rep.jumps_rejected == 0 and rep.plain_steps == 2
    assert 0.0 < rep.max_residual_eta < 1e30
    for f in names + ("gb",):
        assert np.array_equal(a.get(f), b.get(f)), f
    net.close()


@pytest.mark.parametrize("n_eta", [0, 8], ids=["static", "evolving"])
def test_training_reduces_loss_and_jumps_are_checked(nicr, n_eta):
    """Samples from plain trajectories, Adam on the GPU: the training loss falls;
    accepted jumps keep O and Cr totals and stay close to plain stepping; the
    trained network's jumps are markedly more accurate than the same jumps
    without it (a zero head: pure linear extrapolation J u), measured by the
    trapezoid residuals that decide acceptance."""
    shape = (32, 32, 32)
    net = _sur().UNet(shape, max_batch=2, train=True, max_samples=12, seed=3)
    for t in range(4):
        m = _closed_model(nicr, shape, 10 + t, n_eta)
        m.step(40)
        for k in range(3):
            rep = net.record(m, 20)
            assert rep.status == 0
    assert net.samples == 12
    # a recorded sample consumes 1 + jump steps of its trajectory
    net.fit_norm()
    assert net.norm().gain[2] == 0.0     # phi never changes in this workload
    losses = net.train(300, batch=2, lr=1e-3, seed=1)
    first, last = float(np.mean(losses[:10])), float(np.mean(losses[-30:]))
    print(f"training loss {first:.4f} -> {last:.4f}")
    assert np.all(np.isfinite(losses))
    assert last < 0.5 * first
    # surrogate stepping from a fresh state of the same generator
    a = _closed_model(nicr, shape, 99, n_eta)
    b = _closed_model(nicr, shape, 99, n_eta)
    a.step(120)
    b.step(120)
    m0 = [sum(a.stats("o")[0]), sum(a.stats("cr")[0])]
    rep = net.advance(a, 200, jump=20, residual_tol=0.3)
    b.step(200)
    print(f"jumps accepted {rep.jumps_accepted} rejected {rep.jumps_rejected} plain {rep.plain_steps}")
    assert rep.status == 0 and rep.steps_done == 200
    assert rep.plain_steps + 20 * rep.jumps_accepted == 200
    m1 = [sum(a.stats("o")[0]), sum(a.stats("cr")[0])]
    for x, y in zip(m0, m1):
        assert abs(x - y) <= 1e-12 * abs(x)
    for f in ("o", "cr"):
        fa, fb = a.get(f), b.get(f)
        rel = np.linalg.norm(fa - fb) / np.linalg.norm(fb)
        print(f"{f}: rel L2 vs plain {rel:.3e}")
        assert rel < 1e-2
    # the same jumps without the network: a zero head proposes x0 + J u0
    zero = _sur().UNet(shape, seed=1)
    zero.set_norm(net.norm())
    z = _closed_model(nicr, shape, 99, n_eta)
    z.step(120)
    rz = zero.advance(z, 200, jump=20, residual_tol=0.3)
    print(f"largest accepted residuals: trained O {rep.max_residual_o:.3f} Cr {rep.max_residual_cr:.3f}; "
          f"without the network O {rz.max_residual_o:.3f} Cr {rz.max_residual_cr:.3f}")
    assert rep.jumps_accepted >= 8 and rz.jumps_accepted >= 1
    assert rep.max_residual_o < 0.6 * rz.max_residual_o
    assert rep.max_residual_cr < 0.6 * rz.max_residual_cr
    zero.close()
    net.close()


_HALFK_SCRIPT = """
import importlib, sys
import numpy as np
sur = importlib.import_module("ni-cr_oxidation_phase-field_simulation_amd.surrogate")
shape = (16, 32, 32)
net = sur.UNet(shape, max_batch=1, train=False, seed=7)
rng = np.random.default_rng(11)
net.set_params(rng.normal(0, 0.03, size=net.nparam).astype(np.float32))
net.set_norm(phi_mean=0.5, phi_scale=0.3, gb_mean=0.5, gb_scale=0.3, c_mean=[-3.5, 0.2, -7.0, -6.0],
             c_scale=[0.5, 0.02, 0.5, 0.5], gain=[1.0, 1.0, 1.0], jump=50)
sz = (1,) + shape
fields = np.stack([rng.uniform(0, 2e-3, size=sz), 0.2 + 0.01 * rng.normal(size=sz), rng.uniform(0, 1, size=sz),
                   rng.uniform(0, 1, size=sz), 1e-7 * rng.normal(size=sz), 1e-6 * rng.normal(size=sz),
                   1e-4 * rng.normal(size=sz)], axis=1)
np.save(sys.argv[1], net.forward(fields))
net.close()
"""


def test_input_layer_half_k_equals_full_k(tmp_path):
    """The input convolution computes only the 16 input channels that carry
    data (the input tensor's channels 16..31 are zero): the network's output
    must be bitwise that of the 32-channel computation (NICR_CONV_HALFK=0). The
    switch is read once per process, so each variant runs in a child process."""
    import os
    import subprocess
    import sys
    repo = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    script = tmp_path / "halfk.py"
    script.write_text(_HALFK_SCRIPT)
    outs = []
    for flag in ("1", "0"):
        env = dict(os.environ, NICR_CONV_HALFK=flag, PYTHONPATH=repo)
        path = tmp_path / f"out{flag}.npy"
        print(f"[halfk] NICR_CONV_HALFK={flag}", flush=True)
        r = subprocess.run([sys.executable, str(script), str(path)], cwd=repo, env=env, capture_output=True,
                           text=True, timeout=240)
        assert r.returncode == 0, r.stderr[-2000:]
        outs.append(np.load(path))
    assert np.array_equal(outs[0], outs[1])
    assert np.abs(outs[0]).max() > 0
