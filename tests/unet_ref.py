"""CPU float32 reference of the surrogate U-Net (torch), for the GPU kernel checks.

The reference follows the network of include/nicr_nn.h layer by layer and
rounds to bfloat16 exactly where the device kernels store bfloat16 tensors:
activations in the forward pass and their gradients in the backward pass
(`_RoundBF16` rounds both). Weights of the convolutions are bfloat16 values,
biases and the head are float32. Convolutions zero-pad dimension 0 and wrap
dimensions 1 and 2 (periodic Y and Z).
"""
from __future__ import annotations

import numpy as np
import torch
import torch.nn.functional as F

# (cin, cout) per convolution in parameter order
CONVS = [(32, 32), (32, 32), (32, 64), (64, 64), (64, 128), (128, 128), (128, 256), (256, 256),
         (384, 128), (128, 128), (192, 64), (64, 64), (96, 32), (32, 32)]


class _RoundBF16(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x):
        return x.to(torch.bfloat16).to(torch.float32)

    @staticmethod
    def backward(ctx, g):
        return g.to(torch.bfloat16).to(torch.float32)


def rb(x):
    return _RoundBF16.apply(x)


def bf16(a):
    """Round a numpy array to bfloat16 values (kept as float32)."""
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(torch.bfloat16).to(torch.float32).numpy()


def conv3(x, w, b):
    """x (nb, cin, s0, s1, s2); w (27, cout, cin) tap-major; zero pad dim 0, wrap dims 1, 2."""
    cout, cin = w.shape[1], w.shape[2]
    wt = w.reshape(3, 3, 3, cout, cin).permute(3, 4, 0, 1, 2)
    xp = F.pad(x, (1, 1, 1, 1, 0, 0), mode="circular")
    return F.conv3d(xp, wt, b, padding=(1, 0, 0))


def split_params(p):
    """Flat parameter vector -> list of (W (27, cout, cin), b) and the head (W (3, 32), b (3,))."""
    out, off = [], 0
    for cin, cout in CONVS:
        n = 27 * cout * cin
        out.append((p[off:off + n].reshape(27, cout, cin), p[off + n:off + n + cout]))
        off += n + cout
    hw = p[off:off + 96].reshape(3, 32)
    hb = p[off + 96:off + 99]
    assert off + 99 == p.size
    return out, (hw, hb)


def param_count():
    return sum(27 * ci * co + co for ci, co in CONVS) + 99


def up2(x):
    return x.repeat_interleave(2, dim=2).repeat_interleave(2, dim=3).repeat_interleave(2, dim=4)


def forward(params, x0):
    """params: list of torch leaf tensors [(W, b)...] + head (hw, hb); x0 (nb, 32, s0, s1, s2) bf16 values."""
    convs, (hw, hb) = params

    def cr(i, x):
        w, b = convs[i]
        return rb(F.relu(conv3(x, w, b)))

    e1 = cr(1, cr(0, x0))
    p1 = rb(F.avg_pool3d(e1, 2))
    e2 = cr(3, cr(2, p1))
    p2 = rb(F.avg_pool3d(e2, 2))
    e3 = cr(5, cr(4, p2))
    p3 = rb(F.avg_pool3d(e3, 2))
    b4 = cr(7, cr(6, p3))
    u3 = rb(torch.cat([up2(b4), e3], 1))
    d3 = cr(9, cr(8, u3))
    u2 = rb(torch.cat([up2(d3), e2], 1))
    d2 = cr(11, cr(10, u2))
    u1 = rb(torch.cat([up2(d2), e1], 1))
    d1 = cr(13, cr(12, u1))
    return torch.einsum("ncxyz,oc->noxyz", d1, hw) + hb.view(1, 3, 1, 1, 1)


def volume_stats(f):
    """Statistics of one volume (binary64): mean and population std of O and Cr,
    RMS of the updates of O, Cr, phi."""
    out = []
    for c in (0, 1):
        a = np.asarray(f[c], np.float64)
        m = a.sum() / a.size
        out += [m, np.sqrt(max(0.0, (a * a).sum() / a.size - m * m))]
    for c in (4, 5, 6):
        a = np.asarray(f[c], np.float64)
        out.append(np.sqrt((a * a).sum() / a.size))
    return out


def make_input(fields, norm):
    """fields (nb, 7, s0, s1, s2) float32 -> x0 (nb, 32, ...) as the device's input kernel forms it.

    norm: dict with phi_mean, phi_scale, gb_mean, gb_scale, c_mean[4], c_scale[4]."""
    f = np.asarray(fields, np.float32)
    x0 = np.zeros((f.shape[0], 32) + f.shape[2:], np.float32)
    for b in range(f.shape[0]):
        mo, so, mc, sc, ro, rc, rp = volume_stats(f[b])
        mean = [mo, mc, norm["phi_mean"], norm["gb_mean"]]
        inv = [1.0 / so if so > 0 else 0.0, 1.0 / sc if sc > 0 else 0.0, 1.0 / norm["phi_scale"],
               1.0 / norm["gb_scale"]]
        for c in range(4):
            x0[b, c] = ((f[b, c].astype(np.float64) - mean[c]) * inv[c]).astype(np.float32)
        for k, r in enumerate((ro, rc, rp)):
            x0[b, 4 + k] = (f[b, 4 + k].astype(np.float64) * (1.0 / r if r > 0 else 0.0)).astype(np.float32)
        cst = [np.log10(max(so, 1e-300)), mc, np.log10(max(ro, 1e-300)), np.log10(max(rc, 1e-300))]
        for k in range(4):
            x0[b, 7 + k] = np.float32((cst[k] - norm["c_mean"][k]) / norm["c_scale"][k])
    return bf16(x0)


def loss_and_grads(pvec, fields, target, norm):
    """(outputs, L, flat gradient) of L = (1/3) sum (out - target)^2 for the flat parameters pvec."""
    convs, (hw, hb) = split_params(pvec)
    tp = []
    for w, b in convs:
        tp.append((torch.tensor(bf16(w), requires_grad=True), torch.tensor(b, requires_grad=True)))
    th = (torch.tensor(hw, requires_grad=True), torch.tensor(hb, requires_grad=True))
    x0 = torch.from_numpy(make_input(fields, norm))
    out = forward((tp, th), x0)
    t = torch.from_numpy(np.asarray(target, np.float32))
    loss = ((out - t) ** 2).sum() / 3.0
    loss.backward()
    g = []
    for w, b in tp:
        g += [w.grad.numpy().ravel(), b.grad.numpy().ravel()]
    g += [th[0].grad.numpy().ravel(), th[1].grad.numpy().ravel()]
    return out.detach().numpy(), float(loss.detach()), np.concatenate(g)
