"""Synthetic initial states of the workloads' shapes (numpy, host side).

These stand in for states the reference would build with its Model constructor
when a test or a benchmark needs a state the reference CLI cannot produce (noise
on Cr, many grains mapped onto fewer eta fields, smooth interfaces). Every array
is float64 in the reference's storage order: (nx, ny) or (nx, ny, nz), C order.
"""
from __future__ import annotations

import numpy as np

LAYER_O, LAYER_CR = 0.0023, 0.182
OXIDE_O, OXIDE_CR = 0.56300203, 0.35677915


def grid(shape):
    """1-based reference indices i, j, k broadcastable over `shape`."""
    ax = [np.arange(1, n + 1, dtype=np.float64) for n in shape]
    return np.meshgrid(*ax, indexing="ij")


def fill_smooth(shape, n_eta=0):
    """The smooth state of the reference self-tests (mpi3d-slab/NiCr3D_grains.cpp:1092-1101,
    mpi2d/NiCr2D_grains.cpp:1115-1120); the 2D form drops the k terms."""
    g = grid(shape)
    i, j = g[0], g[1]
    k = g[2] if len(shape) == 3 else 0.0
    o = .001 * (1. + .08 * np.sin(.3 * i + .5 * j - .4 * k))
    cr = .19 + .001 * np.cos(.2 * i - .3 * j + .6 * k)
    p = .1 + .01 * np.cos(.1 * i + .4 * j - .2 * k)
    eta = [.2 + .02 * np.cos(.2 * i - .3 * j + .7 * gg + .5 * k) for gg in range(n_eta)]
    return o, cr, p, eta


def oxide_particle(shape, centre=None, radius=4.0, width=1.5):
    """Smooth phi in [0,1] of a spherical (circular) oxide particle, for interface tests."""
    g = grid(shape)
    if centre is None:
        centre = [0.5 * (n + 1) for n in shape]
    r = np.sqrt(sum((gi - ci) ** 2 for gi, ci in zip(g, centre)))
    return 0.5 * (1.0 - np.tanh((r - radius) / width))


def noisy_alloy(shape, cr_mean=0.20, cr_noise=0.005, seed=20260909, layer=True):
    """O = 0 in the bulk, Cr = cr_mean + U(-noise, noise), inlet layer at i = 1.

    The noise is std::uniform_real_distribution<double> over
    std::mt19937_64(seed) drawn in the reference's loop order (the generator
    the reference seeds its grains with, mpi3d-slab/NiCr3D_grains.cpp:425-427),
    taken from the host library so that C++ drivers rebuild the same state.
    The concentration-BC layer then overwrites plane i = 1 as the reference's
    constructor does (mpi2d/NiCr2D_grains.cpp:214-219)."""
    from .host import uniform_noise
    o = np.zeros(shape)
    cr = cr_mean + uniform_noise(shape, -cr_noise, cr_noise, seed)
    p = np.zeros(shape)
    if layer:
        o[0] = LAYER_O
        cr[0] = LAYER_CR
    return o, cr, p


def voronoi_labels(shape, n_seeds, seed=20260909):
    """Nearest-seed labels with periodic Y (and Z), non-periodic X, and the
    distance to the nearest competing seed's bisector, computed in chunks of X planes."""
    rng = np.random.default_rng(seed)
    dims = len(shape)
    seeds = np.stack([rng.uniform(0, n, n_seeds) for n in shape], axis=1)
    if n_seeds >= 2:
        try:
            from scipy.spatial import cKDTree
        except ImportError:
            cKDTree = None
        if cKDTree is not None:
            # k-d tree over the seeds with periodic Y (and Z); X is made
            # non-periodic by a box far longer than every in-mesh distance
            box = np.array([4.0 * sum(shape)] + [float(n) for n in shape[1:]])
            tree = cKDTree(seeds, boxsize=box)
            labels = np.empty(shape, dtype=np.int32)
            dist = np.empty(shape)
            axes = [np.arange(n) + 0.5 for n in shape[1:]]
            for x0 in range(0, shape[0], 16):
                xs = np.arange(x0, min(shape[0], x0 + 16)) + 0.5
                pts = np.stack(np.meshgrid(xs, *axes, indexing="ij"), axis=-1).reshape(-1, dims)
                d, i = tree.query(pts, k=2, workers=-1)
                sl = (slice(x0, x0 + len(xs)),)
                labels[sl] = i[:, 0].reshape((len(xs),) + tuple(shape[1:]))
                dist[sl] = (0.5 * (d[:, 1] - d[:, 0])).reshape((len(xs),) + tuple(shape[1:]))
            return labels, dist, seeds
    labels = np.empty(shape, dtype=np.int32)
    dist = np.empty(shape)
    for x0 in range(shape[0]):
        pts = np.stack(np.meshgrid(*([np.array([x0 + 0.5])] + [np.arange(n) + 0.5 for n in shape[1:]]),
                                   indexing="ij"), axis=-1).reshape(-1, dims)
        d = pts[:, None, :] - seeds[None, :, :]
        for a in range(1, dims):
            n = shape[a]
            d[..., a] -= n * np.round(d[..., a] / n)
        r2 = np.sum(d * d, axis=-1)
        part = np.partition(r2, 1, axis=1)
        lab = np.argmin(r2, axis=1)
        labels[x0] = lab.reshape(shape[1:])
        # bisector distance approximation (r2nd - r1st) / (2 * separation) ~ (sqrt r2 - sqrt r1) / 2
        dist[x0] = (0.5 * (np.sqrt(part[:, 1]) - np.sqrt(part[:, 0]))).reshape(shape[1:])
    return labels, dist, seeds


def gb_mask_from_distance(dist, fwhm=4.0):
    """Gaussian GB mask of the fixed-grain mode (mpi2d:287, mpi3d-slab:463)."""
    return np.exp(-4. * np.log(2.) * dist * dist / (fwhm * fwhm))


def colour_seeds(labels, n_seeds, n_colours):
    """Greedy colouring of the seed adjacency graph (neighbouring grains differ)."""
    adj = [set() for _ in range(n_seeds)]
    for a in range(labels.ndim):
        l0 = labels
        l1 = np.roll(labels, -1, axis=a)
        if a == 0:
            l0, l1 = labels[:-1], labels[1:]
        m = l0 != l1
        for u, v in zip(l0[m].ravel(), l1[m].ravel()):
            adj[u].add(v)
            adj[v].add(u)
    colour = np.full(n_seeds, -1, dtype=np.int32)
    for s in range(n_seeds):
        used = {colour[t] for t in adj[s] if colour[t] >= 0}
        c = 0
        while c in used:
            c += 1
        if c >= n_colours:
            raise ValueError("seed adjacency needs more colours than eta fields")
        colour[s] = c
    return colour


def eta_from_labels(labels, dist, colour, n_eta, ell=1.5, metal=None):
    """Smooth eta fields: each cell's own grain gets weight 1, the nearest competitor
    exp(-2 d / ell); both normalised, then summed per colour (eta index)."""
    shape = labels.shape
    eta = np.zeros((n_eta,) + shape)
    w_own = 1.0 / (1.0 + np.exp(-2.0 * dist / ell))
    own = colour[labels]
    for c in range(n_eta):
        eta[c] = np.where(own == c, w_own, 0.0)
    # spread the remainder onto the colour of the nearest other grain: use the
    # neighbouring cell's colour along the axis of largest label change
    rest = 1.0 - w_own
    nb = own.copy()
    for a in range(labels.ndim):
        for sh in (-1, 1):
            rolled = np.roll(own, sh, axis=a)
            nb = np.where(nb == own, rolled, nb)
    for c in range(n_eta):
        eta[c] += np.where((nb == c) & (nb != own), rest, 0.0)
    eta_sum = eta.sum(axis=0)
    eta /= np.where(eta_sum > 0, eta_sum, 1.0)
    if metal is not None:
        eta *= metal
    return eta


def closed_alloy(shape, n_seeds=16, seed=1, cr_mean=0.20, cr_noise=0.005, o_background=1e-3, o_noise=0.2,
                 n_eta=0):
    """State of the surrogate workloads: Voronoi grains, Cr = cr_mean +
    U(-cr_noise, cr_noise), O = o_background (1 + U(-o_noise, o_noise)), no
    oxide (phi = 0), no inlet layer. With closed boundaries the step conserves
    O and Cr and no cell leaves [0, 1]. n_eta = 0: static grains, returns
    (o, cr, p, gb) with the GB mask of the fixed-grain mode; n_eta > 0: grains
    coloured onto n_eta evolving eta fields, returns (o, cr, p, eta[n_eta])."""
    rng = np.random.default_rng(seed)
    labels, dist, _ = voronoi_labels(shape, n_seeds, seed=seed)
    if n_eta > 0:
        grains = eta_from_labels(labels, dist, colour_seeds(labels, n_seeds, n_eta), n_eta)
    else:
        grains = gb_mask_from_distance(dist)
    del labels, dist
    cr = cr_mean + rng.uniform(-cr_noise, cr_noise, size=shape)
    o = o_background * (1.0 + rng.uniform(-o_noise, o_noise, size=shape))
    p = np.zeros(shape)
    return o, cr, p, grains
