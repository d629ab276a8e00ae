"""Learned surrogate of the step over the C ABI of include/nicr_nn.h (libnicr.so).

`UNet` owns a device network (BF16 MFMA kernels, FP32 master weights) for
volumes of the model's 3D mesh; `Surrogate` records training samples from a
binary64 `Model`, trains the network on the GPU and advances a model with
network jumps that the plain step accepts or rejects. No counterpart in the
reference. There is no CPU path: every call runs the library's kernels.
"""
from __future__ import annotations

import ctypes
import time

import numpy as np

from . import NicrStepReport, load_library

_VOIDP = ctypes.c_void_p
_FPTR = ctypes.POINTER(ctypes.c_float)
_DPTR = ctypes.POINTER(ctypes.c_double)


class UNetNorm(ctypes.Structure):
    """nicr_unet_norm: fixed scales of phi, gb, the volume-statistics channels, and output gains."""
    _fields_ = [("phi_mean", ctypes.c_float), ("phi_scale", ctypes.c_float), ("gb_mean", ctypes.c_float),
                ("gb_scale", ctypes.c_float), ("c_mean", ctypes.c_float * 4), ("c_scale", ctypes.c_float * 4),
                ("gain", ctypes.c_float * 3), ("jump", ctypes.c_int32), ("stat_planes", ctypes.c_int32)]

    def as_dict(self):
        return {"phi_mean": self.phi_mean, "phi_scale": self.phi_scale, "gb_mean": self.gb_mean,
                "gb_scale": self.gb_scale, "c_mean": list(self.c_mean), "c_scale": list(self.c_scale),
                "gain": list(self.gain), "jump": self.jump, "stat_planes": self.stat_planes}


def make_norm(phi_mean=0.0, phi_scale=1.0, gb_mean=0.0, gb_scale=1.0, c_mean=(0.0,) * 4, c_scale=(1.0,) * 4,
              gain=(1.0,) * 3, jump=50, stat_planes=0) -> UNetNorm:
    nm = UNetNorm()
    nm.jump = int(jump)
    nm.stat_planes = int(stat_planes)
    nm.phi_mean, nm.phi_scale, nm.gb_mean, nm.gb_scale = phi_mean, phi_scale, gb_mean, gb_scale
    for k in range(4):
        nm.c_mean[k], nm.c_scale[k] = c_mean[k], c_scale[k]
    for k in range(3):
        nm.gain[k] = gain[k]
    return nm


class SurOptions(ctypes.Structure):
    _fields_ = [("jump", ctypes.c_int32), ("conserve_mass", ctypes.c_int32), ("residual_tol", ctypes.c_double),
                ("min_value", ctypes.c_double), ("max_value", ctypes.c_double), ("settle_steps", ctypes.c_int32),
                ("fallback_steps", ctypes.c_int32), ("min_jump", ctypes.c_int32)]


class SurReport(ctypes.Structure):
    _fields_ = [("steps_done", ctypes.c_int32), ("plain_steps", ctypes.c_int32), ("jumps_accepted", ctypes.c_int32),
                ("jumps_rejected", ctypes.c_int32), ("status", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("max_residual_o", ctypes.c_double), ("max_residual_cr", ctypes.c_double),
                ("min_rejected_residual", ctypes.c_double), ("max_residual_eta", ctypes.c_double),
                ("last_residual", ctypes.c_double * 3),
                ("step", NicrStepReport)]


def _declare(lib):
    if getattr(lib, "_nn_declared", False):
        return
    i, ll, u64, f, d = ctypes.c_int, ctypes.c_longlong, ctypes.c_uint64, ctypes.c_float, ctypes.c_double
    sigs = {
        "nicr_unet_create": [i, i, i, i, i, i, u64, ctypes.POINTER(_VOIDP)],
        "nicr_unet_destroy": [_VOIDP],
        "nicr_unet_param_count": [_VOIDP, ctypes.POINTER(ll)],
        "nicr_unet_device_bytes": [_VOIDP, ctypes.POINTER(ctypes.c_size_t)],
        "nicr_unet_get_params": [_VOIDP, _FPTR],
        "nicr_unet_set_params": [_VOIDP, _FPTR],
        "nicr_unet_get_grads": [_VOIDP, _FPTR],
        "nicr_unet_set_norm": [_VOIDP, ctypes.POINTER(UNetNorm)],
        "nicr_unet_get_norm": [_VOIDP, ctypes.POINTER(UNetNorm)],
        "nicr_unet_forward": [_VOIDP, i, _FPTR, _FPTR],
        "nicr_unet_loss_backward": [_VOIDP, i, _FPTR, _FPTR, _DPTR],
        "nicr_unet_set_timing": [_VOIDP, i],
        "nicr_unet_forward_ms": [_VOIDP, _DPTR],
        "nicr_nn_conv_check": [i, i, i, i, i, i, i, i, _FPTR, _FPTR, _FPTR, _FPTR, _FPTR],
        "nicr_sur_record": [_VOIDP, _VOIDP, i, ctypes.POINTER(NicrStepReport)],
        "nicr_sur_samples": [_VOIDP, ctypes.POINTER(i)],
        "nicr_sur_fit_norm": [_VOIDP],
        "nicr_sur_train": [_VOIDP, i, i, f, u64, _FPTR],
        "nicr_sur_set_augment": [_VOIDP, i],
        "nicr_sur_eval": [_VOIDP, i, i, _DPTR],
        "nicr_sur_advance": [_VOIDP, _VOIDP, i, ctypes.POINTER(SurOptions), ctypes.POINTER(SurReport)],
    }
    for name, args in sigs.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    lib._nn_declared = True


def _lib():
    lib = load_library()
    _declare(lib)
    return lib


def _check(lib, rc):
    if rc != 0:
        raise RuntimeError(lib.nicr_last_error().decode())


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _fp(a):
    return a.ctypes.data_as(_FPTR)


def conv_check(op: int, x, w, b, dy, relu: bool = False):
    """One 3x3x3 convolution of the network's kernels on host arrays.

    x: (nb, s0, s1, s2, cin), dy: (nb, s0, s1, s2, cout), w: (27, cout, cin),
    b: (cout,). op 0 -> y (nb, s0, s1, s2, cout); op 1 -> dx (.., cin); op 2 -> dw (27, cout, cin)."""
    lib = _lib()
    w = _f32(w)
    cout, cin = w.shape[1], w.shape[2]
    ref = x if x is not None else dy
    nb, s0, s1, s2 = ref.shape[:4]
    xs = _f32(x) if x is not None else None
    dys = _f32(dy) if dy is not None else None
    bs = _f32(b) if b is not None else None
    if op == 0:
        out = np.empty((nb, s0, s1, s2, cout), np.float32)
    elif op == 1:
        out = np.empty((nb, s0, s1, s2, cin), np.float32)
    else:
        out = np.empty((27, cout, cin), np.float32)
    _check(lib, lib.nicr_nn_conv_check(op, nb, s0, s1, s2, cin, cout, int(relu),
                                       _fp(xs) if xs is not None else None, _fp(w),
                                       _fp(bs) if bs is not None else None,
                                       _fp(dys) if dys is not None else None, _fp(out)))
    return out


class UNet:
    """Device U-Net for volumes `shape` = (s0, s1, s2), multiples of 8."""

    def __init__(self, shape, max_batch: int = 1, train: bool = False, max_samples: int = 0, seed: int = 1):
        lib = _lib()
        self.lib = lib
        self.shape = tuple(int(s) for s in shape)
        h = _VOIDP()
        _check(lib, lib.nicr_unet_create(*self.shape, int(max_batch), int(bool(train)), int(max_samples), int(seed),
                                         ctypes.byref(h)))
        self.h = h
        n = ctypes.c_longlong(0)
        _check(lib, lib.nicr_unet_param_count(h, ctypes.byref(n)))
        self.nparam = n.value

    @property
    def cells(self) -> int:
        return int(np.prod(self.shape))

    def device_bytes(self) -> int:
        b = ctypes.c_size_t(0)
        _check(self.lib, self.lib.nicr_unet_device_bytes(self.h, ctypes.byref(b)))
        return b.value

    def params(self) -> np.ndarray:
        out = np.empty(self.nparam, np.float32)
        _check(self.lib, self.lib.nicr_unet_get_params(self.h, _fp(out)))
        return out

    def set_params(self, p) -> None:
        p = _f32(p)
        assert p.size == self.nparam
        _check(self.lib, self.lib.nicr_unet_set_params(self.h, _fp(p)))

    def grads(self) -> np.ndarray:
        out = np.empty(self.nparam, np.float32)
        _check(self.lib, self.lib.nicr_unet_get_grads(self.h, _fp(out)))
        return out

    def set_norm(self, norm: UNetNorm = None, **kw) -> None:
        """Set the normalisation (a UNetNorm, or make_norm's keywords)."""
        nm = norm if norm is not None else make_norm(**kw)
        _check(self.lib, self.lib.nicr_unet_set_norm(self.h, ctypes.byref(nm)))

    def norm(self) -> UNetNorm:
        nm = UNetNorm()
        _check(self.lib, self.lib.nicr_unet_get_norm(self.h, ctypes.byref(nm)))
        return nm

    def forward(self, x):
        """x: (nb, 7, s0, s1, s2) float32 (O, Cr, phi, gb, plain updates of O, Cr, phi)
        -> (nb, 3, s0, s1, s2) normalised corrections."""
        x = _f32(x)
        out = np.empty((x.shape[0], 3) + self.shape, np.float32)
        _check(self.lib, self.lib.nicr_unet_forward(self.h, x.shape[0], _fp(x), _fp(out)))
        return out

    def loss_backward(self, x, target) -> float:
        x, t = _f32(x), _f32(target)
        loss = ctypes.c_double(0.0)
        _check(self.lib, self.lib.nicr_unet_loss_backward(self.h, x.shape[0], _fp(x), _fp(t), ctypes.byref(loss)))
        return loss.value

    # ---- training set and training ----
    def record(self, model, jump: int) -> NicrStepReport:
        rep = NicrStepReport()
        _check(self.lib, self.lib.nicr_sur_record(self.h, model._h.h, int(jump), ctypes.byref(rep)))
        return rep

    @property
    def samples(self) -> int:
        n = ctypes.c_int(0)
        _check(self.lib, self.lib.nicr_sur_samples(self.h, ctypes.byref(n)))
        return n.value

    def fit_norm(self) -> None:
        _check(self.lib, self.lib.nicr_sur_fit_norm(self.h))

    def set_augment(self, on: bool) -> None:
        """Train on samples through random Y/Z shifts and mirrors (nicr_sur_set_augment)."""
        _check(self.lib, self.lib.nicr_sur_set_augment(self.h, int(bool(on))))

    def train(self, iters: int, batch: int = 1, lr: float = 1e-3, seed: int = 1) -> np.ndarray:
        losses = np.zeros(max(1, iters), np.float32)
        _check(self.lib, self.lib.nicr_sur_train(self.h, int(iters), int(batch), float(lr), int(seed), _fp(losses)))
        return losses[:iters]

    def eval(self, first: int, count: int) -> float:
        mse = ctypes.c_double(0.0)
        _check(self.lib, self.lib.nicr_sur_eval(self.h, int(first), int(count), ctypes.byref(mse)))
        return mse.value

    # ---- surrogate stepping ----
    def advance(self, model, nsteps: int, jump: int = 50, conserve_mass: bool = True, residual_tol: float = 0.1,
                bounds=(0.0, 1.0), settle_steps: int = 0, fallback_steps: int = 0, min_jump: int = 0) -> SurReport:
        """Advance a binary64 Model by nsteps with accepted network jumps (nicr_sur_advance)."""
        opt = SurOptions(int(jump), int(bool(conserve_mass)), float(residual_tol), float(bounds[0]),
                         float(bounds[1]), int(settle_steps), int(fallback_steps), int(min_jump))
        rep = SurReport()
        _check(self.lib, self.lib.nicr_sur_advance(self.h, model._h.h, int(nsteps), ctypes.byref(opt),
                                                   ctypes.byref(rep)))
        return rep

    def run_loop(self, model, nsteps: int, first: int = 1, next_event=None, jump: int = 50, **kw) -> dict:
        """Steps first .. first + nsteps - 1 of the reference run loop
        (mpi2d/NiCr2D_grains.cpp:1216-1223, mpi3d-slab/NiCr3D_grains.cpp:1365-1372)
        with network jumps: Model::nucleate runs before step n whenever
        n dt + 1e-12 period >= the next event time (as Model.run_loop), and the
        steps between two scans are one advance() call, so no jump crosses a
        scan. kw: advance()'s options. Returns totals over the calls:
        {"steps", "scans", "inserted", "next_event", "plain_steps",
        "jumps_accepted", "jumps_rejected", "max_residual" (O, Cr, eta),
        "min_rejected_residual", "status"}."""
        from . import NUCLEATION
        c = model.c
        period = c.nucleation_period
        nxt = period if next_event is None else next_event
        ksp = c.nucleation == NUCLEATION["ksp"]
        out = {"steps": 0, "scans": 0, "inserted": 0, "plain_steps": 0, "jumps_accepted": 0, "jumps_rejected": 0,
               "max_residual": [0.0, 0.0, 0.0], "min_rejected_residual": None, "status": 0}
        n, last = first, first + nsteps - 1
        while n <= last:
            if ksp and n * c.dt + 1.e-12 * period >= nxt:
                out["inserted"] += model.nucleate()
                out["scans"] += 1
                nxt += period
            m = n
            while ksp and m < last and (m + 1) * c.dt + 1.e-12 * period < nxt:
                m += 1
            if not ksp:
                m = last
            rep = self.advance(model, m - n + 1, jump=jump, **kw)
            out["steps"] += rep.steps_done
            out["plain_steps"] += rep.plain_steps
            out["jumps_accepted"] += rep.jumps_accepted
            out["jumps_rejected"] += rep.jumps_rejected
            out["max_residual"] = [max(a, b) for a, b in zip(out["max_residual"], (
                rep.max_residual_o, rep.max_residual_cr, rep.max_residual_eta))]
            if rep.jumps_rejected:
                mr = rep.min_rejected_residual
                out["min_rejected_residual"] = mr if out["min_rejected_residual"] is None else min(
                    out["min_rejected_residual"], mr)
            if rep.status != 0:
                out["status"] = rep.status
                break
            n = m + 1
        out["next_event"] = nxt
        return out

    def set_timing(self, on: bool) -> None:
        _check(self.lib, self.lib.nicr_unet_set_timing(self.h, int(bool(on))))

    def forward_ms(self) -> float:
        ms = ctypes.c_double(0.0)
        _check(self.lib, self.lib.nicr_unet_forward_ms(self.h, ctypes.byref(ms)))
        return ms.value

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.nicr_unet_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- training pipeline on trajectories of the plain step ----

def closed_model(nicr, shape, state, precision="fp64"):
    """A binary64 3D Model of the surrogate workload (closed boundaries, no Ckp
    correction, no insertion) holding `state` = (o, cr, p, gb) for static
    grains or (o, cr, p, eta[n_eta]) for evolving ones."""
    grains = np.asarray(state[3])
    n_eta = grains.shape[0] if grains.ndim == len(shape) + 1 else 0
    cfg = nicr.make_config(3, nx=shape[0], ny=shape[1], nz=shape[2], grains=True, grain_evolution=n_eta > 0,
                           n_eta=n_eta, bc="closed", ck="off", bounds="legacy", nucleation="off",
                           precision=precision)
    m = nicr.Model(cfg)
    for k, a in zip(("o", "cr", "p"), state[:3]):
        m.set(k, a)
    if n_eta:
        for g in range(n_eta):
            m.set(g, grains[g])
    else:
        m.set("gb", grains)
    return m


def record_inlet_trajectories(net, nicr, n_traj, jump=50, n_struct=16, diameter=63.0, warm=(20, 5000), seed=0,
                              log=None, noise=0.0, noise_frac=0.75, noise_fields=("cr",)):
    """One training sample from each of n_traj plain binary64 trajectories of
    the static-grain inlet workload (the generator of the 128^3 static-grain
    configuration): the reference's Voronoi grains of --grain-diameter-grid
    `diameter` (static GB mask), Cr 0.20 +- U(0.005), O = 0 with the
    concentration-BC inlet on x = 0, legacy Ck, legacy bounds, Ksp scans on the
    reference schedule. Trajectory t uses grain structure t % n_struct and its
    own composition noise, runs a random number of steps in `warm` of the
    reference loop, then records (state, plain update, change over `jump`
    steps). noise: as record_trajectories, for the fields in noise_fields only
    (Cr by default: O is zero ahead of the inlet front, where cell-wise noise
    is clipped into a one-sided source)."""
    from . import workloads as wl
    from . import synthetic as syn
    shape = net.shape
    t0 = time.perf_counter()
    gbs = [wl.static_grain_gb(shape, diameter, seed + 100 + s) for s in range(n_struct)]
    t1 = time.perf_counter()
    steps = 0
    for t in range(n_traj):
        r = np.random.default_rng(seed + 1000 + t)
        o, cr, p = syn.noisy_alloy(shape, 0.20, 0.005, seed=seed + 1000 + t)
        m = wl.inlet_model(nicr, "fp64", (o, cr, p), gb=gbs[t % n_struct])
        w = int(r.integers(warm[0], warm[1]))
        res = m.run_loop(w)
        if res["steps"] != w:
            raise RuntimeError(f"trajectory {t}: a plain step was rejected")
        if noise > 0.0 and r.uniform() < noise_frac:
            c = r.uniform(0.0, noise)
            before = {k: m.get(k) for k in noise_fields}
            m.step(1)
            steps += 1
            for k in noise_fields:
                a = m.get(k)
                amp = c * jump * float(np.sqrt(np.mean((a - before[k]) ** 2)))
                m.set(k, np.clip(a + amp * r.standard_normal(shape), 0.0, 1.0))
        rr = net.record(m, jump)
        if rr.status != 0:
            raise RuntimeError(f"trajectory {t}: a plain step was rejected (status {rr.status})")
        steps += w + 1 + jump
        m.close()
        if log and (t + 1) % 50 == 0 and t + 1 < n_traj:
            log(f"{t + 1} trajectories recorded ({time.perf_counter() - t1:.1f} s)")
    t2 = time.perf_counter()
    if log:
        log(f"{n_struct} grain structures in {t1 - t0:.1f} s; {n_traj} trajectories ({steps} plain steps) "
            f"recorded in {t2 - t1:.1f} s")
    return {"trajectories": n_traj, "plain_steps": steps, "seconds": t2 - t0}


def record_trajectories(net, nicr, n_traj, jump=50, n_struct=16, n_seeds=16, warm=(20, 600), seed=0, log=None,
                        n_eta=0, noise=0.0, noise_frac=0.75):
    """One training sample from each of n_traj plain binary64 trajectories.

    Trajectory t starts from the grain structure t % n_struct (Voronoi, n_seeds
    seeds; static GB mask, or n_eta > 0 evolving eta fields), Cr mean
    U(0.17, 0.23) with U(-0.005, 0.005) noise, O background 10^U(-3.5, -2.5)
    with 20 % noise, takes a random number of plain steps in `warm`, then
    records (state, plain update, change over `jump` steps).

    noise > 0: in a share noise_frac of the trajectories, O and Cr are then
    perturbed cell by cell with Gaussian noise of RMS c * jump * RMS(u), c ~
    U(0, noise), u the field's last plain update, before the sample is
    recorded. A jump leaves errors of this kind in the state, and the plain
    steps of the sample show how the method damps them; without such samples
    the extrapolation jump * u amplifies the stiff ones from jump to jump."""
    from . import synthetic as syn
    shape = net.shape
    t0 = time.perf_counter()
    gbs = []
    for s in range(n_struct):
        labels, dist, _ = syn.voronoi_labels(shape, n_seeds, seed=seed + 100 + s)
        if n_eta > 0:
            gbs.append(syn.eta_from_labels(labels, dist, syn.colour_seeds(labels, n_seeds, n_eta), n_eta))
        else:
            gbs.append(syn.gb_mask_from_distance(dist))
    t1 = time.perf_counter()
    steps = 0
    for t in range(n_traj):
        r = np.random.default_rng(seed + 1000 + t)
        crm, ob = r.uniform(0.17, 0.23), 10 ** r.uniform(-3.5, -2.5)
        cr = crm + r.uniform(-0.005, 0.005, size=shape)
        o = ob * (1.0 + r.uniform(-0.2, 0.2, size=shape))
        m = closed_model(nicr, shape, (o, cr, np.zeros(shape), gbs[t % n_struct]))
        w = int(r.integers(warm[0], warm[1]))
        rep = m.step(w)
        if noise > 0.0 and r.uniform() < noise_frac:
            c = r.uniform(0.0, noise)
            before = {k: m.get(k) for k in ("o", "cr")}
            m.step(1)
            steps += 1
            for k in ("o", "cr"):
                a = m.get(k)
                amp = c * jump * float(np.sqrt(np.mean((a - before[k]) ** 2)))
                m.set(k, np.clip(a + amp * r.standard_normal(shape), 0.0, 1.0))
        rr = net.record(m, jump)
        if rr.status != 0:
            raise RuntimeError(f"trajectory {t}: a plain step was rejected (status {rr.status})")
        steps += w + 1 + jump
        m.close()
        if log and (t + 1) % 50 == 0 and t + 1 < n_traj:
            log(f"{t + 1} trajectories recorded ({time.perf_counter() - t1:.1f} s)")
    t2 = time.perf_counter()
    if log:
        log(f"{n_struct} grain structures in {t1 - t0:.1f} s; {n_traj} trajectories ({steps} plain steps) "
            f"recorded in {t2 - t1:.1f} s")
    return {"trajectories": n_traj, "plain_steps": steps, "seconds": t2 - t0}


def train_schedule(net, iters, batch=2, lr=1e-3, seed=1, log=None):
    """Adam with the learning rate stepped down to 0.3x at 60 % and 0.1x at 85 % of iters."""
    t0 = time.perf_counter()
    done, losses = 0, []
    b1, b2 = int(0.6 * iters), int(0.85 * iters)
    while done < iters:
        nxt = b1 if done < b1 else (b2 if done < b2 else iters)
        cur = lr if done < b1 else (0.3 * lr if done < b2 else 0.1 * lr)
        k = min(250, nxt - done)
        ls = net.train(k, batch=batch, lr=cur, seed=seed + done)
        losses.append(ls)
        done += k
        if log:
            log(f"iter {done}: loss {float(np.mean(ls[-min(50, len(ls)):])):.4f} ({time.perf_counter() - t0:.1f} s)")
    allv = np.concatenate(losses)
    return {"iters": iters, "batch": batch, "seconds": time.perf_counter() - t0,
            "loss_first": float(np.mean(allv[:20])), "loss_last": float(np.mean(allv[-100:]))}
