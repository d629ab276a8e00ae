"""MI355X Ni-Cr phase-field time step: Python mirror of the reference `class Model`.

The reference programs expose their time step through `class Model`
(mpi2d/NiCr2D_grains.cpp:194-711, mpi3d-slab/NiCr3D_grains.cpp:336-978): public
fields `o`, `cr`, `p`, `eta`, the budget `b`, and the members `step()`,
`refresh()`, `nucleate()`, `stamp()` and `sum()`. This module offers the same
names over the C ABI of libnicr.so (include/nicr.h), whose kernels run on the
GPU. There is no CPU fallback: constructing a Model without the built library or
without a HIP device raises.

Field arrays are numpy float64 arrays of shape (nx, ny) in 2D and (nx, ny, nz) in
3D, holding interior cells in the reference's storage order (Y fastest in 2D,
Z fastest in 3D; mpi2d:149, mpi3d-slab:165-167).
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libnicr.so")

# ---- constants mirrored from include/nicr.h ----
ABI_VERSION = 2
BC = {"concentration": 0, "mu": 1, "closed": 2}
CK = {"legacy": 0, "off": 1}
BOUNDS = {"stop": 0, "legacy": 1}
NUCLEATION = {"ksp": 0, "center": 1, "off": 2}
PRECISION = {"fp64": 0, "mixed": 1}
FIELD = {"o": 0, "cr": 1, "p": 2, "gb": 3, "muO": 4, "muCr": 5, "mo": 6, "mc": 7,
         "eta_s2": 8, "eta_pairs": 9}
FIELD_ETA0 = 64
STEP_OK, STEP_NONFINITE, STEP_OUT_OF_BOUNDS, STEP_ETA_RANGE = 0, 1, 2, 3


class NicrConfig(ctypes.Structure):
    """POD mirror of the reference Config (mpi2d:39-64, mpi3d-slab:43-72)."""
    _fields_ = [
        ("abi_version", ctypes.c_int32), ("dim", ctypes.c_int32),
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("steps", ctypes.c_int32), ("output_every", ctypes.c_int32), ("left_edge_sites", ctypes.c_int32),
        ("random_seed", ctypes.c_uint64),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double), ("dt", ctypes.c_double),
        ("grain_diameter_grid", ctypes.c_double), ("gb_fwhm_grid", ctypes.c_double),
        ("gb_factor_o", ctypes.c_double), ("gb_factor_cr", ctypes.c_double),
        ("radius_grid", ctypes.c_double), ("exclusion_gap_grid", ctypes.c_double),
        ("nucleation_period", ctypes.c_double),
        ("mu_res", ctypes.c_double), ("surface_c_ref", ctypes.c_double),
        ("bc", ctypes.c_int32), ("ck", ctypes.c_int32), ("bounds", ctypes.c_int32), ("nucleation", ctypes.c_int32),
        ("grains", ctypes.c_int32), ("oxygen_diffusion", ctypes.c_int32),
        ("grain_evolution", ctypes.c_int32), ("frozen_eta_cache", ctypes.c_int32),
        ("eta_W", ctypes.c_double), ("eta_kappa", ctypes.c_double), ("eta_L", ctypes.c_double),
        ("eta_oxide_penalty", ctypes.c_double),
        ("max_eta_fields", ctypes.c_int32), ("n_eta", ctypes.c_int32), ("precision", ctypes.c_int32),
        ("reserved0", ctypes.c_int32),
        ("eta_K", ctypes.c_double), ("eta_A", ctypes.c_double), ("eta_scale", ctypes.c_double),
        ("layer_o", ctypes.c_double),
    ]


class NicrBudget(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double * 2) for n in
                ("nucO", "nucCr", "ckO", "ckCr", "clipO", "clipCr", "bcO", "bcCr")] + \
               [("nuclei", ctypes.c_int64), ("clipped", ctypes.c_int64)]


class NicrStepReport(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("steps_done", ctypes.c_int32),
                ("i", ctypes.c_int32), ("j", ctypes.c_int32), ("k", ctypes.c_int32), ("grain", ctypes.c_int32),
                ("o", ctypes.c_double), ("cr", ctypes.c_double), ("phi", ctypes.c_double),
                ("eta_value", ctypes.c_double)]


class NicrStats(ctypes.Structure):
    _fields_ = [("sum", ctypes.c_double * 2), ("min", ctypes.c_double), ("max", ctypes.c_double)]


BUDGET_KEYS = ("nucO", "nucCr", "ckO", "ckCr", "clipO", "clipCr", "bcO", "bcCr")

_VOIDP = ctypes.c_void_p
_DPTR = ctypes.POINTER(ctypes.c_double)


def _declare(lib, prefix: str):
    """Attach argtypes for every entry point of the nicr C ABI shape."""
    sigs = {
        "create": ([ctypes.POINTER(NicrConfig), ctypes.POINTER(_VOIDP)], ctypes.c_int),
        "destroy": ([_VOIDP], ctypes.c_int),
        "upload_field": ([_VOIDP, ctypes.c_int, _DPTR], ctypes.c_int),
        "download_field": ([_VOIDP, ctypes.c_int, _DPTR], ctypes.c_int),
        "set_budget": ([_VOIDP, ctypes.POINTER(NicrBudget)], ctypes.c_int),
        "get_budget": ([_VOIDP, ctypes.POINTER(NicrBudget)], ctypes.c_int),
        "refresh": ([_VOIDP], ctypes.c_int),
        "step": ([_VOIDP, ctypes.c_int, ctypes.POINTER(NicrStepReport)], ctypes.c_int),
        "nucleate": ([_VOIDP, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
        "stamp": ([_VOIDP, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
        "center_seed": ([_VOIDP, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
        "last_error": ([], ctypes.c_char_p),
    }
    for name, (args, res) in sigs.items():
        fn = getattr(lib, f"{prefix}_{name}", None)
        if fn is None:
            continue
        fn.argtypes = args
        fn.restype = res


_LIB = None


def load_library(path: str = LIB_PATH):
    """Load libnicr.so. Raises if it has not been built (no fallback path exists)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise RuntimeError(f"libnicr.so is not built ({path}); run __graft_entry__.build() first")
    lib = ctypes.CDLL(path)
    _declare(lib, "nicr")
    lib.nicr_config_default.argtypes = [ctypes.POINTER(NicrConfig), ctypes.c_int]
    lib.nicr_config_derive.argtypes = [ctypes.POINTER(NicrConfig)]
    lib.nicr_field_stats.argtypes = [_VOIDP, ctypes.c_int, ctypes.POINTER(NicrStats)]
    lib.nicr_sum_h_phi.argtypes = [_VOIDP, _DPTR]
    lib.nicr_kernel_times.argtypes = [_VOIDP, ctypes.c_char_p, ctypes.c_size_t]
    lib.nicr_reset_kernel_times.argtypes = [_VOIDP]
    lib.nicr_set_timing.argtypes = [_VOIDP, ctypes.c_int]
    lib.nicr_sync.argtypes = [_VOIDP]
    lib.nicr_device_bytes.argtypes = [_VOIDP, ctypes.POINTER(ctypes.c_size_t)]
    lib.nicr_copy_probe.argtypes = [_VOIDP, ctypes.c_size_t, ctypes.c_int, _DPTR]
    _LIB = lib
    return lib


def make_config(dim: int = 3, **kw) -> NicrConfig:
    """Reference defaults (Config{}) overridden by keyword arguments.

    String options take the reference's spellings: bc="mu", ck="off",
    bounds="legacy", nucleation="center", precision="mixed". Booleans may be
    given as bool. Derived eta coefficients are filled by nicr_config_derive.
    """
    lib = load_library()
    cfg = NicrConfig()
    if lib.nicr_config_default(ctypes.byref(cfg), dim) != 0:
        raise RuntimeError(lib.nicr_last_error().decode())
    maps = {"bc": BC, "ck": CK, "bounds": BOUNDS, "nucleation": NUCLEATION, "precision": PRECISION}
    for k, v in kw.items():
        if k in maps and isinstance(v, str):
            v = maps[k][v]
        if isinstance(v, bool):
            v = int(v)
        setattr(cfg, k, v)
    if "dx" in kw and "dy" not in kw:
        cfg.dy = kw["dx"]
    if "dx" in kw and "dz" not in kw:
        cfg.dz = kw["dx"]
    if dim == 2:
        cfg.nz = 1
    lib.nicr_config_derive(ctypes.byref(cfg))
    return cfg


def budget_to_dict(b: NicrBudget) -> dict:
    d = {k: float(getattr(b, k)[0]) + float(getattr(b, k)[1]) for k in BUDGET_KEYS}
    d["nuclei"] = int(b.nuclei)
    d["clipped"] = int(b.clipped)
    return d


def dict_to_budget(d: dict) -> NicrBudget:
    b = NicrBudget()
    for k in BUDGET_KEYS:
        v = d.get(k, 0.0)
        hi, lo = (v if isinstance(v, (tuple, list)) else (float(v), 0.0))
        getattr(b, k)[0] = hi
        getattr(b, k)[1] = lo
    b.nuclei = int(d.get("nuclei", 0))
    b.clipped = int(d.get("clipped", 0))
    return b


def step_error_message(rep: NicrStepReport, dim: int) -> str:
    """The runtime_error text Model::step throws (2D:573,585-588,352-354; SL:791,803-806,519-522)."""
    cell = f"{rep.i},{rep.j}" + (f",{rep.k}" if dim == 3 else "")
    if rep.status == STEP_NONFINITE:
        return "Nonfinite trial field: reduce dt / inspect the model; no update accepted."
    if rep.status == STEP_OUT_OF_BOUNDS:
        return (f"Out-of-bounds trial at ({cell}): O={rep.o:.16e}, Cr={rep.cr:.16e}, phi={rep.phi:.16e}"
                ". No clipping/update accepted. Reduce dt; if failure persists, investigate model.")
    if rep.status == STEP_ETA_RANGE:
        return (f"Eta trial outside [0,1]: grain={rep.grain} cell=({cell}) value={rep.eta_value:.17g}"
                ". All PDE fields rejected together; eta is never clipped. Check eta coefficients/dt.")
    return ""


class StepRejected(RuntimeError):
    """Raised by Model.step when the trial is rejected (the reference throws std::runtime_error)."""

    def __init__(self, message: str, report: NicrStepReport):
        super().__init__(message)
        self.report = report


class _Handle:
    """Owns one context of a nicr-shaped C ABI (libnicr, or the oracle in tests)."""

    def __init__(self, lib, prefix: str, cfg: NicrConfig):
        self.lib, self.prefix = lib, prefix
        self.cfg = cfg
        h = _VOIDP()
        rc = self._fn("create")(ctypes.byref(cfg), ctypes.byref(h))
        if rc != 0:
            raise RuntimeError(f"{prefix}_create failed: {self._fn('last_error')().decode()}")
        self.h = h

    def _fn(self, name):
        return getattr(self.lib, f"{self.prefix}_{name}")

    def check(self, rc):
        if rc != 0:
            raise RuntimeError(f"{self.prefix} call failed ({rc}): {self._fn('last_error')().decode()}")

    def close(self):
        if getattr(self, "h", None):
            self._fn("destroy")(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Model:
    """GPU-resident mirror of the reference `class Model` time-step interface."""

    def __init__(self, cfg: NicrConfig, _lib=None, _prefix: str = "nicr"):
        lib = _lib if _lib is not None else load_library()
        self._h = _Handle(lib, _prefix, cfg)
        self.c = cfg
        self.dim = cfg.dim
        self.shape = (cfg.nx, cfg.ny) if cfg.dim == 2 else (cfg.nx, cfg.ny, cfg.nz)
        self.n_eta = cfg.n_eta if (cfg.grains and cfg.grain_evolution) else 0

    # ---- fields ----
    def get(self, name) -> np.ndarray:
        fid = FIELD_ETA0 + name if isinstance(name, int) else FIELD[name]
        out = np.empty(self.shape, dtype=np.float64)
        self._h.check(self._h._fn("download_field")(self._h.h, fid, out.ctypes.data_as(_DPTR)))
        return out

    def set(self, name, arr) -> None:
        fid = FIELD_ETA0 + name if isinstance(name, int) else FIELD[name]
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(arr, dtype=np.float64), self.shape))
        self._h.check(self._h._fn("upload_field")(self._h.h, fid, a.ctypes.data_as(_DPTR)))

    o = property(lambda s: s.get("o"), lambda s, v: s.set("o", v))
    cr = property(lambda s: s.get("cr"), lambda s, v: s.set("cr", v))
    p = property(lambda s: s.get("p"), lambda s, v: s.set("p", v))
    gb = property(lambda s: s.get("gb"), lambda s, v: s.set("gb", v))

    def eta(self, g: int) -> np.ndarray:
        return self.get(int(g))

    def set_eta(self, g: int, arr) -> None:
        self.set(int(g), arr)

    # ---- budget ----
    @property
    def b(self) -> dict:
        bb = NicrBudget()
        self._h.check(self._h._fn("get_budget")(self._h.h, ctypes.byref(bb)))
        return budget_to_dict(bb)

    @b.setter
    def b(self, d: dict) -> None:
        bb = dict_to_budget(d)
        self._h.check(self._h._fn("set_budget")(self._h.h, ctypes.byref(bb)))

    def budget_raw(self) -> NicrBudget:
        bb = NicrBudget()
        self._h.check(self._h._fn("get_budget")(self._h.h, ctypes.byref(bb)))
        return bb

    def set_budget_raw(self, bb: NicrBudget) -> None:
        self._h.check(self._h._fn("set_budget")(self._h.h, ctypes.byref(bb)))

    # ---- time step ----
    def refresh(self) -> None:
        self._h.check(self._h._fn("refresh")(self._h.h))

    def step(self, nsteps: int = 1) -> NicrStepReport:
        rep = NicrStepReport()
        self._h.check(self._h._fn("step")(self._h.h, int(nsteps), ctypes.byref(rep)))
        if rep.status != STEP_OK:
            raise StepRejected(step_error_message(rep, self.dim), rep)
        return rep

    def try_step(self, nsteps: int = 1) -> NicrStepReport:
        rep = NicrStepReport()
        self._h.check(self._h._fn("step")(self._h.h, int(nsteps), ctypes.byref(rep)))
        return rep

    def run_loop(self, nsteps: int, first: int = 1, next_event: float | None = None) -> dict:
        """Steps first .. first + nsteps - 1 of the reference run loop without
        output (mpi2d/NiCr2D_grains.cpp:1216-1223, mpi3d-slab/NiCr3D_grains.cpp:1365-1372):
        with Ksp nucleation, Model::nucleate runs before step n whenever
        n dt + 1e-12 period >= the next event time, which then advances by one
        period. The steps between events are queued as one nicr_step call.
        Returns {"steps", "scans", "inserted", "next_event", "report"}; stops at a
        rejected step like the reference (report.status != 0)."""
        c = self.c
        period = c.nucleation_period
        nxt = period if next_event is None else next_event
        ksp = c.nucleation == NUCLEATION["ksp"]
        done = scans = inserted = 0
        n = first
        last = first + nsteps - 1
        rep = NicrStepReport()
        while n <= last:
            if ksp and n * c.dt + 1.e-12 * period >= nxt:
                inserted += self.nucleate()
                scans += 1
                nxt += period
            # the steps before the next event (event k fires before step n_k)
            m = n
            while ksp and m < last and (m + 1) * c.dt + 1.e-12 * period < nxt:
                m += 1
            if not ksp:
                m = last
            rep = self.try_step(m - n + 1)
            done += rep.steps_done
            if rep.status != STEP_OK:
                break
            n = m + 1
        return {"steps": done, "scans": scans, "inserted": inserted, "next_event": nxt, "report": rep}

    # ---- insertion ----
    def nucleate(self) -> int:
        n = ctypes.c_int(0)
        self._h.check(self._h._fn("nucleate")(self._h.h, ctypes.byref(n)))
        return n.value

    def stamp(self, x: int, y: int, z: int = 0) -> int:
        n = ctypes.c_int(0)
        self._h.check(self._h._fn("stamp")(self._h.h, int(x), int(y), int(z), ctypes.byref(n)))
        return n.value

    def initialize_center_oxide(self) -> int:
        """The centre seed (Model::initialize_center_oxide); returns the cell count.
        Raises RuntimeError with the reference's message when it is refused."""
        n = ctypes.c_int(0)
        self._h.check(self._h._fn("center_seed")(self._h.h, ctypes.byref(n)))
        return n.value

    # ---- diagnostics (GPU reductions) ----
    def stats(self, name: str) -> tuple:
        st = NicrStats()
        self._h.check(self._h.lib.nicr_field_stats(self._h.h, FIELD[name], ctypes.byref(st)))
        return (st.sum[0], st.sum[1]), st.min, st.max

    def sum_h_phi(self) -> tuple:
        out = (ctypes.c_double * 2)()
        self._h.check(self._h.lib.nicr_sum_h_phi(self._h.h, out))
        return out[0], out[1]

    def set_timing(self, on: bool) -> None:
        self._h.lib.nicr_set_timing(self._h.h, int(on))

    def reset_kernel_times(self) -> None:
        self._h.lib.nicr_reset_kernel_times(self._h.h)

    def kernel_times(self) -> dict:
        buf = ctypes.create_string_buffer(1024)
        self._h.lib.nicr_kernel_times(self._h.h, buf, 1024)
        out = {}
        for item in buf.value.decode().split(";"):
            if item:
                name, rest = item.split("=")
                ms, cnt = rest.split(":")
                out[name] = (float(ms), int(cnt))
        return out

    def sync(self) -> None:
        self._h.check(self._h.lib.nicr_sync(self._h.h))

    def copy_probe(self, nbytes: int, reps: int = 10) -> float:
        """Mean ms of a plain device copy of nbytes (bandwidth reference)."""
        ms = ctypes.c_double(0.0)
        self._h.check(self._h.lib.nicr_copy_probe(self._h.h, int(nbytes), int(reps), ctypes.byref(ms)))
        return ms.value

    def device_bytes(self) -> int:
        n = ctypes.c_size_t(0)
        self._h.lib.nicr_device_bytes(self._h.h, ctypes.byref(n))
        return n.value

    def close(self) -> None:
        self._h.close()


@dataclass
class InitialState:
    """Fields produced by Model::Model on the host (Voronoi, eta, centre seed)."""
    o: np.ndarray
    cr: np.ndarray
    p: np.ndarray
    gb: np.ndarray
    grain: np.ndarray
    eta: np.ndarray | None
    budget: dict
    initialO: float
    initialCr: float
    center_seed_cells: int
    seeds: np.ndarray
    volumes: np.ndarray


class NicrBatchInstance(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("layer_o", ctypes.c_double)]


class Ensemble:
    """Independent 2D instances stepped together on the device (nicr_batch_*).

    `cfg` is the shared 2D configuration (physics, dt, BCs; binary64, no
    evolving grains); `instances` is a list of (nx, ny, layer_o). Each instance
    steps bitwise as its own Model would, with its own rejection and budgets.
    """

    def __init__(self, cfg: NicrConfig, instances):
        lib = load_library()
        for name, args in (("create", [ctypes.POINTER(NicrConfig), ctypes.POINTER(NicrBatchInstance), ctypes.c_int,
                                       ctypes.POINTER(_VOIDP)]),
                           ("destroy", [_VOIDP]),
                           ("info", [_VOIDP, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_size_t)]),
                           ("upload_field", [_VOIDP, ctypes.c_int, ctypes.c_int, _DPTR]),
                           ("download_field", [_VOIDP, ctypes.c_int, ctypes.c_int, _DPTR]),
                           ("get_budget", [_VOIDP, ctypes.c_int, ctypes.POINTER(NicrBudget)]),
                           ("set_budget", [_VOIDP, ctypes.c_int, ctypes.POINTER(NicrBudget)]),
                           ("step", [_VOIDP, ctypes.c_int, ctypes.POINTER(NicrStepReport)]),
                           ("set_tiling", [_VOIDP, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
                           ("autotune", [_VOIDP, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                         ctypes.POINTER(ctypes.c_int), _DPTR])):
            fn = getattr(lib, f"nicr_batch_{name}")
            fn.argtypes = args
            fn.restype = ctypes.c_int
        self.lib = lib
        self.cfg = cfg
        self.shapes = [(int(nx), int(ny)) for nx, ny, _ in instances]
        arr = (NicrBatchInstance * len(instances))(*[NicrBatchInstance(int(nx), int(ny), float(lo))
                                                      for nx, ny, lo in instances])
        h = _VOIDP()
        self._check(lib.nicr_batch_create(ctypes.byref(cfg), arr, len(instances), ctypes.byref(h)))
        self.h = h
        self.n = len(instances)

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(self.lib.nicr_last_error().decode())

    @property
    def cells(self) -> int:
        n = ctypes.c_longlong(0)
        self._check(self.lib.nicr_batch_info(self.h, ctypes.byref(n), None))
        return n.value

    def get(self, inst: int, name: str) -> np.ndarray:
        out = np.empty(self.shapes[inst], dtype=np.float64)
        self._check(self.lib.nicr_batch_download_field(self.h, inst, FIELD[name], out.ctypes.data_as(_DPTR)))
        return out

    def set(self, inst: int, name: str, arr) -> None:
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(arr, dtype=np.float64), self.shapes[inst]))
        self._check(self.lib.nicr_batch_upload_field(self.h, inst, FIELD[name], a.ctypes.data_as(_DPTR)))

    def budget(self, inst: int) -> dict:
        bb = NicrBudget()
        self._check(self.lib.nicr_batch_get_budget(self.h, inst, ctypes.byref(bb)))
        return budget_to_dict(bb)

    def budget_raw(self, inst: int) -> NicrBudget:
        bb = NicrBudget()
        self._check(self.lib.nicr_batch_get_budget(self.h, inst, ctypes.byref(bb)))
        return bb

    def set_budget_raw(self, inst: int, bb: NicrBudget) -> None:
        self._check(self.lib.nicr_batch_set_budget(self.h, inst, ctypes.byref(bb)))

    def step(self, n: int):
        """n steps of every instance; returns the per-instance NicrStepReport list."""
        reps = (NicrStepReport * self.n)()
        self._check(self.lib.nicr_batch_step(self.h, int(n), reps))
        return list(reps)

    def set_tiling(self, tx: int, grid: int = 0, graph: bool = True) -> None:
        self._check(self.lib.nicr_batch_set_tiling(self.h, int(tx), int(grid), int(bool(graph))))

    def autotune(self, steps: int = 5):
        """Pick the fastest tiling on dry steps; returns (tx, grid, ms per step)."""
        tx, grid, ms = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_double(0.0)
        self._check(self.lib.nicr_batch_autotune(self.h, int(steps), ctypes.byref(tx), ctypes.byref(grid),
                                                 ctypes.byref(ms)))
        return tx.value, grid.value, ms.value

    def close(self) -> None:
        if self.h:
            self.lib.nicr_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
