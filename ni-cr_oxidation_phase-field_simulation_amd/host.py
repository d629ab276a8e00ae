"""ctypes access to libnicr_host.so: the initial state of Model::Model and the writers.

The drop-in programs (nicr2d, nicr3d) use this library directly from C++; this
module exposes the same code to Python so tests and tools can build the
reference's initial state (Voronoi grains, eta fields, centre seed;
mpi2d/NiCr2D_grains.cpp:209-327, mpi3d-slab/NiCr3D_grains.cpp:356-494) and
write the reference's output files from any state.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import (FIELD, FIELD_ETA0, NicrBudget, NicrConfig, PKG_DIR, budget_to_dict, load_library)

HOST_LIB = os.path.join(PKG_DIR, "libnicr_host.so")
_HLIB = None
_DPTR = ctypes.POINTER(ctypes.c_double)


def host_library():
    global _HLIB
    if _HLIB is None:
        load_library()   # libnicr.so first (the host library links it)
        if not os.path.exists(HOST_LIB):
            raise RuntimeError(f"{HOST_LIB} is not built; run __graft_entry__.build() first")
        lib = ctypes.CDLL(HOST_LIB)
        lib.nicr_host_create.argtypes = [ctypes.POINTER(NicrConfig), ctypes.c_char_p, ctypes.c_char_p,
                                         ctypes.POINTER(ctypes.c_void_p)]
        lib.nicr_host_last_error.restype = ctypes.c_char_p
        for name in ("nicr_host_destroy", "nicr_host_n_eta", "nicr_host_n_seeds", "nicr_host_center_seed_cells"):
            getattr(lib, name).argtypes = [ctypes.c_void_p]
        lib.nicr_host_get.argtypes = [ctypes.c_void_p, ctypes.c_int, _DPTR]
        lib.nicr_host_set_field.argtypes = [ctypes.c_void_p, ctypes.c_int, _DPTR]
        lib.nicr_host_budget.argtypes = [ctypes.c_void_p, ctypes.POINTER(NicrBudget), _DPTR]
        lib.nicr_host_set_budget.argtypes = [ctypes.c_void_p, ctypes.POINTER(NicrBudget)]
        lib.nicr_host_seeds.argtypes = [ctypes.c_void_p, _DPTR]
        lib.nicr_host_uniform_noise.argtypes = [ctypes.c_ulonglong, ctypes.c_size_t, ctypes.c_double,
                                                ctypes.c_double, _DPTR]
        lib.nicr_host_write.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p]
        lib.nicr_host_parse.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_char_p),
                                        ctypes.POINTER(NicrConfig), ctypes.c_char_p, ctypes.c_size_t]
        _HLIB = lib
    return _HLIB


def uniform_noise(shape, lo: float, hi: float, seed: int = 20260909) -> np.ndarray:
    """std::uniform_real_distribution<double>(lo, hi) over std::mt19937_64(seed),
    drawn in the reference's loop order (i, then j, then k = storage order)."""
    lib = host_library()
    out = np.empty(shape, dtype=np.float64)
    if lib.nicr_host_uniform_noise(int(seed), out.size, float(lo), float(hi), out.ctypes.data_as(_DPTR)) != 0:
        raise RuntimeError(lib.nicr_host_last_error().decode())
    return out


OPTION_KEYS = ("bc", "ck", "bounds", "nucleation", "out", "eta_output", "output_format", "output_precision",
               "compression_level", "precision", "self_test")


def parse_args(dim: int, args):
    """The drop-in programs' argument parser. Returns (config, string options).

    Raises RuntimeError with the reference's message on invalid input; returns
    (None, {"early_exit": text}) for --help / --version."""
    lib = host_library()
    argv = ["nicr%dd" % dim] + list(args)
    arr = (ctypes.c_char_p * len(argv))(*[a.encode() for a in argv])
    cfg = NicrConfig()
    buf = ctypes.create_string_buffer(16384)
    rc = lib.nicr_host_parse(dim, len(argv), arr, ctypes.byref(cfg), buf, len(buf))
    if rc < 0:
        raise RuntimeError(lib.nicr_host_last_error().decode())
    if rc == 1:
        return None, {"early_exit": buf.value.decode()}
    return cfg, dict(zip(OPTION_KEYS, buf.value.decode().split("\n")))


class HostState:
    """The reference's constructed initial state, computed on the host."""

    def __init__(self, cfg: NicrConfig, bc: str, nucleation: str):
        self.lib = host_library()
        self.cfg = cfg
        h = ctypes.c_void_p()
        if self.lib.nicr_host_create(ctypes.byref(cfg), bc.encode(), nucleation.encode(), ctypes.byref(h)) != 0:
            raise RuntimeError(self.lib.nicr_host_last_error().decode())
        self.h = h
        self.shape = (cfg.nx, cfg.ny) if cfg.dim == 2 else (cfg.nx, cfg.ny, cfg.nz)
        self.n_eta = self.lib.nicr_host_n_eta(h)

    def get(self, name) -> np.ndarray:
        fid = 32 if name == "grain" else (FIELD_ETA0 + name if isinstance(name, int) else FIELD[name])
        out = np.empty(self.shape)
        if self.lib.nicr_host_get(self.h, fid, out.ctypes.data_as(_DPTR)) != 0:
            raise RuntimeError(self.lib.nicr_host_last_error().decode())
        return out

    def set(self, name, arr) -> None:
        fid = FIELD_ETA0 + name if isinstance(name, int) else FIELD[name]
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(arr, dtype=np.float64), self.shape))
        if self.lib.nicr_host_set_field(self.h, fid, a.ctypes.data_as(_DPTR)) != 0:
            raise RuntimeError(self.lib.nicr_host_last_error().decode())

    def budget(self):
        b = NicrBudget()
        init = (ctypes.c_double * 4)()
        self.lib.nicr_host_budget(self.h, ctypes.byref(b), init)
        return b, (init[0], init[1]), (init[2], init[3])

    def set_budget(self, b: NicrBudget) -> None:
        self.lib.nicr_host_set_budget(self.h, ctypes.byref(b))

    def seeds(self) -> np.ndarray:
        n = self.lib.nicr_host_n_seeds(self.h)
        out = np.empty((n, 3))
        self.lib.nicr_host_seeds(self.h, out.ctypes.data_as(_DPTR))
        return out

    def center_seed_cells(self) -> int:
        return self.lib.nicr_host_center_seed_cells(self.h)

    def write(self, what: str, path: str, step: int = 0, options: str = "") -> None:
        code = {"metadata": 0, "log": 1, "grain_log": 2, "vtk": 3, "vti": 4}[what]
        if self.lib.nicr_host_write(self.h, code, path.encode(), step, options.encode()) != 0:
            raise RuntimeError(self.lib.nicr_host_last_error().decode())

    def upload_to(self, model) -> None:
        """Copy the initial state into a device Model (or the oracle)."""
        for f in ("o", "cr", "p", "gb"):
            model.set(f, self.get(f))
        for g in range(self.n_eta):
            model.set(g, self.get(g))
        b, _, _ = self.budget()
        model.set_budget_raw(b)

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.nicr_host_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
