"""The C ABI libraries load and export every entry point their headers declare,
and the ctypes mirrors of the ABI structs have the C layout (CPU only: no call
touches a device)."""
import ctypes
import os
import re
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "ni-cr_oxidation_phase-field_simulation_amd")
DECL = re.compile(r"^\s*(?:const\s+char\s*\*|void\s*\*?|int)\s+(nicr_\w+)\s*\(", re.M)


def _declared(header):
    with open(os.path.join(REPO, "include", header)) as f:
        return sorted(set(DECL.findall(f.read())))


@pytest.mark.parametrize("header,lib", [("nicr.h", "libnicr.so"), ("nicr_nn.h", "libnicr.so"),
                                        ("nicr_host.h", "libnicr_host.so")])
def test_library_exports_every_declared_symbol(header, lib):
    names = _declared(header)
    assert len(names) > 10
    so = ctypes.CDLL(os.path.join(PKG, lib))
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, f"{lib} lacks {missing}"


def test_ctypes_struct_layout_matches_c(nicr, tmp_path):
    src = tmp_path / "sizes.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "nicr.h"\n'
                   'int main(void){printf("%zu %zu %zu %zu %zu %zu\\n", sizeof(nicr_config), '
                   'offsetof(nicr_config, layer_o), sizeof(nicr_budget), sizeof(nicr_step_report), '
                   'sizeof(nicr_stats), sizeof(nicr_batch_instance));return 0;}\n')
    exe = tmp_path / "sizes"
    subprocess.run(["gcc", "-std=c99", f"-I{os.path.join(REPO, 'include')}", str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = [ctypes.sizeof(nicr.NicrConfig), nicr.NicrConfig.layer_o.offset, ctypes.sizeof(nicr.NicrBudget),
            ctypes.sizeof(nicr.NicrStepReport), ctypes.sizeof(nicr.NicrStats),
            ctypes.sizeof(nicr.NicrBatchInstance)]
    assert got == want


def test_surrogate_struct_layout_matches_c(nicr, tmp_path):
    """nicr_nn.h structs against the ctypes mirrors in surrogate.py."""
    import importlib
    sur = importlib.import_module("ni-cr_oxidation_phase-field_simulation_amd.surrogate")
    src = tmp_path / "nn_sizes.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "nicr_nn.h"\n'
                   'int main(void){printf("%zu %zu %zu %zu %zu %zu %zu\\n", sizeof(nicr_unet_norm), '
                   'offsetof(nicr_unet_norm, jump), sizeof(nicr_sur_options), offsetof(nicr_sur_options, settle_steps), '
                   'sizeof(nicr_sur_report), offsetof(nicr_sur_report, last_residual), '
                   'offsetof(nicr_sur_report, step));return 0;}\n')
    exe = tmp_path / "nn_sizes"
    subprocess.run(["gcc", "-std=c99", f"-I{os.path.join(REPO, 'include')}", str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = [ctypes.sizeof(sur.UNetNorm), sur.UNetNorm.jump.offset, ctypes.sizeof(sur.SurOptions),
            sur.SurOptions.settle_steps.offset, ctypes.sizeof(sur.SurReport), sur.SurReport.last_residual.offset,
            sur.SurReport.step.offset]
    assert got == want


def test_grain_entry_points_reject_bad_arguments_before_the_device():
    """nicr_grain_distances and nicr_grain_terms check their arguments before any
    HIP call (the error path runs without a GPU)."""
    so = ctypes.CDLL(os.path.join(PKG, "libnicr.so"))
    so.nicr_last_error.restype = ctypes.c_char_p
    fn = so.nicr_grain_distances
    fn.argtypes = [ctypes.c_int] * 4 + [ctypes.c_void_p] * 4
    seeds = (ctypes.c_double * 3)(1.0, 1.0, 1.0)
    owner = (ctypes.c_int32 * 8)()
    db = (ctypes.c_double * 8)()
    assert fn(0, 2, 2, 1, seeds, owner, db, None) == -1
    assert b"bad argument" in so.nicr_last_error()
    assert fn(2, 2, 2, 0, seeds, owner, db, None) == -1
    assert fn(2, 2, 2, 1, None, owner, db, None) == -1
    gt = so.nicr_grain_terms
    gt.argtypes = [ctypes.c_void_p] * 4
    ext = (ctypes.c_double * 5)()
    cnt = (ctypes.c_int64 * 4)()
    assert gt(None, db, ext, cnt) == -1
    assert b"null argument" in so.nicr_last_error()


@pytest.mark.parametrize("change, message", [
    ({"abi_version": 1}, b"ABI version mismatch"),
    ({"dim": 4}, b"dim must be 2 or 3"),
    ({"nx": 0}, b"mesh dimensions must be positive"),
    ({"nz": -3}, b"mesh dimensions must be positive"),
    ({"precision": 7}, b"unknown precision"),
    ({"n_eta": -1}, b"bad n_eta"),
    # 2^20 x 2^12 x 2^12 cells: more 256-cell tiles than 32-bit tile indices hold
    ({"nx": 1 << 20, "ny": 1 << 12, "nz": 1 << 12}, b"mesh too large for 32-bit tile indexing"),
])
def test_create_rejects_bad_configurations_before_the_device(nicr, change, message):
    """nicr_create validates the configuration on the host, before any HIP call
    (the error path runs without a GPU): NICR_E_INVALID (-1), the reason in
    nicr_last_error, and no context."""
    lib = nicr.load_library()
    cfg = nicr.make_config(3, nx=16, ny=16, nz=16)
    for k, v in change.items():
        setattr(cfg, k, v)
    ctx = ctypes.c_void_p()
    assert lib.nicr_create(ctypes.byref(cfg), ctypes.byref(ctx)) == -1
    assert message in lib.nicr_last_error()
    assert not ctx.value
