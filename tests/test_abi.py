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


@pytest.mark.parametrize("header,lib", [("nicr.h", "libnicr.so"), ("nicr_host.h", "libnicr_host.so")])
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
