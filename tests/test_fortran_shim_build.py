"""The Fortran drop-in compiles in both of its configurations (CPU only).

Built on its own, LEQ_BICGS keeps ITER_TOT in the library; built inside the
MFIX tree with -DMFIX_LEQSOL_ITER_TOT it also adds each solve's count to
leqsol's ITER_TOT(VNO) (model/leq_bicgs.f:458, 491;
model/leqsol_mod.f:17).  The second configuration is compiled here against a
module that declares only the one array the shim uses, with MFIX's name and
type, which is what the shim needs from leqsol.
"""
import shutil
import subprocess
from pathlib import Path

import pytest

from mfix_amd import build

SHIM = Path(build.FSRC) / "leq_bicgs_shim.f90"


@pytest.mark.parametrize("in_tree", [False, True], ids=["standalone", "mfix_tree"])
def test_shim_compiles(tmp_path, in_tree):
    fc = build._flang()
    if fc is None:
        pytest.skip("no flang")
    flags = [fc, "-cpp", "-c", "-module-dir", str(tmp_path)]
    if in_tree:
        mod = tmp_path / "leqsol_iface.f90"
        mod.write_text("module leqsol\n  implicit none\n  integer :: iter_tot(10) = 0\nend module leqsol\n")
        subprocess.run(flags + [str(mod), "-o", str(tmp_path / "leqsol.o")], check=True, capture_output=True)
        flags.append("-DMFIX_LEQSOL_ITER_TOT")
    out = subprocess.run(flags + [str(SHIM), "-o", str(tmp_path / "shim.o")], capture_output=True, text=True)
    assert out.returncode == 0, out.stderr
    nm = shutil.which("nm")
    if nm:
        syms = subprocess.run([nm, str(tmp_path / "shim.o")], capture_output=True, text=True).stdout
        assert "leq_bicgs_" in syms
        assert ("leqsol" in syms) == in_tree
