"""The Fortran drop-in: an unmodified ScaLAPACK-style caller of PDGESV
(tests/fortran/pdgesv_drop_in.f90) compiled with flang and linked against
libscalapack_amd + the 1 x 1 BLACS shim instead of libscalapack.

CPU: it compiles and links, and every routine it calls resolves to the two
libraries.  GPU: it runs and reports a scaled residual below 16.
"""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
SRC = ROOT / "tests" / "fortran" / "pdgesv_drop_in.f90"
PKG = ROOT / "scalapack_amd"
FLANG = shutil.which("flang") or "/opt/rocm/llvm/bin/flang"


def _build(tmp_path: Path) -> Path:
    if not Path(FLANG).exists():
        pytest.skip("flang not available")
    if not (PKG / "libscalapack_amd.so").exists():
        pytest.skip("library not built")
    exe = tmp_path / "drop_in"
    subprocess.run([FLANG, "-O2", str(SRC), f"-L{PKG}", "-lscalapack_amd", "-lsla_blacs1x1",
                    f"-Wl,-rpath,{PKG}", "-o", str(exe)], check=True, capture_output=True,
                   text=True)
    return exe


def test_fortran_drop_in_links(tmp_path):
    exe = _build(tmp_path)
    undef = subprocess.run(["nm", "-D", "--undefined-only", str(exe)], check=True,
                           capture_output=True, text=True).stdout
    wanted = {"pdgesv_", "blacs_pinfo_", "blacs_get_", "blacs_gridinit_", "blacs_gridinfo_",
              "blacs_gridexit_", "blacs_exit_"}
    names = {ln.split()[-1] for ln in undef.splitlines() if ln.strip()}
    assert wanted <= names
    provided = set()
    for lib in ("libscalapack_amd.so", "libsla_blacs1x1.so"):
        out = subprocess.run(["nm", "-D", "--defined-only", str(PKG / lib)], check=True,
                             capture_output=True, text=True).stdout
        provided |= {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
    assert wanted <= provided


@pytest.mark.gpu
def test_fortran_drop_in_runs(tmp_path):
    exe = _build(tmp_path)
    print("[gpu] Fortran PDGESV drop-in N=300 NRHS=2", flush=True)
    proc = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert proc.returncode == 0, proc.stdout + proc.stderr
    assert proc.stdout.strip().startswith("OK"), proc.stdout
