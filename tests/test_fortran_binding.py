"""The Fortran iso_c_binding module (hydra_amd/fortran/hydra_amd_c.f90)
compiles with amdflang, links against the library and reaches it: without a
device hyd_init's error comes back through the binding; on a GPU the small
run is compared with the Python binding's."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FC = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
LIBDIR = os.path.join(ROOT, "hydra_amd")


def build(tmp_path):
    if not os.path.exists(FC):
        pytest.skip("amdflang not available")
    if not os.path.exists(os.path.join(LIBDIR, "libhydra_amd.so")):
        pytest.skip("library not built")
    mod = os.path.join(ROOT, "hydra_amd", "fortran", "hydra_amd_c.f90")
    prog = os.path.join(ROOT, "tests", "fortran", "check_binding.f90")
    exe = str(tmp_path / "check_binding")
    subprocess.run([FC, "-J", str(tmp_path), "-c", mod, "-o", str(tmp_path / "m.o")], check=True)
    subprocess.run([FC, "-J", str(tmp_path), prog, str(tmp_path / "m.o"), "-L" + LIBDIR,
                    "-lhydra_amd", "-Wl,-rpath," + LIBDIR, "-o", exe], check=True)
    return exe


def test_binding_reports_errors_without_device(tmp_path):
    import torch
    if torch.cuda.device_count() > 0:
        pytest.skip("a device is present")
    exe = build(tmp_path)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120).stdout
    assert "init rc=-4" in out and "no HIP device" in out


@pytest.mark.gpu
def test_binding_runs_on_device(tmp_path):
    from hydra_amd import Hydra
    exe = build(tmp_path)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120).stdout
    assert "nptq=" in out, out
    ng = 32
    ix, iy = np.meshgrid(np.arange(ng), np.arange(ng))
    qq = 4 * np.pi * np.cos(2 * np.pi * ix / ng) * np.cos(2 * np.pi * iy / ng)
    h = Hydra(ng, ncontq=20)
    h.set_field(qq)
    h.run(10)
    n, nc = h.counts()
    t = h.scalar("t")
    h.close()
    f = dict(re.findall(r"(\w+)=\s*(\S+)", out))
    assert int(f["nptq"]) == n and int(f["nq"]) == nc
    assert float(f["t"]) == t  # 17 significant digits round-trip a double
