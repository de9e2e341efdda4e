"""The Fortran iso_c_binding module compiles with amdflang and links against
liboqph (CPU); on a GPU the Fortran driver's routec_fock_jk result matches
the oracle for the same basis and density."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from helpers import ROOT

LIBDIR = os.path.join(ROOT, "openqp_amd", "_lib")
FDIR = os.path.join(ROOT, "openqp_amd", "fortran")
FC = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"


def _build(tmp):
    for src, obj in (("oqph_bridge.F90", "bridge.o"), ("bridge_driver.F90", "drv.o")):
        subprocess.check_call([FC, "-c", os.path.join(FDIR, src), "-o", os.path.join(tmp, obj)], cwd=tmp)
    exe = os.path.join(tmp, "drv")
    subprocess.check_call([FC, os.path.join(tmp, "drv.o"), os.path.join(tmp, "bridge.o"), f"-L{LIBDIR}", "-loqph",
                           f"-Wl,-rpath,{LIBDIR}", "-o", exe], cwd=tmp)
    return exe


@pytest.mark.skipif(not os.path.exists(FC) or not os.path.exists(os.path.join(LIBDIR, "liboqph.so")),
                    reason="amdflang or liboqph missing")
def test_fortran_bridge_compiles_and_links(tmp_path):
    assert os.path.exists(_build(str(tmp_path)))


@pytest.mark.gpu
def test_fortran_driver_matches_oracle(tmp_path, oracle):
    from openqp_amd import basis as B
    exe = _build(str(tmp_path))
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120).stdout
    print(out, flush=True)
    info = int(out.split("info=")[1].split()[0])
    assert info == 0
    vals = {}
    for ln in out.splitlines():
        if ")=" in ln:
            k, v = ln.split("=")
            vals[k.strip()] = float(v)
    f_drv = np.array([vals[f"f({i})"] for i in range(1, 22)])
    finc = np.array([vals[f"finc({i})"] for i in range(1, 22)])
    fmx = np.array([vals[f"fmx({i})"] for i in range(1, 22)])
    assert int(out.split("inc_rc=")[1].split()[0]) == 0
    assert int(out.split("mixed_rc=")[1].split()[0]) == 0
    ex = np.array([10.0, 0.8, 0.9, 0.5])
    am = np.array([0, 0, 1, 0], np.int32)
    cc = np.array([1.0 / np.sqrt(B.gauss_norm(2 * e, l)) for e, l in zip(ex, am)])
    bas = B.Basis("drv", am, np.zeros(4, np.int32), np.array([0, 0, 0, 1], np.int32), np.ones(4, np.int32),
                  np.arange(4, dtype=np.int32), np.array([0, 1, 2, 5], np.int32), ex, cc,
                  np.array([[0, 0, 0], [0, 0, 1.4]], np.float64), np.array([8.0, 1.0]))
    d = 0.01 * np.arange(1, 22)
    f, _ = oracle.fock_jk(bas, d)
    # every element of the seam build, the incremental build and the mixed build
    assert np.abs(f_drv - f).max() < 1e-10
    assert np.abs(finc - f).max() < 1e-10
    assert np.abs(fmx - f).max() < 1e-6
