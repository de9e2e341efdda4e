"""The no-allocation RHF build (oqph_fock_jk_noalloc, through the C ABI)
against the allocating device build (oqph_fock_jk_device) of the same session
and density.  Two calls in a row, the second on a smaller system (a second
session on the same GPU, which reuses the device's static scratch), then the
first system again at a new density (new work lists uploaded into the
scratch).  The two entry points run the same kernels: G agrees to 1e-12
elementwise (FP64 atomics sum in a different order) with identical quartet
counts; the first system is also checked against the oracle (1e-10)."""
import sys

import numpy as np
import pytest

from helpers import DeviceBuffer, small_system, density_for

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _pair(fb, dp):
    d = DeviceBuffer(dp.size)
    f_ref = DeviceBuffer(dp.size)
    f_na = DeviceBuffer(dp.size)
    d.upload(dp)
    f_na.upload(np.full(dp.size, 7.0))  # the entry point must overwrite every element
    st_ref = fb.fock_jk_device(d.address, f_ref.address)
    st_na = fb.fock_jk_noalloc(d.address, f_na.address)
    out = f_ref.download(), f_na.download(), st_ref, st_na
    for b in (d, f_ref, f_na):
        b.free()
    return out


def test_noalloc_matches_device_build(oracle):
    from openqp_amd.lib import FockBuilder
    big = small_system("cc-pvdz", nw=4, seed=7)
    small = small_system("6-31g(d)", nw=2, seed=8)
    fa = FockBuilder(big)
    fa.prepare()
    fs = FockBuilder(small)
    fs.prepare()
    cases = [("(H2O)4/cc-pVDZ", fa, big, 20, 11), ("(H2O)2/6-31G(d)", fs, small, 10, 12),
             ("(H2O)4/cc-pVDZ again, new density", fa, big, 20, 13)]
    for name, fb, bas, nocc, seed in cases:
        _say("no-alloc case", name, "nbf", bas.nbf, "seed", seed)
        _, dp = density_for(oracle, bas, nocc, seed=seed)
        f_ref, f_na, st_ref, st_na = _pair(fb, dp)
        err = float(np.abs(f_na - f_ref).max())
        _say("  max |f_noalloc - f_device| =", err, "quartets", st_na.quartets_computed, st_ref.quartets_computed)
        assert err <= 1e-12
        assert st_na.quartets_computed == st_ref.quartets_computed
        assert st_na.quartets_screened == st_ref.quartets_screened
        if bas is big:
            fo, _ = oracle.fock_jk(bas, dp)
            assert np.abs(f_na - fo).max() <= 1e-10
    fa.close()
    fs.close()


def test_noalloc_needs_prepared_session():
    """An unprepared session declines (code 4) before touching the device."""
    from openqp_amd.lib import FockBuilder, OqphError
    fb = FockBuilder(small_system("6-31g(d)", nw=1))
    n = fb.bas.nbf * (fb.bas.nbf + 1) // 2
    d, f = DeviceBuffer(n), DeviceBuffer(n)
    d.upload(np.zeros(n))
    with pytest.raises(OqphError, match="not prepared"):
        fb.fock_jk_noalloc(d.address, f.address)
    d.free()
    f.free()
    fb.close()
