"""Drop-in boundary: the tier-2 per-routine entry points and the Fortran
modules that replace the reference's m_gaussel, m_bcycle and m_advs.

CPU: the replacement modules (cress_amd/fortran/dropin/) compile against
the reference's own module files (the ones oracle/_ref builds from the
reference sources), i.e. they fit the reference tree with the same module,
generic and dummy names.
GPU: cress_gaussel, cress_advs and cress_bcycle give bitwise the oracle's
s_gaussel, s_advs and s_bcycle results on random arrays (the oracle being
pinned bitwise against the reference's routines in test_oracle_ref.py).
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from cress_amd import build as hb
from cress_amd import host
from oracle import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REFMOD = os.path.join(ROOT, "oracle", "_ref", "obj32")
DROPIN = os.path.join(ROOT, "cress_amd", "fortran", "dropin")
NI, NJ, NK = 13, 11, 14


@pytest.mark.skipif(not os.path.isdir(REFMOD) or not os.path.exists(hb.FLANG),
                    reason="reference modules or flang absent")
@pytest.mark.parametrize("name", ["gaussel", "bcycle", "advs"])
def test_dropin_module_compiles_in_reference_tree(name, tmp_path):
    moddir = os.path.join(hb.LIBDIR, "mod")
    if not os.path.exists(os.path.join(moddir, "m_cress_hip.mod")):
        pytest.skip("m_cress_hip not built")
    r = subprocess.run([hb.FLANG, "-c", "-ffp-contract=off", "-I", REFMOD, "-I", moddir, "-module-dir",
                        str(tmp_path), os.path.join(DROPIN, name + ".f90"), "-o",
                        str(tmp_path / (name + ".o"))], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr


def _rnd(rng, lo=-1.0, hi=1.0):
    return np.asfortranarray(rng.uniform(lo, hi, (NI + 2, NJ + 2, NK)).astype(np.float32))


def _ptr(a):
    assert a.flags["F_CONTIGUOUS"] and a.dtype == np.float32
    return a.ctypes.data_as(C.c_void_p)


@pytest.mark.gpu
def test_tier2_gaussel_matches_oracle():
    rng = np.random.default_rng(7)
    rr, tt, ff = _rnd(rng), _rnd(rng), _rnd(rng)
    ss = _rnd(rng, 3.0, 5.0)
    g = [x.copy(order="F") for x in (rr, ss, tt, ff)]
    o = [x.copy(order="F") for x in (rr, ss, tt, ff)]
    host.check(host.lib().cress_gaussel(2, NI - 2, 2, NJ - 2, 3, NK - 2, NI, NJ, NK, *[_ptr(a) for a in g]),
               "cress_gaussel")
    c = po.make_cfg(NI, NJ, NK, dx=700.0, dy=900.0, dz=350.0)
    po.Oracle(32).or_gaussel(c, 2, NI - 2, 2, NJ - 2, 3, NK - 2, *o)
    np.testing.assert_array_equal(g[3], o[3])
    np.testing.assert_array_equal(g[1], o[1])


@pytest.mark.gpu
@pytest.mark.parametrize("advopt", [1, 2, 3])
def test_tier2_advs_matches_oracle(advopt):
    rng = np.random.default_rng(20 + advopt)
    ux, vx, wx, s = (_rnd(rng) for _ in range(4))
    cfg = host.make_config(NI, NJ, NK, 700.0, 900.0, 350.0, 4.0, 1.0, advopt=advopt, cphopt=0, sthopt=0)
    g = _rnd(rng)
    o = [g.copy(order="F")] + [_rnd(rng) for _ in range(4)]
    host.check(host.lib().cress_advs(C.byref(cfg), _ptr(ux), _ptr(vx), _ptr(wx), _ptr(s), _ptr(g)),
               "cress_advs")
    c = po.make_cfg(NI, NJ, NK, dx=700.0, dy=900.0, dz=350.0, advopt=advopt)
    po.Oracle(32).or_advs(c, ux, vx, wx, s, *o)
    # the routine writes sfrc over the interior only
    np.testing.assert_array_equal(g[2:NI - 1, 2:NJ - 1, 1:NK - 2], o[0][2:NI - 1, 2:NJ - 1, 1:NK - 2])


@pytest.mark.gpu
def test_tier2_bcycle_matches_oracle():
    rng = np.random.default_rng(31)
    cfg = host.make_config(NI, NJ, NK, 700.0, 900.0, 350.0, 4.0, 1.0, cphopt=0)
    c = po.make_cfg(NI, NJ, NK, dx=700.0, dy=900.0, dz=350.0)
    ora = po.Oracle(32)
    for args in ((2, 1, NI - 2, NI - 1, 2, 1, NJ - 2, NJ - 1), (4, 0, NI - 3, NI + 1, 3, 0, NJ - 3, NJ)):
        a = _rnd(rng)
        b = a.copy(order="F")
        host.check(host.lib().cress_bcycle(C.byref(cfg), *args, NK, _ptr(a)), "cress_bcycle")
        ora.or_bcycle(c, *args, NK, b)
        np.testing.assert_array_equal(a, b)
