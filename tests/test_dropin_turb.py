"""The turbulence drop-in modules (cress_amd/fortran/dropin/turbdrv.f90,
turbflx.f90, turbs.f90, forcetke.f90) called as the reference calls them.

The same harness entry points (oracle/ref_harness.f90: ref_turbdrv,
ref_turbflx, ref_turbs, which set the name tables and call s_turbdrv,
s_turbflx, s_turbs with the reference's own argument lists) run twice: in
oracle/_ref/libcressrefmpi32.so, the reference's routines, and in
libcressrefdrop32.so, where the drop-in modules replace m_turbdrv,
m_turbflx and m_turbs and their C calls (cress_turbdrv, cress_turbflx,
cress_turbs) are answered on the CPU by the oracle (oracle/dropin_cpu.c),
itself pinned bitwise against these reference routines
(test_oracle_turb.py).  Every array the drop-ins return must be bitwise
the reference's: this checks what they pass across the boundary (the
Fortran-to-C argument order, fmois -> moist, the options read with
getiname/getrname, the spacings).  CPU only; the GPU side of the same entry
points is checked against the oracle in test_gpu_turb.py."""
import os

import numpy as np
import pytest

from oracle import pyoracle as po
from oracle import refmpi
from test_oracle_turb import DRV_OUT, DTB, INTERIOR, NI, NJ, NK, cfg, drv_inputs
from turb_inputs import tke_inputs

pytestmark = pytest.mark.skipif(not (refmpi.have() and os.path.exists(refmpi.LIB_DROPIN)),
                                reason="oracle/_ref drop-in part not built")

DRV_IN = ("jcb", "jcb8u", "jcb8v", "jcb8w", "pbr", "ptbr", "rbr", "up", "vp", "wp", "ppp", "ptpp", "qvp", "tkep",
          "qallp")
# what the drop-in returns (t11, t22, t12, t31, t32 are the reference's work arrays)
TAKEN = ["ufrc", "vfrc", "wfrc", "rkh", "rkv", "priv", "ssq", "nsq8w", "t13", "t23", "t33"]


def fld():
    return np.zeros((NI + 2, NJ + 2, NK), np.float32, order="F")


def both(c, name, args):
    a = refmpi.calls(c, [(name, [x.copy(order="F") if isinstance(x, np.ndarray) else x for x in args])])[0]
    b = refmpi.calls(c, [(name, [x.copy(order="F") if isinstance(x, np.ndarray) else x for x in args])],
                     lib=refmpi.LIB_DROPIN)[0]
    return a, b


@pytest.mark.parametrize("wbc,bbc", [(-1, 2), (7, 3)])
@pytest.mark.parametrize("moist,cphopt", [(0, 0), (1, 2)])
def test_turbdrv_dropin(moist, cphopt, wbc, bbc):
    c = cfg(wbc, cphopt, bbc)
    d, outs = drv_inputs(51 + cphopt + (wbc > 0) + 3 * bbc + 7 * moist)
    args = [moist, DTB, NI, NJ, NK] + [d[n] for n in DRV_IN] + [outs[n] for n in DRV_OUT] + [fld() for _ in range(3)]
    a, b = both(c, "ref_turbdrv", args)
    off = 5 + len(DRV_IN)
    for n in TAKEN:
        x, y = a[off + DRV_OUT.index(n)], b[off + DRV_OUT.index(n)]
        if n in INTERIOR:
            x, y = x[1:NI, 1:NJ, 0:NK - 1], y[1:NI, 1:NJ, 0:NK - 1]
        np.testing.assert_array_equal(y, x, err_msg=n)
    for n in ("ufrc", "vfrc", "wfrc"):
        assert not np.array_equal(b[off + DRV_OUT.index(n)], outs[n])


@pytest.mark.parametrize("wbc", [-1, 7])
def test_turbflx_turbs_dropins(wbc):
    c = cfg(wbc, 0, 2)
    rng = np.random.default_rng(30 + (wbc > 0))
    f = lambda lo, hi: np.asfortranarray(rng.uniform(lo, hi, (NI + 2, NJ + 2, NK)).astype(np.float32))
    jcb, s, kh, kv, kw = f(0.9, 1.1), f(290.0, 310.0), f(0.0, 50.0), f(0.0, 50.0), f(0.0, 20.0)
    h = [f(-1.0, 1.0) for _ in range(3)]
    a, b = both(c, "ref_turbflx", [NI, NJ, NK, jcb, s, kh, kv, kw] + h + [fld() for _ in range(4)])
    for q in range(3):
        np.testing.assert_array_equal(b[8 + q], a[8 + q], err_msg="h%d" % (q + 1))
        assert not np.array_equal(b[8 + q], h[q])
    ju, jv, sf = f(0.9, 1.1), f(0.9, 1.1), f(-1.0, 1.0)
    a, b = both(c, "ref_turbs", [NI, NJ, NK, ju, jv] + [x * 5.0 for x in h] + [sf] + [fld() for _ in range(4)])
    np.testing.assert_array_equal(b[8], a[8], err_msg="sfrc")
    assert not np.array_equal(b[8], sf)


@pytest.mark.parametrize("wbc,vsp", [(-1, 0), (7, 1)])
def test_forcetke_dropin(wbc, vsp):
    """s_forcetke through the drop-in: tkefrc, rkh, rkv and rkv8s bitwise
    the reference's (the CPU answer carries the reference's libm exp/log)."""
    c = po.make_cfg(NI, NJ, NK, dx=700.0, dy=900.0, dz=350.0, wbc=wbc, cphopt=0, vspopt=2 * vsp, ksp0=(2, 8),
                    vspvar=b"ooooooooo" + (b"o" if vsp else b"x"), gwave=30.0)
    d, outs = tke_inputs(NI, NJ, NK, 71 + (wbc > 0) + 2 * vsp)
    args = ([8, NI, NJ, NK] + [d[n] for n in ("jcb", "jcb8u", "jcb8v", "jcb8w")] + [fld()] +
            [d[n] for n in ("rst", "rstxu", "rstxv", "rstxwc", "tke", "tkep", "rbct", "priv", "nsq8w", "ssq")] +
            [outs["rkh"], outs["rkv"], outs["tkefrc"], fld(), outs["rkv8s"]] + [fld() for _ in range(3)] +
            [np.zeros((NI, NJ), np.float32, order="F")])
    a, b = both(c, "ref_forcetke", args)
    for n, q in (("rkh", 19), ("rkv", 20), ("tkefrc", 21), ("rkv8s", 23)):
        np.testing.assert_array_equal(b[q], a[q], err_msg=n)
        assert not np.array_equal(b[q], outs[n]), n
