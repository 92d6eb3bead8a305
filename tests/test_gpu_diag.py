"""The FP64 mass and energy diagnostics on the GPU (cress_diag): the device
reductions equal an FP64 numpy restatement (tests/diag_ref.py) of the same
downloaded fields, repeat bit for bit, and close the water budget of a cold
start: with the winds at rest the forward step's water changes only by what
microphysics moves to the surface, and the leapfrog's precipitation
accumulation books half of the 2*dtb fallout per step (upwqp.f90:199-202,
dtp/2), so W(step 1) + 2 P(step 1) = W(0) up to one term: the reference
books the surface rate as the mean of the fluxes at levels 1 and 2, while
the column loses the level-2 flux.  That leaves the same residual in the
FP32 and FP64 oracle (9e-7 of W on the test case), which the GPU must
reproduce."""
import numpy as np
import pytest

from cress_amd import cases, host
from oracle import pyoracle as po
import diag_ref as dr

pytestmark = pytest.mark.gpu


def _get(s):
    return lambda name, lv: s.download(name, lv)


def _base(s):
    return {k: s.download(k) for k in ("rst", "jcb", "pbr", "ptbr")}


@pytest.mark.parametrize("name,dims,steps", [("bubble", (24, 22, 30), 3), ("supercell", (20, 18, 24), 2)])
def test_diag_equals_fp64_numpy(name, dims, steps):
    case = cases.build(name, *dims, cphopt=2)
    s = host.Solver(case.cfg)
    case.load(s)
    s.step(steps, first=True)
    s.sync()
    for lv in (0, 1, 2):
        got = s.diag(lv)
        ref = dr.diag(case.cfg, _base(s), _get(s), lv, s.download("prwtr"), s.download("price"))
        for i, k in enumerate(host.DIAG_NAMES):
            tol = 1e-11 * max(abs(ref[i]), 1e-300) if i < 9 else 0.0
            assert abs(got[k] - ref[i]) <= tol, (lv, k, got[k], ref[i])
        again = s.diag(lv)
        assert again == got
    s.close()


def test_cold_start_water_budget():
    case = cases.build("bubble", 24, 22, 30, cphopt=2)
    s = host.Solver(case.cfg)
    case.load(s)
    d0 = s.diag(1)
    w0 = d0["vapour"] + d0["liquid"] + d0["ice"]
    assert w0 > 0 and d0["precip"] == 0.0
    s.step(1, first=True)
    s.sync()
    d1 = s.diag(1)
    w1 = d1["vapour"] + d1["liquid"] + d1["ice"]
    resid = (w1 + 2.0 * d1["precip"] - w0) / w0
    m = po.Model(case, 32)
    get = lambda name, lv: m.get(name, lv)
    o0 = dr.diag(case.cfg, m.base, get, 1, m.prwtr, m.price)
    m.step(1, first=True)
    o1 = dr.diag(case.cfg, m.base, get, 1, m.prwtr, m.price)
    oresid = (o1[:3].sum() + 2.0 * o1[3] - o0[:3].sum()) / o0[:3].sum()
    print("W0 %.10e W1 %.10e P %.6e residual %.3e (oracle %.3e)" % (w0, w1, d1["precip"], resid, oresid))
    assert d1["precip"] > 0.0
    assert abs(resid) < 5e-6
    assert abs(resid - oresid) < 2e-8
    assert abs(d1["precip"] - o1[3]) <= 1e-5 * o1[3]
    s.close()
