"""GPU parity of the dynamics of the large step (cphopt=0: u, v, w, p', theta',
qv; no microphysics) against the oracle, through the C ABI.

Bar: bitwise equality with the FP32 oracle over the interior
(i=2..ni-2, j=2..nj-2, k=2..nk-2) of every prognostic, after several large
steps including the ctime=0 forward step; and max relative error <= 1e-5
against the FP64 oracle, err = max|x - x64| / max(max|x64|, floor).
"""
import numpy as np
import pytest

from cress_amd import cases, host
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

FIELDS = ["u", "v", "w", "pp", "ptp", "qv"]


def interior(a, cfg):
    return a[2:cfg.ni - 1, 2:cfg.nj - 1, 1:cfg.nk - 2]


def run_pair(case, nsteps):
    s = host.Solver(case.cfg)
    case.load(s)
    s.step(nsteps, first=True)
    m = po.Model(case, 32)
    m.step(nsteps, first=True)
    return s, m


@pytest.mark.parametrize("name,dims,nsteps", [
    ("bubble", (12, 10, 12), 4),
    ("vortex", (16, 14, 20), 3),
])
def test_dynamics_bitwise_vs_fp32_oracle(name, dims, nsteps):
    case = cases.build(name, *dims, cphopt=0)
    s, m = run_pair(case, nsteps)
    for lv in (0, 1):
        for f in FIELDS:
            g = s.download(f, lv)
            o = m.get(f, lv)
            np.testing.assert_array_equal(interior(g, case.cfg), interior(o, case.cfg),
                                          err_msg="%s level %d" % (f, lv))
    s.close()


def test_dynamics_advopt_variants():
    for advopt in (1, 2):
        case = cases.build("bubble", 12, 10, 12, cphopt=0, advopt=advopt)
        s, m = run_pair(case, 2)
        for f in FIELDS:
            np.testing.assert_array_equal(interior(s.download(f, 1), case.cfg),
                                          interior(m.get(f, 1), case.cfg), err_msg=f)
        s.close()


def test_dynamics_vs_fp64_oracle():
    case = cases.build("bubble", 12, 10, 12, cphopt=0)
    s = host.Solver(case.cfg)
    case.load(s)
    s.step(10, first=True)
    m = po.Model(case, 64)
    m.step(10, first=True)
    for f in FIELDS:
        g = interior(s.download(f, 1), case.cfg).astype(np.float64)
        o = interior(m.get(f, 1), case.cfg)
        err = np.abs(g - o).max() / max(np.abs(o).max(), 1e-12)
        assert err <= 1e-5, (f, err)
    s.close()


def test_resting_atmosphere_stays_at_rest():
    """Invariant: a horizontally uniform hydrostatic base state with zero
    perturbations produces zero tendencies (advection of a constant is 0,
    no buoyancy, no pressure gradient)."""
    case = cases.build("bubble", 12, 10, 12, cphopt=0)
    for k in case.init:
        if k != "qv":
            case.init[k][...] = 0.0
    case.init["qv"][...] = case.base["qvbr"]
    s = host.Solver(case.cfg)
    case.load(s)
    s.step(3, first=True)
    for f in ("u", "v", "w", "pp", "ptp"):
        assert np.abs(interior(s.download(f, 1), case.cfg)).max() == 0.0, f
    s.close()
