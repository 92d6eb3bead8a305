"""GPU parity with the 4th-order numerical smoothing (smtopt=2 horizontal,
3 horizontal and vertical) in the forcing, against the oracle (whose
smoothing is pinned bitwise against the reference's smoo4* routines),
through the C ABI.  Bars: bitwise for the dynamics (cphopt=0) on periodic
and open domains; with the cold-rain microphysics err <= 1e-5 relative
after two steps."""
import numpy as np
import pytest

from cress_amd import cases, host
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

DYN = ["u", "v", "w", "pp", "ptp", "qv"]
ALL = DYN + ["qc", "qr", "qi", "qs", "qg", "nci"]


def box(a, f, cfg):
    ih = cfg.ni + 1 if f == "u" else cfg.ni
    jh = cfg.nj + 1 if f == "v" else cfg.nj
    kh = cfg.nk if f == "w" else cfg.nk - 1
    return a[1:ih, 1:jh, :kh]


def run_pair(case, nsteps):
    s = host.Solver(case.cfg)
    case.load(s)
    s.step(nsteps, first=True)
    m = po.Model(case, 32)
    m.step(nsteps, first=True)
    return s, m


@pytest.mark.parametrize("wbc", [-1, 7])
@pytest.mark.parametrize("smtopt", [2, 3])
def test_smoothing_dynamics_bitwise(smtopt, wbc):
    case = cases.build("bubble", 12, 10, 14, cphopt=0, smtopt=smtopt, wbc=wbc, gwave=30.0)
    s, m = run_pair(case, 3)
    for f in DYN:
        np.testing.assert_array_equal(box(s.download(f, 1), f, case.cfg), box(m.get(f, 1), f, case.cfg),
                                      err_msg=f)
    # the smoothing is active: the result differs from the unsmoothed run
    plain = cases.build("bubble", 12, 10, 14, cphopt=0, wbc=wbc, gwave=30.0)
    m0 = po.Model(plain, 32)
    m0.step(3, first=True)
    assert not np.array_equal(m0.get("ptp", 1), m.get("ptp", 1))
    s.close()


def test_smoothing_with_microphysics():
    case = cases.build("supercell", 16, 14, 20, cphopt=2, smtopt=3)
    s, m = run_pair(case, 2)
    for f in ALL:
        a = np.asarray(box(s.download(f, 1), f, case.cfg), np.float64)
        b = np.asarray(box(m.get(f, 1), f, case.cfg), np.float64)
        err = np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)
        assert err <= 1e-5, (f, err)
    s.close()
