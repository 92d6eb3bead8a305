"""The ensemble entry points (cress_ens_*): members of uneven sizes, periodic
and open, with even and odd substep counts, stepped through captured
hipGraph cycles give bitwise the results of stepping each member alone with
cress_large_step -- checked after every call, with graph replays, eager
steps that do not fill a cycle, downloads and recaptures in between -- and
eight uneven members stepped together match the oracle run on each alone."""
import numpy as np
import pytest

from cress_amd import cases, host
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

FIELDS = ["u", "v", "w", "pp", "ptp", "qv", "qc", "qr", "qi", "qs", "qg", "nci"]


def members():
    return [
        cases.build("bubble", 12, 10, 24, cphopt=2),                       # nsstp 5: cycle of 6
        cases.build("supercell", 16, 14, 16, cphopt=2, wbc=7, gwave=30.0),  # open sides
        cases.build("vortex", 20, 18, 16, cphopt=0),
    ]


def same(a, b):
    for lv in (0, 1):
        for f in FIELDS:
            np.testing.assert_array_equal(a.download(f, lv), b.download(f, lv), err_msg="%s level %d" % (f, lv))
    for f in ("prwtr", "price"):
        np.testing.assert_array_equal(a.download(f), b.download(f), err_msg=f)


def test_ensemble_is_bitwise_the_members_alone():
    cs = members()
    ens_s, ref_s = [], []
    for c in cs:
        for lst in (ens_s, ref_s):
            s = host.Solver(c.cfg)
            c.load(s)
            lst.append(s)
    ens = host.Ensemble(ens_s)
    plan = [(1, True), (13, False), (6, False), (2, False), (6, False)]
    for n, first in plan:
        ens.step(n, first=first)
        ens.sync()
        for s in ref_s:
            s.step(1, first=True) if first else s.step(n)
        for a, b in zip(ens_s, ref_s):
            same(a, b)
    ens.close()
    for s in ens_s + ref_s:
        s.close()


def test_ensemble_rejects_a_member_twice():
    c = cases.build("bubble", 12, 10, 12, cphopt=0)
    s = host.Solver(c.cfg)
    with pytest.raises(host.CressError):
        host.Ensemble([s, s])
    s.close()


def uneven_members():
    """Eight members of different sizes, sides, substep counts and physics."""
    return [
        cases.build("bubble", 12, 10, 16, cphopt=0),
        cases.build("bubble", 14, 12, 20, cphopt=2),                        # nsstp 5
        cases.build("supercell", 16, 14, 16, cphopt=0, wbc=7, gwave=30.0),
        cases.build("supercell", 12, 16, 18, cphopt=2, wbc=7, gwave=30.0),
        cases.build("vortex", 20, 18, 16, cphopt=0),
        cases.build("vortex", 13, 11, 14, cphopt=2),
        cases.build("bubble", 18, 14, 12, cphopt=0, advopt=2),
        cases.build("supercell", 14, 14, 22, cphopt=0),
    ]


def test_uneven_members_match_the_oracle():
    """Eight uneven members stepped together (a cold-start step, then one
    graph cycle of six steps) against the oracle run on each member alone:
    dry members bitwise equal to the FP32 oracle over the interior; members
    with microphysics within the large step's microphysics bar (err vs the
    FP64 oracle <= max(1e-5, 2 x the FP32 oracle's own err))."""
    cs = uneven_members()
    sols = []
    for c in cs:
        s = host.Solver(c.cfg)
        c.load(s)
        sols.append(s)
    ens = host.Ensemble(sols)
    ens.step(1, first=True)
    ens.step(6)
    ens.sync()
    nsteps = 7
    for c, s in zip(cs, sols):
        cfg = c.cfg
        cut = lambda a: a[2:cfg.ni - 1, 2:cfg.nj - 1, 1:cfg.nk - 2]
        m32 = po.Model(c, 32)
        m32.step(nsteps, first=True)
        if cfg.cphopt == 0:
            for f in ("u", "v", "w", "pp", "ptp", "qv"):
                np.testing.assert_array_equal(cut(s.download(f, 1)), cut(m32.get(f, 1)),
                                              err_msg="%s %dx%dx%d %s" % (c.name, cfg.ni, cfg.nj, cfg.nk, f))
        else:
            m64 = po.Model(c, 64)
            m64.step(nsteps, first=True)
            for f in FIELDS:
                x64 = cut(m64.get(f, 1)).astype(np.float64)
                sc = max(np.abs(x64).max(), 1e-30)
                eg = np.abs(cut(s.download(f, 1)) - x64).max() / sc
                eo = np.abs(cut(m32.get(f, 1)) - x64).max() / sc
                assert eg <= max(1e-5, 2.0 * eo), (c.name, f, eg, eo)
    ens.close()
    for s in sols:
        s.close()
