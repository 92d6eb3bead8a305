"""The ensemble entry points (cress_ens_*): members of uneven sizes, periodic
and open, with even and odd substep counts, stepped through captured
hipGraph cycles give bitwise the results of stepping each member alone with
cress_large_step -- checked after every call, with graph replays, eager
steps that do not fill a cycle, downloads and recaptures in between."""
import numpy as np
import pytest

from cress_amd import cases, host

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
