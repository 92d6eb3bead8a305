"""The Fortran drop-in modules run inside the reference's own solver.

oracle/_ref/libcressrefdrop32.so is the reference's s_intgdrv and everything
it calls (SrcOrig built with flang and MPICH, one rank), with the drop-in
modules of cress_amd/fortran/dropin/ in place of the reference's m_gaussel,
m_advs, m_advuvw and m_advp; their tier-2 C calls (cress_gaussel,
cress_advs, cress_advuvw, cress_advp) are answered on the CPU by the oracle
(oracle/dropin_cpu.c), whose routines are pinned bitwise against the
reference's.  Whole large steps must then stay bitwise those of the
unmodified reference: this checks everything the drop-ins pass across the
boundary -- the argument order, the options they resolve from the
reference's name tables (getiname/getrname), the grid spacings and the
array extents -- as the reference's own callers (stepwi, forcept, forceqv,
forceq, forceuvw, forcep) use them.  CPU only; the GPU's side of the same
entry points is checked against the oracle in test_dropin.py.
"""
import os

import numpy as np
import pytest

from cress_amd import cases
from oracle import refmpi

pytestmark = pytest.mark.skipif(not (refmpi.have() and os.path.exists(refmpi.LIB_DROPIN)),
                                reason="oracle/_ref drop-in part not built")


@pytest.mark.parametrize("name,kw", [
    ("bubble", dict(cphopt=2)),
    ("bubble", dict(cphopt=0, advopt=2)),
    ("vortex", dict(cphopt=2, advopt=1)),
])
def test_dropins_leave_the_reference_step_unchanged(name, kw):
    case = cases.build(name, 12, 10, 14, **kw)
    ref = refmpi.intgdrv(case, 3)
    drop = refmpi.intgdrv(case, 3, lib=refmpi.LIB_DROPIN)
    sc = case.cfg
    for s, (a, b) in enumerate(zip(ref, drop)):
        for n, f in enumerate(refmpi.LEV):
            la, lb = a[0][..., n], b[0][..., n]
            if sc.advopt == 1:
                # s_shift2nd is skipped for advopt=1 (shift2nd.f90:266-268):
                # the second halo ring keeps what the reference's scratch
                # reuse left there, which the drop-ins (writing no scratch)
                # do not mimic; no 2nd-order stencil reads it
                ie = sc.ni + 1 if f == "u" else sc.ni
                je = sc.nj + 1 if f == "v" else sc.nj
                la, lb = la[1:ie, 1:je], lb[1:ie, 1:je]
            np.testing.assert_array_equal(lb, la, err_msg="step %d %s" % (s + 1, f))
        np.testing.assert_array_equal(b[1], a[1], err_msg="step %d prwtr" % (s + 1))
        np.testing.assert_array_equal(b[2], a[2], err_msg="step %d price" % (s + 1))
    # the drop-ins did run: the prognostics moved
    assert np.abs(ref[-1][0][..., 0] - case.init["u"]).max() > 0.0
