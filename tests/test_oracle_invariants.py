"""Invariants of the whole large step that pin the oracle's update and driver
routines the reference cannot be built for (stepuv, stepwi, steppi, steps,
the drivers and fallblk's sub-stepping reach MPI in the reference, see
DESIGN): a wrong index, a swapped x/y operand or a flipped sign in them
breaks one of these.

* mirror symmetry in x and in y: a state symmetric about the domain centre
  (scalars s(i) = s(nx+3-i), staggered u(i) = -u(nx+4-i)) stays symmetric;
* x/y transposition: on a square grid, s(i,j) = s(j,i) and u(i,j) = v(j,i)
  are kept;
* periodic translation: shifting the initial state by m cells shifts the
  result by m cells, bitwise;
* boundedness and positivity over twenty steps of the moist bubble.
All with cold-rain microphysics on (cphopt=2) and the moist seeding.
"""
import numpy as np
import pytest

from cress_amd import cases
from oracle import pyoracle as po

SCALARS = ["w", "pp", "ptp", "qv", "qc", "qr", "qi", "qs", "qg", "nci"]


def run(case, prec, nsteps):
    m = po.Model(case, prec)
    m.step(nsteps, first=True)
    return m


@pytest.mark.parametrize("smtopt", [0, 3])
def test_mirror_symmetry_x_and_y(smtopt):
    nx, ny = 12, 10
    case = cases.build("bubble", nx, ny, 16, cphopt=2, noise=0.0, smtopt=smtopt)
    m = run(case, 64, 6)
    ix = np.arange(2, nx + 2)          # interior i (Fortran index = array index)
    jy = np.arange(2, ny + 2)
    for f in SCALARS:
        a = m.get(f, 1)[:, :, 1:-2]
        sc = max(np.abs(a).max(), 1e-30)
        dx = np.abs(a[ix][:, jy] - a[nx + 3 - ix][:, jy]).max() / sc
        dy = np.abs(a[:, jy][ix] - a[:, ny + 3 - jy][ix]).max() / sc
        assert dx <= 1e-12 and dy <= 1e-12, (f, dx, dy)
    u, v = m.get("u", 1)[:, :, 1:-2], m.get("v", 1)[:, :, 1:-2]
    iu = np.arange(2, nx + 3)
    jv = np.arange(2, ny + 3)
    su, sv = np.abs(u).max(), np.abs(v).max()
    assert np.abs(u[iu][:, jy] + u[nx + 4 - iu][:, jy]).max() <= 1e-12 * su
    assert np.abs(u[ix][:, jy] - u[ix][:, ny + 3 - jy]).max() <= 1e-12 * su
    assert np.abs(v[ix][:, jv] + v[ix][:, ny + 4 - jv]).max() <= 1e-12 * sv
    assert np.abs(v[ix][:, jy] - v[nx + 3 - ix][:, jy]).max() <= 1e-12 * sv
    # the flow is not trivially zero
    assert su > 1e-3 and np.abs(m.get("w", 1)).max() > 1e-2


def test_transpose_symmetry():
    n = 12
    case = cases.build("bubble", n, n, 16, cphopt=2, noise=0.0)
    m = run(case, 64, 6)
    r = np.arange(2, n + 2)
    for f in SCALARS:
        a = m.get(f, 1)[:, :, 1:-2][r][:, r]
        assert np.abs(a - a.transpose(1, 0, 2)).max() <= 1e-12 * max(np.abs(a).max(), 1e-30), f
    u, v = m.get("u", 1)[:, :, 1:-2], m.get("v", 1)[:, :, 1:-2]
    ru = np.arange(2, n + 3)
    uu = u[ru][:, r]
    vv = v[r][:, ru]
    assert np.abs(uu - vv.transpose(1, 0, 2)).max() <= 1e-12 * np.abs(uu).max()


def _roll_x(a, m, ni):
    """Shift the periodic interior i=2..ni-2 by m cells and re-wrap the halos
    (i = 0, 1, ni-1, ni, ni+1) from it."""
    n = ni - 3
    out = a.copy(order="F")
    interior = np.roll(a[2:ni - 1], m, axis=0)
    full = np.arange(ni + 2)
    src = (full - 2) % n
    out[:] = interior[src]
    return np.asfortranarray(out)


def test_periodic_translation_is_exact():
    case = cases.build("bubble", 12, 10, 12, cphopt=2)
    ni = case.cfg.ni
    m1 = run(case, 32, 3)
    shifted = cases.build("bubble", 12, 10, 12, cphopt=2)
    shifted.init = {k: _roll_x(v, 3, ni) for k, v in shifted.init.items()}
    m2 = run(shifted, 32, 3)
    for f in ["u", "v"] + SCALARS:
        a = _roll_x(m1.get(f, 1), 3, ni)
        np.testing.assert_array_equal(a[2:ni - 1, 2:-3, 1:-2], m2.get(f, 1)[2:ni - 1, 2:-3, 1:-2], err_msg=f)


def test_bounded_and_positive_over_twenty_steps():
    case = cases.build("bubble", 12, 10, 24, cphopt=2)
    m = run(case, 32, 20)
    for f in ("u", "v", "w"):
        assert np.isfinite(m.get(f, 1)).all() and np.abs(m.get(f, 1)).max() < 30.0, f
    for f in ("qv", "qc", "qr", "qi", "qs", "qg", "nci"):
        assert (m.get(f, 1) >= 0.0).all(), f
    assert np.abs(m.get("ptp", 1)).max() < 10.0


# ---- open lateral boundaries (wbc=ebc=sbc=nbc=7) ----
# The radiative fills are mirror images of each other side to side (west
# uses c<0 and s(2)-s(1), east c>0 and s(ni-2)-s(ni-1)), so the symmetries
# above hold on an open domain too, now including the boundary rows
# i=1, ni-1 (scalars) and i=1, ni (u), which the fills prognose.

def test_open_mirror_symmetry_x_and_y():
    nx, ny = 12, 10
    case = cases.build("bubble", nx, ny, 16, cphopt=2, noise=0.0, wbc=7, gwave=30.0)
    m = run(case, 64, 6)
    ni, nj = nx + 3, ny + 3
    ix = np.arange(1, ni)              # i=1..ni-1 incl. the open boundary rows
    jy = np.arange(1, nj)
    for f in SCALARS:
        a = m.get(f, 1)[:, :, 1:-2]
        sc = max(np.abs(a).max(), 1e-30)
        dx = np.abs(a[ix][:, jy] - a[ni - ix][:, jy]).max() / sc
        dy = np.abs(a[:, jy][ix] - a[:, nj - jy][ix]).max() / sc
        assert dx <= 1e-12 and dy <= 1e-12, (f, dx, dy)
    u, v = m.get("u", 1)[:, :, 1:-2], m.get("v", 1)[:, :, 1:-2]
    iu = np.arange(1, ni + 1)
    jv = np.arange(1, nj + 1)
    su, sv = np.abs(u).max(), np.abs(v).max()
    assert np.abs(u[iu][:, jy] + u[ni + 1 - iu][:, jy]).max() <= 1e-12 * su
    assert np.abs(u[iu][:, jy] - u[iu][:, nj - jy]).max() <= 1e-12 * su
    assert np.abs(v[ix][:, jv] + v[ix][:, nj + 1 - jv]).max() <= 1e-12 * sv
    assert np.abs(v[ix][:, jv] - v[ni - ix][:, jv]).max() <= 1e-12 * sv
    # the boundary rows moved: the fills are not identity maps
    assert np.abs(m.get("pp", 1)[1, 2:nj - 1, 2:-3]).max() > 0.0
    assert su > 1e-3 and np.abs(m.get("w", 1)).max() > 1e-2


def test_open_transpose_symmetry():
    n = 12
    case = cases.build("bubble", n, n, 16, cphopt=2, noise=0.0, wbc=7, gwave=30.0)
    m = run(case, 64, 6)
    r = np.arange(1, n + 3)            # i, j = 1..ni-1 incl. the boundary rows
    for f in SCALARS:
        a = m.get(f, 1)[:, :, 1:-2][r][:, r]
        assert np.abs(a - a.transpose(1, 0, 2)).max() <= 1e-12 * max(np.abs(a).max(), 1e-30), f
    u, v = m.get("u", 1)[:, :, 1:-2], m.get("v", 1)[:, :, 1:-2]
    ru = np.arange(1, n + 4)
    uu = u[ru][:, r]
    vv = v[r][:, ru]
    assert np.abs(uu - vv.transpose(1, 0, 2)).max() <= 1e-12 * np.abs(uu).max()


def test_open_phase_speeds_follow_the_step():
    """phasev runs at the end of each step with that step's dtb, dts
    (intgdrv.f90:1023-1055); the first step runs with zero speeds."""
    case = cases.build("bubble", 12, 10, 12, cphopt=2, wbc=7, gwave=30.0)
    m = po.Model(case, 32)
    assert not m.cp.any()
    m.step(1, first=True)
    c = case.cfg
    ni, nj, nk = c.ni, c.nj, c.nk
    ucpx = m.cp[:2 * nj * nk].reshape((nj, nk, 2), order="F")
    dtb, dts = np.float32(0.5 * c.dtbig), np.float32(0.5 * c.dtsml)
    g = np.float32(c.gwave) * dtb * (np.float32(1.0) / np.float32(c.dx))
    assert ucpx[0, 1, 0] == max(-g, np.float32(-1.0)) * (dts / dtb)
    m.step(1)
    dtb, dts = np.float32(c.dtbig), np.float32(c.dtsml)
    g = np.float32(c.gwave) * dtb * (np.float32(1.0) / np.float32(c.dx))
    assert ucpx[0, 1, 1] == min(g, np.float32(1.0)) * (dts / dtb)


def test_open_bounded_and_positive_over_twenty_steps():
    case = cases.build("bubble", 12, 10, 24, cphopt=2, wbc=7, gwave=30.0)
    m = run(case, 32, 20)
    for f in ("u", "v", "w"):
        assert np.isfinite(m.get(f, 1)).all() and np.abs(m.get(f, 1)).max() < 30.0, f
    for f in ("qv", "qc", "qr", "qi", "qs", "qg"):
        assert (m.get(f, 1)[1:-1, 1:-1, :-1] >= 0.0).all(), f
    assert np.abs(m.get("ptp", 1)).max() < 10.0
