"""Launch shapes and the autotuner (cress_autotune, cress_set_launch): the
results of the large step do not depend on the tile shapes, halo layout,
tile order or occupancy cap of any kernel family -- the same bits after
several steps with the default shapes, with every shape set to another
candidate, and with the autotuner's choice -- and the autotuner returns
valid shapes, reuses them for a grid it has seen, and may run between
steps without changing the model's state."""
import numpy as np
import pytest

from cress_amd import cases, host

pytestmark = pytest.mark.gpu

FIELDS = ["u", "v", "w", "pp", "ptp", "qv", "qc", "qr", "qi", "qs", "qg", "nci"]
OTHER = {"ptuv_by": 4, "ptuv_slots": 1, "ptuv_order": 2, "col_by": 1, "col_lds": 0, "fs_by": 8, "fs_order": 1,
         "ff_by": 2, "ff_order": 1}


def _run(case, nsteps, shapes=None, tune=False, tune_between=False):
    s = host.Solver(case.cfg)
    case.load(s)
    if shapes:
        s.set_launch(shapes)
    chosen = s.autotune() if tune else None
    s.step(1, first=True)
    if tune_between:
        chosen = s.autotune()
    s.step(nsteps - 1)
    s.sync()
    out = {f: s.download(f, 1) for f in FIELDS}
    out["prwtr"] = s.download("prwtr")
    s.close()
    return out, chosen


@pytest.mark.parametrize("name,dims", [("bubble", (24, 22, 30)), ("supercell", (20, 18, 24))])
def test_results_do_not_depend_on_launch_shapes(name, dims):
    case = cases.build(name, *dims, cphopt=2)
    ref, _ = _run(case, 3)
    other, _ = _run(case, 3, shapes=OTHER)
    tuned, chosen = _run(case, 3, tune=True)
    between, _ = _run(case, 3, tune_between=True)
    print(name, "autotuned shapes", chosen)
    for k in ref:
        np.testing.assert_array_equal(other[k], ref[k], err_msg=k)
        np.testing.assert_array_equal(tuned[k], ref[k], err_msg=k)
        np.testing.assert_array_equal(between[k], ref[k], err_msg=k)
    assert chosen["ptuv_by"] in (4, 8, 16) and chosen["col_by"] in (1, 2, 4)


def test_set_launch_rejects_bad_shapes():
    case = cases.build("bubble", 12, 10, 12, cphopt=0)
    s = host.Solver(case.cfg)
    bad = dict(OTHER, ptuv_by=3)
    with pytest.raises(host.CressError):
        s.set_launch(bad)
    assert s.get_launch()["ptuv_by"] == 16
    s.close()
