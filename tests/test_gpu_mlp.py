"""The learned microphysics on the GPU: the matrix-core kernels against the
FP64 oracle (oracle/mlp.py, itself pinned by finite differences), training,
the recorded pairs against the conservation laws srcblk obeys, and a large
step with the network against the plain scheme."""
import numpy as np
import pytest

from cress_amd import cases, host, learned
from oracle import mlp as om

pytestmark = pytest.mark.gpu


def _raw_inputs(rng, n):
    """Plausible raw features: T, p, rho, qv, qc, qr, qi, qs, qg."""
    x = np.empty((n, 9))
    x[:, 0] = rng.uniform(220.0, 305.0, n)
    x[:, 1] = rng.uniform(2.0e4, 1.0e5, n)
    x[:, 2] = rng.uniform(0.3, 1.2, n)
    x[:, 3] = rng.uniform(0.0, 2.0e-2, n)
    x[:, 4:] = rng.uniform(0.0, 2.0e-3, (n, 5))
    return x.astype(np.float32)


def _norm(rng):
    nrm = np.zeros(om.NNORM, np.float32)
    nrm[:9] = [260.0, 6.0e4, 0.75, 1.0e-2, 1.0e-3, 1.0e-3, 1.0e-3, 1.0e-3, 1.0e-3]
    nrm[16:25] = 1.0 / np.array([25.0, 2.3e4, 0.26, 5.8e-3, 5.8e-4, 5.8e-4, 5.8e-4, 5.8e-4, 5.8e-4])
    nrm[32:39] = rng.normal(0.0, 1e-4, 7)
    nrm[48:64] = 1.0
    nrm[48:55] = [1e-3, 1e-6, 1e-6, 1e-6, 1e-7, 1e-7, 1e-7]
    nrm[64:73] = -np.inf
    nrm[80:89] = np.inf
    return nrm


def test_forward_kernel_matches_oracle():
    """The fused inference kernel equals the oracle's forward pass with the
    same BF16 roundings to FP32 summation error; a transposed fragment or a
    wrong bias row would be off by O(1)."""
    rng = np.random.default_rng(3)
    m = host.Mlp(capacity=1024, max_batch=4096)
    m.init(11)
    p, _ = m.get()
    p = p.astype(np.float64)
    p[om.OB1:om.OW2] = rng.normal(0, 0.1, om.H)
    p[om.OB4:om.OB4 + 7] = rng.normal(0, 0.1, 7)
    nrm = _norm(rng)
    m.set(p.astype(np.float32), nrm)
    raw = _raw_inputs(rng, 1000)                 # not a multiple of 16
    got = m.forward(raw).astype(np.float64)
    x = om.normalise(raw, nrm.astype(np.float64))
    y = om.forward_bf16(p.astype(np.float32).astype(np.float64), x)[:, :7]
    ref = y * nrm[48:55] + nrm[32:39]
    yn_got = (got - nrm[32:39]) / nrm[48:55]
    err = np.abs(yn_got - y).max() / np.abs(y).max()
    print("forward: max |y| %.3g, rel err %.2e" % (np.abs(y).max(), err))
    assert err <= 2e-2, err
    # and against the unrounded FP64 network: the BF16 roundings' size
    y64 = om.forward(p.astype(np.float32).astype(np.float64), x)[0][:, :7]
    assert np.abs(yn_got - y64).max() / np.abs(y64).max() <= 5e-2
    m.close()


def test_training_gradient_matches_oracle():
    """The training kernels' gradient (BF16 matrix cores, FP32 sums) against
    the FP64 oracle's on the BF16-rounded weights and inputs: within 3% of
    each block's norm."""
    rng = np.random.default_rng(4)
    m = host.Mlp(capacity=1024, max_batch=4096)
    m.init(12)
    p, _ = m.get()
    nrm = _norm(rng)
    m.set(p, nrm)
    n = 2048
    raw = _raw_inputs(rng, n)
    tgt = (rng.normal(0, 1, (n, 7)) * nrm[48:55] + nrm[32:39]).astype(np.float32)
    g, loss = m.grad(raw, tgt)
    x = om.bf16(om.normalise(raw, nrm.astype(np.float64)))
    t = (tgt.astype(np.float64) - nrm[32:39]) / nrm[48:55]
    pw = p.astype(np.float64).copy()
    for lo, hi in ((om.OW1, om.OB1), (om.OW2, om.OB2), (om.OW3, om.OB3), (om.OW4, om.OB4)):
        pw[lo:hi] = om.bf16(pw[lo:hi])
    lref, gref = om.loss_grad(pw, x, t)
    assert abs(loss - lref) <= 2e-2 * lref, (loss, lref)
    blocks = {"W1": (om.OW1, om.OB1), "b1": (om.OB1, om.OW2), "W2": (om.OW2, om.OB2), "b2": (om.OB2, om.OW3),
              "W3": (om.OW3, om.OB3), "b3": (om.OB3, om.OW4), "W4": (om.OW4, om.OB4), "b4": (om.OB4, om.NPARAM)}
    for name, (lo, hi) in blocks.items():
        e = np.linalg.norm(g[lo:hi] - gref[lo:hi]) / np.linalg.norm(gref[lo:hi])
        print("grad %s rel err %.2e" % (name, e))
        assert e <= 3e-2, (name, e)
    m.close()


def test_training_fits_a_smooth_map():
    """Adam on the GPU drives the loss of a smooth synthetic map down by two
    orders of magnitude, and the trained forward kernel reproduces it."""
    rng = np.random.default_rng(5)
    n = 1 << 16
    raw = _raw_inputs(rng, n)
    xs = (raw - raw.mean(0)) / raw.std(0)
    tgt = np.stack([np.sin(xs[:, 0]) + 0.5 * xs[:, 1] * xs[:, 2], np.tanh(xs[:, 3] - xs[:, 4]),
                    xs[:, 5] ** 2 - 1.0, np.cos(xs[:, 6] + xs[:, 7]), xs[:, 8] * xs[:, 0],
                    np.abs(xs[:, 2]) - 0.8, 0.3 * xs[:, 1]], 1).astype(np.float32)
    m = host.Mlp(capacity=n, max_batch=8192)
    m.init(13)
    m.put_samples(raw, tgt)
    m.fit_norm(0.05)
    l0 = m.grad(raw[:4096], tgt[:4096])[1]
    losses = learned.train(m, 1500, batch=4096, lr=3e-3)
    l1 = m.grad(raw[:4096], tgt[:4096])[1]
    print("loss %.3e -> %.3e (last batches %.3e)" % (l0, l1, losses[-20:].mean()))
    assert l1 < 2e-2 * l0
    pred = m.forward(raw[:4096])
    assert np.sqrt(np.mean((pred - tgt[:4096]) ** 2)) < 0.15 * np.sqrt(np.mean(tgt[:4096] ** 2))
    m.close()


@pytest.fixture(scope="module")
def bubble_pairs():
    case = cases.build("bubble", 24, 22, 30, cphopt=2)
    m = host.Mlp(capacity=1 << 21, max_batch=16384)
    n = learned.record(m, case, 8)
    return case, m, n


def test_recorded_pairs_obey_srcblk_conservation(bubble_pairs):
    """Every recorded target conserves total water (srcblk only moves water
    between species; the max(0, .) clamps of newblk aside) and carries the
    latent heating of its phase changes, cp dT = Lv (-dqv) + Lf dq_ice
    (newblk.f90:416-417): the pairs are what srcblk did, in the right order
    and units."""
    case, m, n = bubble_pairs
    assert n > 1000
    inp, tgt = m.get_samples(0, min(n, 200000))
    inp, tgt = inp.astype(np.float64), tgt.astype(np.float64)
    dq = tgt[:, 1:]
    # each target is a difference of two binary32 values of the future level
    # over the microphysics step (2 dtb >= dtbig): its rounding is at most
    # ulp(q)/dt per species, so the sum may miss zero by the species'
    # rounding and no more where newblk's max(0, .) clamps did not act
    dt = case.cfg.dtbig
    q = inp[:, 3:] + np.abs(dq) * 2.0 * dt
    floor = 4.0 * 2.0 ** -24 * q.sum(1) / dt
    resid = np.abs(dq.sum(1))
    ok = resid <= floor
    print("water: %.4f of pairs within the rounding floor, median residual/floor %.2g"
          % (ok.mean(), np.median(resid / floor)))
    assert ok.mean() > 0.95
    t = inp[:, 0]
    lv = 2.50078e6 * np.exp((0.167 + 3.67e-4 * t) * np.log(273.16 / t))
    lf = 3.34e5 + (4.218e3 - 2.106e3) * (t - 273.16)
    heat = (lv * (-dq[:, 0]) + lf * dq[:, 3:].sum(1)) / 1004.0
    big = np.abs(tgt[:, 0]) > 1e-4 * np.abs(tgt[:, 0]).max()
    rel = np.abs(heat[big] - tgt[big, 0]) / np.abs(tgt[big, 0])
    print("pairs %d, water resid median %.1e, heating rel err median %.1e" % (n, np.median(resid), np.median(rel)))
    assert np.median(rel) < 1e-2


def test_network_step_against_plain(bubble_pairs):
    """The large step with the trained network: finite, non-negative mixing
    ratios, total water within 2% of the plain run's, and the trained
    network closer to the plain scheme (accumulated precipitation and q
    RMSE) than the same network untrained."""
    case, m, n = bubble_pairs
    m.fit_norm(0.05)
    nsteps = 12
    s_plain = learned.run(case, nsteps)
    m.init(21)
    s_un = learned.run(case, nsteps, m, 2)
    sc_un = learned.score(s_un, s_plain)
    learned.train(m, 1500, batch=8192, lr=2e-3)
    s_net = learned.run(case, nsteps, m, 2)
    sc = learned.score(s_net, s_plain)
    cfg = case.cfg
    rst = learned.interior(case.base["rst"], cfg).astype(np.float64)
    for s in (s_net,):
        for q in learned.QS:
            a = s.download(q, 1)
            assert np.isfinite(a).all() and (learned.interior(a, cfg) >= 0).all(), q
    tw = lambda s: sum((rst * learned.interior(s.download(q, 1), cfg)).sum() for q in learned.QS)
    assert abs(tw(s_net) - tw(s_plain)) <= 0.02 * tw(s_plain)
    print("untrained", {k: "%.3g" % v for k, v in sc_un.items()})
    print("trained  ", {k: "%.3g" % v for k, v in sc.items()})
    err = lambda d: sum(d["rmse_" + q] / max(d["scale_" + q], 1e-30) for q in ("qc", "qr", "qi", "qs", "qg"))
    assert err(sc) < err(sc_un)
    for s in (s_plain, s_un, s_net):
        s.close()
