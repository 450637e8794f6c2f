"""The oracle's BH (oracle/bh.py) against the reference's own BH fixtures and what the
mathematics fixes: two epochs on the seeded far-field scenario (two probe modes, after
scale_obj_probe) with one batch of all 40 frames (conjugate gradient across the epochs) and
with batches of 16 (steepest descent), every step's β, α, ⟨g, D⟩ and H(D, D) to 1e-13 and the
object and probe changes of each epoch (bh_{full,mini}_*, written by the reference's
refs/f64/bh_step.py; ptyrs/tests/engine_bh.rs:164-243); the gradient against central
differences of the cost and the Hessian against central differences of the gradient, and its
symmetry (ptyrs/tests/engine_bh.rs:95-162)."""

import numpy as np
import pytest

from conftest import max_rel, table
from oracle import scenario as sc
from oracle.bh import Bh, BhBatch, dot
from oracle.ptycho import llk, obj_probe_to_psi


def state():
    p, _ = sc.parity_start(nprobe=2)
    return p


def run(batch_size, nepoch=2):
    p = state()
    bh = Bh(batch_size, probe_rho=0.01, seed=7)
    obj, probe = p.obj, p.probe
    steps, changes = [], []
    for _ in range(nepoch):
        o, pr, s = bh.epoch(p.addr, p.iobs, p.optics, obj, probe)
        steps.append(s)
        changes.append((o - obj, pr - probe))
        obj, probe = o, pr
    return steps, changes


@pytest.mark.parametrize("batch_size,name", [(40, "bh_full"), (16, "bh_mini")])
def test_epochs_match_fixtures(batch_size, name):
    steps, changes = run(batch_size)
    rows = table(f"{name}_steps_f64.txt")
    got = [(e, s) for e, ss in enumerate(steps) for s in ss]
    assert len(got) == len(rows)
    for k, ((epoch, (frames, beta, alpha, top, bottom)), r) in enumerate(zip(got, rows)):
        assert [k, epoch, len(frames)] == [int(r[0]), int(r[1]), int(r[2])]
        assert frames == [int(f) for f in r[7:]]
        if r[3] == 0.0:
            assert beta == 0.0
        else:
            assert abs(beta - r[3]) < 1e-13 * abs(r[3]), (k, beta, r[3])
        for got_v, want in ((alpha, r[4]), (top, r[5]), (bottom, r[6])):
            assert abs(got_v - want) < 1e-13 * abs(want), (k, got_v, want)
    obj_rows = np.array(table(f"{name}_obj_f64.txt"))
    probe_rows = np.array(table(f"{name}_probe_f64.txt"))
    for e, (do, dp) in enumerate(changes):
        want = obj_rows[:, 2 + 2 * e] + 1j * obj_rows[:, 3 + 2 * e]
        got_o = do[0][obj_rows[:, 0].astype(int), obj_rows[:, 1].astype(int)]
        assert max_rel(np.r_[got_o.real, got_o.imag], np.r_[want.real, want.imag]) < 1e-13
        want = probe_rows[:, 3 + 2 * e] + 1j * probe_rows[:, 4 + 2 * e]
        idx = probe_rows[:, :3].astype(int)
        got_p = dp[idx[:, 0], idx[:, 1], idx[:, 2]]
        assert max_rel(np.r_[got_p.real, got_p.imag], np.r_[want.real, want.imag]) < 5e-12


def _cost(p, views, obj, probe):
    """F of the views without the amplitude offset: a quarter of the Euclidian LLK sum."""
    frames = sorted({v[0] for v in views})
    psi = p.optics.exit_forward(obj_probe_to_psi(None, views, obj, probe, p.optics.quad))
    return llk(np.stack([p.iobs[f] for f in frames]), p.addr.nmode, psi).euclidian / 4.0


def _direction(like, rng):
    scale = np.abs(like).max()
    return (rng.uniform(-0.5, 0.5, like.shape) + 1j * rng.uniform(-0.5, 0.5, like.shape)) * scale


def test_gradient_and_hessian_match_central_differences():
    p = state()
    views = p.addr.views()[:16 * p.addr.nmode]
    b = BhBatch(views, p.iobs, p.optics, p.obj, p.probe)
    g = b.grads()
    rng = np.random.default_rng(3)
    d_o, d_p = _direction(p.obj, rng), _direction(p.probe, rng)
    h = 1e-5
    fo = (_cost(p, views, p.obj + h * d_o, p.probe) -
          _cost(p, views, p.obj - h * d_o, p.probe)) / (2 * h)
    fp = (_cost(p, views, p.obj, p.probe + h * d_p) -
          _cost(p, views, p.obj, p.probe - h * d_p)) / (2 * h)
    # the gradient carries the 1e-8 amplitude offset, the differences do not
    assert abs(dot(g, (d_o, 0 * p.probe)) - fo) < 1e-7 * abs(fo)
    assert abs(dot(g, (0 * p.obj, d_p)) - fp) < 1e-7 * abs(fp)
    u = (_direction(p.obj, rng), _direction(p.probe, rng))
    v = (_direction(p.obj, rng), _direction(p.probe, rng))
    h_uv, h_vu = b.hessian(u, v), b.hessian(v, u)
    assert abs(h_uv - h_vu) < 1e-13 * abs(h_uv)

    def du(t):
        bt = BhBatch(views, p.iobs, p.optics, p.obj + t * v[0], p.probe + t * v[1])
        return dot(bt.grads(), u)

    t = 1e-6
    fd = (du(t) - du(-t)) / (2 * t)
    assert abs(h_uv - fd) < 1e-7 * abs(fd)


def test_hessian_is_bilinear_and_the_step_minimises_the_quadratic_model():
    # H(u, v) is linear in each argument; along D the model F + t·⟨g, D⟩ + t²/2·H(D, D) is least
    # at t = −⟨g, D⟩/H(D, D), the step BH takes with one batch of every frame.
    p = state()
    views = p.addr.views()
    b = BhBatch(views, p.iobs, p.optics, p.obj, p.probe)
    rng = np.random.default_rng(4)
    u, v, w = [(_direction(p.obj, rng), _direction(p.probe, rng)) for _ in range(3)]
    lin = b.hessian((u[0] + 2 * w[0], u[1] + 2 * w[1]), v)
    assert abs(lin - b.hessian(u, v) - 2 * b.hessian(w, v)) < 1e-12 * abs(lin)
    g = b.grads()
    d = (-g[0], -g[1])
    alpha = -dot(g, d) / b.hessian(d, d)
    model = [t * dot(g, d) + 0.5 * t * t * b.hessian(d, d) for t in (0.9 * alpha, alpha,
                                                                      1.1 * alpha)]
    assert model[1] < model[0] and model[1] < model[2]
    # and the cost falls along D over the step (the amplitude cost is near-quadratic there)
    f0 = _cost(p, views, p.obj, p.probe)
    assert _cost(p, views, p.obj + alpha * d[0], p.probe + alpha * d[1]) < f0


def test_batch_size_is_checked_and_masked_pixels_are_not_read():
    with pytest.raises(ValueError):
        Bh(0)
    p = state()
    iobs = np.array(p.iobs)
    iobs[3, 5, 7] = np.nan
    b = BhBatch(p.addr.views(), iobs, p.optics, p.obj, p.probe)
    masked = np.isnan(b.root)
    assert masked.sum() == 1 + np.isnan(p.iobs).sum()
    blk = b.td.reshape(b.root.shape[:1] + (p.addr.nmode,) + b.root.shape[1:])
    assert np.all(blk.transpose(1, 0, 2, 3)[:, masked] == 0)


@pytest.mark.parametrize("case", range(4))
@pytest.mark.parametrize("method,batch_size", [(0, 16), (1, 6)])
def test_near_field_epochs_match_fixtures(case, method, batch_size):
    # pty-chi's planar near field on the near-field LSQML state (24² frames, two probe modes,
    # 16 frames; angular spectrum, Fresnel, angular spectrum with an evanescent band, Fresnel
    # back), probe ρ 1, two epochs: steps to 1e-13 (β 1e-11), changes to 5e-12
    # (bh_near_*; ptyrs/tests/engine_bh_near.rs:103-176).
    from test_oracle_lsqml import state as near_state
    addr, obj0, probe0, cases = near_state()
    optics, iobs, _, _ = cases[case]
    bh = Bh(batch_size, probe_rho=1.0, seed=7)
    obj, probe, got = obj0, probe0, []
    for e in range(2):
        obj, probe, s = bh.epoch(addr, iobs, optics, obj, probe)
        got += [(e, x) for x in s]
    rows = [r for r in table("bh_near_steps_f64.txt") if r[0] == case and r[1] == method]
    assert len(got) == len(rows)
    for k, ((epoch, (frames, beta, alpha, top, bottom)), r) in enumerate(zip(got, rows)):
        assert [k, epoch, len(frames)] == [int(r[2]), int(r[3]), int(r[4])]
        assert frames == [int(f) for f in r[9:]]
        if r[5] == 0.0:
            assert beta == 0.0
        else:
            assert abs(beta - r[5]) < 1e-11 * abs(r[5])
        for got_v, want in ((alpha, r[6]), (top, r[7]), (bottom, r[8])):
            assert abs(got_v - want) < 1e-13 * abs(want), (k, got_v, want)
    for now, start, name, lead in ((obj, obj0, "bh_near_obj_f64.txt", 4),
                                   (probe, probe0, "bh_near_probe_f64.txt", 5)):
        rows = np.array([r for r in table(name) if r[0] == case and r[1] == method])
        d = (now - start).ravel()
        assert max_rel(np.stack([d.real, d.imag], axis=1).ravel(),
                       rows[:, lead:lead + 2].ravel()) < 5e-12


@pytest.mark.parametrize("batch_size,name,sampling", [
    (40, "bh_pos_full", "patch"), (16, "bh_pos_mini", "patch"),
    (40, "bh_full_object_full", "full_object"), (16, "bh_full_object_mini", "full_object")])
def test_offsets_match_fixtures(batch_size, name, sampling):
    # Refined offsets (probe ρ 0.01, position ρ 2, seed 7, two epochs from the fixture's
    # offsets) under PATCH (the reference's model of its steps, refs/f64/bh_pos.py) and under
    # FullObject (pty-chi's FULL_OBJECT BH, refs/py/ptychi_full_object.py): steps, the object
    # and probe changes over the run relative to the start's largest value, the offsets' changes
    # relative to their largest, all to 3e-12 (ptyrs/tests/engine_bh_pos.rs:61-184).
    p = state()
    rows = np.array(table(f"{name}_offsets_f64.txt"))
    d0 = rows[:, 1:3].copy()
    bh = Bh(batch_size, probe_rho=0.01, seed=7, position_rho=2.0, offsets=d0, refine=True,
            sampling=sampling)
    obj, probe, got = p.obj, p.probe, []
    for e in range(2):
        obj, probe, s = bh.epoch(p.addr, p.iobs, p.optics, obj, probe)
        got += [(e, x) for x in s]
    want = table(f"{name}_steps_f64.txt")
    assert len(got) == len(want)
    for k, ((epoch, (frames, beta, alpha, top, bottom)), r) in enumerate(zip(got, want)):
        assert [k, epoch, len(frames)] == [int(r[0]), int(r[1]), int(r[2])]
        assert frames == [int(f) for f in r[7:]]
        if r[3] == 0.0:
            assert beta == 0.0
        for got_v, w in ((beta, r[3]), (alpha, r[4]), (top, r[5]), (bottom, r[6])):
            if w != 0.0:
                assert abs(got_v - w) < 3e-12 * abs(w), (k, got_v, w)
    for now, start, fname, lead in ((obj, p.obj, f"{name}_obj_f64.txt", 2),
                                    (probe, p.probe, f"{name}_probe_f64.txt", 3)):
        r = np.array(table(fname))
        d = (now - start).ravel()
        got_c = np.stack([d.real, d.imag], axis=1).ravel()
        want_c = r[:, lead:lead + 2].ravel()
        assert np.abs(got_c - want_c).max() < 3e-12 * np.abs(start).max()
    dd = (bh.offsets - d0).ravel()
    want_d = (rows[:, 3:5] - d0).ravel()
    assert np.abs(dd - want_d).max() < 3e-12 * np.abs(want_d).max()


def test_patch_offsets_gradient_and_hessian_match_central_differences():
    # The offsets' gradient and the Hessian over object, probe and offsets against central
    # differences of the cost and of the gradient (ptyrs/src/engine/bh.rs:960-1010), held and
    # refined offsets; at zero offsets the batch is the unmoved one.
    p = state()
    nb = 8
    views = p.addr.views()[:nb * p.addr.nmode]
    rng = np.random.default_rng(11)
    d0 = rng.uniform(-0.4, 0.4, (nb, 2))

    def cost(obj, probe, d):
        b = BhBatch(views, p.iobs, p.optics, obj, probe, d, False)
        keep = ~np.isnan(b.root)
        a = b.amp - 1e-8
        return float((((a - b.root) ** 2)[keep]).sum() +
                     (2e-8 * b.root * np.log(b.amp))[keep].sum())

    b = BhBatch(views, p.iobs, p.optics, p.obj, p.probe, d0, True)
    g = b.grads()
    w = (_direction(p.obj, rng), _direction(p.probe, rng), rng.uniform(-1, 1, (nb, 2)))
    h = 1e-5
    fd = (cost(p.obj + h * w[0], p.probe + h * w[1], d0 + h * w[2]) -
          cost(p.obj - h * w[0], p.probe - h * w[1], d0 - h * w[2])) / (2 * h)
    assert abs(dot(g, w) - fd) < 1e-7 * abs(fd)
    u = (_direction(p.obj, rng), _direction(p.probe, rng), rng.uniform(-1, 1, (nb, 2)))
    v = (_direction(p.obj, rng), _direction(p.probe, rng), rng.uniform(-1, 1, (nb, 2)))
    h_uv, h_vu = b.hessian(u, v), b.hessian(v, u)
    assert abs(h_uv - h_vu) < 1e-12 * abs(h_uv)

    def du(t):
        bt = BhBatch(views, p.iobs, p.optics, p.obj + t * v[0], p.probe + t * v[1],
                     d0 + t * v[2], True)
        return dot(bt.grads(), u)

    t = 1e-6
    fd = (du(t) - du(-t)) / (2 * t)
    assert abs(h_uv - fd) < 1e-6 * abs(fd)
    # held offsets: no offsets' part, the same waves
    bh_held = BhBatch(views, p.iobs, p.optics, p.obj, p.probe, d0, False)
    assert bh_held.grads()[2] is None
    assert np.allclose(bh_held.psi, b.psi, rtol=0, atol=0)
    zero = BhBatch(views, p.iobs, p.optics, p.obj, p.probe, np.zeros((nb, 2)), False)
    plain = BhBatch(views, p.iobs, p.optics, p.obj, p.probe)
    assert np.abs(zero.psi - plain.psi).max() < 1e-12 * np.abs(plain.psi).max()
    gz, gp = zero.grads(), plain.grads()
    for x, y in zip(gz[:2], gp[:2]):
        assert np.abs(x - y).max() < 1e-12 * np.abs(y).max()


@pytest.mark.parametrize("case", range(3))
@pytest.mark.parametrize("method,batch_size", [(0, 16), (1, 6)])
def test_multislice_epochs_match_fixtures(case, method, batch_size):
    # Three slices with unequal gaps on the multislice LSQML state (angular-spectrum exit,
    # Fresnel, far field), probe ρ 1, three epochs: steps to 2e-13, every slice's and the
    # probe's change to 2e-12 (bh_multislice_*, the step rule by autograd on the reference's
    # forward model; ptyrs/tests/engine_bh_multislice.rs:114-185).
    from test_oracle_multislice import state as ms_state
    addr, objects, probe0, cases = ms_state()
    optics, iobs = cases[case][0], cases[case][1]
    obj0 = np.stack(objects)
    bh = Bh(batch_size, probe_rho=1.0, seed=7)
    obj, probe, got = obj0, probe0, []
    for e in range(3):
        obj, probe, s = bh.epoch(addr, iobs, optics, obj, probe)
        got += [(e, x) for x in s]
    rows = [r for r in table("bh_multislice_steps_f64.txt") if r[0] == case and r[1] == method]
    assert len(got) == len(rows)
    for k, ((epoch, (frames, beta, alpha, top, bottom)), r) in enumerate(zip(got, rows)):
        assert [k, epoch] == [int(r[2]), int(r[3])]
        assert frames == [int(f) for f in r[10:]]
        if r[5] == 0.0:
            assert beta == 0.0
        else:
            assert abs(beta - r[5]) < 2e-13 * abs(r[5]), (k, beta, r[5])
        for got_v, want in ((alpha, r[6]), (top, r[7]), (bottom, r[8])):
            assert abs(got_v - want) < 2e-13 * abs(want), (k, got_v, want)
    for now, start, name in ((obj, obj0, "bh_multislice_obj_f64.txt"),
                             (probe, probe0, "bh_multislice_probe_f64.txt")):
        r = np.array([t for t in table(name) if t[0] == case and t[1] == method])
        d = (now - start).ravel()
        assert max_rel(np.stack([d.real, d.imag], axis=1).ravel(), r[:, 5:7].ravel()) < 2e-12


def test_multislice_full_object_gradient_and_hessian_match_central_differences():
    # Three slices read through their sampled windows at held FullObject offsets: the slices'
    # and the probe's gradient against central differences of the cost, the Hessian against
    # central differences of the gradient and symmetric; at zero offsets the batch is the
    # plain multislice one (pinned by the bh_multislice_* fixtures).
    from oracle.bh import BhSlices
    from test_oracle_multislice import state as ms_state
    addr, objects, probe, cases = ms_state()
    optics, iobs = cases[0][0], cases[0][1]
    obj = np.stack(objects)
    rng = np.random.default_rng(21)
    obj = obj * np.exp(1j * rng.uniform(-0.3, 0.3, obj.shape))
    nb = 6
    views = addr.views()[:nb * addr.nmode]
    d0 = rng.uniform(-0.45, 0.45, (nb, 2))

    def cost(o, p):
        b = BhSlices(views, iobs, optics, o, p, d0)
        keep = ~np.isnan(b.root)
        a = b.amp - 1e-8
        return float((((a - b.root) ** 2)[keep]).sum() +
                     (2e-8 * b.root * np.log(b.amp))[keep].sum())

    b = BhSlices(views, iobs, optics, obj, probe, d0)
    g = b.grads()
    w = (_direction(obj, rng), _direction(probe, rng))
    h = 1e-5
    fd = (cost(obj + h * w[0], probe + h * w[1]) - cost(obj - h * w[0], probe - h * w[1])) / (2 * h)
    assert abs(dot(g, w) - fd) < 1e-7 * abs(fd)
    u = (_direction(obj, rng), _direction(probe, rng))
    v = (_direction(obj, rng), _direction(probe, rng))
    h_uv, h_vu = b.hessian(u, v), b.hessian(v, u)
    assert abs(h_uv - h_vu) < 1e-12 * abs(h_uv)

    def du(t):
        return dot(BhSlices(views, iobs, optics, obj + t * v[0], probe + t * v[1], d0).grads(), u)

    t = 1e-6
    fd = (du(t) - du(-t)) / (2 * t)
    assert abs(h_uv - fd) < 1e-6 * abs(fd)
    zero = BhSlices(views, iobs, optics, obj, probe, np.zeros((nb, 2)))
    plain = BhSlices(views, iobs, optics, obj, probe)
    assert np.abs(zero.psi - plain.psi).max() < 1e-12 * np.abs(plain.psi).max()
    for x, y in zip(zero.grads()[:2], plain.grads()[:2]):
        assert np.abs(x - y).max() < 1e-12 * np.abs(y).max()
    assert abs(zero.hessian(u, v) - plain.hessian(u, v)) < 1e-12 * abs(plain.hessian(u, v))
    # the offsets move what the slices read
    assert np.abs(b.psi - plain.psi).max() > 1e-3 * np.abs(plain.psi).max()


def test_multislice_refined_full_object_offsets_match_central_differences():
    # Refined FullObject offsets through three slices: the offsets' gradient (every slice's
    # reads' gradient against the reads' derivatives) and the Hessian over slices, probe and
    # offsets against central differences of the cost and of the gradient; held, the same waves.
    from oracle.bh import BhSlices
    from test_oracle_multislice import state as ms_state
    addr, objects, probe, cases = ms_state()
    optics, iobs = cases[0][0], cases[0][1]
    rng = np.random.default_rng(22)
    obj = np.stack(objects) * np.exp(1j * rng.uniform(-0.3, 0.3, np.stack(objects).shape))
    nb = 6
    views = addr.views()[:nb * addr.nmode]
    d0 = rng.uniform(-0.45, 0.45, (nb, 2))

    def cost(o, p, d):
        b = BhSlices(views, iobs, optics, o, p, d)
        keep = ~np.isnan(b.root)
        a = b.amp - 1e-8
        return float((((a - b.root) ** 2)[keep]).sum() +
                     (2e-8 * b.root * np.log(b.amp))[keep].sum())

    b = BhSlices(views, iobs, optics, obj, probe, d0, True)
    g = b.grads()
    w = (_direction(obj, rng), _direction(probe, rng), rng.uniform(-1, 1, (nb, 2)))
    h = 1e-5
    fd = (cost(obj + h * w[0], probe + h * w[1], d0 + h * w[2]) -
          cost(obj - h * w[0], probe - h * w[1], d0 - h * w[2])) / (2 * h)
    assert abs(dot(g, w) - fd) < 1e-7 * abs(fd)
    wd = (0 * w[0], 0 * w[1], w[2])   # the offsets' part alone
    fd = (cost(obj, probe, d0 + h * w[2]) - cost(obj, probe, d0 - h * w[2])) / (2 * h)
    assert abs(dot(g, wd) - fd) < 1e-6 * abs(fd)
    u = (_direction(obj, rng), _direction(probe, rng), rng.uniform(-1, 1, (nb, 2)))
    v = (_direction(obj, rng), _direction(probe, rng), rng.uniform(-1, 1, (nb, 2)))
    h_uv, h_vu = b.hessian(u, v), b.hessian(v, u)
    assert abs(h_uv - h_vu) < 1e-12 * abs(h_uv)

    def du(t):
        bt = BhSlices(views, iobs, optics, obj + t * v[0], probe + t * v[1], d0 + t * v[2], True)
        return dot(bt.grads(), u)

    t = 1e-6
    fd = (du(t) - du(-t)) / (2 * t)
    assert abs(h_uv - fd) < 1e-6 * abs(fd)
    held = BhSlices(views, iobs, optics, obj, probe, d0)
    assert held.grads()[2] is None and np.array_equal(held.psi, b.psi)
