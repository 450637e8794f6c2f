"""The seeded far-field scenario (oracle/scenario.py) and the far-field fixture families it
unlocks, on the oracle.

The scenario restates the reference's simulation (`common::dm_scenario`, ChaCha8 Poisson
counts); it is pinned by the fixtures' own totals: the scaled start's Poisson LLK sum of
`ml_grad_values_f64.txt` at 1e-12 (ptyrs/tests/engine_reconstruct.rs:244-256) can only match
with every count drawn as the reference drew it. On that start: DM and RAAR against PyNX,
rPIE against pty-chi (ptyrs/tests/engine_reconstruct.rs:161-236, LLK records at 1e-9), the ML
object gradient against the independent f64 model (1e-12), and LSQML's step and three epochs
against pty-chi (ptyrs/tests/engine_lsqml.rs:197-302, steps at 1e-13/1e-12, changes at 1e-12).
"""

import functools
import math

import numpy as np
import pytest

from conftest import max_rel, table
from oracle import scenario as sc
from oracle.lsqml import AMPLITUDE, POISSON, Lsqml, illumination


@functools.lru_cache(maxsize=None)
def start(nprobe):
    return sc.parity_start(nprobe=nprobe)


def fresh(nprobe):
    """A copy of the scaled start that a run may change."""
    p, inputs = start(nprobe)
    counts, y, x, obj, probe = inputs
    q = sc.far_field_ptycho(counts, y, x, obj, probe, *sc.GEOMETRY)
    q.obj, q.probe = p.obj.copy(), p.probe.copy()
    return q


def test_chacha_core_matches_the_published_keystream():
    # RFC 8439 A.1 test vector 1: ChaCha20, all-zero key and nonce, block 0.
    want = ("76b8e0ada0f13d90405d6ae55386bd28bdd219b8a08ded1aa836efcc8b770dc7"
            "da41597c5157488d7724e03fb8d84a376a43b8f41518a11cc387b669b2ee6586")
    got = sc.chacha_blocks(bytes(32), 0, 1, 20)[0].astype("<u4").tobytes().hex()
    assert got == want


def test_poisson_is_seeded_and_masks():
    # ptyrs/src/base/sim.rs:187-197
    lam = np.array([0.0, 5.0, 100.0, 1e3]).reshape(1, 2, 2)
    mask = np.array([False, False, False, True]).reshape(2, 2)
    a, b = sc.poisson(lam, 7, mask), sc.poisson(lam, 7, mask)
    assert np.array_equal(a, b) and a.ravel()[0] == 0 and a.ravel()[3] == 65535
    with pytest.raises(ValueError):
        sc.poisson(np.array([1e6]).reshape(1, 1, 1), 1)


def test_poisson_sampler_has_the_poisson_moments():
    # Both branches (Knuth below λ = 12, rejection above): mean and variance λ.
    for lam in (3.0, 40.0):
        v = sc.poisson(np.full((1, 100, 100), lam), 11).astype(float)
        assert abs(v.mean() - lam) < 4 * math.sqrt(lam / v.size)
        assert abs(v.var() / lam - 1.0) < 0.05


def test_probe_modes_are_normalised_and_orthogonal():
    # ptyrs/src/base/sim.rs:167-185
    p = sc.sim_probe([32, 32], 3)
    power = [(np.abs(p[k]) ** 2).sum() for k in range(3)]
    assert abs(sum(power) - 1.0) < 1e-12 and abs(power[0] / power[1] - 4.0) < 1e-9
    for a, b in ((0, 1), (0, 2), (1, 2)):
        assert abs((p[a] * np.conj(p[b])).sum()) < 1e-6


def test_fermat_spacing_and_raster_layout():
    # ptyrs/src/base/sim.rs:199-215
    y, x = sc.fermat(200, 1.0)
    d = np.hypot(y[:, None] - y[None], x[:, None] - x[None])
    np.fill_diagonal(d, np.inf)
    assert 0.7 < d.min(axis=1).mean() < 1.3
    ry, rx = sc.raster(2, 3, 2.0)
    assert ry.tolist() == [-1.0, -1.0, -1.0, 1.0, 1.0, 1.0]
    assert rx.tolist() == [-2.0, 0.0, 2.0, -2.0, 0.0, 2.0]


def test_object_amplitude_and_phase_ranges():
    # ptyrs/src/base/sim.rs:217-223
    a = np.abs(sc.sim_object([64, 64]))
    assert a.min() >= 0.45 and a.max() <= 0.96


def test_scenario_counts_are_the_references():
    # The scaled start's Poisson LLK sum and nb_obs (ptyrs/tests/engine_reconstruct.rs:244-256).
    nb_obs, want = table("ml_grad_values_f64.txt")[0]
    p, _ = start(2)
    assert p.nb_obs == int(nb_obs)
    got = p.llk().poisson * p.nb_obs
    assert abs(got - want) <= 1e-12 * abs(want), (got, want)


def test_obj_grad_matches_ml_grad_reference():
    # ptyrs/tests/engine_reconstruct.rs:244-275
    p, _ = start(2)
    g = p.grads(True, False)[0]
    rows = np.array(table("ml_grad_f64.txt"))
    assert g.shape[0] == 1 and len(rows) == g.shape[1] * g.shape[2]
    got = g[0][rows[:, 0].astype(int), rows[:, 1].astype(int)]
    want = rows[:, 2] + 1j * rows[:, 3]
    assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()


def history_matches(history, lag, name):
    rows = table(name)
    assert len(history) == len(rows)
    for (cycle, _, l), r in zip(history, rows):
        assert cycle == r[0] + lag
        for a, b in zip((l.poisson, l.gaussian, l.euclidian, l.nb_photons_calc), r[1:]):
            assert abs(a - b) < 1e-9 * abs(b), (cycle, a, b)


@pytest.mark.parametrize("algorithm", ["dm", "raar"])
def test_projection_llk_matches_pynx(algorithm):
    # ptyrs/tests/engine_reconstruct.rs:184-223: record 0 is the start, the rest lag one cycle.
    p = fresh(1)
    first = p.llk()
    getattr(p, algorithm)(21, calc_llk=1)
    assert p.history[0][0] == 0 and p.history[0][2].poisson == first.poisson
    history_matches(p.history[1:], 1, f"{algorithm}_pynx_f64.txt")


RPIE_PTYCHI = dict(obj_step=1.0, probe_step=1.0, obj_alpha=0.5, probe_alpha=0.5, batch_size=4,
                   calc_llk=1, seed=7)  # ptyrs/tests/common/ptychi/mod.rs:8-16


def test_rpie_llk_matches_ptychi():
    # ptyrs/tests/engine_reconstruct.rs:225-236
    p = fresh(1)
    p.rpie(20, **RPIE_PTYCHI)
    history_matches(p.history, 0, "rpie_ptychi_f64.txt")


def _changes(obj, probe, obj0, probe0, obj_file, probe_file, tol=1e-12):
    for got, file, lead in ((obj - obj0, obj_file, 2), (probe - probe0, probe_file, 3)):
        rows = np.array(table(file))
        g = np.stack([got.real.ravel(), got.imag.ravel()], axis=1).ravel()
        err = max_rel(g, rows[:, lead:lead + 2].ravel())
        assert err < tol, (file, err)


@pytest.mark.parametrize("cost,name", [(AMPLITUDE, "gaussian"), (POISSON, "poisson")])
def test_lsqml_step_matches_ptychi(cost, name):
    # One step on frames 0..20, the illumination as preconditioner
    # (ptyrs/tests/engine_lsqml.rs:197-283).
    p = fresh(2)
    run = Lsqml(20, cost, 0)
    run.preconditioner = illumination(p.views_all, p.obj.shape, p.probe)
    views = p.views_all[:20 * p.addr.nmode]
    o2, p2, st = run.step(views, p.iobs, p.optics, p.obj, p.probe)
    rows = table(f"lsqml_{name}_steps_f64.txt")
    assert st["frames"] == list(range(20)) and [int(r[0]) for r in rows] == list(range(20))
    for key, col, tol in (("a_far", 1, 1e-13), ("a_obj", 2, 1e-12), ("a_probe", 3, 1e-12)):
        assert max_rel(st[key], [r[col] for r in rows]) < tol, key
    _changes(o2, p2, p.obj, p.probe, f"lsqml_{name}_obj_f64.txt", f"lsqml_{name}_probe_f64.txt")


@pytest.mark.parametrize("cost,name", [(AMPLITUDE, "gaussian"), (POISSON, "poisson")])
def test_lsqml_epochs_match_ptychi(cost, name):
    # Three epochs of batches of 16, seed 7 (ptyrs/tests/engine_lsqml.rs:285-341).
    p = fresh(2)
    run = Lsqml(16, cost, 7)
    rows = table(f"lsqml_epochs_{name}_steps_f64.txt")
    obj, probe, nxt = p.obj, p.probe, 0
    for e in range(3):
        obj, probe, steps = run.epoch(p.addr, p.iobs, p.optics, obj, probe)
        for st in steps:
            part = rows[nxt:nxt + len(st["frames"])]
            nxt += len(st["frames"])
            assert all(r[0] == e for r in part) and st["frames"] == [int(r[1]) for r in part]
            for key, col, tol in (("a_far", 2, 1e-13), ("a_obj", 3, 1e-12),
                                  ("a_probe", 4, 1e-12)):
                assert max_rel(st[key], [r[col] for r in part]) < tol, (e, key)
    assert nxt == len(rows)
    _changes(obj, probe, p.obj, p.probe, f"lsqml_epochs_{name}_obj_f64.txt",
             f"lsqml_epochs_{name}_probe_f64.txt")


# rPIE with refined position offsets under PATCH (ptyrs/tests/engine_rpie_pos.rs:58-160):
# dm_scenario(32, 40, 2, 2e5, 4.0) after scale_obj_probe, from the reference's start offsets.
RPIE_POS = {
    "step": dict(obj_step=0.9, probe_step=0.8, obj_alpha=0.3, probe_alpha=0.2, batch_size=20,
                 calc_llk=1, seed=7, nepoch=1),
    "epochs": dict(obj_step=1.0, probe_step=1.0, obj_alpha=0.5, probe_alpha=0.5, batch_size=4,
                   calc_llk=1, seed=7, nepoch=3),
}


@pytest.mark.parametrize("case", ["step", "epochs"])
def test_rpie_positions_match_ptychi(case):
    kw = dict(RPIE_POS[case])
    nepoch = kw.pop("nepoch")
    p = fresh(2)
    rows = np.array(table(f"rpie_pos_{case}_offsets_f64.txt"))
    start_off = rows[:, 1:3]
    p.set_position(start_off, refine=True)
    o0, p0 = p.obj.copy(), p.probe.copy()
    p.rpie(nepoch, **kw)
    _changes(p.obj, p.probe, o0, p0, f"rpie_pos_{case}_obj_f64.txt",
             f"rpie_pos_{case}_probe_f64.txt")
    err = max_rel((p.offsets - start_off).ravel(), (rows[:, 3:5] - start_off).ravel())
    assert err < 3e-12, err


def test_rpie_positions_held_and_zero_are_plain_rpie():
    # Zero offsets held: the moved waves are P·q up to FFT rounding, so the run is plain rPIE.
    a, b = fresh(2), fresh(2)
    b.set_position(np.zeros((len(b.iobs), 2)), refine=False)
    a.rpie(2, **RPIE_PTYCHI)
    b.rpie(2, **RPIE_PTYCHI)
    for x, y in ((b.obj, a.obj), (b.probe, a.probe)):
        assert np.abs(x - y).max() < 1e-12 * np.abs(y).max()
    assert np.all(b.offsets == 0.0)


def test_ml_positions_match_the_model():
    # ML with refined offsets, four cycles of object and probe (ptyrs/tests/engine_ml_pos.rs:
    # 51-117): changes against max|start| at 1e-13, offset changes at 1e-12 of their max.
    p = fresh(2)
    rows = np.array(table("ml_pos_offsets_f64.txt"))
    off0 = rows[:, 1:3]
    p.set_position(off0, refine=True)
    o0, p0 = p.obj.copy(), p.probe.copy()
    p.ml(4, update_object=True, update_probe=True, calc_llk=1, zero_phase_ramp=False)
    for x, x0, file, lead in ((p.obj, o0, "ml_pos_obj_f64.txt", 2),
                              (p.probe, p0, "ml_pos_probe_f64.txt", 3)):
        r = np.array(table(file))
        got = np.stack([(x - x0).real.ravel(), (x - x0).imag.ravel()], axis=1).ravel()
        scale = max(np.abs(x0.real).max(), np.abs(x0.imag).max())
        err = np.abs(got - r[:, lead:lead + 2].ravel()).max() / scale
        assert err < 1e-13, (file, err)
    want = (rows[:, 3:5] - off0).ravel()
    err = np.abs((p.offsets - off0).ravel() - want).max() / np.abs(want).max()
    assert err < 1e-12, err


# The same runs under FULL_OBJECT (ptyrs/tests/engine_rpie_pos.rs:178-226): each view's window
# Fourier-interpolated at its corner plus its offset; "odd" is the step on the scaled start's
# 76 x 80 plane less its first row and column (both sides odd), each window one pixel nearer
# the origin and so on the same pixels.
@pytest.mark.parametrize("case", ["step", "epochs", "odd"])
def test_rpie_full_object_positions_match_ptychi(case):
    kw = dict(RPIE_POS["step" if case == "odd" else case])
    nepoch = kw.pop("nepoch")
    p = fresh(2)
    if case == "odd":
        _, inputs = start(2)
        counts, y, x, _, _ = inputs
        p = sc.far_field_ptycho(counts, y, x, p.obj[:, 1:, 1:].copy(), p.probe.copy(),
                                *sc.GEOMETRY)
        assert p.obj.shape[-2:] == (75, 79)
    p.set_sampling("full_object")
    rows = np.array(table(f"rpie_full_object_{case}_offsets_f64.txt"))
    start_off = rows[:, 1:3]
    p.set_position(start_off, refine=True)
    o0, p0 = p.obj.copy(), p.probe.copy()
    p.rpie(nepoch, **kw)
    _changes(p.obj, p.probe, o0, p0, f"rpie_full_object_{case}_obj_f64.txt",
             f"rpie_full_object_{case}_probe_f64.txt")
    err = max_rel((p.offsets - start_off).ravel(), (rows[:, 3:5] - start_off).ravel())
    assert err < 3e-12, err


# AP, DM and RAAR under PATCH offsets (ptyrs/src/engine/ptycho.rs:605-755 with
# Shifted::psi_to_obj_probe, ptyrs/src/engine/features.rs:653-720): at zero offsets held the
# moved waves are P·q and the probe's conjugate-gradient step is the plain blend (its step
# length is 1 when nothing moves), so the runs are the plain ones to the transforms' rounding.
@pytest.mark.parametrize("algorithm", ["ap", "dm", "raar"])
def test_projections_under_zero_offsets_are_the_plain_ones(algorithm):
    a, b = fresh(2), fresh(2)
    b.set_position(np.zeros((len(b.iobs), 2)), refine=False)
    kw = dict(calc_llk=1, zero_phase_ramp=False, probe_inertia=0.01)
    getattr(a, algorithm)(4, **kw)
    getattr(b, algorithm)(4, **kw)
    for x, y in ((b.obj, a.obj), (b.probe, a.probe)):
        assert np.abs(x - y).max() < 1e-11 * np.abs(y).max(), algorithm
    assert max_rel([r[2].poisson for r in b.history], [r[2].poisson for r in a.history]) < 1e-12


@pytest.mark.parametrize("algorithm", ["ap", "dm", "raar"])
def test_projections_refine_offsets_toward_the_truth(algorithm):
    # The scenario's own positions are the truth: offsets started off it by up to 0.3 px
    # (zero mean: the step leaves the frames' mean, a shift of the whole object, alone) move
    # back toward it.
    p = fresh(2)
    d0 = np.random.default_rng(5).uniform(-0.3, 0.3, (len(p.iobs), 2))
    d0 -= d0.mean(axis=0)
    p.set_position(d0.copy(), refine=True)
    getattr(p, algorithm)(6, calc_llk=1, zero_phase_ramp=False, probe_inertia=0.01)
    assert np.sqrt((p.offsets ** 2).mean()) < np.sqrt((d0 ** 2).mean())
    llks = [r[2].poisson for r in p.history]
    assert llks[-1] < llks[0]


def test_moved_probe_step_minimises_along_its_direction():
    # The probe's update under offsets (ptyrs/src/engine/features.rs:653-720) is one
    # conjugate-gradient step of J(P') = Σ_views ‖ψ − W(P')·O_win‖² + reg·‖P' − P‖², reg =
    # inertia·max Σ|O_win|², W(P') the moved waves: J is quadratic along the step, so the step's
    # length α must be J's minimiser there (J(P + t·ΔP) ≥ J(P + ΔP) for every t) and J must fall.
    from oracle.base import to_natural
    from oracle.ptycho import apply_amplitude
    p = fresh(2)
    d0 = np.random.default_rng(8).uniform(-0.4, 0.4, (len(p.iobs), 2))
    p.set_position(d0 - d0.mean(axis=0), refine=False)
    nmode = p.addr.nmode
    p.psi = p.optics.exit_adjoint(apply_amplitude(p.iobs, nmode,
                                                  p.optics.exit_forward(p.calc_psi())))
    views = p.views_all
    ny, nx = p.probe.shape[-2:]
    win = np.stack([p.obj[o, oy:oy + ny, ox:ox + nx] for (_, _, _, o, oy, ox) in views])
    norm = np.zeros(p.probe.shape)
    for k, v in enumerate(views):
        norm[v[2]] += np.abs(win[k]) ** 2
    inertia = 0.05
    reg = inertia * norm.max()
    psi_n = to_natural(p.psi)
    p0 = p.probe.copy()

    def cost(probe):
        keep, p.probe = p.probe, probe
        w = p.moved_waves(views)
        p.probe = keep
        return float((np.abs(psi_n - w * win) ** 2).sum() + reg * (np.abs(probe - p0) ** 2).sum())

    p._moved_psi_to_obj_probe(None, inertia)
    step = p.probe - p0
    assert np.abs(step).max() > 1e-6 * np.abs(p0).max()
    j1 = cost(p0 + step)
    assert j1 < cost(p0)
    for t in (0.8, 0.98, 1.02, 1.25):
        assert cost(p0 + t * step) > j1
    # Per mode too: scaling one mode's step alone also raises J.
    for m in range(p0.shape[0]):
        s = step.copy()
        s[m] *= 1.1
        assert cost(p0 + s) > j1


def test_ml_full_object_zero_offsets_is_plain_ml():
    # Under FullObject the views read the object sampled at corner plus offset; at zero offsets
    # the samples are the windows themselves, so ML is the plain run to rounding
    # (ptyrs/src/engine/features.rs:865-958, :978-985).
    a, b = fresh(2), fresh(2)
    b.set_position(np.zeros((len(b.iobs), 2)), refine=False)
    b.set_sampling("full_object")
    for p in (a, b):
        p.ml(3, update_object=True, update_probe=True, calc_llk=1, zero_phase_ramp=False)
    for x, y in ((b.obj, a.obj), (b.probe, a.probe)):
        assert np.abs(x - y).max() < 1e-10 * np.abs(y).max()
    assert max_rel([r[2].poisson for r in b.history], [r[2].poisson for r in a.history]) < 1e-12


def test_ml_full_object_gradient_matches_central_differences():
    # The gradients under FullObject (the reads' gradient placed by the sampling's adjoint, the
    # probe's through the sampled windows) against central differences of the Poisson LLK sum.
    from oracle.ptycho import llk
    p = fresh(2)
    d0 = np.random.default_rng(12).uniform(-0.4, 0.4, (len(p.iobs), 2))
    p.set_position(d0, refine=False)
    p.set_sampling("full_object")
    go, gp = p.grads(True, True)
    rng = np.random.default_rng(13)
    do = (rng.normal(size=p.obj.shape) + 1j * rng.normal(size=p.obj.shape)) * 0.01
    dp = (rng.normal(size=p.probe.shape) + 1j * rng.normal(size=p.probe.shape)) * 0.01 * \
        np.abs(p.probe).max()

    def cost(t):
        o0, p0 = p.obj, p.probe
        p.obj, p.probe = o0 + t * do, p0 + t * dp
        psi = p.optics.exit_forward(p.calc_psi())
        p.obj, p.probe = o0, p0
        return llk(p.iobs, p.addr.nmode, psi).poisson

    h = 1e-5
    fd = (cost(h) - cost(-h)) / (2 * h)
    an = float((np.conj(go) * do).real.sum() + (np.conj(gp) * dp).real.sum())
    assert abs(an - fd) < 1e-6 * abs(fd)


def test_ml_full_object_refines_offsets_toward_the_truth():
    p = fresh(2)
    d0 = np.random.default_rng(14).uniform(-0.3, 0.3, (len(p.iobs), 2))
    d0 -= d0.mean(axis=0)
    p.set_position(d0.copy(), refine=True)
    p.set_sampling("full_object")
    p.ml(6, update_object=True, update_probe=True, calc_llk=1, zero_phase_ramp=False)
    assert np.sqrt((p.offsets ** 2).mean()) < np.sqrt((d0 ** 2).mean())
    llks = [r[2].poisson for r in p.history]
    assert llks[-1] < llks[0]


@pytest.mark.parametrize("algorithm", ["ap", "dm", "raar"])
def test_projections_full_object_zero_offsets_are_the_plain_ones(algorithm):
    # Under FullObject at zero offsets the reads are the windows, AᴴA is the diagonal D, the
    # first PCG step is the blend on unheld pixels and the rest move nothing beyond rounding
    # (ptyrs/src/engine/features.rs:726-734).
    a, b = fresh(2), fresh(2)
    b.set_position(np.zeros((len(b.iobs), 2)), refine=False)
    b.set_sampling("full_object")
    kw = dict(calc_llk=1, zero_phase_ramp=False, probe_inertia=0.01)
    getattr(a, algorithm)(3, **kw)
    getattr(b, algorithm)(3, **kw)
    for x, y in ((b.obj, a.obj), (b.probe, a.probe)):
        assert np.abs(x - y).max() < 1e-9 * np.abs(y).max(), algorithm
    assert max_rel([r[2].poisson for r in b.history], [r[2].poisson for r in a.history]) < 1e-11


def test_full_object_object_solve_descends_its_quadratic():
    # solve_obj's PCG steps lower J(O') = ‖psi − A·O'‖² + reg·‖O' − O‖² at sub-pixel offsets,
    # where AᴴA couples the pixels; more steps lower it further.
    from oracle.base import to_natural
    from oracle.ptycho import apply_amplitude, solve_obj
    from oracle.sampling import Sampler, view_offsets
    import oracle.ptycho as op
    p = fresh(2)
    d0 = np.random.default_rng(15).uniform(-0.45, 0.45, (len(p.iobs), 2))
    p.set_position(d0, refine=False)
    p.set_sampling("full_object")
    nm = p.addr.nmode
    psi = p.optics.exit_adjoint(apply_amplitude(p.iobs, nm, p.optics.exit_forward(p.calc_psi())))
    views = p.views_all
    sm = Sampler(p.obj.shape[-2:], p.probe.shape[-2:], views, view_offsets(views, p.offsets))
    w = np.stack([p.probe[v[2]] * p.quad for v in views])
    psi_n = to_natural(psi)
    d = sm.diag(w.real ** 2 + w.imag ** 2, p.obj.shape[0])
    reg = 0.01 * d.max()

    def J(o):
        return float((np.abs(psi_n - w * sm.windows(o)) ** 2).sum() +
                     reg * (np.abs(o - p.obj) ** 2).sum())

    keep = op.OBJ_PCG
    try:
        op.OBJ_PCG = 1
        j1 = J(solve_obj(sm, w, psi_n, p.obj, 0.01))
        op.OBJ_PCG = 10
        j10 = J(solve_obj(sm, w, psi_n, p.obj, 0.01))
    finally:
        op.OBJ_PCG = keep
    assert j10 < j1 < J(p.obj)


def test_full_object_object_solve_converges_to_its_minimiser():
    # Run long, solve_obj's PCG reaches the minimiser of J(O') = ‖psi − A·O'‖² + reg·‖O' − O‖²,
    # J written with the windows alone (no adjoint): along random directions J's slope there
    # (exact by central differences, J being quadratic) vanishes against its slope at O, and J
    # rises both ways. A wrong adjoint, diagonal or conjugation leaves a stationary point of
    # another quadratic.
    from oracle.base import to_natural
    from oracle.ptycho import apply_amplitude, solve_obj
    from oracle.sampling import Sampler, view_offsets
    import oracle.ptycho as op
    p = fresh(2)
    d0 = np.random.default_rng(17).uniform(-0.45, 0.45, (len(p.iobs), 2))
    p.set_position(d0, refine=False)
    p.set_sampling("full_object")
    nm = p.addr.nmode
    psi = p.optics.exit_adjoint(apply_amplitude(p.iobs, nm, p.optics.exit_forward(p.calc_psi())))
    views = p.views_all
    sm = Sampler(p.obj.shape[-2:], p.probe.shape[-2:], views, view_offsets(views, p.offsets))
    w = np.stack([p.probe[v[2]] * p.quad for v in views])
    psi_n = to_natural(psi)
    reg = 0.01 * sm.diag(w.real ** 2 + w.imag ** 2, p.obj.shape[0]).max()

    def J(o):
        return float((np.abs(psi_n - w * sm.windows(o)) ** 2).sum() +
                     reg * (np.abs(o - p.obj) ** 2).sum())

    keep = op.OBJ_PCG
    try:
        op.OBJ_PCG = 300
        x = solve_obj(sm, w, psi_n, p.obj, 0.01)
    finally:
        op.OBJ_PCG = keep
    rng = np.random.default_rng(18)
    for _ in range(3):
        dlt = (rng.standard_normal(p.obj.shape) + 1j * rng.standard_normal(p.obj.shape)) * 0.01
        at_x = (J(x + dlt) - J(x - dlt)) / 2
        at_o = (J(p.obj + dlt) - J(p.obj - dlt)) / 2
        assert abs(at_x) < 1e-7 * abs(at_o), (at_x, at_o)
        assert J(x + dlt) > J(x) and J(x - dlt) > J(x)


@pytest.mark.parametrize("algorithm", ["ap", "dm"])
def test_projections_full_object_refine_offsets_toward_the_truth(algorithm):
    p = fresh(2)
    d0 = np.random.default_rng(16).uniform(-0.3, 0.3, (len(p.iobs), 2))
    d0 -= d0.mean(axis=0)
    p.set_position(d0.copy(), refine=True)
    p.set_sampling("full_object")
    getattr(p, algorithm)(6, calc_llk=1, zero_phase_ramp=False, probe_inertia=0.01)
    assert np.sqrt((p.offsets ** 2).mean()) < np.sqrt((d0 ** 2).mean())
