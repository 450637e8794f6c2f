"""The data model of the step, pinned by the reference's own unit-test values on both the FP64
oracle (`oracle/base.py`, `oracle/optics.py`, `oracle/ptycho.py`) and the product's host path
(`ptyrs_amd.base`, the C ABI's host-side preparation): view corners and their rounding, the
address table, the object shape, PyNX position preprocessing, observed counts, the far-field
phase, and the operators' hand values, dot tests and consistent-pair invariant.

Reference tests restated: ptyrs/src/base/addr.rs:151-214, ptyrs/src/base/positions.rs:102-143,
ptyrs/src/base/counts.rs:140-198, ptyrs/src/engine/obs.rs:73-107,
ptyrs/src/engine/ops.rs:383-502. No device is used.
"""

import math

import numpy as np
import pytest

import ptyrs_amd as pa
from ptyrs_amd import base as pb
from oracle import base as ob
from oracle import ptycho as op
from oracle.optics import Optics

# (obj, frame, shift) -> corner, ptyrs/src/base/addr.rs:155-170: (40 - 16) // 2 = 12.
CORNERS = [((40, 16, 0.5), 12), ((40, 16, 1.5), 14), ((40, 16, -0.4), 12),
           ((40, 16, -12.5), 0), ((40, 16, 12.0), 24), ((40, 16, -12.6), None),
           ((40, 16, 12.6), None), ((40, 16, -20.0), None), ((40, 16, 30.0), None),
           ((40, 16, math.nan), None), ((15, 16, 0.0), None)]


@pytest.mark.parametrize("args,want", CORNERS)
def test_view_corner_rounds_half_even_and_rejects_windows_outside(args, want):
    assert ob.view_corner(*args) == want
    assert pb.view_corner(*args) == want


def test_view_corner_ties_go_to_even_on_both_sides():
    # Every half-integer shift from -12.5 to 12.5: round half to even, not half away.
    for k in range(-25, 26, 2):
        s = k / 2.0
        want = int(np.round(12 + s))
        assert want % 2 == 0
        assert ob.view_corner(40, 16, s) == want
        assert pb.view_corner(40, 16, s) == want


def test_table_rejects_a_frame_outside_the_object():
    # (20 - 8) // 2 = 6: a shift of 6 puts the window on the last row (addr.rs:172-182).
    ok = ob.AddrTable([0.0, 6.0], [0.0, 0.0], [1.0, 1.0], [8, 8], [20, 20], 1, 1)
    assert list(ok.oy) == [6, 12] and list(ok.ox) == [6, 6]
    with pytest.raises(ValueError):
        ob.AddrTable([0.0, 6.6], [0.0, 0.0], [1.0, 1.0], [8, 8], [20, 20], 1, 1)
    with pytest.raises(ValueError):
        ob.AddrTable([0.0], [0.0], [1.0, 1.0], [8, 8], [20, 20], 0, 1)


def test_table_orders_frame_major_and_maps_modes():
    # addr.rs:199-207
    t = ob.AddrTable([0.0, 2.0], [0.0, -3.0], [1.0, 1.0], [8, 8], [20, 20], 2, 1)
    v = t.views()
    assert len(v) == 4
    assert v[2][:4] == (1, 0, 0, 0)
    assert (v[3][1], v[3][2]) == (1, 1)
    assert (v[2][4], v[2][5]) == (8, 3)


def test_obj_shape_matches_pynx_formula():
    # ny = 2*(|ceil(-3.2)|+1) + 16 + 8 = 32, nx = 2*(ceil(5.1)+1) + 16 + 8 = 38 -> 40
    # (addr.rs:209-214)
    assert ob.required_obj_shape([-3.2, 1.0], [0.0, 5.1], [1.0, 1.0], [16, 16], 8) == [32, 40]
    assert pb.required_obj_shape([-3.2, 1.0], [0.0, 5.1], [1.0, 1.0], [16, 16], 8) == [32, 40]


def test_required_shape_holds_every_view_and_its_neighbour():
    # addr.rs:184-197
    y, x = [-3.2, 1.0, 9.49], [0.0, 5.1, -7.5]
    for margin in (0, 1, 4, 16):
        for f in (ob.required_obj_shape, pb.required_obj_shape):
            obj = f(y, x, [1.0, 1.0], [16, 16], margin)
            for axis, vals in ((0, y), (1, x)):
                for d in vals:
                    c = ob.view_corner(obj[axis], 16, d)
                    assert c is not None and c + 16 < obj[axis], (margin, axis, d)


def _snapped(pixels, px):
    v = [p * px for p in pixels]
    oy, ox = ob.pynx_preprocess(v, v, [px, px])
    gy, gx = pb.pynx_preprocess(v, v, [px, px])
    np.testing.assert_array_equal(gy, oy)
    np.testing.assert_array_equal(gx, ox)
    return np.asarray(ox) / px


def test_near_integer_positions_snap_to_pixels():
    # positions.rs:114-121
    np.testing.assert_allclose(_snapped([0.0, 2.02, 4.0, 6.0], 1e-8), [-3, -1, 1, 3], atol=1e-9)


def test_shared_subpixel_offset_is_removed_then_snapped():
    # centred fractions 0.50, 0.51, 0.50, 0.49: one shared offset of 0.5 px (positions.rs:123-130)
    np.testing.assert_allclose(_snapped([0.0, 3.01, 6.0, 8.99], 1e-8), [-5, -2, 1, 4], atol=1e-9)


def test_scattered_fractions_are_only_centred():
    # positions.rs:132-143
    x = [0.0, 1.3, 2.71, 4.05]
    mean = sum(x) / 4.0
    got = _snapped(x, 1.0)
    np.testing.assert_allclose(got, [v - mean for v in x], atol=1e-12)


def test_near_integer_fractions_are_recentred_in_metres():
    # PyNX recentres the fraction with the position in metres, not pixels, which the
    # reference keeps (positions.rs:86-93). Centred pixels -2, -0.94, 0, 0.94, 2: in metres the
    # recentred fractions are all ~1e-8 and snap; in pixels they would spread by 0.12 and stay.
    got = _snapped([0.0, 1.06, 2.0, 2.94, 4.0], 1e-8)
    np.testing.assert_allclose(got, [-2, -1, 0, 1, 2], atol=1e-9)


def test_preprocess_refuses_non_finite_results():
    with pytest.raises(ValueError):
        pb.pynx_axis([0.0, 1e-8], math.inf)
    with pytest.raises(ValueError):
        pb.pynx_axis([2e10], 1e-8)


def _unit_scan(y, x, n=16, counts=None):
    # wavelength·distance / (n·detector_pixel) = 1: positions in metres are object pixels.
    c = np.zeros((len(y), n, n), np.uint16) if counts is None else counts
    return pa.Scan(c, y, x, 1.0, float(n), 1.0)


def test_scan_corners_are_the_oracles_table_with_ties_and_edges():
    # Unsnapped positions (scattered fractions) that land on half pixels and on both edges of
    # a 40-pixel object: the product's preprocessing and corners equal the oracle's.
    y = [0.5, -1.3, 0.8, 11.9, -11.5 - 0.4 / 3]
    x = [1.5, 2.25, -0.7, -12.4 - 0.2 / 5, 9.6]
    yy, xx = ob.pynx_preprocess(y, x, [1.0, 1.0])
    t = ob.AddrTable(yy, xx, [1.0, 1.0], [16, 16], [40, 40], 1, 1)
    got = pb.scan_corners(_unit_scan(y, x), [40, 40])
    np.testing.assert_array_equal(got[:, 0], t.oy)
    np.testing.assert_array_equal(got[:, 1], t.ox)
    # The same through the engine's own corner rule on each axis.
    for f in range(len(y)):
        assert got[f, 0] == pb.view_corner(40, 16, yy[f])
        assert got[f, 1] == pb.view_corner(40, 16, xx[f])


def test_scan_corners_refuse_a_frame_outside_the_object():
    with pytest.raises(ValueError, match="leaves the object"):
        pb.scan_corners(_unit_scan([-3.6, 3.6], [0.0, 0.0], n=8), [14, 14])
    got = pb.scan_corners(_unit_scan([-3.0, 3.0], [0.0, 0.0], n=8), [14, 14])
    assert got.tolist() == [[0, 3], [6, 3]]


def test_shift_moves_centre_to_origin_and_masks():
    # obs.rs:79-90 on the oracle (4x6 frames).
    assert ob.shift_index(0, 4, 6) == 2 * 6 + 3
    assert ob.shift_index(ob.shift_index(7, 4, 6), 4, 6) == 7
    raw = np.arange(24, dtype=np.uint16).reshape(1, 4, 6)
    raw[0, 2, 3] = ob.MASKED
    iobs, nb = ob.observed(raw)
    assert np.isnan(iobs[0].ravel()[0]) and iobs[0].ravel()[1] == 16.0 and nb == 23
    # The product's device counts (square frames): shifted index j reads detector pixel
    # shift(j), masked pixels read -1.
    raw = np.arange(36, dtype=np.uint16).reshape(1, 6, 6)
    raw[0, 3, 3] = ob.MASKED
    got, nb, total = pb.observed(_unit_scan([0.0], [0.0], n=6, counts=raw))
    assert nb == 35 and total == float(raw.sum() - ob.MASKED)
    flat = got[0].ravel()
    assert flat[0] == -1.0
    for j in range(1, 36):
        assert flat[j] == raw[0].ravel()[ob.shift_index(j, 6, 6)], j


def test_counts_total_is_observed_total():
    # obs.rs:92-106: masked pixels and full counts take the total past 2^24.
    i = np.arange(64 * 32 * 32)
    raw = np.where(i % 7 == 0, ob.MASKED, np.where(i % 7 == 1, ob.MAX_COUNT, i % 4093))
    raw = raw.astype(np.uint16).reshape(64, 32, 32)
    iobs, nb = ob.observed(raw)
    exact = float(raw[raw != ob.MASKED].astype(np.uint64).sum())
    assert exact > 2.0 ** 24
    assert float(np.nansum(iobs)) == exact
    _, nb2, total = pb.observed(_unit_scan(np.zeros(64), np.zeros(64), n=32, counts=raw))
    assert nb2 == nb and total == exact


def test_u16_and_float_counts_give_the_same_device_counts():
    # counts.rs:3-7 and the float path: below 65535 the float counts (negative = masked) reduce
    # to the u16 counts (65535 = masked) bit for bit.
    rng = np.random.default_rng(3)
    raw = rng.integers(0, 65535, size=(5, 16, 16)).astype(np.uint16)
    raw[rng.random(raw.shape) < 0.05] = ob.MASKED
    raw[0, 0, 0] = ob.MAX_COUNT
    f = np.where(raw == ob.MASKED, -1.0, raw.astype(np.float64)).astype(np.float32)
    y, x = np.linspace(-3, 3, 5), np.zeros(5)
    a = pb.observed(_unit_scan(y, x, counts=raw))
    b = pb.observed(_unit_scan(y, x, counts=f))
    np.testing.assert_array_equal(a[0], b[0])
    assert a[1:] == b[1:]
    bad = f.copy()
    bad[0, 1, 1] = np.inf
    with pytest.raises(ValueError):
        pb.observed(_unit_scan(y, x, counts=bad))


def test_quad_phase_is_one_at_centre_and_symmetric():
    # ops.rs:409-419: geometry λ = 1e-10, z = 2, 5e-5 pixels, 8x8 frames.
    px = 1e-10 * 2.0 / (8 * 5e-5)
    want = complex(math.cos(math.pi / 2e-10 * 2.0 * px * px), math.sin(math.pi / 2e-10 * 2.0 * px * px))
    for q in (Optics.far_field(1e-10, 2.0, 5e-5, [8, 8]).quad, pb.quad_phase(1e-10, 2.0, 5e-5, [8, 8])):
        assert q[4, 4] == 1.0
        assert abs(q[3, 4] - q[5, 4]) < 1e-15
        assert abs(q[3, 5] - want) < 1e-14
    np.testing.assert_allclose(pb.quad_phase(1e-10, 2.0, 5e-5, [8, 8]),
                               Optics.far_field(1e-10, 2.0, 5e-5, [8, 8]).quad, atol=1e-15)


def _wave(n, seed):
    # ops.rs:390-397
    t = np.arange(n) * 0.731 + seed
    return np.sin(t) + 0.3 + 1j * np.cos(1.7 * t)


def _setup():
    # Three frames, two probe modes, overlapping views, 8x6 frames in a 14x12 object
    # (ops.rs:399-408).
    addr = ob.AddrTable([0.0, 2.0, -3.0], [1.0, -2.0, 2.0], [1.0, 1.0], [8, 6], [14, 12], 2, 1)
    obj = _wave(168, 0.2).reshape(1, 14, 12)
    probe = _wave(96, 1.9).reshape(2, 8, 6)
    quad = Optics.far_field(1e-10, 1.0, 1e-4, [8, 6]).quad
    return addr, obj, probe, quad


def _dot_error(fwd, adj, nx, ny, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, nx) + 1j * rng.uniform(-1, 1, nx)
    y = rng.uniform(-1, 1, ny) + 1j * rng.uniform(-1, 1, ny)
    ax, ahy = fwd(x), adj(y)
    lhs, rhs = np.vdot(y, ax), np.vdot(ahy, x)
    scale = max(np.linalg.norm(ax) * np.linalg.norm(y), np.linalg.norm(x) * np.linalg.norm(ahy))
    return abs(lhs - rhs) / scale


def test_back_projections_are_adjoints_of_obj_probe_to_psi():
    # ops.rs:420-438, DOT_TOL = 1e-10 (linop.rs:42); a missing conjugate fails.
    addr, obj, probe, quad = _setup()
    v = addr.views()
    npsi = len(v) * 48
    obj_fwd = lambda x: op.obj_probe_to_psi(addr, v, x.reshape(obj.shape), probe, quad).ravel()
    obj_adj = lambda y: op.accumulate_obj(v, obj.shape, probe, quad, y.reshape(len(v), 8, 6))[0].ravel()
    pr_fwd = lambda x: op.obj_probe_to_psi(addr, v, obj, x.reshape(probe.shape), quad).ravel()
    pr_adj = lambda y: op.accumulate_probe(v, obj, probe.shape, quad, y.reshape(len(v), 8, 6))[0].ravel()
    assert _dot_error(obj_fwd, obj_adj, obj.size, npsi, 1) < 1e-10
    assert _dot_error(pr_fwd, pr_adj, probe.size, npsi, 2) < 1e-10
    wrong = lambda y: np.conj(obj_adj(np.conj(y)))
    assert _dot_error(obj_fwd, wrong, obj.size, npsi, 1) > 1e-3


def test_psi_to_obj_probe_keeps_a_consistent_pair():
    # psi = P·O·q exactly: both updates return the same object and probe (ops.rs:440-455).
    addr, obj, probe, quad = _setup()
    v = addr.views()
    psi = op.obj_probe_to_psi(addr, v, obj, probe, quad)
    o2, p2 = op.psi_to_obj_probe(v, quad, psi, obj, probe, 0.01, 0.001)
    assert np.abs(o2 - obj).max() < 1e-12 and np.abs(p2 - probe).max() < 1e-12


def test_apply_amplitude_sets_modulus_and_skips_masked():
    # ops.rs:461-486
    raw = np.array([(i * 7) % 50 for i in range(48)], np.uint16)
    raw[20] = ob.MASKED
    iobs, _ = ob.observed(raw.reshape(2, 4, 6))
    before = _wave(2 * 2 * 24, 0.7).reshape(4, 4, 6)
    psi = op.apply_amplitude(iobs, 2, before)
    o = iobs.reshape(2, 24)
    b, p = before.reshape(2, 2, 24), psi.reshape(2, 2, 24)
    for f in range(2):
        for j in range(24):
            if np.isnan(o[f, j]):
                assert p[f, 0, j] == b[f, 0, j] and p[f, 1, j] == b[f, 1, j]
                continue
            c = abs(p[f, 0, j]) ** 2 + abs(p[f, 1, j]) ** 2
            assert abs(c - o[f, j]) < 1e-12 * max(o[f, j], 1.0)
            r = p[f, 0, j] / b[f, 0, j]
            assert abs(r.imag) < 1e-12 and abs(p[f, 1, j] / b[f, 1, j] - r) < 1e-9


def test_llk_matches_hand_values_and_skips_masked():
    # One frame, one mode, 2x2: iobs [4, 0, MASKED, 9] in shifted order (ops.rs:487-502).
    raw = np.array([9, ob.MASKED, 0, 4], np.uint16).reshape(1, 2, 2)
    iobs, nb = ob.observed(raw)
    flat = iobs.ravel()
    assert flat[0] == 4.0 and flat[1] == 0.0 and np.isnan(flat[2]) and flat[3] == 9.0
    psi = (1j * np.array([2.0, 1.0, 5.0, 2.0])).reshape(1, 2, 2)
    l = op.llk(iobs, 1, psi)
    p = (4.0 - 4.0) + 1.0 + (4.0 - 9.0 + 9.0 * math.log(9.0 / 4.0))
    assert abs(l.poisson - p) < 1e-14
    assert abs(l.gaussian - (0.0 + 1.0 + 25.0 / 10.0)) < 1e-14
    assert abs(l.euclidian - 4.0 * (0.0 + 1.0 + 1.0)) < 1e-14
    assert l.nb_photons_calc == 4.0 + 1.0 + 4.0
    n = l.normalised(nb)
    assert n.poisson * 3.0 == l.poisson and n.nb_photons_calc == 9.0
