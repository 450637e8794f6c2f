"""Parity at the benchmark workloads' own shapes, on the GPU through the C ABI.

1. Plain rPIE, 64² probe, 256² object, 10×10 raster, 1e6 photons, 50 epochs: against the FP64
   oracle on the same seeded inputs (complex128 to FP64 rounding, complex64 to 1e-4 on the LLK).
2. The headline shape (256² × 4 modes, 1024² object, 1024 positions): one rPIE epoch of both
   precisions against the C++ FP64 restatement of the reference CPU rPIE (the oracle's
   cpu_rpie, pinned to the numpy oracle) on the same inputs and corners; complex64 against
   complex128 after 8 rPIE and 4 LSQML epochs.
3. ML's cost gradient, 128² × 2 modes, 512² object × 3 slices (exp(iφ − μ), φ ≤ 0.3 rad,
   μ = 0.05φ), 10 keV, 10 nm pixels, 2 µm slices, 400 positions: against the oracle's adjoint
   and against five-point central differences (step 1e-4) of the GPU cost in complex128 and
   in complex64; a few torch Adam steps through the autograd Function lower the cost.
4. Sixteen independent scans of uneven sizes (128² and 256² probes, 1–4 modes, 200–3000
   Fermat positions at half the probe's FWHM, 1e6–1e8 photons): their rPIE epochs run all at
   once (ptyrs_group_rpie) give the bits of each scan run alone, and lower every LLK.
5. The large multislice shape (512² × 4 modes, 4096² object × 8 slices at 1 µm, 10 keV, 5 nm
   pixels) with all 5000 Fermat positions perturbed by N(0, 3 px), slice waves checkpointed
   every 2 slices: LSQML with refined positions in complex64 and complex128 lowers the LLK and
   the position error alike, the two objects within 1e-4 NRMSE per slice.
Data for 3–5 come from the GPU model itself (`intensities`): Poisson noise drawn on the host for
3, the model intensities rounded to whole counts for 4 and 5 (the property tested does not
depend on the noise, and sampling 10⁹ counts on the host would dominate the test).
"""

import math

import numpy as np
import pytest

import ptyrs_amd as pa
from conftest import max_rel
from ptyrs_amd import sim
from test_gpu_rpie import rel_l2

pytestmark = pytest.mark.gpu


def _scan(d):
    return pa.Scan(d["counts"], d["y"], d["x"], d["wavelength"], d["distance"],
                   d["detector_pixel"])


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_rpie_64_probe_100_positions_50_epochs_matches_oracle(dtype):
    from oracle.ptycho import far_field_ptycho
    d = sim.scan("small")
    g = pa.GpuPtycho(_scan(d), d["start_obj"], d["start_probe"], dtype=dtype)
    g.rpie(50, obj_step=1.0, probe_step=1.0, obj_alpha=0.1, probe_alpha=0.1, calc_llk=1)
    o = far_field_ptycho(d["counts"].astype(np.float64), d["y"], d["x"], d["start_obj"],
                         d["start_probe"], d["wavelength"], d["distance"], d["detector_pixel"])
    o.rpie(50, calc_llk=1)
    want = [r[2].poisson for r in o.history]
    tol = {"c128": 1e-9, "c64": 1e-4}[dtype]
    print("rpie 64²", dtype, max_rel(g.history["poisson"], want), rel_l2(g.object, o.obj),
          rel_l2(g.probe, o.probe))
    assert max_rel(g.history["poisson"], want) < tol
    assert rel_l2(g.object, o.obj) < tol * 30
    assert rel_l2(g.probe, o.probe) < tol * 30


def aligned_nrmse(a, b, weight):
    w = weight / weight.sum()
    ny, nx = a.shape[-2:]
    yy, xx = np.mgrid[:ny, :nx]
    ph = np.angle(a[0] * np.conj(b[0]))
    A = np.stack([np.ones(ny * nx), yy.ravel(), xx.ravel()], axis=1) * np.sqrt(w.ravel())[:, None]
    coef = np.linalg.lstsq(A, ph.ravel() * np.sqrt(w.ravel()), rcond=None)[0]
    a2 = a * np.exp(-1j * (coef[1] * yy + coef[2] * xx))
    a2 = a2 * np.exp(-1j * np.angle((w * a2 * np.conj(b)).sum()))
    return float(np.sqrt((w * np.abs(a2 - b) ** 2).sum() / (w * np.abs(b) ** 2).sum()))


def test_headline_one_rpie_epoch_against_the_fp64_oracle():
    from oracle import cpu
    from oracle.base import AddrTable, pynx_preprocess
    from oracle.optics import Optics
    d = sim.scan("headline")
    n = d["n"]
    optics = Optics.far_field(d["wavelength"], d["distance"], d["detector_pixel"], [n, n])
    py_, px_ = pynx_preprocess(d["y"], d["x"], optics.pixel)
    addr = AddrTable(py_, px_, optics.pixel, [n, n], d["start_obj"].shape[-2:], 4, 1)
    corners = np.c_[addr.oy, addr.ox]
    c = cpu.CpuRpie(d["counts"], corners, d["start_obj"], d["start_probe"], optics.quad,
                    threads=0)
    c.rpie(1)
    want_obj, want_probe = c.obj[None], c.probe
    weight = np.zeros(want_obj.shape[-2:])
    p2 = (np.abs(want_probe) ** 2).sum(axis=0)
    for oy, ox in corners:
        weight[oy:oy + n, ox:ox + n] += p2
    change = np.linalg.norm(want_obj - d["start_obj"])
    for dtype, tol in (("c128", 1e-9), ("c64", 1e-4)):
        g = pa.GpuPtycho(_scan(d), d["start_obj"], d["start_probe"], dtype=dtype)
        assert np.array_equal(g.corners, corners)
        g.rpie(1)
        o, p = g.object, g.probe
        e_obj = float(np.linalg.norm(o - want_obj) / change)
        e_probe = rel_l2(p, want_probe)
        e_nrmse = aligned_nrmse(o, want_obj, weight)
        print("headline vs FP64 oracle", dtype, e_obj, e_probe, e_nrmse)
        assert e_nrmse < tol and e_probe < tol and e_obj < 10 * tol


def test_headline_256_probe_4_modes_fp32_against_fp64():
    d = sim.scan("headline")
    scan = _scan(d)
    out = {}
    for dtype in ("c64", "c128"):
        g = pa.GpuPtycho(scan, d["start_obj"], d["start_probe"], dtype=dtype)
        g.rpie(8)
        r = (g.object, g.probe, g.llk()["poisson"])
        g.lsqml(4, 0, 100, "poisson", 7)
        out[dtype] = r + (g.object, g.llk()["poisson"], g.corners, g.probe)
    a, b = out["c64"], out["c128"]
    n = d["n"]
    ov = np.zeros(a[0].shape[-2:])
    p2 = (np.abs(b[1]) ** 2).sum(axis=0)
    for oy, ox in b[5]:
        ov[oy:oy + n, ox:ox + n] += p2
    print("headline rpie", aligned_nrmse(a[0], b[0], ov), abs(a[2] - b[2]) / abs(b[2]),
          "lsqml", aligned_nrmse(a[3], b[3], ov), abs(a[4] - b[4]) / abs(b[4]))
    assert aligned_nrmse(a[0], b[0], ov) < 1e-4
    assert abs(a[2] - b[2]) < 1e-6 * abs(b[2])
    assert aligned_nrmse(a[3], b[3], ov) < 1e-4
    assert abs(a[4] - b[4]) < 1e-6 * abs(b[4])


# ---- multislice gradient: 128² × 2 modes, 512² × 3 slices, 400 positions -------------------------

LAMBDA_10KEV = 1.2398e-10


def slice_object(side, nslice, seed, phase_max):
    """Slices exp(iφ − 0.05φ), φ Gaussian-filtered noise in [0, phase_max]; each slice the same
    field rolled by a different offset (one filtered field serves every slice)."""
    phi = np.angle(sim.random_object(side, side, seed, amp=(1.0, 1.0), phase=(0.0, phase_max)))
    out = [np.roll(phi, (37 * s, 53 * s), axis=(-2, -1)) for s in range(nslice)]
    return np.stack([np.exp(1j * p - 0.05 * p) for p in out])  # [nslice, 1, side, side]


def far_geometry(n, pixel):
    dpix = 75e-6
    return LAMBDA_10KEV, pixel * n * dpix / LAMBDA_10KEV, dpix


def gradient_workload():
    n, side, nslice, npos = 128, 512, 3, 400
    pixel = 10e-9
    lam, z, dpix = far_geometry(n, pixel)
    probe = sim.hermite_gauss_probe(n, powers=(0.8, 0.2)) * 300.0
    objects = slice_object(side, nslice, 31, 0.3)
    fwhm = 2.355 * n / 12
    y, x = sim.fermat_positions(npos, 0.3 * fwhm, pixel)
    optics = pa.Optics.far_field(lam, z, dpix, [n, n]).with_slices(
        [pa.Wavefield([n, n], lam, [pixel, pixel], 2e-6)] * (nslice - 1))
    return n, objects, probe, y, x, optics, (lam, z, dpix)


def counts_from_model(scan_args, objects, probe, optics, dtype, seed, photons=None):
    """Poisson counts of the GPU model's intensities at the truth."""
    y, x, geom, n = scan_args
    zero = np.zeros((len(y), n, n), dtype=np.float32)
    g = pa.GpuPtycho(pa.Scan(zero, y, x, *geom), objects, probe, optics=optics, dtype=dtype)
    inten = g.intensities()
    if photons is not None:
        inten *= photons / inten.sum(axis=(1, 2)).mean()
    return np.random.default_rng(seed).poisson(inten).astype(np.float32), g.corners


def test_multislice_gradient_128_probe_3_slices_400_positions():
    torch = pytest.importorskip("torch")
    from ptyrs_amd.autograd import cost
    from oracle.base import observed
    from oracle.lsqml import cost_grads
    from oracle.optics import Optics, Wavefield
    n, objects, probe, y, x, optics, geom = gradient_workload()
    counts, corners = counts_from_model((y, x, geom, n), objects, probe, optics, "c128", 41)
    rng = np.random.default_rng(43)
    start = objects * np.exp(1j * rng.uniform(-0.05, 0.05, objects.shape))
    g = pa.GpuPtycho(pa.Scan(counts, y, x, *geom), start, probe * 0.9, optics=optics,
                     dtype="c128")
    v, go, gp = g.grad("poisson")
    # The oracle's adjoint on the same state and corners (two probe modes, one object mode).
    lam, z, dpix = geom
    o_opt = Optics.far_field(lam, z, dpix, [n, n]).with_slices(
        [Wavefield([n, n], lam, [10e-9, 10e-9], 2e-6)] * 2)
    views = [(f, m, m, 0, int(corners[f, 0]), int(corners[f, 1]))
             for f in range(len(y)) for m in range(2)]
    iobs, _ = observed(counts.astype(np.float64))
    wv, wo, wp = cost_grads(o_opt, views, iobs, 2, list(start), probe * 0.9)
    assert abs(v - wv) <= 1e-9 * abs(wv)
    for s in range(3):
        assert rel_l2(go[s], wo[s]) < 1e-8, s
    assert rel_l2(gp, wp) < 1e-8
    # Central differences (step 1e-4, five points) of the GPU cost.
    def f(objs, pr):
        g.set_object(objs)
        g.set_probe(pr)
        return g.grad("poisson")[0]
    h = 1e-4
    ts = (h, -h, 2 * h, -2 * h)
    stencil = lambda v: (8 * (v[0] - v[1]) - (v[2] - v[3])) / (12 * h)
    p0 = probe * 0.9
    dirs = []  # (slice or None for the probe, direction, FP64 cost at the stencil's points)
    for s in range(3):
        dd = np.zeros_like(start)
        dd[s] = (rng.uniform(-0.5, 0.5, start[s].shape) + 1j * rng.uniform(-0.5, 0.5, start[s].shape))
        dirs.append((s, dd, [f(start + t * dd, p0) for t in ts]))
    dp = (rng.uniform(-0.5, 0.5, probe.shape) + 1j * rng.uniform(-0.5, 0.5, probe.shape)) * 30
    dirs.append((None, dp, [f(start, p0 + t * dp) for t in ts]))
    inner = lambda a, b: float((np.conj(a) * b).real.sum())
    for s, d, vals in dirs:
        got = inner(go if s is not None else gp, d)
        assert abs(got - stencil(vals)) <= 1e-5 * abs(stencil(vals)), s
    # complex64 gradients against complex128 on the same state.
    g32 = pa.GpuPtycho(pa.Scan(counts, y, x, *geom), start, p0, optics=optics)
    v32, go32, gp32 = g32.grad("poisson")
    assert abs(v32 - v) <= 1e-5 * abs(v)
    assert rel_l2(go32, go) < 1e-4 and rel_l2(gp32, gp) < 1e-4
    # The complex64 adjoint against the central differences, in the same directions: those of
    # the complex128 cost (the derivative to FP64 rounding) within 1e-4, and those of the
    # complex64 cost itself within 1e-4 plus what its own rounding can move a five-point
    # difference at step 1e-4: (8(|e₁| + |e₂|) + |e₃| + |e₄|)/(12h), e_i its cost's difference
    # from the complex128 cost at the stencil's points.
    def f32(objs, pr):
        g32.set_object(objs)
        g32.set_probe(pr)
        return g32.grad("poisson")[0]
    errs64, errs32, allow = [], [], []
    for s, d, vals in dirs:
        want = stencil(vals)
        got = inner(go32 if s is not None else gp32, d)
        v32s = [f32(start + t * d, p0) if s is not None else f32(start, p0 + t * d) for t in ts]
        e = [abs(a - b) for a, b in zip(v32s, vals)]
        errs64.append(abs(got - want) / abs(want))
        errs32.append(abs(got - stencil(v32s)) / abs(want))
        allow.append(1e-4 + (8 * (e[0] + e[1]) + e[2] + e[3]) / (12 * h) / abs(want))
    print("complex64 adjoint against complex128 central differences", errs64)
    print("complex64 adjoint against complex64 central differences", errs32, "allowed", allow)
    assert max(errs64) <= 1e-4
    assert all(a <= b for a, b in zip(errs32, allow))
    g32.set_object(start)
    g32.set_probe(p0)
    # A few Adam steps on the slices through the autograd Function lower the cost.
    ot = torch.tensor(start, dtype=torch.complex128, requires_grad=True)
    pt = torch.tensor(probe * 0.9, dtype=torch.complex128)
    opt = torch.optim.Adam([ot], lr=5e-3)
    first = None
    for _ in range(5):
        opt.zero_grad()
        loss = cost(g32, ot, pt)
        first = loss.item() if first is None else first
        loss.backward()
        opt.step()
    assert cost(g32, ot.detach(), pt).item() < first


# ---- sixteen independent scans of uneven sizes -------------------------------------------------------

def sixteen_scans(seed=404):
    """The sixteen scans' inputs: per scan (n, modes, positions, photons, scan, start, probe)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(16):
        n = int(rng.choice([128, 256]))
        nmode = int(rng.integers(1, 5))
        npos = int(rng.integers(200, 3001))
        photons = float(10 ** rng.uniform(6, 8))
        pixel = sim.pixel_size(n)
        y, x = sim.fermat_positions(npos, 0.5 * 2.355 * n / 12, pixel)
        side = (int(np.ceil(np.abs(np.r_[y, x]).max() / pixel)) * 2 + n + 64 + 3) // 4 * 4
        obj = sim.random_object(side, side, seed + 1 + k)
        probe = sim.hermite_gauss_probe(n, powers=(0.7, 0.15, 0.1, 0.05)[:nmode])
        geom = (sim.WAVELENGTH, sim.DISTANCE, sim.DETECTOR_PIXEL)
        truth = pa.GpuPtycho(pa.Scan(np.zeros((npos, n, n), np.float32), y, x, *geom), obj,
                             probe)
        inten = truth.intensities()
        del truth
        counts = np.rint(inten * (photons / inten.sum(axis=(1, 2)).mean())).astype(np.float32)
        del inten
        start = np.exp(1j * np.random.default_rng(seed + 200 + k).uniform(-1, 1, obj.shape))
        scale = math.sqrt(photons / (np.abs(probe) ** 2).sum())
        out.append((n, nmode, npos, photons, pa.Scan(counts, y, x, *geom), start, probe * scale))
        print(f"scan {k}: {n}² x {nmode} modes, {npos} positions", flush=True)
    return out


def test_sixteen_uneven_scans_all_at_once_match_each_alone():
    scans = sixteen_scans()
    assert sum(s[2] for s in scans) > 16 * 1000 and min(s[2] for s in scans) >= 200
    alone = []
    for n, nmode, npos, photons, scan, start, probe in scans:
        g = pa.GpuPtycho(scan, start, probe)
        l0 = g.llk()["poisson"]
        g.rpie(2, calc_llk=1)
        alone.append((l0, g.object, g.probe, g.history["poisson"]))
        del g
    group = [pa.GpuPtycho(s[4], s[5], s[6]) for s in scans]
    pa.rpie_group(group, 2, {"calc_llk": 1})
    for k, (g, (l0, o, p, h)) in enumerate(zip(group, alone)):
        assert np.array_equal(g.object, o) and np.array_equal(g.probe, p), k
        assert np.array_equal(g.history["poisson"], h), k
        assert np.all(np.isfinite(h)) and h[-1] < l0, (k, scans[k][:4], l0, list(h))


# ---- 512² × 4 modes, 4096² × 8 slices, refined positions -----------------------------------------

def test_large_multislice_with_5000_refined_positions():
    n, side, nslice, npos = 512, 4096, 8, 5000
    pixel = 5e-9
    lam, z, dpix = far_geometry(n, pixel)
    probe = sim.hermite_gauss_probe(n) * 1e3
    objects = slice_object(side, nslice, 51, 0.1)
    y, x = sim.fermat_positions(npos, 25.0, pixel)
    optics = pa.Optics.far_field(lam, z, dpix, [n, n]).with_slices(
        [pa.Wavefield([n, n], lam, [pixel, pixel], 1e-6)] * (nslice - 1))
    truth = pa.GpuPtycho(pa.Scan(np.zeros((npos, n, n), np.float32), y, x, lam, z, dpix),
                         objects, probe, optics=optics)
    counts = np.rint(truth.intensities()).astype(np.float32)
    del truth
    print("large multislice data ready", flush=True)
    scan = pa.Scan(counts, y, x, lam, z, dpix)
    d0 = np.random.default_rng(55).normal(0.0, 3.0, (npos, 2))
    rms0 = float(np.sqrt((d0 ** 2).mean()))
    res = {}
    for dtype in ("c64", "c128"):
        g = pa.GpuPtycho(scan, objects, probe, optics=optics, dtype=dtype)
        g.set_checkpoint(2)  # the slice waves checkpointed every 2 slices in the adjoint
        assert g.sampling == "patch"
        g.set_features(position=d0)
        l0 = g.llk()["poisson"]
        g.lsqml(3, 1, 100, "poisson", 7)
        h = g.history["poisson"]
        d = g.features["position"]
        rms = float(np.sqrt((d ** 2).mean()))
        print("large multislice", dtype, l0, list(h), rms0, rms, flush=True)
        assert np.all(np.isfinite(h)) and h[-1] < l0 and rms < rms0
        res[dtype] = (g.objects, d, h, g.corners, g.probe)
    a, b = res["c64"], res["c128"]
    weight = np.zeros(a[0].shape[-2:])
    p2 = (np.abs(b[4]) ** 2).sum(axis=0)
    for oy, ox in b[3]:
        weight[oy:oy + n, ox:ox + n] += p2
    # Both runs start from the same state, so only a global phase can separate them.
    w = weight / weight.sum()

    def phase_nrmse(u, v):
        u = u * np.exp(-1j * np.angle((w * u * np.conj(v)).sum()))
        return float(np.sqrt((w * np.abs(u - v) ** 2).sum() / (w * np.abs(v) ** 2).sum()))
    nrmse = [phase_nrmse(a[0][s, 0], b[0][s, 0]) for s in range(nslice)]
    print("large multislice c64 against c128 per slice", nrmse,
          float(np.abs(a[1] - b[1]).max()), max_rel(a[2], b[2]))
    assert max(nrmse) < 1e-4
    assert np.abs(a[1] - b[1]).max() < 1e-2
    assert max_rel(a[2], b[2]) < 1e-5
