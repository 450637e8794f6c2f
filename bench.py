"""Benchmark of the per-position reconstruction step on one MI355X.

Workload (the headline configuration): rPIE, 256×256 probe with 4 incoherent Hermite–Gauss
modes, 1024×1024 object, 32×32 raster of 1024 positions (24 px step, ±2 px jitter), Poisson
data at 1e7 photons per pattern, rPIE's default parameters (RpieParams: steps 1, alphas 0.1,
batches of 100 positions, ptyrs/src/engine/ptycho.rs:145-186). A step is one rPIE iteration (epoch)
over all 1024 positions; `value` is positions·iterations per second of the complex-float
(FP32) engine with the scan resident in HBM. The plain complex-double (FP64) engine runs the
same epochs from the same start; `nrmse` compares the two objects after phase-ramp and
global-phase alignment on the illuminated region, and `nrmse.vs_oracle` one epoch of each
precision against the FP64 CPU restatement (oracle/cpu_rpie.cpp) from the same start. The line
also carries LSQML's and BH's rates and kernel rooflines and AP, DM, RAAR and ML rates on the
same scan.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
    python bench.py --algorithm lsqml|bh|ap|dm|raar|ml [--steps K]   # one algorithm alone
    python bench.py --multi-scan [--steps K]   # 16 uneven scans: sequential vs grouped epochs
    python bench.py --large-multislice [--steps K]   # 512² x 4, 4096² x 8 slices, refined positions
    python bench.py --copy-probe     # a plain 2 GiB device copy: the PMC calibration kernel and
                                     # the achievable bandwidth (run as a child by default)

Prints one JSON line.
"""

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12          # MI355X HBM3E, bytes/s
BATCH = 100               # RpieParams::default().batch_size
N, M, NPOS = 256, 4, 1024
# HBM bytes per dispatch of the dominant kernels at this workload, from rocprofv3 FETCH_SIZE and
# WRITE_SIZE passes scaled by the copy probe's readings (the file gives the commands).
TRAFFIC = os.path.join(ROOT, "profiles", "traffic_headline.txt")


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` (a kernel kind) from the committed counter summary."""
    try:
        with open(TRAFFIC) as f:
            for line in f:
                p = line.split()
                if len(p) == 3 and p[0] == kernel:
                    return float(p[1]) + float(p[2])
    except OSError:
        pass
    return None


def copy_probe(gib=2, reps=20):
    """A plain device copy of `gib` GiB (float32), the kernel whose FETCH_SIZE and WRITE_SIZE
    calibrate the counters (its bytes are known), and the achievable-bandwidth reference:
    returns its figures (GB/s counts the bytes read plus written)."""
    import torch
    n = gib * (1 << 30) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t) / reps
    del src, dst
    torch.cuda.empty_cache()
    return {"copy_gib": gib, "read_bytes": n * 4, "write_bytes": n * 4,
            "ms": round(dt * 1e3, 4), "gb_per_s": round(2 * n * 4 / dt / 1e9, 1)}


def copy_bandwidth():
    """`copy_probe` in a child process (torch's HIP runtime and the engine's are separate
    libraries, one per process), before this process touches the GPU; None if it fails."""
    import subprocess
    try:
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--copy-probe"],
                           capture_output=True, text=True, timeout=300)
        return json.loads(r.stdout.strip().splitlines()[-1]) if r.returncode == 0 else None
    except (subprocess.SubprocessError, ValueError, IndexError):
        return None


def aligned_nrmse(a, b, weight):
    """NRMSE of `a` against `b` after removing a linear phase ramp and a global phase
    (least squares, weighted by `weight`)."""
    w = weight / weight.sum()
    r = a * np.conj(b)
    # Phase ramp: weighted least-squares plane fit of the angle of a·conj(b).
    ny, nx = r.shape[-2:]
    yy, xx = np.mgrid[:ny, :nx]
    ph = np.angle(r[0])
    A = np.stack([np.ones(ny * nx), yy.ravel(), xx.ravel()], axis=1) * np.sqrt(w.ravel())[:, None]
    coef = np.linalg.lstsq(A, ph.ravel() * np.sqrt(w.ravel()), rcond=None)[0]
    a2 = a * np.exp(-1j * (coef[1] * yy + coef[2] * xx))
    g = (w * a2 * np.conj(b)).sum()
    a2 = a2 * np.exp(-1j * np.angle(g))
    return float(np.sqrt((w * np.abs(a2 - b) ** 2).sum() / (w * np.abs(b) ** 2).sum()))


def illumination(corners, probe, shape):
    """Σ over views of Σ_m |P_m|² placed at the views' corners: the NRMSE weight."""
    n = probe.shape[-1]
    ov = np.zeros(shape[-2:])
    p2 = (np.abs(probe) ** 2).sum(axis=0)
    for oy, ox in corners:
        ov[oy:oy + n, ox:ox + n] += p2
    return ov


def phase_nrmse(a, b, weight):
    """NRMSE of `a` against `b` after a global phase (weighted by `weight`): for two runs from
    the same start, where no phase ramp separates them."""
    w = weight / weight.sum()
    g = (w * a * np.conj(b)).sum()
    a2 = a * np.exp(-1j * np.angle(g))
    return float(np.sqrt((w * np.abs(a2 - b) ** 2).sum() / (w * np.abs(b) ** 2).sum()))


def cpu_threads():
    """The host threads this process may use: OMP_NUM_THREADS if set, else its CPU affinity."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    return int(env) if env.isdigit() and int(env) > 0 else len(os.sched_getaffinity(0))


def cpu_baseline(d, batch):
    """The C++ FP64 restatement of the reference's CPU rPIE (oracle/cpu_rpie.cpp, checked
    against the numpy oracle) on the same workload: one full epoch (every position, batches of
    `batch`, OpenMP over the host threads), and three batches on one thread."""
    from oracle import cpu
    from oracle.base import AddrTable, pynx_preprocess
    from oracle.optics import Optics
    optics = Optics.far_field(d["wavelength"], d["distance"], d["detector_pixel"], [N, N])
    py_, px_ = pynx_preprocess(d["y"], d["x"], optics.pixel)
    addr = AddrTable(py_, px_, optics.pixel, [N, N], d["start_obj"].shape[-2:], M, 1)
    corners = np.c_[addr.oy, addr.ox]
    threads = cpu_threads()
    c = cpu.CpuRpie(d["counts"], corners, d["start_obj"], d["start_probe"], optics.quad,
                    threads=threads)
    t = time.perf_counter()
    c.rpie(1, batch_size=batch)
    dt = time.perf_counter() - t
    oracle_state = dict(obj=c.obj[None].copy(), probe=c.probe.copy(), corners=corners)
    c1 = cpu.CpuRpie(d["counts"], corners, d["start_obj"], d["start_probe"], optics.quad,
                     threads=1)
    t = time.perf_counter()
    for s in range(3):
        c1.batch(range(s * batch, (s + 1) * batch))
    dt1 = time.perf_counter() - t
    return dict(value=round(NPOS / dt, 1), unit="positions·iter/s", cores=threads, kind="port",
                sample=f"one rPIE epoch ({NPOS} positions, batches of {batch}) at 256²×4 modes, "
                       f"1024² object, FP64, C++ with OpenMP on {threads} threads: {dt:.2f} s",
                single_thread=dict(value=round(3 * batch / dt1, 1), cores=1,
                                   sample=f"three rPIE batches of {batch} positions on one "
                                          f"thread: {dt1:.2f} s")), oracle_state


def against_oracle(pa, scan, d, params, state):
    """Object and probe NRMSE of one rPIE epoch of each GPU precision against the FP64 CPU
    restatement's epoch from the same start, corners and batches (the object after phase-ramp
    and global-phase alignment on the illuminated region, the probe after a global phase)."""
    weight = illumination(state["corners"], state["probe"], state["obj"].shape)
    out = {"oracle": "oracle/cpu_rpie.cpp (FP64), one epoch", "epochs": 1}
    for dtype in ("c64", "c128"):
        g = pa.GpuPtycho(scan, d["start_obj"], d["start_probe"], dtype=dtype)
        if not np.array_equal(g.corners, state["corners"]):
            raise RuntimeError("engine and oracle corners differ")
        g.rpie(1, **params)
        pw = np.ones(state["probe"].shape[-2:])
        out[dtype] = {"object": aligned_nrmse(g.object, state["obj"], weight),
                      "probe": float(np.sqrt(sum(phase_nrmse(g.probe[m], state["probe"][m], pw) ** 2
                                                 for m in range(M)) / M)),
                      "probe_rel_l2": float(np.linalg.norm(g.probe - state["probe"])
                                            / np.linalg.norm(state["probe"]))}
        del g
    return out


# Algorithmic bytes per position of each timed launch kind (c bytes per complex, r per real):
# the reads and writes each kernel's arithmetic needs, counted once.
def rpie_bytes(c, r):
    n = N * N
    return {
        "pass_a": n * c + M * n * c,                 # O window in, rows-transformed Ψ out
        "pass_b": 2 * M * n * c + 4 * n,             # Ψ in and out, float counts in
        # Ψ' and the overlaps D_m in, O window in, the mode-summed object term out
        "pass_c": M * n * c + M * n * r + n * c + n * c,
        "obj_update": n * c,                         # the summed term in (object RMW amortised)
    }


def lsqml_bytes(c, r):
    n = N * N
    return {
        "pass_a": n * c + M * n * c,
        # icalc pass: Ψ and counts in, icalc out; Poisson pass: Ψ and counts in, Ψ out
        "pass_b": 3 * M * n * c + 8 * n + n * r,
        "pass_c": 2 * M * n * c + n * c,             # Ψ' and the O window in, χ out
        "obj_dir": M * n * c,                        # χ of every view over the tile in
        "dots": M * n * c + 2 * n * c,               # χ, the O and dO windows in
    }


def bh_bytes(c, r, cg):
    """BH's algorithmic bytes per frame and kernel kind; `cg`: the Hessian passes carry the
    previous direction too (two fields)."""
    n = N * N
    nf = 2 if cg else 1
    return {
        "pass_a": n * c + M * n * c,
        "pass_b": 3 * M * n * c + 4 * n,             # Ψ in and out, U·Ψ kept, counts in
        "pass_c": 2 * M * n * c + n * c,             # Ψ' and the O window in, r out
        "obj_dir": M * n * c,                        # r of every view over the tile in
        # the O, u_O (and v_O) windows and r in, L_u (and L_v) row-transformed out
        "hess_a": (1 + nf) * n * c + (1 + nf) * M * n * c,
        "hess_b": (nf + 1) * M * n * c + 4 * n,      # L fields and U·Ψ in, counts in
    }


def kernel_roofline(stats, per_frame, positions):
    """Roofline of the launch kind with the largest total time: its algorithmic bytes per
    launch (`per_frame` bytes × the positions one launch covers, `positions` over all its
    launches) over its mean launch time from HIP events."""
    kern = max(stats, key=lambda k: stats[k][1])
    launches, ms = stats[kern]
    frames_per_launch = positions / max(launches, 1)
    alg_bytes = per_frame.get(kern, 0) * frames_per_launch
    avg_s = ms / 1e3 / max(launches, 1)
    achieved = alg_bytes / avg_s / 1e9 if alg_bytes else None
    return {"kernel": kern, "bound": "hbm",
            "achieved": round(achieved, 1) if achieved else None,
            "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": round(achieved * 1e9 / HBM_PEAK, 4) if achieved else None,
            "alg_bytes_per_launch": alg_bytes, "avg_launch_ms": round(avg_s * 1e3, 4)}


def algorithm_rates(pa, scan, d, steps=3, algorithms=("ap", "dm", "raar", "ml"),
                    full_object=False):
    """AP, DM, RAAR and ML (their default parameters, no LLK records) on the same scan:
    positions·cycle/s of `steps` cycles after one warm-up cycle, complex64 and complex128.
    With `full_object` every frame reads its object window sampled at its corner plus a held
    sub-pixel offset (FullObject sampling; offsets uniform in ±0.45 px, seed 5)."""
    out = {"steps": steps}
    if full_object:
        out["sampling"] = "full_object, held offsets uniform in ±0.45 px"
    offs = np.random.default_rng(5).uniform(-0.45, 0.45, (NPOS, 2))
    for alg in algorithms:
        r = {}
        for dtype in ("c64", "c128"):
            g = pa.GpuPtycho(scan, d["start_obj"], d["start_probe"], dtype=dtype)
            if full_object:
                g.set_sampling("full_object")
                g.set_features(position=offs, hold=("position",))
            getattr(g, alg)(1)
            g.synchronize()
            t = time.perf_counter()
            getattr(g, alg)(steps)
            g.synchronize()
            dt = time.perf_counter() - t
            r["value" if dtype == "c64" else "fp64"] = round(NPOS * steps / dt, 1)
            del g
        r["speedup_fp32_vs_fp64"] = round(r["value"] / r["fp64"], 3)
        if alg in ("dm", "raar") and not full_object:
            # GpuOptions::psi_f16: the carried exit waves as f16 pairs.
            g = pa.GpuPtycho(scan, d["start_obj"], d["start_probe"], psi_f16=True)
            getattr(g, alg)(1)
            g.synchronize()
            t = time.perf_counter()
            getattr(g, alg)(steps)
            g.synchronize()
            r["psi_f16"] = round(NPOS * steps / (time.perf_counter() - t), 1)
            del g
        out[alg] = r
        print(f"[bench] {alg}: {r}", file=sys.stderr, flush=True)
    return out


def bh_rates(pa, scan, d, steps=3):
    """BH (pty-chi's bilinear-Hessian method, probe ρ 0.01) on the same scan: batches of 100
    (steepest descent) and one batch of all 1024 frames (conjugate gradient from the second
    epoch), positions·iter/s of `steps` epochs after one warm-up, complex64 and complex128; the
    conjugate-gradient run's kernel times and the roofline of its dominant kernel."""
    out = {"probe_rho": 0.01, "steps": steps}
    for name, batch in (("batches_100", 100), ("cg_all_frames", NPOS)):
        r = {}
        for dtype in ("c64", "c128"):
            g = pa.GpuPtycho(scan, d["start_obj"], d["start_probe"], dtype=dtype)
            g.bh(1, 0, batch, 0.01, 7)
            g.synchronize()
            t = time.perf_counter()
            g.bh(steps, 0, batch, 0.01, 7)
            g.synchronize()
            dt = time.perf_counter() - t
            r["value" if dtype == "c64" else "fp64"] = round(NPOS * steps / dt, 1)
            r["llk_poisson_" + dtype] = g.llk()["poisson"]
            if dtype == "c64" and batch == NPOS:
                g.profile(True)
                g.bh(2, 0, batch, 0.01, 7)
                g.synchronize()
                stats = g.kernel_stats()
                g.profile(False)
                r["kernels_ms_per_step"] = {k: round(v[1] / 2, 4) for k, v in stats.items()
                                            if v[0]}
                # the second epoch of the call is a conjugate-gradient one, the first is not:
                # the roofline counts the two-field Hessian passes for both.
                r["roofline"] = kernel_roofline(stats, bh_bytes(8, 4, True), NPOS * 2)
            del g
        out[name] = r
        print(f"[bench] bh {name}: {r}", file=sys.stderr, flush=True)
    return out


def lsqml_rates(pa, scan, d, a, steps=5):
    """LSQML (Poisson, the same batches) on the same scan: positions·iter/s of `steps` epochs
    after one warm-up, complex64 and complex128, and the LLK of both after them."""
    out = {"cost": "poisson", "batch": a.batch, "steps": steps}
    for dtype in ("c64", "c128"):
        g = pa.GpuPtycho(scan, d["start_obj"], d["start_probe"], dtype=dtype)
        g.lsqml(1, 0, a.batch, "poisson", 7)
        g.synchronize()
        t = time.perf_counter()
        g.lsqml(steps, 0, a.batch, "poisson", 7)
        g.synchronize()
        dt = time.perf_counter() - t
        out["value" if dtype == "c64" else "fp64"] = round(NPOS * steps / dt, 1)
        out["llk_poisson_" + dtype] = g.llk()["poisson"]
        if dtype == "c64":
            g.profile(True)
            g.lsqml(1, 0, a.batch, "poisson", 7)
            g.synchronize()
            stats = g.kernel_stats()
            g.profile(False)
            out["kernels_ms_per_step"] = {k: round(v[1], 4) for k, v in stats.items() if v[0]}
            out["roofline"] = kernel_roofline(stats, lsqml_bytes(8, 4), NPOS)
    return out


def multi_scan(epochs=3, nscan=16, seed=404):
    """Sixteen independent seeded scans of uneven sizes (probe 128² or 256², 1–4 Hermite–Gauss
    modes, 200–3000 positions on a Fermat spiral at half the probe's FWHM, objects covering
    them, Poisson data at 1e6–1e8 photons from the engine's own model), rPIE epochs of all of
    them: one engine after another (each call waits for its device work), then all at once
    (ptyrs_group_rpie, every engine's epochs on its own stream). Prints one JSON line."""
    import ptyrs_amd as pa
    from ptyrs_amd import sim
    rng = np.random.default_rng(seed)
    scans, t0 = [], time.perf_counter()
    for k in range(nscan):
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
        inten *= photons / inten.sum(axis=(1, 2)).mean()
        counts = np.random.default_rng(seed + 100 + k).poisson(inten).astype(np.float32)
        del inten
        start = np.exp(1j * np.random.default_rng(seed + 200 + k).uniform(-1, 1, obj.shape))
        scale = math.sqrt(photons / (np.abs(probe) ** 2).sum())
        g = pa.GpuPtycho(pa.Scan(counts, y, x, *geom), start, probe * scale)
        scans.append(dict(engine=g, n=n, modes=nmode, positions=npos, photons=photons))
        print(f"[bench] scan {k}: {n}² x {nmode} modes, {npos} positions", file=sys.stderr,
              flush=True)
    print(f"[bench] 16 scans ready in {time.perf_counter() - t0:.1f} s", file=sys.stderr,
          flush=True)
    engines = [s["engine"] for s in scans]
    total = sum(s["positions"] for s in scans)
    for g in engines:  # warm-up: workspaces and epoch graphs
        g.rpie(1)
    t = time.perf_counter()
    for g in engines:
        g.rpie(epochs)
    seq = time.perf_counter() - t
    pa.rpie_group(engines, 1)
    t = time.perf_counter()
    pa.rpie_group(engines, epochs)
    grp = time.perf_counter() - t
    # Every engine's epoch captured in one graph (a branch per engine), replayed per epoch.
    pa.rpie_group(engines, 1, one_graph=True)
    t = time.perf_counter()
    pa.rpie_group(engines, epochs, one_graph=True)
    one = time.perf_counter() - t
    # The autotuner's choice of how many scans share the device (one epoch per candidate),
    # then `epochs` epochs at that width.
    from ptyrs_amd import tune
    tuned = tune.tune_group(engines, cache=False)
    t = time.perf_counter()
    tune.run_waves(engines, epochs, width=tuned["width"], one_graph=tuned["one_graph"])
    tun = time.perf_counter() - t
    best = min(grp, tun)
    out = {"metric": "aggregate positions·iter/s, 16 independent scans of uneven sizes (rPIE)",
           "value": round(total * epochs / best, 1), "unit": "positions·iter/s",
           "sequential": round(total * epochs / seq, 1), "speedup": round(seq / best, 3),
           "all_at_once": round(total * epochs / grp, 1),
           "one_graph": round(total * epochs / one, 1),
           "tuned": {"width": tuned["width"], "one_graph": tuned["one_graph"],
                     "value": round(total * epochs / tun, 1),
                     "epoch_seconds": {k: round(v, 4) for k, v in tuned["seconds"].items()}},
           "epochs": epochs, "positions": total, "dtype": "c64", "data": "synthetic",
           "launch": "ptyrs_group_rpie: each engine's epoch a hipGraph replay on its own stream",
           "scans": [{k: s[k] for k in ("n", "modes", "positions")} for s in scans]}
    print(json.dumps(out), flush=True)


def large_multislice(epochs=2, npos=5000, seed=505):
    """The large multislice configuration: 512² probe × 4 modes, 4096² object × 8 slices
    (exp(iφ − 0.05φ), φ ≤ 0.1 rad) at 1 µm, 10 keV, 5 nm pixels, `npos` positions on a Fermat
    spiral (25 px apart) whose model positions start perturbed by N(0, 3 px); Poisson data from
    the engine's own model at the true positions. LSQML (Poisson, batches of 100) with every
    position refined, slice waves checkpointed every 2 slices, `epochs` epochs in complex64 and
    in complex128 from the same start. Prints one JSON line: positions·iter/s of both, the
    position RMSE against the truth before and after, the LLK, and the complex64 object's NRMSE
    against the complex128 one (aligned, illuminated region)."""
    import ptyrs_amd as pa
    from ptyrs_amd import sim
    n, side, nslice, pixel = 512, 4096, 8, 5e-9
    lam = 1.2398e-10
    dpix = 75e-6
    z = pixel * n * dpix / lam
    probe = sim.hermite_gauss_probe(n) * 1e3
    phi = np.angle(sim.random_object(side, side, seed, amp=(1.0, 1.0), phase=(0.0, 0.1)))
    objects = np.stack([np.exp(1j * np.roll(phi, (37 * k, 53 * k), axis=(-2, -1))
                               - 0.05 * np.roll(phi, (37 * k, 53 * k), axis=(-2, -1)))
                        for k in range(nslice)])
    del phi
    y, x = sim.fermat_positions(npos, 25.0, pixel)
    optics = pa.Optics.far_field(lam, z, dpix, [n, n]).with_slices(
        [pa.Wavefield([n, n], lam, [pixel, pixel], 1e-6)] * (nslice - 1))
    t0 = time.perf_counter()
    truth = pa.GpuPtycho(pa.Scan(np.zeros((npos, n, n), np.float32), y, x, lam, z, dpix),
                         objects, probe, optics=optics)
    inten = truth.intensities()
    del truth
    counts = np.empty(inten.shape, np.float32)
    rng = np.random.default_rng(seed + 1)
    for f0 in range(0, npos, 250):
        counts[f0:f0 + 250] = rng.poisson(inten[f0:f0 + 250])
    del inten
    print(f"[bench] large multislice data ready in {time.perf_counter() - t0:.1f} s", file=sys.stderr,
          flush=True)
    d0 = np.random.default_rng(seed + 2).normal(0.0, 3.0, (npos, 2))
    scan = pa.Scan(counts, y, x, lam, z, dpix)
    out = {"metric": "positions/sec/iter, multislice LSQML with position refinement "
                     "(512² x 4 modes, 4096² x 8 slices, 5000 positions)",
           "unit": "positions·iter/s", "epochs": epochs, "data": "synthetic",
           "position_rmse_start_px": float(np.sqrt((d0 ** 2).mean()))}
    res = {}
    for dtype in ("c64", "c128"):
        g = pa.GpuPtycho(scan, objects, probe, optics=optics, dtype=dtype)
        g.set_checkpoint(2)
        g.set_features(position=d0)
        l0 = g.llk()["poisson"]
        g.synchronize()
        t = time.perf_counter()
        g.lsqml(epochs, 0, 100, "poisson", 7)
        g.synchronize()
        dt = time.perf_counter() - t
        dpos = g.features["position"]
        res[dtype] = g
        out[dtype] = {"value": round(npos * epochs / dt, 1), "seconds": round(dt, 3),
                      "llk_start": l0, "llk_end": g.llk()["poisson"],
                      "position_rmse_px": float(np.sqrt((dpos ** 2).mean())),
                      "wave_bytes": g.checkpoint[1]}
        print(f"[bench] large multislice {dtype}: {epochs} epochs in {dt:.2f} s", file=sys.stderr,
              flush=True)
    o32, o64 = res["c64"].objects, res["c128"].objects
    weight = illumination(res["c128"].corners, res["c128"].probe, o64.shape[-2:])
    out["nrmse_objects"] = [phase_nrmse(o32[k, 0], o64[k, 0], weight) for k in range(nslice)]
    out["value"] = out["c64"]["value"]
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--steps64", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe", action="store_true")
    ap.add_argument("--multi-scan", action="store_true",
                    help="the sixteen-scan workload instead: sequential vs grouped epochs")
    ap.add_argument("--large-multislice", action="store_true",
                    help="the large multislice workload with position refinement instead")
    ap.add_argument("--no-copy", action="store_true",
                    help="skip the child copy-bandwidth process (counter passes)")
    ap.add_argument("--batch", type=int, default=BATCH, help="rPIE batch size (positions)")
    ap.add_argument("--full-object", action="store_true",
                    help="with --algorithm ap|dm|raar|ml: FullObject sampling at held offsets")
    ap.add_argument("--algorithm", choices=("lsqml", "bh", "ap", "dm", "raar", "ml"), default=None,
                    help="time one other algorithm alone on the headline scan (for profiling)")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the reconstruction step is not sharded across "
              "GPUs (--gpus 1)", file=sys.stderr)
        sys.exit(2)
    if a.copy_probe:
        print(json.dumps(copy_probe()), flush=True)
        return
    if a.multi_scan:
        multi_scan(a.steps)
        return
    if a.large_multislice:
        large_multislice(a.steps)
        return

    copy = None if (a.no_copy or a.algorithm) else copy_bandwidth()

    import ptyrs_amd as pa
    from ptyrs_amd import sim

    t0 = time.perf_counter()
    d = sim.scan("headline")
    print(f"[bench] synthetic scan ready in {time.perf_counter() - t0:.1f} s", file=sys.stderr,
          flush=True)
    scan = pa.Scan(d["counts"], d["y"], d["x"], d["wavelength"], d["distance"],
                   d["detector_pixel"])
    params = dict(batch_size=a.batch, obj_step=1.0, probe_step=1.0, obj_alpha=0.1, probe_alpha=0.1,
                  seed=7)
    if a.algorithm == "lsqml":
        print(json.dumps({"workload": "headline scan", "lsqml": lsqml_rates(pa, scan, d, a, a.steps)}),
              flush=True)
        return
    if a.algorithm == "bh":
        print(json.dumps({"workload": "headline scan", "bh": bh_rates(pa, scan, d, a.steps)}),
              flush=True)
        return
    if a.algorithm is not None:
        out = algorithm_rates(pa, scan, d, a.steps, (a.algorithm,), a.full_object)
        print(json.dumps({"workload": "headline scan", **out}), flush=True)
        return

    # Timed: captured epochs (hipGraph replays), no per-kernel events. Then a few more epochs
    # launched eagerly under HIP events on the engine's stream for the per-kernel times.
    results = {}
    prof_steps = min(a.steps, 3)
    for dtype, steps in (("c64", a.steps), ("c128", a.steps64)):
        g = pa.GpuPtycho(scan, d["start_obj"], d["start_probe"], dtype=dtype)
        g.rpie(a.warmup, **params)
        g.synchronize()
        t = time.perf_counter()
        g.rpie(steps, **params)
        g.synchronize()
        dt = time.perf_counter() - t
        stats = None
        if dtype == "c64":
            g.profile(True)
            g.rpie(prof_steps, **params)
            g.synchronize()
            stats = g.kernel_stats()
            g.profile(False)
        results[dtype] = dict(engine=g, seconds=dt, steps=steps, stats=stats)
        print(f"[bench] {dtype}: {steps} epochs in {dt:.3f} s", file=sys.stderr, flush=True)

    # Same number of epochs from the same start for the NRMSE comparison.
    g32, g64 = results["c64"]["engine"], results["c128"]["engine"]
    extra = a.steps + prof_steps - a.steps64
    if extra > 0:
        g64.rpie(extra, **params)
    o32, o64 = g32.object, g64.object
    weight = illumination(g64.corners, g64.probe, o64.shape)
    nrmse_obj = aligned_nrmse(o32, o64, weight)
    p32, p64 = g32.probe, g64.probe
    nrmse_probe = float(np.linalg.norm(p32 - p64) / np.linalg.norm(p64))
    llk32, llk64 = g32.llk()["poisson"], g64.llk()["poisson"]

    r32, r64 = results["c64"], results["c128"]
    pos32 = NPOS * r32["steps"] / r32["seconds"]
    pos64 = NPOS * r64["steps"] / r64["seconds"]

    # Dominant kernel: the largest total time under HIP events over the timed region.
    stats = r32["stats"]
    roof = kernel_roofline(stats, rpie_bytes(8, 4), NPOS * prof_steps)
    roof["traffic"] = measured_traffic(roof["kernel"]) if a.batch == BATCH else None
    achieved = roof["achieved"]

    out = {
        "metric": "positions/sec/iter at 256² probe ×4 modes: FP32 vs FP64 plain at fixed NRMSE",
        "value": round(pos32, 1),
        "unit": "positions·iter/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(1e3 * r32["seconds"] / r32["steps"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic",
        "config": {"workload": "rPIE epoch, 256x256 probe x 4 modes, 1024 positions (32x32 raster), "
                               f"1024x1024 object, batch {a.batch}, 1e7 photons/pattern",
                   "global_batch": a.batch, "positions": NPOS, "probe": [M, N, N],
                   "object": [1024, 1024]},
        "fp64": {"value": round(pos64, 1), "ms_per_step": round(1e3 * r64["seconds"] / r64["steps"], 3),
                 "steps": r64["steps"]},
        "speedup_fp32_vs_fp64": round(pos32 / pos64, 3),
        "nrmse": {"object": nrmse_obj, "probe": nrmse_probe,
                  "epochs": a.warmup + a.steps + prof_steps,
                  "llk_poisson_c64": llk32, "llk_poisson_c128": llk64},
        "epoch_launch": "hipGraph replay (kernel times from eager epochs under HIP events)",
        "kernels_ms_per_step": {k: round(v[1] / prof_steps, 4) for k, v in stats.items()},
        "roofline": roof,
    }
    if copy:
        out["roofline"]["copy_gb_per_s"] = copy["gb_per_s"]
        if achieved:
            out["roofline"]["frac_of_copy"] = round(achieved / copy["gb_per_s"], 4)
    for k in ("c64", "c128"):
        del results[k]["engine"]
    del g, g32, g64
    out["lsqml"] = lsqml_rates(pa, scan, d, a)
    out["algorithms"] = algorithm_rates(pa, scan, d)
    out["bh"] = bh_rates(pa, scan, d)
    if not a.no_cpu:
        out["cpu_baseline"], state = cpu_baseline(d, a.batch)
        out["nrmse"]["vs_oracle"] = against_oracle(pa, scan, d, params, state)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
