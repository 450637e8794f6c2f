"""The reference's seeded far-field test scenario, restated: ChaCha8 Poisson sampling, the
simulated object and probe, Fermat and raster scans, the far-field simulation, and the parity
start (`dm_scenario`) its f64 fixtures were written from.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py). Third-party arithmetic restated from the
published algorithms of the pinned crates (Cargo.lock): rand_core 0.6.4 `seed_from_u64`
(PCG32 output steps), rand_chacha 0.3.1 `ChaCha8Rng` (D. J. Bernstein's ChaCha with 8 rounds,
64-bit block counter and 64-bit stream id, four blocks per refill), rand 0.8 `Standard` f64
(53 high bits), rand_distr 0.4.3 `Poisson` (Knuth's product method below λ = 12, Numerical
Recipes' Cauchy rejection with its Lanczos `log_gamma` above) and `Cauchy` (tan(π·u)). The
reference's call sites: ptyrs/src/base/sim.rs:18-160, ptyrs/src/engine/sim.rs:1-116,
ptyrs/tests/common/engine/scenario.rs:20-77.
"""

import math

import numpy as np

from .base import AddrTable, pynx_preprocess, required_obj_shape, shift_index
from .optics import fft2
from .ptycho import far_field_ptycho, obj_probe_to_psi

M64 = (1 << 64) - 1
M32 = (1 << 32) - 1


def seed_from_u64(state):
    """rand_core 0.6 `SeedableRng::seed_from_u64`: 32 seed bytes from PCG32 output steps, each
    after advancing `state` by the PCG multiplier and increment."""
    mul, inc = 6364136223846793005, 11634580027462260723
    out = bytearray()
    for _ in range(8):
        state = (state * mul + inc) & M64
        xorshifted = (((state >> 18) ^ state) >> 27) & M32
        rot = state >> 59
        x = ((xorshifted >> rot) | (xorshifted << ((32 - rot) & 31))) & M32
        out += x.to_bytes(4, "little")
    return bytes(out)


def chacha_blocks(key, counter, nblocks, rounds):
    """`nblocks` ChaCha keystream blocks (16 little-endian u32 words each) from block `counter`
    on, stream id 0: the state is the constants "expand 32-byte k", the key's 8 words, a
    64-bit counter (words 12, 13) and a 64-bit stream id (words 14, 15)."""
    k = np.frombuffer(key, dtype="<u4").astype(np.uint32)
    ctr = np.arange(counter, counter + nblocks, dtype=np.uint64)
    s = np.zeros((16, nblocks), dtype=np.uint32)
    s[0:4] = np.array([0x61707865, 0x3320646E, 0x79622D32, 0x6B206574], np.uint32)[:, None]
    s[4:12] = k[:, None]
    s[12] = (ctr & np.uint64(M32)).astype(np.uint32)
    s[13] = (ctr >> np.uint64(32)).astype(np.uint32)
    x = s.copy()

    def rotl(v, c):
        return (v << np.uint32(c)) | (v >> np.uint32(32 - c))

    def qr(a, b, c, d):
        x[a] += x[b]; x[d] = rotl(x[d] ^ x[a], 16)
        x[c] += x[d]; x[b] = rotl(x[b] ^ x[c], 12)
        x[a] += x[b]; x[d] = rotl(x[d] ^ x[a], 8)
        x[c] += x[d]; x[b] = rotl(x[b] ^ x[c], 7)

    for _ in range(rounds // 2):
        qr(0, 4, 8, 12); qr(1, 5, 9, 13); qr(2, 6, 10, 14); qr(3, 7, 11, 15)
        qr(0, 5, 10, 15); qr(1, 6, 11, 12); qr(2, 7, 8, 13); qr(3, 4, 9, 14)
    return (x + s).T.copy()  # [nblocks, 16]


class ChaCha8:
    """rand_chacha 0.3.1 `ChaCha8Rng::seed_from_u64(seed)` read through `next_u64`: its block
    buffer holds four consecutive blocks, and at every buffer edge `next_u64` still takes two
    consecutive words of the keystream, low word first, so the u64 stream is the keystream's
    word pairs in order."""

    CHUNK = 4096  # blocks generated at a time

    def __init__(self, seed):
        self.key = seed_from_u64(seed & M64)
        self.counter = 0
        self.buf = np.empty(0, dtype=np.uint64)
        self.pos = 0

    def next_u64(self):
        if self.pos >= self.buf.size:
            w = chacha_blocks(self.key, self.counter, self.CHUNK, 8).ravel().astype(np.uint64)
            self.counter += self.CHUNK
            self.buf = w[0::2] | (w[1::2] << np.uint64(32))
            self.pos = 0
        v = int(self.buf[self.pos])
        self.pos += 1
        return v

    def uniform(self):
        """rand 0.8 `Standard` for f64: the 53 high bits over 2^53, in [0, 1)."""
        return (self.next_u64() >> 11) * (1.0 / (1 << 53))


_LG = (76.18009172947146, -86.50532032941677, 24.01409824083091, -1.231739572450155,
       0.1208650973866179e-2, -0.5395239384953e-5)


def log_gamma(x):
    """rand_distr 0.4.3 `utils::log_gamma`: the Lanczos series of Numerical Recipes' `gammln`."""
    tmp = x + 5.5
    log = (x + 0.5) * math.log(tmp) - tmp
    a = 1.000000000190015
    denom = x
    for c in _LG:
        denom = denom + 1.0
        a = a + c / denom
    return log + math.log(2.5066282746310005 * a / x)


class Poisson:
    """rand_distr 0.4.3 `Poisson::new(λ)` and its `sample`."""

    def __init__(self, lam):
        if not lam > 0.0:
            raise ValueError("Poisson needs λ > 0")
        self.lam = lam
        self.exp_lambda = math.exp(-lam)
        self.log_lambda = math.log(lam)
        self.sqrt_2lambda = math.sqrt(2.0 * lam)
        self.magic_val = lam * self.log_lambda - log_gamma(1.0 + lam)

    def sample(self, rng):
        if self.lam < 12.0:
            # Knuth: the count of uniforms whose running product stays above e^-λ.
            result = 1.0
            p = rng.uniform()
            while p > self.exp_lambda:
                p = p * rng.uniform()
                result = result + 1.0
            return result - 1.0
        while True:
            while True:
                comp_dev = math.tan(math.pi * rng.uniform())
                result = self.sqrt_2lambda * comp_dev + self.lam
                if result >= 0.0:
                    break
            result = math.floor(result)
            check = 0.9 * (1.0 + comp_dev * comp_dev) * math.exp(
                result * self.log_lambda - log_gamma(1.0 + result) - self.magic_val)
            if rng.uniform() <= check:
                return result


def poisson(intensity, seed, detector_mask=None):
    """`base::sim::poisson`: every value of `intensity` (`[nframe, ny, nx]`, detector order)
    Poisson-sampled in order from one ChaCha8 stream seeded by `seed`, 0 where λ <= 0, MASKED
    (65535) where `detector_mask` (`[ny, nx]`) is set. ptyrs/src/base/sim.rs:127-160"""
    inten = np.asarray(intensity, dtype=np.float64)
    plane = inten.shape[1] * inten.shape[2]
    mask = None if detector_mask is None else np.asarray(detector_mask, bool).ravel()
    rng = ChaCha8(seed)
    out = np.empty(inten.size, dtype=np.uint16)
    for i, lam in enumerate(inten.ravel().tolist()):
        if mask is not None and mask[i % plane]:
            out[i] = 65535
            continue
        v = Poisson(lam).sample(rng) if lam > 0.0 else 0.0
        if v > 65534:
            raise ValueError(f"count {v} at {i} past the u16 path")
        out[i] = int(v)
    return out.reshape(inten.shape)


def fermat(n, step):
    """`sim::fermat`: r_k = (step/√π)·√k, θ_k = k·π(3 − √5); (y, x) = r(sin θ, cos θ).
    ptyrs/src/base/sim.rs:30-48"""
    golden = math.pi * (3.0 - math.sqrt(5.0))
    c = step / math.sqrt(math.pi)
    y, x = [], []
    for k in range(n):
        r, t = c * math.sqrt(float(k)), float(k) * golden
        y.append(r * math.sin(t))
        x.append(r * math.cos(t))
    return np.array(y), np.array(x)


def raster(ny, nx, step):
    """`sim::raster`: y-major, centred on 0. ptyrs/src/base/sim.rs:13-28"""
    y = [(iy - (ny - 1.0) / 2.0) * step for iy in range(ny) for _ in range(nx)]
    x = [(ix - (nx - 1.0) / 2.0) * step for _ in range(ny) for ix in range(nx)]
    return np.array(y), np.array(x)


def _smoothstep(e0, e1, v):
    t = min(max((v - e0) / (e1 - e0), 0.0), 1.0)
    return t * t * (3.0 - 2.0 * t)


def sim_object(shape):
    """`sim::object`: a Siemens star and three Gaussian features in phase and amplitude,
    `[1, ny, nx]`. ptyrs/src/base/sim.rs:55-83"""
    ny, nx = shape
    s = min(ny, nx) / 2.0

    def blob(u, v, cu, cv, w):
        return math.exp(-((u - cu) * (u - cu) + (v - cv) * (v - cv)) / (2.0 * w * w))

    out = np.empty((ny, nx), dtype=np.complex128)
    for iy in range(ny):
        for ix in range(nx):
            u, v = (ix - nx / 2.0) / s, (iy - ny / 2.0) / s
            r = math.hypot(u, v)
            star = (0.5 * (1.0 + math.tanh(3.0 * math.cos(12.0 * math.atan2(v, u))))
                    * _smoothstep(0.08, 0.12, r) * (1.0 - _smoothstep(0.55, 0.6, r)))
            phase = star + 0.6 * blob(u, v, 0.55, -0.45, 0.12) - 0.4 * blob(u, v, -0.5, 0.5, 0.1)
            amp = 0.95 - 0.3 * blob(u, v, -0.45, -0.55, 0.15) - 0.15 * star
            out[iy, ix] = complex(amp * math.cos(phase), amp * math.sin(phase))
    return out[None]


def _hermite(n, t):
    return 1.0 if n == 0 else (2.0 * t if n == 1 else 4.0 * t * t - 2.0)


def sim_probe(frame, nmode):
    """`sim::probe`: Hermite–Gauss orders (0,0), (1,0), (0,1), (1,1), (2,0), (0,2) on one
    envelope σ = n/12 with phase 1.5r² + 0.3k, mode k of power 4^-k, total 1.
    ptyrs/src/base/sim.rs:93-125"""
    orders = [(0, 0), (1, 0), (0, 1), (1, 1), (2, 0), (0, 2)]
    if not 1 <= nmode <= len(orders):
        raise ValueError("probe modes must be 1..=6")
    ny, nx = frame
    sigma = min(ny, nx) / 12.0
    weights = [0.25 ** k for k in range(nmode)]
    wsum = sum(weights)
    out = np.empty((nmode, ny, nx), dtype=np.complex128)
    for k, (a, b) in enumerate(orders[:nmode]):
        plane = np.empty(ny * nx, dtype=np.complex128)
        for iy in range(ny):
            for ix in range(nx):
                u, v = (ix - nx / 2.0) / sigma, (iy - ny / 2.0) / sigma
                r2 = u * u + v * v
                amp = _hermite(a, u) * _hermite(b, v) * math.exp(-r2 / 2.0)
                ph = 1.5 * r2 + 0.3 * k
                plane[iy * nx + ix] = complex(amp * math.cos(ph), amp * math.sin(ph))
        norm = math.sqrt(sum(c.real * c.real + c.imag * c.imag for c in plane.tolist()))
        scale = math.sqrt(weights[k] / wsum) / norm
        out[k] = (plane * scale).reshape(ny, nx)
    return out


def pixel_size(wavelength, distance, detector_pixel, frame):
    lz = wavelength * distance
    return [lz / (frame[0] * detector_pixel), lz / (frame[1] * detector_pixel)]


def simulate(wavelength, distance, detector_pixel, y, x, obj, probe, photons, seed,
             detector_mask=None):
    """`engine::sim::simulate`: `Σ_modes |FT(P·O·q)|²` per frame in detector order, scaled so
    the mean frame holds `photons`, Poisson-sampled with `seed`. Returns the u16 counts and the
    scale. ptyrs/src/engine/sim.rs:62-116"""
    frame = list(probe.shape[-2:])
    pixel = pixel_size(wavelength, distance, detector_pixel, frame)
    py_, px_ = pynx_preprocess(y, x, pixel)
    addr = AddrTable(py_, px_, pixel, frame, obj.shape[-2:], probe.shape[0], obj.shape[0])
    from .optics import Optics
    quad = Optics.far_field(wavelength, distance, detector_pixel, frame).quad
    psi = fft2(obj_probe_to_psi(addr, addr.views(), obj, probe, quad))
    ny, nx = frame
    n = ny * nx
    nmode = addr.nmode
    far = (psi.real ** 2 + psi.imag ** 2).reshape(addr.nframe, nmode, n).sum(axis=1)
    idx = np.array([shift_index(i, ny, nx) for i in range(n)])
    inten = far[:, idx]
    scale = photons / (inten.sum() / addr.nframe)
    inten = (inten * scale).reshape(addr.nframe, ny, nx)
    return poisson(inten, seed, detector_mask), scale


GEOMETRY = (1.5e-10, 1.0, 75e-6)


def dm_scenario(frame, npos, nprobe, photons, step):
    """The parity scenario (`dm_scenario`): `npos` Fermat positions `step` object pixels apart,
    rounded to f32 as CXI stores them, the simulated object covering them (margin 16), the
    simulated probe, counts drawn with seed 1; the start object at unit modulus with xorshift
    phases in [-0.5, 0.5), the start probe the truth with its defocus phase cut from 1.5 r²/σ²
    to 1.0. Returns (counts u16, y, x, start object, start probe).
    ptyrs/tests/common/engine/scenario.rs:20-77"""
    wl, z, dpix = GEOMETRY
    pixel = pixel_size(wl, z, dpix, [frame, frame])
    fy, fx = fermat(npos, step * pixel[1])
    y = fy.astype(np.float32).astype(np.float64)
    x = fx.astype(np.float32).astype(np.float64)
    py_, px_ = pynx_preprocess(y, x, pixel)
    shape = required_obj_shape(py_, px_, pixel, [frame, frame], 16)
    truth_obj = sim_object(shape)
    truth_probe = sim_probe([frame, frame], nprobe)
    counts, _ = simulate(wl, z, dpix, y, x, truth_obj, truth_probe, photons, 1)
    state = 0x9E3779B97F4A7C15
    ph = np.empty(truth_obj.size)
    for i in range(truth_obj.size):
        state ^= (state << 13) & M64
        state ^= state >> 7
        state ^= (state << 17) & M64
        ph[i] = (state >> 11) / float(1 << 53) - 0.5
    start_obj = (np.cos(ph) + 1j * np.sin(ph)).reshape(truth_obj.shape)
    sigma = frame / 12.0
    i = np.arange(frame * frame)
    u = ((i % frame) - frame / 2.0) / sigma
    w = ((i // frame) - frame / 2.0) / sigma
    cut = np.array([complex(math.cos(-0.5 * (a * a + b * b)), math.sin(-0.5 * (a * a + b * b)))
                    for a, b in zip(u.tolist(), w.tolist())]).reshape(frame, frame)
    start_probe = truth_probe * cut[None]
    return counts, y, x, start_obj, start_probe


def parity_start(nprobe=1, frame=32, npos=40, photons=2e5, step=4.0):
    """The oracle's `Ptycho` on the scenario after `scale_obj_probe`
    (ptyrs/tests/engine_reconstruct.rs:153-159), with the scenario's inputs."""
    counts, y, x, obj, probe = dm_scenario(frame, npos, nprobe, photons, step)
    wl, z, dpix = GEOMETRY
    p = far_field_ptycho(counts, y, x, obj, probe, wl, z, dpix)
    p.scale_obj_probe()
    return p, (counts, y, x, obj, probe)
