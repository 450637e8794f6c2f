"""Data model of the oracle: the FFT shift map, view corners and the address table, PyNX
position preprocessing, observed intensities and seeded random batches.

Test infrastructure (see oracle/__init__.py). Each function cites the reference it restates.
"""

import numpy as np

MASKED = 65535  # ptyrs/src/base/counts.rs:3-7
MAX_COUNT = 65534
U64 = (1 << 64) - 1


def shift_index(j, ny, nx):
    """Natural index of shifted index `j` (fftshift on both axes, its own inverse for even
    sides). ptyrs/src/engine/obs.rs:3-11"""
    y, x = j // nx, j % nx
    return ((y + ny // 2) % ny) * nx + (x + nx // 2) % nx


def to_shifted(nat):
    """Natural plane(s) `[..., ny, nx]` to the shifted layout: `out[j] = nat[shift(j)]`."""
    ny, nx = nat.shape[-2:]
    return np.roll(nat, (-(ny // 2), -(nx // 2)), axis=(-2, -1))


def to_natural(sh):
    """Inverse of `to_shifted` (the same map for even sides)."""
    ny, nx = sh.shape[-2:]
    return np.roll(sh, (ny // 2, nx // 2), axis=(-2, -1))


def pixel_size_object(wavelength, distance, detector_pixel, frame):
    """`[py, px] = λz / (n · detector_pixel)`. ptyrs/src/base/geometry.rs:118-126"""
    lz = wavelength * distance
    return [lz / (frame[0] * detector_pixel), lz / (frame[1] * detector_pixel)]


def _rem_euclid(a, b):
    r = np.fmod(a, b)
    return np.where(r < 0, r + abs(b), r)


def _pynx_axis(v, px):
    """One axis of PyNX position preprocessing. ptyrs/src/base/positions.rs:76-100"""
    v = np.asarray(v, dtype=np.float64)
    if v.size == 0:
        return v.copy()
    n = float(v.size)
    mean = v.sum() / n
    pos = v - mean
    frac = _rem_euclid(pos / px, 1.0)
    if np.all(np.abs(np.minimum(frac, 1.0 - frac)) < 0.1):
        frac = _rem_euclid(pos + 0.5, 1.0) - 0.5
    fmean = frac.sum() / n
    if np.all(np.abs(frac - fmean) <= 0.02 + 1e-5 * abs(fmean)):
        pos = np.round((pos - fmean * px) / px) * px
    return pos


def pynx_preprocess(y, x, pixel):
    """PyNX `PtychoData` preprocessing: centre, then snap a shared sub-pixel offset.
    ptyrs/src/base/positions.rs:51-63"""
    return _pynx_axis(y, pixel[0]), _pynx_axis(x, pixel[1])


def view_corner(obj, frame, shift):
    """`(obj − frame)//2 + shift` rounded half-to-even, None outside the object.
    ptyrs/src/base/addr.rs:118-127"""
    if obj < frame:
        return None
    span = obj - frame
    if not np.isfinite(shift):
        return None
    c = float(np.round((span // 2) + shift))
    if 0.0 <= c <= span:
        return int(c)
    return None


def required_obj_shape(y, x, pixel, frame, margin):
    """PyNX `calc_obj_shape(..., multiple=4)`. ptyrs/src/base/addr.rs:129-149"""
    def axis(v, p, n):
        v = np.asarray(v, dtype=np.float64)
        ext = max([abs(np.ceil(t / p)) + 1.0 for t in v] + [0.0])
        length = int(2.0 * ext) + n + margin
        return (length + 3) // 4 * 4
    return [axis(y, pixel[0], frame[0]), axis(x, pixel[1], frame[1])]


class AddrTable:
    """Per-view corners: view `k = f * nmode + m`, probe mode `m % nprobe`, object mode
    `m // nprobe`. ptyrs/src/base/addr.rs:20-81"""

    def __init__(self, y, x, pixel, frame, obj, nprobe, nobj):
        if nprobe == 0 or nobj == 0:
            raise ValueError("nprobe and nobj must be >= 1")
        if obj[0] < frame[0] or obj[1] < frame[1]:
            raise ValueError(f"object {obj} smaller than frame {frame}")
        self.nframe = len(y)
        self.nprobe, self.nobj = nprobe, nobj
        self.nmode = nprobe * nobj
        self.frame, self.obj = list(frame), list(obj)
        oy, ox = [], []
        for f, (py, px) in enumerate(zip(y, x)):
            cy = view_corner(obj[0], frame[0], py / pixel[0])
            cx = view_corner(obj[1], frame[1], px / pixel[1])
            if cy is None or cx is None:
                raise ValueError(f"frame {f} leaves object {obj}")
            oy.append(cy)
            ox.append(cx)
        self.oy = np.array(oy, dtype=np.int64)
        self.ox = np.array(ox, dtype=np.int64)

    def views(self):
        """(frame, mode, probe, obj, oy, ox) rows in table order."""
        rows = []
        for f in range(self.nframe):
            for m in range(self.nmode):
                rows.append((f, m, m % self.nprobe, m // self.nprobe,
                             int(self.oy[f]), int(self.ox[f])))
        return rows


def observed(counts):
    """Counts `[nframe, ny, nx]` (detector order; `MASKED` or negative = masked) to the
    shifted float intensities the operators read, NaN where masked, and `nb_obs`.
    ptyrs/src/engine/obs.rs:13-35 (u16 path) with the float-counts extension: a value
    above MAX_COUNT is kept as a float rather than refused."""
    c = np.asarray(counts)
    if c.dtype == np.uint16:
        mask = c == MASKED
    else:
        mask = c < 0
    iobs = np.where(mask, np.nan, c.astype(np.float64))
    iobs = to_shifted(iobs)
    return iobs, int((~mask).sum())


def _mix(z):
    """SplitMix64's output function. ptyrs/src/engine/batch.rs:7-12"""
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & U64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & U64
    return z ^ (z >> 31)


def random_batches(nframe, size, seed, epoch):
    """Fisher–Yates from the end with SplitMix64 draws from `seed ^ mix(epoch)`, cut into
    chunks of `size`, each sorted. ptyrs/src/engine/batch.rs:14-38"""
    if size <= 0:
        raise ValueError("batch size must be positive")
    state = (seed ^ _mix(epoch & U64)) & U64
    order = list(range(nframe))
    for i in range(nframe - 1, 0, -1):
        state = (state + 0x9E3779B97F4A7C15) & U64
        r = _mix(state)
        j = (r * (i + 1)) >> 64
        order[i], order[j] = order[j], order[i]
    return [sorted(order[s:s + size]) for s in range(0, nframe, size)]
