"""Host-side scan preparation (ptyrs::base), the steps `GpuPtycho` runs before it touches a
device, through the C ABI of libptyrs_hip.so: PyNX position preprocessing, view corners, the
object shape a scan needs, the far-field quadratic phase and the counts as the device reads
them. None of these calls a device.
"""

import ctypes

import numpy as np

from . import _check, _ptr, _scan_optics_c, library


def pynx_axis(v, pixel):
    """`Positions::pynx_preprocess` along one axis (ptyrs/src/base/positions.rs:51-100):
    positions `v` in metres centred on their mean, and snapped to whole pixels when they
    share one sub-pixel offset."""
    a = np.ascontiguousarray(v, dtype=np.float64)
    out = np.empty_like(a)
    _check(library().ptyrs_pynx_preprocess(_ptr(a, ctypes.c_double), a.size, float(pixel),
                                           _ptr(out, ctypes.c_double)))
    return out


def pynx_preprocess(y, x, pixel):
    """Both axes; `pixel` is `[py, px]` (ptyrs/src/base/positions.rs:51-63)."""
    return pynx_axis(y, pixel[0]), pynx_axis(x, pixel[1])


def view_corner(obj, frame, shift):
    """`view_corner` (ptyrs/src/base/addr.rs:118-127): `(obj − frame)//2 + shift` rounded half
    to even, or None when the window leaves the object."""
    c = ctypes.c_int64()
    _check(library().ptyrs_view_corner(int(obj), int(frame), float(shift), ctypes.byref(c)))
    return None if c.value < 0 else int(c.value)


def required_obj_shape(y, x, pixel, frame, margin):
    """`required_obj_shape` (ptyrs/src/base/addr.rs:129-149) of positions `y`, `x` (metres),
    `pixel` `[py, px]`, `frame` `[ny, nx]`: `[ny, nx]` of the object."""
    y = np.ascontiguousarray(y, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    if y.shape != x.shape:
        raise ValueError(f"{y.size} y positions vs {x.size} x positions")
    out = (ctypes.c_size_t * 2)()
    _check(library().ptyrs_required_obj_shape(
        _ptr(y, ctypes.c_double), _ptr(x, ctypes.c_double), y.size, float(pixel[0]),
        float(pixel[1]), int(frame[0]), int(frame[1]), int(margin), out))
    return [out[0], out[1]]


def scan_corners(scan, obj_shape, optics=None):
    """The view corners `[nframe, 2]` (oy, ox) `GpuPtycho` builds for `scan` in an object of
    `obj_shape` `[nyo, nxo]` (AddrTable::new over the preprocessed positions,
    ptyrs/src/base/addr.rs:35-81)."""
    sc, oc = _scan_optics_c(scan, optics)
    out = np.empty((scan.counts.shape[0], 2), dtype=np.int64)
    _check(library().ptyrs_scan_corners(
        ctypes.byref(sc), ctypes.byref(oc) if oc is not None else None, int(obj_shape[0]),
        int(obj_shape[1]), _ptr(out, ctypes.c_int64), out.size))
    return out


def observed(scan, optics=None):
    """The counts as the device reads them (Observed::from_counts,
    ptyrs/src/engine/obs.rs:13-35): `[nframe, ny, nx]` float32 in the exit transform's DFT
    index order (for the far field the fftshifted detector frames), -1 where masked, with
    `nb_obs` and the unmasked total."""
    sc, oc = _scan_optics_c(scan, optics)
    out = np.empty(scan.counts.shape, dtype=np.float32)
    nb, tot = ctypes.c_size_t(), ctypes.c_double()
    _check(library().ptyrs_scan_observed(
        ctypes.byref(sc), ctypes.byref(oc) if oc is not None else None,
        _ptr(out, ctypes.c_float), out.size, ctypes.byref(nb), ctypes.byref(tot)))
    return out, nb.value, tot.value


def quad_phase(wavelength, distance, detector_pixel, frame):
    """`quad_phase` (ptyrs/src/engine/ops.rs:11-33): the far-field phase
    `exp(iπ/(λz)·(dx²px² + dy²py²))`, `[ny, nx]` complex in the natural layout."""
    out = np.empty(tuple(int(f) for f in frame), dtype=np.complex128)
    _check(library().ptyrs_quad_phase(float(wavelength), float(distance), float(detector_pixel),
                                      out.shape[0], out.shape[1],
                                      _ptr(out.view(np.float64), ctypes.c_double), out.size * 2))
    return out


_FAULT_EXITS = {"launch": 0, "capture": 1}


def arm_fault(exit, pass_=0):
    """Test hook of the abort gate (ptyrs/src/fault.rs:62-114): `exit` ("launch": each kernel
    launch of an engine call; "capture": each graph capture) fails at its hit after `pass_`
    others, once, as a device failure raised before anything is enqueued. None disarms.
    Process-wide."""
    code = -1 if exit is None else _FAULT_EXITS[exit]
    _check(library().ptyrs_test_arm_fault(code, int(pass_)))


def fault_fired():
    """Whether the armed exit has failed."""
    f = ctypes.c_int()
    _check(library().ptyrs_test_fault_fired(ctypes.byref(f)))
    return bool(f.value)
