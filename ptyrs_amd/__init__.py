"""ptyrs_amd: the per-position ptychographic reconstruction step on an MI355X, with the Python
surface of ptyrs-py's GPU engine (ptyrs-py/src/gpu.rs, ptyrs-py/src/lib.rs).

The work runs in libptyrs_hip.so (hand-written HIP kernels for gfx950) through the C ABI of
include/ptyrs_hip.h. There is no CPU fallback: without the built library every engine call
raises.

    scan = Scan(counts, y, x, wavelength, distance, detector_pixel)
    g = GpuPtycho(scan, obj, probe, dtype="c64")
    g.rpie(10, batch_size=64, calc_llk=1)
    g.object, g.probe, g.history, g.llk()

Errors follow ptyrs-py (ptyrs-py/src/lib.rs:27-50, :52-108): bad input is a ValueError, a
device failure a RuntimeError, an unknown keyword a TypeError.
"""

import ctypes
import math
import os
import pathlib

import numpy as np

__all__ = ["Scan", "Wavefield", "Optics", "GpuPtycho", "library", "MASKED"]

MASKED = 65535
_HERE = pathlib.Path(__file__).resolve().parent
_LIB_PATH = _HERE / "libptyrs_hip.so"

_OK, _EINVAL, _EDEVICE, _EABORTED = 0, 1, 2, 3
_C64, _C128 = 0, 1
_ALGORITHMS = ["ap", "dm", "raar", "rpie", "ml", "bh", "lsqml"]


class _ScanC(ctypes.Structure):
    _fields_ = [("nframe", ctypes.c_size_t), ("ny", ctypes.c_size_t), ("nx", ctypes.c_size_t),
                ("counts_u16", ctypes.POINTER(ctypes.c_uint16)),
                ("counts_f32", ctypes.POINTER(ctypes.c_float)),
                ("y", ctypes.POINTER(ctypes.c_double)), ("x", ctypes.POINTER(ctypes.c_double)),
                ("wavelength", ctypes.c_double), ("distance", ctypes.c_double),
                ("detector_pixel", ctypes.c_double),
                ("corners", ctypes.POINTER(ctypes.c_int64))]


class _OpticsC(ctypes.Structure):
    _fields_ = [("free_space", ctypes.c_int), ("wavelength", ctypes.c_double),
                ("pixel_y", ctypes.c_double), ("pixel_x", ctypes.c_double),
                ("distance", ctypes.c_double), ("ngap", ctypes.c_size_t),
                ("gaps", ctypes.POINTER(ctypes.c_double))]


class _OptionsC(ctypes.Structure):
    _fields_ = [("stack_size", ctypes.c_size_t), ("graph", ctypes.c_int),
                ("llk_slots", ctypes.c_size_t), ("fused", ctypes.c_int),
                ("psi_f16", ctypes.c_int), ("dtype", ctypes.c_int)]


class _RpieC(ctypes.Structure):
    _fields_ = [("has_obj_step", ctypes.c_int), ("obj_step", ctypes.c_double),
                ("has_probe_step", ctypes.c_int), ("probe_step", ctypes.c_double),
                ("obj_alpha", ctypes.c_double), ("probe_alpha", ctypes.c_double),
                ("batch_size", ctypes.c_size_t), ("calc_llk", ctypes.c_size_t),
                ("seed", ctypes.c_uint64)]


class _ApC(ctypes.Structure):
    _fields_ = [("has_obj_inertia", ctypes.c_int), ("obj_inertia", ctypes.c_double),
                ("has_probe_inertia", ctypes.c_int), ("probe_inertia", ctypes.c_double),
                ("calc_llk", ctypes.c_size_t), ("zero_phase_ramp", ctypes.c_int)]


class _DmC(ctypes.Structure):
    _fields_ = [("has_obj_inertia", ctypes.c_int), ("obj_inertia", ctypes.c_double),
                ("has_probe_inertia", ctypes.c_int), ("probe_inertia", ctypes.c_double),
                ("alpha", ctypes.c_double), ("loop_obj_probe", ctypes.c_size_t),
                ("calc_llk", ctypes.c_size_t), ("zero_phase_ramp", ctypes.c_int)]


class _RaarC(ctypes.Structure):
    _fields_ = [("has_obj_inertia", ctypes.c_int), ("obj_inertia", ctypes.c_double),
                ("has_probe_inertia", ctypes.c_int), ("probe_inertia", ctypes.c_double),
                ("beta", ctypes.c_double), ("loop_obj_probe", ctypes.c_size_t),
                ("calc_llk", ctypes.c_size_t), ("zero_phase_ramp", ctypes.c_int)]


class _MlC(ctypes.Structure):
    _fields_ = [("update_object", ctypes.c_int), ("update_probe", ctypes.c_int),
                ("calc_llk", ctypes.c_size_t), ("zero_phase_ramp", ctypes.c_int)]


class _LsqmlC(ctypes.Structure):
    _fields_ = [("batch_size", ctypes.c_size_t), ("cost", ctypes.c_int), ("seed", ctypes.c_uint64)]


class _BhC(ctypes.Structure):
    _fields_ = [("batch_size", ctypes.c_size_t), ("probe_rho", ctypes.c_double),
                ("position_rho", ctypes.c_double), ("seed", ctypes.c_uint64)]


class _BhStepC(ctypes.Structure):
    _fields_ = [("first", ctypes.c_size_t), ("count", ctypes.c_size_t),
                ("beta", ctypes.c_double), ("alpha", ctypes.c_double),
                ("top", ctypes.c_double), ("bottom", ctypes.c_double)]


class _OprC(ctypes.Structure):
    _fields_ = [("k", ctypes.c_size_t), ("eigenmodes", ctypes.POINTER(ctypes.c_double)),
                ("weights", ctypes.POINTER(ctypes.c_double)),
                ("intensity", ctypes.POINTER(ctypes.c_double)),
                ("refine_weights", ctypes.c_int), ("refine_intensity", ctypes.c_int)]


class _LaunchC(ctypes.Structure):
    _fields_ = [("stack_size", ctypes.c_size_t), ("pass_c_blocks", ctypes.c_int),
                ("chi_blocks", ctypes.c_int)]


class _StepC(ctypes.Structure):
    _fields_ = [("batch", ctypes.c_size_t), ("frame", ctypes.c_size_t), ("a_far", ctypes.c_double),
                ("a_obj", ctypes.c_double), ("a_probe", ctypes.c_double)]


class _RecordC(ctypes.Structure):
    _fields_ = [("cycle", ctypes.c_size_t), ("algorithm", ctypes.c_int),
                ("poisson", ctypes.c_double), ("gaussian", ctypes.c_double),
                ("euclidian", ctypes.c_double), ("nb_photons_calc", ctypes.c_double)]


_lib = None

# (name, argtypes, restype) of every entry point of include/ptyrs_hip.h.
_P = ctypes.c_void_p
_D = ctypes.POINTER(ctypes.c_double)
_SZ = ctypes.c_size_t
_SIGNATURES = [
    ("ptyrs_engine_create", [ctypes.POINTER(_ScanC), ctypes.POINTER(_OpticsC), _D, _SZ, _SZ, _SZ,
                             _D, _SZ, ctypes.POINTER(_OptionsC), ctypes.c_int,
                             ctypes.POINTER(_P)], ctypes.c_int),
    ("ptyrs_engine_rpie", [_P, ctypes.POINTER(_RpieC), _SZ], ctypes.c_int),
    ("ptyrs_engine_ap", [_P, ctypes.POINTER(_ApC), _SZ], ctypes.c_int),
    ("ptyrs_engine_dm", [_P, ctypes.POINTER(_DmC), _SZ], ctypes.c_int),
    ("ptyrs_engine_raar", [_P, ctypes.POINTER(_RaarC), _SZ], ctypes.c_int),
    ("ptyrs_engine_ml", [_P, ctypes.POINTER(_MlC), _SZ], ctypes.c_int),
    ("ptyrs_engine_lsqml", [_P, ctypes.POINTER(_LsqmlC), _SZ, _SZ], ctypes.c_int),
    ("ptyrs_engine_lsqml_opr", [_P, ctypes.POINTER(_LsqmlC), _SZ, _SZ, ctypes.POINTER(_OprC)],
     ctypes.c_int),
    ("ptyrs_engine_bh", [_P, ctypes.POINTER(_BhC), _SZ, _SZ], ctypes.c_int),
    ("ptyrs_engine_bh_steps", [_P, ctypes.POINTER(_BhStepC), _SZ, ctypes.POINTER(_SZ)],
     ctypes.c_int),
    ("ptyrs_engine_bh_frames", [_P, ctypes.POINTER(_SZ), _SZ], ctypes.c_int),
    ("ptyrs_engine_set_launch", [_P, ctypes.POINTER(_LaunchC)], ctypes.c_int),
    ("ptyrs_engine_launch", [_P, ctypes.POINTER(_LaunchC)], ctypes.c_int),
    ("ptyrs_engine_lsqml_steps", [_P, ctypes.POINTER(_StepC), _SZ, ctypes.POINTER(_SZ)],
     ctypes.c_int),
    ("ptyrs_engine_lsqml_slice_steps", [_P, _D, _SZ], ctypes.c_int),
    ("ptyrs_engine_set_position", [_P, _D, _SZ, ctypes.c_int], ctypes.c_int),
    ("ptyrs_engine_position", [_P, _D, _SZ, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    ("ptyrs_engine_sampling", [_P, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    ("ptyrs_engine_set_sampling", [_P, ctypes.c_int], ctypes.c_int),
    ("ptyrs_engine_grad", [_P, ctypes.c_int, _D, _D, _SZ, _D, _SZ], ctypes.c_int),
    ("ptyrs_engine_intensities", [_P, _D, _SZ], ctypes.c_int),
    ("ptyrs_engine_live_reset", [_P], ctypes.c_int),
    ("ptyrs_engine_live_push", [_P, ctypes.POINTER(ctypes.c_uint16),
                                ctypes.POINTER(ctypes.c_float), _SZ], ctypes.c_int),
    ("ptyrs_engine_live_frames", [_P, ctypes.POINTER(_SZ), ctypes.POINTER(_SZ)], ctypes.c_int),
    ("ptyrs_group_rpie", [ctypes.POINTER(_P), _SZ, ctypes.POINTER(_RpieC), _SZ], ctypes.c_int),
    ("ptyrs_group_rpie_graph", [ctypes.POINTER(_P), _SZ, ctypes.POINTER(_RpieC), _SZ],
     ctypes.c_int),
    ("ptyrs_engine_set_checkpoint", [_P, _SZ], ctypes.c_int),
    ("ptyrs_engine_checkpoint", [_P, ctypes.POINTER(_SZ), ctypes.POINTER(_SZ)], ctypes.c_int),
    ("ptyrs_engine_llk", [_P, ctypes.POINTER(_RecordC)], ctypes.c_int),
    ("ptyrs_engine_llk_noalloc", [_P, ctypes.POINTER(_RecordC)], ctypes.c_int),
    ("ptyrs_engine_object", [_P, _D, _SZ], ctypes.c_int),
    ("ptyrs_engine_probe", [_P, _D, _SZ], ctypes.c_int),
    ("ptyrs_engine_set_object", [_P, _D, _SZ], ctypes.c_int),
    ("ptyrs_engine_set_probe", [_P, _D, _SZ], ctypes.c_int),
    ("ptyrs_engine_history", [_P, ctypes.POINTER(_RecordC), _SZ, ctypes.POINTER(_SZ)],
     ctypes.c_int),
    ("ptyrs_engine_cycle", [_P, ctypes.POINTER(_SZ)], ctypes.c_int),
    ("ptyrs_engine_corners", [_P, ctypes.POINTER(ctypes.c_int64), _SZ], ctypes.c_int),
    ("ptyrs_engine_synchronize", [_P], ctypes.c_int),
    ("ptyrs_engine_profile", [_P, ctypes.c_int], ctypes.c_int),
    ("ptyrs_engine_kernel_stats", [_P, ctypes.c_char_p, ctypes.POINTER(_SZ), _D], ctypes.c_int),
    ("ptyrs_kernel_names", [], ctypes.c_char_p),
    ("ptyrs_engine_destroy", [_P], None),
    ("ptyrs_last_error", [ctypes.c_char_p, _SZ], _SZ),
    ("ptyrs_test_arm_fault", [ctypes.c_int, _SZ], ctypes.c_int),
    ("ptyrs_test_fault_fired", [ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    ("ptyrs_pynx_preprocess", [_D, _SZ, ctypes.c_double, _D], ctypes.c_int),
    ("ptyrs_view_corner", [_SZ, _SZ, ctypes.c_double, ctypes.POINTER(ctypes.c_int64)],
     ctypes.c_int),
    ("ptyrs_required_obj_shape", [_D, _D, _SZ, ctypes.c_double, ctypes.c_double, _SZ, _SZ, _SZ,
                                  ctypes.POINTER(_SZ)], ctypes.c_int),
    ("ptyrs_scan_corners", [ctypes.POINTER(_ScanC), ctypes.POINTER(_OpticsC), _SZ, _SZ,
                            ctypes.POINTER(ctypes.c_int64), _SZ], ctypes.c_int),
    ("ptyrs_scan_observed", [ctypes.POINTER(_ScanC), ctypes.POINTER(_OpticsC),
                             ctypes.POINTER(ctypes.c_float), _SZ, ctypes.POINTER(_SZ), _D],
     ctypes.c_int),
    ("ptyrs_quad_phase", [ctypes.c_double, ctypes.c_double, ctypes.c_double, _SZ, _SZ, _D, _SZ],
     ctypes.c_int),
]


def library():
    """The loaded libptyrs_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        path = str(_LIB_PATH)
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: build it with `python -m ptyrs_amd.build`")
        lib = ctypes.CDLL(path)
        for name, args, res in _SIGNATURES:
            f = getattr(lib, name)
            f.argtypes, f.restype = args, res
        _lib = lib
    return _lib


def _check(status):
    if status == _OK:
        return
    lib = library()
    buf = ctypes.create_string_buffer(4096)
    lib.ptyrs_last_error(buf, len(buf))
    msg = buf.value.decode(errors="replace")
    if status == _EINVAL:
        raise ValueError(msg)
    raise RuntimeError(msg)


def _overlay(defaults, kw, what):
    """Each keyword replaces the default of the field of the same name; an unknown keyword is a
    TypeError (ptyrs-py/src/lib.rs:52-108)."""
    out = dict(defaults)
    unknown = [k for k in kw if k not in defaults]
    if unknown:
        raise TypeError(f"unknown parameters {unknown} for {what}")
    out.update(kw)
    return out


def _ptr(a, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


class Scan:
    """A far-field scan (ptyrs-py/src/lib.rs:430-463): counts `[nframe, ny, nx]` in detector
    order, uint16 with 65535 masked or float with negative values masked (float counts may exceed
    65534, which the uint16 path cannot hold); positions `y`, `x` in metres; geometry in metres.
    `corners` (`[nframe, 2]` integer (oy, ox)) gives the view corners directly, an address table
    (ptyrs/src/base/addr.rs:20-81) in place of the positions' PyNX preprocessing."""

    def __init__(self, counts, y, x, wavelength, distance, detector_pixel, corners=None):
        c = np.asarray(counts)
        if c.ndim != 3:
            raise TypeError("counts must be a 3-D uint16 or float array")
        if c.dtype == np.uint16:
            self.counts = np.ascontiguousarray(c)
        elif np.issubdtype(c.dtype, np.floating):
            if not np.all(np.isfinite(c)):
                raise ValueError("counts must be finite")
            self.counts = np.ascontiguousarray(c, dtype=np.float32)
        else:
            raise TypeError("counts must be a 3-D uint16 or float array")
        self.y = np.ascontiguousarray(y, dtype=np.float64)
        self.x = np.ascontiguousarray(x, dtype=np.float64)
        if self.y.shape != (c.shape[0],) or self.x.shape != (c.shape[0],):
            raise ValueError(f"{self.y.size} y and {self.x.size} x positions for {c.shape[0]} frames")
        for name, v in (("wavelength", wavelength), ("distance", distance),
                        ("detector_pixel", detector_pixel)):
            if not (math.isfinite(v) and v > 0):
                raise ValueError(f"{name} must be finite and > 0, got {v}")
        self.wavelength, self.distance, self.detector_pixel = wavelength, distance, detector_pixel
        self.corners = None
        if corners is not None:
            self.corners = np.ascontiguousarray(corners, dtype=np.int64)
            if self.corners.shape != (c.shape[0], 2):
                raise ValueError(f"corners must be [{c.shape[0]}, 2], got {self.corners.shape}")

    @property
    def frame(self):
        return list(self.counts.shape[1:])

    def _c(self):
        s = _ScanC()
        s.nframe, s.ny, s.nx = self.counts.shape
        if self.counts.dtype == np.uint16:
            s.counts_u16 = _ptr(self.counts, ctypes.c_uint16)
        else:
            s.counts_f32 = _ptr(self.counts, ctypes.c_float)
        s.y, s.x = _ptr(self.y, ctypes.c_double), _ptr(self.x, ctypes.c_double)
        s.wavelength, s.distance, s.detector_pixel = (self.wavelength, self.distance,
                                                      self.detector_pixel)
        if self.corners is not None:
            s.corners = _ptr(self.corners, ctypes.c_int64)
        return s


class Wavefield:
    """One free-space propagation (ptyrs-py/src/cpu.rs:44-57): `[ny, nx]` planes (both even),
    `wavelength`, source-plane `pixel` `[py, px]` and signed `distance` in metres."""

    def __init__(self, shape, wavelength, pixel, distance):
        if len(shape) != 2 or any(int(s) <= 0 or int(s) % 2 for s in shape):
            raise ValueError(f"plane shape {shape} must be even and non-zero")
        if not (math.isfinite(wavelength) and wavelength > 0):
            raise ValueError("wavelength must be finite and > 0")
        if not all(math.isfinite(p) and p > 0 for p in pixel):
            raise ValueError("pixel must be finite and > 0")
        if math.isnan(distance) or distance == 0.0 or distance == -math.inf:
            raise ValueError("distance must be non-zero and not -inf")
        self.shape = [int(s) for s in shape]
        self.wavelength, self.pixel, self.distance = wavelength, list(pixel), distance


class Optics:
    """The exit-wave model (ptyrs-py/src/cpu.rs:77-135): PyNX's far field or pty-chi's
    planar free-space model of one Wavefield, through `len(gaps) + 1` object slices."""

    def __init__(self, free_space, wavefield=None, far=None, gaps=()):
        self._free_space, self.wavefield, self._far = free_space, wavefield, far
        self.gaps = list(gaps)

    @staticmethod
    def far_field(wavelength, distance, detector_pixel, frame):
        return Optics(False, far=(wavelength, distance, detector_pixel, list(frame)))

    @staticmethod
    def free_space(w):
        return Optics(True, wavefield=w)

    def with_slices(self, gaps):
        """This exit model through `len(gaps) + 1` object slices, gap `s` an angular-spectrum
        Wavefield from slice `s` to `s + 1` (ptyrs-py/src/cpu.rs:122-133)."""
        for w in gaps:
            if not isinstance(w, Wavefield):
                raise TypeError("gaps must be Wavefields")
            if not math.isfinite(w.distance):
                raise ValueError("angular spectrum needs a finite distance")
        return Optics(self._free_space, self.wavefield, self._far, gaps)

    def nslice(self):
        return len(self.gaps) + 1

    def _c(self, scan):
        o = _OpticsC()
        if self.gaps:
            for w in self.gaps:
                if w.shape != scan.frame:
                    raise ValueError(f"slice gap plane {w.shape} is not the frame {scan.frame}")
            self._gaps_c = np.ascontiguousarray(
                [[w.wavelength, w.pixel[0], w.pixel[1], w.distance] for w in self.gaps],
                dtype=np.float64)
            o.ngap, o.gaps = len(self.gaps), _ptr(self._gaps_c, ctypes.c_double)
        if self._free_space:
            w = self.wavefield
            if w.shape != scan.frame:
                raise ValueError(f"propagator plane {w.shape} differs from frame {scan.frame}")
            o.free_space, o.wavelength = 1, w.wavelength
            o.pixel_y, o.pixel_x, o.distance = w.pixel[0], w.pixel[1], w.distance
            return o
        lam, z, dpix, frame = self._far
        if frame != scan.frame:
            raise ValueError(f"optics frame {frame} differs from the scan's {scan.frame}")
        # The far field of a geometry is the scan's own when they agree.
        o.free_space = 0
        return o, (lam, z, dpix)


def _scan_optics_c(scan, optics):
    """The C structs of a scan under `optics` (None: the far field of the scan geometry)."""
    sc = scan._c()
    oc = None
    if optics is not None:
        r = optics._c(scan)
        if isinstance(r, tuple):
            oc, (lam, z, dpix) = r
            sc.wavelength, sc.distance, sc.detector_pixel = lam, z, dpix
        else:
            oc = r
    return sc, oc


_GPU_OPTIONS = dict(stack_size=None, graph=True, llk_slots=8, fused=True, psi_f16=False,
                    dtype="c64")
_AP_PARAMS = dict(obj_inertia=0.01, probe_inertia=None, calc_llk=0, zero_phase_ramp=True)
_DM_PARAMS = dict(obj_inertia=0.01, probe_inertia=0.001, alpha=0.02, loop_obj_probe=1, calc_llk=0,
                  zero_phase_ramp=True)
_RAAR_PARAMS = dict(obj_inertia=0.01, probe_inertia=0.001, beta=0.9, loop_obj_probe=1, calc_llk=0,
                    zero_phase_ramp=True)
_ML_PARAMS = dict(update_object=True, update_probe=False, calc_llk=0, zero_phase_ramp=True)
_RPIE_PARAMS = dict(obj_step=1.0, probe_step=1.0, obj_alpha=0.1, probe_alpha=0.1,
                    batch_size=100, calc_llk=0, seed=7)


class GpuPtycho:
    """The GPU reconstruction state of a scan (ptyrs-py/src/gpu.rs:24-267), on HIP device
    `device`, under `optics` in place of the far field of the scan's geometry if given;
    keywords are GpuOptions fields (stack_size, graph, llk_slots, fused, psi_f16) plus `dtype`,
    "c64" (complex float, the default) or "c128" (complex double, the plain FP64 path).
    `object` is complex `[nobj, nyo, nxo]`, or `[nslice, nobj, nyo, nxo]` with a slice per slice
    of `optics` (more than one runs LSQML and the LLK alone, of one mode), `probe` complex
    `[nprobe, ny, nx]`."""

    def __init__(self, scan, object, probe, device=0, optics=None, **options):
        opts = _overlay(_GPU_OPTIONS, options, "GpuOptions")
        if opts["dtype"] not in ("c64", "c128"):
            raise ValueError(f"dtype must be \"c64\" or \"c128\", got {opts['dtype']!r}")
        if opts["stack_size"] is not None and int(opts["stack_size"]) < 1:
            raise ValueError("stack_size must be >= 1")
        if int(opts["llk_slots"]) < 1:
            raise ValueError("llk_slots must be >= 1")
        lib = library()
        obj = np.ascontiguousarray(object, dtype=np.complex128)
        nslice = optics.nslice() if optics is not None else 1
        if obj.ndim == 3:
            obj = obj[None]
        if obj.ndim != 4:
            raise TypeError("object must be complex128 [nmode, nyo, nxo] or "
                            "[nslice, nmode, nyo, nxo]")
        if obj.shape[0] != nslice:
            raise ValueError(f"{obj.shape[0]} object slices for {nslice} slices of the optics")
        pr = np.ascontiguousarray(probe, dtype=np.complex128)
        if pr.ndim != 3:
            raise TypeError("probe must be complex128 [nmode, ny, nx]")
        if list(pr.shape[1:]) != scan.frame:
            raise ValueError(f"probe plane {list(pr.shape[1:])} differs from frame {scan.frame}")
        sc, oc = _scan_optics_c(scan, optics)
        oo = _OptionsC()
        oo.stack_size = int(opts["stack_size"] or 0)
        oo.graph = int(bool(opts["graph"]))
        oo.llk_slots = int(opts["llk_slots"])
        oo.fused, oo.psi_f16 = int(bool(opts["fused"])), int(bool(opts["psi_f16"]))
        oo.dtype = _C64 if opts["dtype"] == "c64" else _C128
        h = ctypes.c_void_p()
        _check(lib.ptyrs_engine_create(
            ctypes.byref(sc), ctypes.byref(oc) if oc is not None else None,
            _ptr(obj.view(np.float64), ctypes.c_double), obj.shape[1], obj.shape[2], obj.shape[3],
            _ptr(pr.view(np.float64), ctypes.c_double), pr.shape[0], ctypes.byref(oo), device,
            ctypes.byref(h)))
        self._h = h
        self._objs_shape, self._probe_shape = obj.shape, pr.shape
        self._obj_shape = obj.shape if nslice > 1 else obj.shape[1:]
        self.nslice = nslice
        self.nframe = scan.counts.shape[0]
        self.dtype = opts["dtype"]
        self._inputs = (scan, device, optics, dict(options))

    def _clone(self):
        """A new engine on the same scan, device, optics and options, starting from this
        engine's current object, probe, position offsets, sampling and launch shapes (its own
        history and cycle count start empty)."""
        scan, device, optics, options = self._inputs
        obj = self.objects if self.nslice > 1 else self.object
        g = GpuPtycho(scan, obj, self.probe, device, optics, **options)
        f = self.features
        if f["position"] is not None:
            g.set_features(position=f["position"], hold=f["hold"])
        g.set_sampling(self.sampling)
        g.set_launch(**self.launch)
        return g

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.ptyrs_engine_destroy(h)
            self._h = None

    @staticmethod
    def _inertias(c, p):
        c.has_obj_inertia = p["obj_inertia"] is not None
        c.obj_inertia = p["obj_inertia"] if p["obj_inertia"] is not None else 0.0
        c.has_probe_inertia = p["probe_inertia"] is not None
        c.probe_inertia = p["probe_inertia"] if p["probe_inertia"] is not None else 0.0
        if int(p["calc_llk"]) < 0:
            raise ValueError("calc_llk must be >= 0")
        c.calc_llk, c.zero_phase_ramp = int(p["calc_llk"]), int(bool(p["zero_phase_ramp"]))

    def ap(self, ncycle, **params):
        """`ncycle` AP cycles; keywords are ApParams fields (ptyrs/src/engine/ptycho.rs:48-69)."""
        p = _overlay(_AP_PARAMS, params, "ApParams")
        c = _ApC()
        self._inertias(c, p)
        _check(library().ptyrs_engine_ap(self._h, ctypes.byref(c), int(ncycle)))

    def dm(self, ncycle, **params):
        """`ncycle` DM cycles; keywords are DmParams fields (ptyrs/src/engine/ptycho.rs:71-95)."""
        p = _overlay(_DM_PARAMS, params, "DmParams")
        c = _DmC()
        self._inertias(c, p)
        c.alpha, c.loop_obj_probe = float(p["alpha"]), int(p["loop_obj_probe"])
        _check(library().ptyrs_engine_dm(self._h, ctypes.byref(c), int(ncycle)))

    def raar(self, ncycle, **params):
        """`ncycle` RAAR cycles; keywords are RaarParams fields (ptyrs/src/engine/ptycho.rs:97-121)."""
        p = _overlay(_RAAR_PARAMS, params, "RaarParams")
        c = _RaarC()
        self._inertias(c, p)
        c.beta, c.loop_obj_probe = float(p["beta"]), int(p["loop_obj_probe"])
        _check(library().ptyrs_engine_raar(self._h, ctypes.byref(c), int(ncycle)))

    def ml(self, ncycle, **params):
        """`ncycle` ML cycles; keywords are MlParams fields (ptyrs/src/engine/ptycho.rs:123-143)."""
        p = _overlay(_ML_PARAMS, params, "MlParams")
        c = _MlC()
        c.update_object, c.update_probe = int(bool(p["update_object"])), int(bool(p["update_probe"]))
        if int(p["calc_llk"]) < 0:
            raise ValueError("calc_llk must be >= 0")
        c.calc_llk, c.zero_phase_ramp = int(p["calc_llk"]), int(bool(p["zero_phase_ramp"]))
        _check(library().ptyrs_engine_ml(self._h, ctypes.byref(c), int(ncycle)))

    def rpie(self, nepoch, **params):
        """`nepoch` rPIE epochs; keywords are RpieParams fields (ptyrs/src/engine/ptycho.rs:145-186);
        a step of None leaves that array fixed."""
        _check(library().ptyrs_engine_rpie(self._h, ctypes.byref(self._rpie_c(params)),
                                           int(nepoch)))

    @staticmethod
    def _rpie_c(params):
        p = _overlay(_RPIE_PARAMS, params, "RpieParams")
        c = _RpieC()
        c.has_obj_step = p["obj_step"] is not None
        c.obj_step = p["obj_step"] if p["obj_step"] is not None else 0.0
        c.has_probe_step = p["probe_step"] is not None
        c.probe_step = p["probe_step"] if p["probe_step"] is not None else 0.0
        c.obj_alpha, c.probe_alpha = p["obj_alpha"], p["probe_alpha"]
        if int(p["batch_size"]) < 0 or int(p["calc_llk"]) < 0:
            raise ValueError("batch_size and calc_llk must be >= 0")
        c.batch_size, c.calc_llk, c.seed = int(p["batch_size"]), int(p["calc_llk"]), int(p["seed"])
        return c

    def lsqml(self, nepoch, calc_llk, batch_size, cost, seed, **kw):
        """`nepoch` LSQML epochs from a fresh LSQML (`cost` "amplitude" or "poisson"), an LLK
        record every `calc_llk` epochs and after the last (ptyrs-py/src/gpu.rs:150-192). With
        `opr=(eigenmodes [k, ny, nx], weights [nframe, k], intensity [nframe] or None)` the
        epochs model that orthogonal probe relaxation of probe mode 0 (`hold` may name
        "weights" and "intensity") and return its `(eigenmodes, weights, intensity)` as they
        leave them; None without `opr`."""
        opr = kw.pop("opr", None)
        hold = list(kw.pop("hold", ()))
        if kw:
            raise TypeError(f"unexpected keyword arguments {sorted(kw)}")
        if cost not in ("amplitude", "poisson"):
            raise ValueError(f'cost must be "amplitude" or "poisson", got {cost!r}')
        for h in hold:
            if h not in ("weights", "intensity"):
                raise ValueError(f'hold takes "weights" and "intensity", got {h!r}')
        c = _LsqmlC()
        c.batch_size, c.cost, c.seed = int(batch_size), 0 if cost == "amplitude" else 1, int(seed)
        if opr is None:
            if hold:
                raise ValueError("hold names OPR values but no opr is given")
            _check(library().ptyrs_engine_lsqml(self._h, ctypes.byref(c), int(nepoch),
                                                int(calc_llk)))
            return None
        eig, w, inten = opr
        eig = np.array(eig, dtype=np.complex128, order="C")
        w = np.array(w, dtype=np.float64, order="C")
        if eig.ndim != 3 or eig.shape[1:] != self._probe_shape[1:] or eig.shape[0] < 1:
            raise ValueError(f"eigenmodes must be [k, ny, nx] for k >= 1, got {eig.shape}")
        if w.shape != (self.nframe, eig.shape[0]):
            raise ValueError(f"weights must be [nframe, k] = {(self.nframe, eig.shape[0])}, "
                             f"got {w.shape}")
        if inten is not None:
            inten = np.array(inten, dtype=np.float64, order="C")
            if inten.shape != (self.nframe,):
                raise ValueError(f"{inten.size} main-mode weights for {self.nframe} frames")
        elif "intensity" in hold:
            raise ValueError("hold names the intensity but none is given")
        v = _OprC()
        v.k = eig.shape[0]
        v.eigenmodes = _ptr(eig, ctypes.c_double)
        v.weights = _ptr(w, ctypes.c_double)
        v.intensity = _ptr(inten, ctypes.c_double) if inten is not None else None
        v.refine_weights = int("weights" not in hold)
        v.refine_intensity = int("intensity" not in hold)
        _check(library().ptyrs_engine_lsqml_opr(self._h, ctypes.byref(c), int(nepoch),
                                                int(calc_llk), ctypes.byref(v)))
        return eig, w, inten

    def bh(self, nepoch, calc_llk, batch_size, probe_rho, seed, *, position_rho=0.1):
        """`nepoch` BH epochs from a fresh BH (pty-chi's bilinear-Hessian method on the amplitude
        cost), an LLK record every `calc_llk` epochs and after the last; one batch of every
        frame runs conjugate gradient across the call's epochs (ptyrs-py/src/gpu.rs:124-144)."""
        c = _BhC()
        c.batch_size, c.probe_rho = int(batch_size), float(probe_rho)
        c.position_rho, c.seed = float(position_rho), int(seed)
        _check(library().ptyrs_engine_bh(self._h, ctypes.byref(c), int(nepoch), int(calc_llk)))

    def bh_steps(self):
        """The last BH epoch's steps as `Bh.epoch` returns them, dicts of `frames`, `beta`,
        `alpha`, `top`, `bottom`; none before a BH call (ptyrs-py/src/gpu.rs:146-149)."""
        lib = library()
        n = ctypes.c_size_t()
        _check(lib.ptyrs_engine_bh_steps(self._h, None, 0, ctypes.byref(n)))
        recs = (_BhStepC * max(n.value, 1))()
        _check(lib.ptyrs_engine_bh_steps(self._h, recs, n.value, ctypes.byref(n)))
        nf = self.nframe if n.value else 0
        frames = np.zeros(max(nf, 1), dtype=np.uintp)
        _check(lib.ptyrs_engine_bh_frames(
            self._h, frames.ctypes.data_as(ctypes.POINTER(ctypes.c_size_t)), nf))
        return [dict(frames=[int(f) for f in frames[r.first:r.first + r.count]], beta=r.beta,
                     alpha=r.alpha, top=r.top, bottom=r.bottom) for r in recs[:n.value]]

    def lsqml_steps(self):
        """The last LSQML epoch's steps as dicts of `frames`, `a_far`, `a_obj`
        (`[nslice, frames]`) and `a_probe`, one per batch (ptyrs-py/src/lib.rs:236-248)."""
        lib = library()
        n = ctypes.c_size_t()
        _check(lib.ptyrs_engine_lsqml_steps(self._h, None, 0, ctypes.byref(n)))
        recs = (_StepC * max(n.value, 1))()
        _check(lib.ptyrs_engine_lsqml_steps(self._h, recs, n.value, ctypes.byref(n)))
        a_obj = np.empty((n.value, self.nslice))
        _check(lib.ptyrs_engine_lsqml_slice_steps(
            self._h, _ptr(a_obj, ctypes.c_double), a_obj.size))
        out = []
        for i, r in enumerate(recs[:n.value]):
            if not out or out[-1]["batch"] != r.batch:
                out.append(dict(batch=r.batch, frames=[], a_far=[], a_obj=[], a_probe=[]))
            d = out[-1]
            d["frames"].append(r.frame)
            d["a_far"].append(r.a_far)
            d["a_obj"].append(a_obj[i])
            d["a_probe"].append(r.a_probe)
        return [dict(frames=np.array(d["frames"]), a_far=np.array(d["a_far"]),
                     a_obj=np.array(d["a_obj"]).T, a_probe=np.array(d["a_probe"]))
                for d in out]

    def grad(self, cost="poisson"):
        """The cost of the current state over every frame and its gradients, the adjoint of
        the exit-wave model on the device: `(value, objects_grad, probe_grad)` with value the
        Poisson LLK sum (`cost="poisson"`, the Poisson NLL up to a constant) or
        Σ(√icalc − √iobs)² (`"amplitude"`), and the gradients 2·∂L/∂O* of every slice
        (`[nslice, nobj, nyo, nxo]`) and 2·∂L/∂P* — torch's gradients of a real loss of complex
        tensors (ptyrs/src/engine/ptycho.rs:492-561, multislice.rs:266-442)."""
        if cost not in ("poisson", "amplitude"):
            raise ValueError(f"cost must be \"poisson\" or \"amplitude\", got {cost!r}")
        go = np.empty(self._objs_shape, dtype=np.complex128)
        gp = np.empty(self._probe_shape, dtype=np.complex128)
        v = ctypes.c_double()
        _check(library().ptyrs_engine_grad(
            self._h, 1 if cost == "poisson" else 0, ctypes.byref(v),
            _ptr(go.view(np.float64), ctypes.c_double), go.size * 2,
            _ptr(gp.view(np.float64), ctypes.c_double), gp.size * 2))
        return v.value, go, gp

    def intensities(self):
        """The model's detector intensities Σ_m |U ψ_m|² of every landed frame,
        `[nframe, ny, nx]`, detector order (the counts' layout)."""
        out = np.empty((self.live_frames[0],) + tuple(self._probe_shape[1:]))
        _check(library().ptyrs_engine_intensities(self._h, _ptr(out, ctypes.c_double), out.size))
        return out

    def live_reset(self):
        """Live ingestion (LiveAp, ptyrs/src/live/recon.rs:184-266): no frame of the planned
        scan is landed; every algorithm, the LLK and the intensities run over the frames
        landed so far. An engine starts with every planned frame landed."""
        _check(library().ptyrs_engine_live_reset(self._h))

    def live_push(self, counts):
        """Lands the next planned frames: `counts` `[k, ny, nx]` in detector order, uint16
        (65535 masked) or float (negative masked); ValueError past the planned frames."""
        c = np.asarray(counts)
        if c.ndim != 3 or tuple(c.shape[1:]) != tuple(self._probe_shape[1:]):
            raise ValueError("counts must be [k, ny, nx] frames of the scan's shape")
        u16 = f32 = None
        if c.dtype == np.uint16:
            c = np.ascontiguousarray(c)
            u16 = _ptr(c, ctypes.c_uint16)
        elif np.issubdtype(c.dtype, np.floating):
            c = np.ascontiguousarray(c, dtype=np.float32)
            f32 = _ptr(c, ctypes.c_float)
        else:
            raise TypeError("counts must be uint16 or float")
        _check(library().ptyrs_engine_live_push(self._h, u16, f32, c.shape[0]))

    @property
    def live_frames(self):
        """(landed, planned) frames."""
        a, b = _SZ(), _SZ()
        _check(library().ptyrs_engine_live_frames(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def set_checkpoint(self, every):
        """Keeps the wave entering every `every`-th slice for the multislice adjoint and
        recomputes the others on the way back (0 or 1: keep every slice's, the reference's
        behaviour); results are bit-identical."""
        if int(every) < 0:
            raise ValueError("checkpoint interval must be >= 0")
        _check(library().ptyrs_engine_set_checkpoint(self._h, int(every)))

    @property
    def checkpoint(self):
        """`(every, wave_bytes)`: the checkpoint interval and the device bytes the kept incident
        waves hold."""
        k, b = ctypes.c_size_t(), ctypes.c_size_t()
        _check(library().ptyrs_engine_checkpoint(self._h, ctypes.byref(k), ctypes.byref(b)))
        return k.value, b.value

    def llk(self):
        """Normalised PyNX LLK of the current object and probe."""
        r = _RecordC()
        _check(library().ptyrs_engine_llk(self._h, ctypes.byref(r)))
        return dict(poisson=r.poisson, gaussian=r.gaussian, euclidian=r.euclidian,
                    nb_photons_calc=r.nb_photons_calc)

    def llk_noalloc(self):
        """The same LLK computed without allocating device memory (ptyrs_engine_llk_noalloc:
        the work buffers live in one static 64 MiB array per GPU; calls overlapping on one
        GPU race)."""
        r = _RecordC()
        _check(library().ptyrs_engine_llk_noalloc(self._h, ctypes.byref(r)))
        return dict(poisson=r.poisson, gaussian=r.gaussian, euclidian=r.euclidian,
                    nb_photons_calc=r.nb_photons_calc)

    @property
    def objects(self):
        """The object's slices `[nslice, nobj, nyo, nxo]` in the order the wave crosses them."""
        out = np.empty(self._objs_shape, dtype=np.complex128)
        _check(library().ptyrs_engine_object(self._h, _ptr(out.view(np.float64), ctypes.c_double),
                                             out.size * 2))
        return out

    @property
    def object(self):
        """The object; of a multislice object its first slice."""
        return self.objects[0]

    @property
    def probe(self):
        out = np.empty(self._probe_shape, dtype=np.complex128)
        _check(library().ptyrs_engine_probe(self._h, _ptr(out.view(np.float64), ctypes.c_double),
                                            out.size * 2))
        return out

    def set_object(self, obj):
        """Replaces the object: `[nobj, nyo, nxo]`, of a multislice object every slice
        `[nslice, nobj, nyo, nxo]`."""
        a = np.ascontiguousarray(obj, dtype=np.complex128)
        if a.shape != self._obj_shape:
            raise ValueError(f"object {a.shape} differs from {self._obj_shape}")
        _check(library().ptyrs_engine_set_object(self._h, _ptr(a.view(np.float64), ctypes.c_double),
                                                 a.size * 2))

    def set_features(self, *, scale=None, background=None, position=None, hold=()):
        """As `Ptycho.set_features` (ptyrs-py/src/gpu.rs:199-213, lib.rs:313-345): `position`
        `[nframe, 2]` offsets `[dy, dx]` in object pixels, refined by LSQML unless `hold` names
        "position"; left out, none are modelled. This engine models no intensity scale or
        background."""
        hold = list(hold)
        for h in hold:
            if h not in ("scale", "background", "position"):
                raise ValueError(f"hold names {h!r}, not scale, background or position")
            if {"scale": scale, "background": background, "position": position}[h] is None:
                raise ValueError(f"hold names {h!r}, which is not given")
        if scale is not None or background is not None:
            raise ValueError("this GPU engine models no intensity scale or background")
        lib = library()
        if position is None:
            _check(lib.ptyrs_engine_set_position(self._h, None, 0, 0))
            return
        d = np.ascontiguousarray(position, dtype=np.float64)
        if d.ndim != 2 or d.shape[1] != 2:
            raise ValueError(f"position offsets must be [nframe, 2], got {list(d.shape)}")
        if d.shape[0] != self.nframe:
            raise ValueError(f"{d.shape[0]} position offsets for {self.nframe} frames")
        _check(lib.ptyrs_engine_set_position(self._h, _ptr(d, ctypes.c_double), d.size,
                                             int("position" not in hold)))

    @property
    def features(self):
        """As `Ptycho.features`: `scale`, `background`, `position` (None where not modelled)
        and `hold`, the names of those held."""
        out = np.empty((self.nframe, 2))
        state = ctypes.c_int()
        _check(library().ptyrs_engine_position(self._h, _ptr(out, ctypes.c_double), out.size,
                                               ctypes.byref(state)))
        position = out if state.value else None
        return dict(scale=None, background=None, position=position,
                    hold=["position"] if state.value == 1 else [])

    @property
    def launch(self):
        """The launch shapes that change results only by rounding: `stack_size` (frames per
        stack of AP, DM, RAAR, ML and the LLK pass, 0 the default), `pass_c_blocks` and
        `chi_blocks` (workgroups per launch rPIE's and LSQML's pass C aim at)."""
        c = _LaunchC()
        _check(library().ptyrs_engine_launch(self._h, ctypes.byref(c)))
        return dict(stack_size=c.stack_size, pass_c_blocks=c.pass_c_blocks,
                    chi_blocks=c.chi_blocks)

    def set_launch(self, **kw):
        """Sets any of `launch`'s keys (0 restores a default)."""
        unknown = set(kw) - {"stack_size", "pass_c_blocks", "chi_blocks"}
        if unknown:
            raise TypeError(f"unexpected keyword arguments {sorted(unknown)}")
        cur = self.launch
        cur.update(kw)
        c = _LaunchC()
        c.stack_size = int(cur["stack_size"])
        c.pass_c_blocks, c.chi_blocks = int(cur["pass_c_blocks"]), int(cur["chi_blocks"])
        _check(library().ptyrs_engine_set_launch(self._h, ctypes.byref(c)))

    @property
    def sampling(self):
        """`"patch"` or `"full_object"` (Sampling::auto of the optics, object and frame until
        set)."""
        s = ctypes.c_int()
        _check(library().ptyrs_engine_sampling(self._h, ctypes.byref(s)))
        return "patch" if s.value == 0 else "full_object"

    def set_sampling(self, sampling):
        if sampling not in ("patch", "full_object"):
            raise ValueError(f"sampling must be \"patch\" or \"full_object\", got {sampling!r}")
        _check(library().ptyrs_engine_set_sampling(self._h, 0 if sampling == "patch" else 1))

    def set_probe(self, probe):
        a = np.ascontiguousarray(probe, dtype=np.complex128)
        if a.shape != self._probe_shape:
            raise ValueError(f"probe {a.shape} differs from {self._probe_shape}")
        _check(library().ptyrs_engine_set_probe(self._h, _ptr(a.view(np.float64), ctypes.c_double),
                                                a.size * 2))

    @property
    def history(self):
        """LLK records as columns: cycle, algorithm and the Llk fields (ptyrs-py/src/lib.rs:395-411)."""
        lib = library()
        n = ctypes.c_size_t()
        _check(lib.ptyrs_engine_history(self._h, None, 0, ctypes.byref(n)))
        recs = (_RecordC * max(n.value, 1))()
        _check(lib.ptyrs_engine_history(self._h, recs, n.value, ctypes.byref(n)))
        rs = recs[:n.value]
        return dict(cycle=np.array([r.cycle for r in rs], dtype=np.int64),
                    algorithm=[_ALGORITHMS[r.algorithm] for r in rs],
                    poisson=np.array([r.poisson for r in rs]),
                    gaussian=np.array([r.gaussian for r in rs]),
                    euclidian=np.array([r.euclidian for r in rs]),
                    nb_photons_calc=np.array([r.nb_photons_calc for r in rs]))

    @property
    def cycle(self):
        c = ctypes.c_size_t()
        _check(library().ptyrs_engine_cycle(self._h, ctypes.byref(c)))
        return c.value

    @property
    def corners(self):
        out = np.empty((self.nframe, 2), dtype=np.int64)
        _check(library().ptyrs_engine_corners(self._h, _ptr(out, ctypes.c_int64), out.size))
        return out

    def synchronize(self):
        _check(library().ptyrs_engine_synchronize(self._h))

    def profile(self, enable=True):
        """Times every launch with HIP events on the engine's stream from now on."""
        _check(library().ptyrs_engine_profile(self._h, int(bool(enable))))

    def kernel_stats(self):
        """{kernel: (launches, total ms)} of the launches timed since `profile()`."""
        lib = library()
        out = {}
        for name in lib.ptyrs_kernel_names().decode().split(","):
            n, ms = ctypes.c_size_t(), ctypes.c_double()
            _check(lib.ptyrs_engine_kernel_stats(self._h, name.encode(), ctypes.byref(n),
                                                 ctypes.byref(ms)))
            out[name] = (n.value, ms.value)
        return out


def rpie_group(engines, nepoch, params=None, one_graph=False):
    """`nepoch` rPIE epochs of several independent scans at once (ptyrs_group_rpie): each
    engine's epochs as `GpuPtycho.rpie` runs them, the engines side by side on the device, each
    epoch a graph replay on the engine's own stream — or, with `one_graph`, every engine's epoch
    in one graph launch (ptyrs_group_rpie_graph). `params` is one dict of RpieParams fields for
    all, or one dict per engine."""
    engines = list(engines)
    if params is None or isinstance(params, dict):
        params = [params or {}] * len(engines)
    if len(params) != len(engines):
        raise ValueError(f"{len(params)} parameter sets for {len(engines)} engines")
    hs = (ctypes.c_void_p * max(len(engines), 1))(*[g._h.value for g in engines])
    cs = (_RpieC * max(len(engines), 1))(*[GpuPtycho._rpie_c(p) for p in params])
    f = library().ptyrs_group_rpie_graph if one_graph else library().ptyrs_group_rpie
    _check(f(hs, len(engines), cs, int(nepoch)))
