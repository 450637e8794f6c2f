"""The C-ABI library: it loads, exports every entry point include/ptyrs_hip.h declares, and
refuses bad input before it touches a device (no compute calls: this runs without a GPU)."""

import re

import numpy as np
import pytest

import ptyrs_amd as pa
from conftest import ROOT


def declared():
    text = (ROOT / "include" / "ptyrs_hip.h").read_text()
    return sorted(set(re.findall(r"\b(ptyrs_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_entry_point():
    lib = pa.library()
    names = declared()
    assert "ptyrs_engine_create" in names and "ptyrs_engine_rpie" in names
    for name in names:
        assert hasattr(lib, name), name
    assert set(pa.library().ptyrs_kernel_names().decode().split(",")) >= {"pass_a", "pass_b", "pass_c"}


def _scan(n=32, nframe=3):
    counts = np.zeros((nframe, n, n), dtype=np.uint16)
    y = np.array([0.0, 1e-8, -1e-8])[:nframe]
    return pa.Scan(counts, y, y, 1.5e-10, 1.0, 75e-6)


def test_bad_input_is_a_value_error_before_any_device_call():
    obj = np.ones((1, 64, 64), complex)
    probe = np.ones((1, 32, 32), complex)
    with pytest.raises(ValueError, match="not supported"):
        counts = np.zeros((3, 34, 34), np.uint16)
        pa.GpuPtycho(pa.Scan(counts, [0, 0, 0], [0, 0, 0], 1.5e-10, 1.0, 75e-6), obj,
                     np.ones((1, 34, 34), complex))
    far = pa.Scan(np.zeros((3, 32, 32), np.uint16), [0.0, 3e-7, -3e-7], [0.0, 0.0, 0.0],
                  1.5e-10, 1.0, 75e-6)  # ±4.8 object pixels
    with pytest.raises(ValueError, match="leaves the object"):
        pa.GpuPtycho(far, np.ones((1, 36, 36), complex), probe)
    with pytest.raises(ValueError):
        pa.GpuPtycho(_scan(), obj, np.ones((1, 16, 16), complex))
    with pytest.raises(ValueError, match="dtype"):
        pa.GpuPtycho(_scan(), obj, probe, dtype="f16")


def test_unknown_keywords_are_type_errors():
    with pytest.raises(TypeError, match="unknown"):
        pa.GpuPtycho(_scan(), np.ones((1, 64, 64), complex), np.ones((1, 32, 32), complex),
                     nonsense=1)


def test_scan_rejects_non_finite_counts_and_bad_geometry():
    with pytest.raises(ValueError):
        pa.Scan(np.full((1, 4, 4), np.nan), [0.0], [0.0], 1e-10, 1.0, 1e-4)
    with pytest.raises(ValueError):
        pa.Scan(np.zeros((1, 4, 4)), [0.0], [0.0], -1e-10, 1.0, 1e-4)
    with pytest.raises(TypeError):
        pa.Scan(np.zeros((4, 4)), [0.0], [0.0], 1e-10, 1.0, 1e-4)


def test_gpu_options_refusals_happen_before_any_device_call():
    obj = np.ones((1, 64, 64), complex)
    probe = np.ones((1, 32, 32), complex)
    # GpuOptions (ptyrs/src/engine/gpu/mod.rs:61-108): stack_size Some(0) and llk_slots 0 are
    # refused (psi_f16's refusals: test_psi_f16_needs_fused_passes_and_complex64).
    with pytest.raises(ValueError, match="stack_size"):
        pa.GpuPtycho(_scan(), obj, probe, stack_size=0)
    with pytest.raises(ValueError, match="llk_slots"):
        pa.GpuPtycho(_scan(), obj, probe, llk_slots=0)
    with pytest.raises(ValueError, match="65535 views"):
        pa.GpuPtycho(pa.Scan(np.zeros((20000, 32, 32), np.uint16), np.zeros(20000),
                             np.zeros(20000), 1.5e-10, 1.0, 75e-6),
                     obj, np.ones((4, 32, 32), complex), stack_size=20000)


def test_non_finite_float_counts_are_refused_by_the_library():
    from ptyrs_amd import base as pb
    c = np.zeros((1, 16, 16), np.float32)
    s = pa.Scan(c, [0.0], [0.0], 1.0, 16.0, 1.0)
    s.counts[0, 2, 2] = np.inf  # past the Scan's own check
    with pytest.raises(ValueError, match="finite"):
        pb.observed(s)


def test_the_fault_hook_arms_and_disarms_without_a_device():
    from ptyrs_amd import base as pb
    pb.arm_fault("launch", 3)
    assert not pb.fault_fired()
    pb.arm_fault(None)
    with pytest.raises(KeyError):
        pb.arm_fault("nowhere")
    assert pa.library().ptyrs_test_arm_fault(7, 0) == 1


def test_psi_f16_needs_fused_passes_and_complex64():
    # GpuOptions::psi_f16 without fused passes is refused (ptyrs/tests/engine_gpu_f16.rs:139);
    # its f16 pairs hold complex64 waves. Refused before any device call.
    obj = np.ones((1, 64, 64), complex)
    probe = np.ones((1, 32, 32), complex)
    with pytest.raises(ValueError, match="psi_f16 needs"):
        pa.GpuPtycho(_scan(), obj, probe, psi_f16=True, fused=False)
    with pytest.raises(ValueError, match="c64"):
        pa.GpuPtycho(_scan(), obj, probe, psi_f16=True, dtype="c128")
