"""The C ABI of libgpu_tracking.so without a GPU: the library loads, exports
every function include/gpu_tracking.h declares, its structs have the layout
the Python binding assumes, and the parameter / column / error handling
follows the reference (lib.rs:73-91, execute_gpu.rs:40-65,197-226).
No compute call is made here."""
import ctypes
import os
import re
import struct
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gpu_tracking.h")


@pytest.fixture(scope="module")
def lib():
    from tracking_optimizations_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        from tracking_optimizations_amd import _build
        _build.build(verbose=False)
    return _lib.load()


def _declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gt_[a-z0-9_]+)\s*\(", text)))


def test_exports_every_declared_function(lib):
    from tracking_optimizations_amd import _lib
    declared = _declared_functions()
    assert len(declared) >= 25
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\sT\s(gt_\w+)", out))
    missing = [f for f in declared if f not in exported]
    assert not missing, missing
    assert sorted(_lib.EXPORTS) == declared


def test_struct_layout_matches_ctypes(tmp_path):
    """Compile a C probe against the header and compare sizes/offsets."""
    from tracking_optimizations_amd._lib import GtFrameDesc, GtParams, GtTiming, GtTuning
    fields = {"gt_params": GtParams, "gt_timing": GtTiming, "gt_frame_desc": GtFrameDesc,
              "gt_tuning": GtTuning}
    src = ['#include <stdio.h>', '#include <stddef.h>', '#include "gpu_tracking.h"', "int main(void){"]
    for cname, cls in fields.items():
        src.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            src.append(f'printf("{cname}.{fname} %zu\\n", offsetof({cname}, {fname}));')
    src.append("return 0;}")
    c = tmp_path / "probe.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(c), "-o", str(exe)], check=True)
    got = dict(l.split() for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                 check=True).stdout.splitlines())
    for cname, cls in fields.items():
        assert int(got[cname]) == ctypes.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert int(got[f"{cname}.{fname}"]) == getattr(cls, fname).offset, (cname, fname)


def test_python_defaults(lib):
    """batch_rust's Option defaults, lib.rs:77-91."""
    from tracking_optimizations_amd._lib import GtParams
    p = GtParams()
    assert lib.gt_params_default(ctypes.byref(p), 9) == 0
    assert (p.diameter, p.separation, p.smoothing_size, p.max_iterations) == (9, 10, 9, 10)
    assert (p.minmass, p.noise_size, p.percentile) == (0.0, 1.0, 64.0)
    assert abs(p.threshold - 1 / 255) < 1e-9 and p.maxsize == float("inf")
    assert (p.characterize, p.filter_close, p.cpu_processed, p.preprocess) == (0, 1, 0, 1)
    assert not any([p.has_maxsize, p.has_threshold, p.has_invert, p.has_percentile, p.has_topn,
                    p.has_preprocess, p.has_search_range, p.has_memory, p.has_sig_radius,
                    p.has_bg_radius, p.has_gap_radius])


def test_rust_defaults(lib):
    """impl Default for TrackingParams, execute_gpu.rs:40-65."""
    from tracking_optimizations_amd._lib import GtParams
    p = GtParams()
    assert lib.gt_params_rust_default(ctypes.byref(p)) == 0
    assert (p.diameter, p.separation, p.smoothing_size, p.max_iterations) == (9, 11, 9, 10)
    assert (p.cpu_processed, p.characterize, p.filter_close) == (1, 0, 1)


@pytest.mark.parametrize("kw,status", [
    (dict(maxsize=3.0), 2), (dict(threshold=0.1), 2), (dict(invert=True), 2),
    (dict(topn=5), 2), (dict(preprocess=False), 2), (dict(search_range=5.0), 0),
    (dict(sig_radius=3.0), 0), (dict(bg_radius=8.0), 1), (dict(sig_radius=3.0, bg_radius=8.0), 0),
    (dict(noise_size=0.0), 1), (dict(percentile=101.0), 1), (dict(percentile=-1.0), 1),
    (dict(smoothing_size=0), 1), (dict(percentile=64.0), 0), (dict(), 0),
])
def test_validate(lib, kw, status):
    from tracking_optimizations_amd.gpu_tracking import make_params
    p = make_params(9, **kw)
    assert lib.gt_params_validate(ctypes.byref(p), 0, 0) == status
    if status:
        assert lib.gt_last_error().decode()


def test_validate_frame_shape(lib):
    from tracking_optimizations_amd.gpu_tracking import make_params
    p = make_params(9, smoothing_size=11)
    assert lib.gt_params_validate(ctypes.byref(p), 64, 64) == 0
    assert lib.gt_params_validate(ctypes.byref(p), 4, 64) == 1   # h < smoothing_size/2
    assert lib.gt_params_validate(ctypes.byref(p), 5, 5) == 0


def test_column_names_order(lib):
    from tracking_optimizations_amd.gpu_tracking import column_names, make_params
    names = lambda **kw: [c[0] for c in column_names(make_params(9, **kw))]
    assert names() == ["frame", "mass", "y", "x"]
    assert names(cpu_processed=True) == ["frame", "mass", "y", "x", "Rg", "ecc", "signal"]
    assert names(characterize=True, cpu_processed=True) == [
        "frame", "mass", "y", "x", "Rg", "ecc", "signal", "Rg_gpu", "raw_mass_gpu",
        "signal_gpu", "ecc_gpu"]
    assert names(search_range=3.0, sig_radius=2.0, bg_radius=5.0) == [
        "frame", "mass", "y", "x", "particle", "raw_mass", "raw_bg_median", "raw_mass_corrected"]
    dt = dict(column_names(make_params(9, search_range=1.0)))
    assert dt["frame"] == "int" and dt["particle"] == "int" and dt["mass"] == "float"


def test_no_device_behaviour(lib):
    """Without a GPU the library reports it; nothing aborts."""
    from tracking_optimizations_amd import _lib
    from tracking_optimizations_amd.gpu_tracking import make_params
    if _lib.device_count() > 0:
        pytest.skip("a device is present")
    h = ctypes.c_void_p()
    assert lib.gt_context_create(0, ctypes.byref(h)) == _lib.GT_E_NO_DEVICE
    out = ctypes.c_void_p()
    p = make_params(9)
    assert lib.gt_locate_u16(None, None, 0, 8, 8, ctypes.byref(p), 0, ctypes.byref(out)) == \
        _lib.GT_E_INVALID


def test_link_argument_checks(lib):
    """gt_link rejects a missing context before touching a device."""
    from tracking_optimizations_amd import _lib
    ids = (ctypes.c_uint64 * 1)()
    rows = (ctypes.c_float * 4)()
    assert lib.gt_link(None, rows, 1, 4, 5.0, 0, ids) == _lib.GT_E_INVALID
    assert "context" in lib.gt_last_error().decode()
    st = _lib.GtLinkStats()
    assert lib.gt_context_link_stats(None, ctypes.byref(st)) == _lib.GT_E_INVALID


def test_python_api_errors_before_device():
    from tracking_optimizations_amd import GtInvalid, GtNotImplemented, batch
    v = np.zeros((2, 32, 32), np.float32)
    with pytest.raises(GtNotImplemented):
        batch(v, 9, maxsize=3)
    with pytest.raises(GtNotImplemented):
        batch(v, 9, invert=True)
    with pytest.raises(GtInvalid):
        batch(v, 9, bg_radius=4.0)
    with pytest.raises(GtInvalid):
        batch(np.zeros((2, 32, 64), np.float32)[:, :, ::2], 9)
    with pytest.raises(GtInvalid):
        batch(np.zeros((2, 32, 32), np.float64), 9)


def test_batch_file_rejects_unknown_extension():
    from tracking_optimizations_amd import GtInvalid, batch_file
    with pytest.raises(GtInvalid):
        batch_file("movie.avi", 9)
