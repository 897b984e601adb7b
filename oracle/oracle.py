"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: this module is the checker for the HIP locate path.
Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg may
import it; the product package never does.

The oracle restates the reference's locate pipeline in C (see gt_oracle.h for
the file:line citations and the arithmetic contract).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("diameter", ctypes.c_uint32),
        ("minmass", ctypes.c_float),
        ("separation", ctypes.c_uint32),
        ("noise_size", ctypes.c_float),
        ("smoothing_size", ctypes.c_uint32),
        ("max_iterations", ctypes.c_uint32),
        ("has_percentile", ctypes.c_int32),
        ("percentile", ctypes.c_float),
        ("characterize", ctypes.c_int32),
        ("cpu_processed", ctypes.c_int32),
        ("literal", ctypes.c_int32),
    ]


class OrcGeom(ctypes.Structure):
    _fields_ = [
        ("k", ctypes.c_int32),
        ("R", ctypes.c_int32),
        ("band", ctypes.c_int32),
        ("rint", ctypes.c_int32),
        ("kd", ctypes.c_int32),
        ("a", ctypes.c_int32),
        ("r2", ctypes.c_float),
    ]


def build() -> str:
    """Compile the oracle with its Makefile (idempotent)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def use_native_build(out_dir: str) -> str:
    """Compile the oracle for this machine's CPU (ARCH=-march=native) into
    out_dir and load that build instead of oracle/_build's: the CPU baseline
    of bench.py runs on the host it is timed on.  Returns the library path."""
    global _lib, _LIB_PATH
    out_dir = os.path.abspath(out_dir)
    subprocess.run(["make", "-s", "-B", "-C", _HERE, "ARCH=-march=native", f"OUT={out_dir}"],
                   check=True)
    _LIB_PATH = os.path.join(out_dir, "liboracle.so")
    _lib = None
    return _LIB_PATH


def use_library(path: str) -> None:
    """Load the oracle from another build of the same sources (for instance
    the AddressSanitizer build of `make asan`)."""
    global _lib, _LIB_PATH
    _LIB_PATH = os.path.abspath(path)
    _lib = None


def set_threads(n: int) -> int:
    """Threads over the rows of one frame (1: serial); returns the previous."""
    return lib().orc_set_threads(int(n))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        fp = ctypes.POINTER(ctypes.c_float)
        dp = ctypes.POINTER(ctypes.c_double)
        i64p = ctypes.POINTER(ctypes.c_int64)
        u8p = ctypes.POINTER(ctypes.c_uint8)
        L.orc_geometry.argtypes = [ctypes.POINTER(OrcParams), ctypes.POINTER(OrcGeom)]
        L.orc_gauss_weights.argtypes = [ctypes.c_int32, ctypes.c_float, fp, fp]
        L.orc_bandpass_f32.argtypes = [fp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                       ctypes.c_float, fp]
        L.orc_bandpass_f64.argtypes = [dp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                       ctypes.c_double, dp]
        L.orc_masses_centers_f32.argtypes = [fp, ctypes.c_int32, ctypes.c_int32,
                                             ctypes.c_uint32, fp, fp]
        L.orc_is_max_map.argtypes = [fp, ctypes.c_int32, ctypes.c_int32, ctypes.c_uint32, u8p]
        L.orc_percentile_threshold.argtypes = [fp, ctypes.c_size_t, ctypes.c_float]
        L.orc_percentile_threshold.restype = ctypes.c_double
        L.orc_locate_frame.argtypes = [fp, ctypes.c_int32, ctypes.c_int32,
                                       ctypes.POINTER(OrcParams), ctypes.c_int32,
                                       ctypes.c_int32, dp, i64p, ctypes.c_int64, i64p, dp]
        L.orc_ncols.argtypes = [ctypes.POINTER(OrcParams)]
        L.orc_ncols.restype = ctypes.c_int32
        L.orc_candidates_frame.argtypes = [fp, ctypes.c_int32, ctypes.c_int32,
                                           ctypes.POINTER(OrcParams), i64p, fp, fp, fp,
                                           ctypes.c_int64, i64p]
        L.orc_dedup_bruteforce.argtypes = [fp, fp, fp, i64p, ctypes.c_int64,
                                           ctypes.c_uint32, u8p]
        i32p = ctypes.POINTER(ctypes.c_int32)
        L.orc_adjoint_particle.argtypes = [fp, ctypes.c_int32, ctypes.c_int32,
                                           ctypes.POINTER(OrcParams), ctypes.c_int32,
                                           ctypes.c_int64, dp, i32p, i32p, dp, i32p]
        L.orc_particle_f64.argtypes = [dp, ctypes.c_int32, ctypes.c_int32,
                                       ctypes.POINTER(OrcParams), ctypes.c_int64, dp, i32p, i32p]
        L.orc_link_rows.argtypes = [fp, ctypes.c_int64, ctypes.c_int32, ctypes.c_float,
                                    ctypes.c_int64, ctypes.POINTER(ctypes.c_uint64)]
        L.orc_link_rows.restype = ctypes.c_int
        L.orc_char_tables.argtypes = [ctypes.c_int32, fp, fp, fp, ctypes.c_int32, i32p]
        L.orc_char_tables.restype = ctypes.c_int32
        L.orc_characterise_at.argtypes = [fp, fp, ctypes.c_int32, ctypes.c_int32,
                                          ctypes.POINTER(OrcParams), ctypes.c_int32,
                                          ctypes.c_int32, ctypes.c_float, fp, fp]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_set_threads.restype = ctypes.c_int
        L.orc_set_mutation.argtypes = [ctypes.c_int]
        L.orc_set_mutation.restype = ctypes.c_int
        _lib = L
    return _lib


def link_rows_c(rows: np.ndarray, search_range: float, memory: int = 0) -> np.ndarray:
    """The C restatement of link_rust (gt_linker.c): uint64 id per row."""
    a = np.ascontiguousarray(rows, np.float32)
    if a.ndim != 2 or a.shape[1] < 4:
        raise ValueError("rows must be n x >= 4 float32")
    ids = np.zeros(a.shape[0], np.uint64)
    rc = lib().orc_link_rows(_ptr(a, ctypes.c_float), a.shape[0], a.shape[1],
                             float(search_range), int(memory),
                             ids.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
    if rc != 0:
        raise MemoryError("orc_link_rows failed")
    return ids


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


@dataclass
class Params:
    """Locate parameters, Python defaults of lib.rs:77-91."""
    diameter: int
    minmass: float = 0.0
    separation: int | None = None
    noise_size: float = 1.0
    smoothing_size: int | None = None
    max_iterations: int = 10
    percentile: float | None = None
    characterize: bool = False
    cpu_processed: bool = False
    literal: bool = False

    def c(self) -> OrcParams:
        return OrcParams(
            diameter=self.diameter,
            minmass=self.minmass,
            separation=self.diameter + 1 if self.separation is None else self.separation,
            noise_size=self.noise_size,
            smoothing_size=self.diameter if self.smoothing_size is None else self.smoothing_size,
            max_iterations=self.max_iterations,
            has_percentile=0 if self.percentile is None else 1,
            percentile=0.0 if self.percentile is None else self.percentile,
            characterize=int(self.characterize),
            cpu_processed=int(self.cpu_processed),
            literal=int(self.literal),
        )


def geometry(p: Params) -> OrcGeom:
    g = OrcGeom()
    cp = p.c()
    if lib().orc_geometry(ctypes.byref(cp), ctypes.byref(g)) != 0:
        raise ValueError("invalid parameters")
    return g


def gauss_weights(k: int, noise_size: float):
    w = np.zeros(2 * (k // 2) + 1, np.float32)
    norm = ctypes.c_float()
    lib().orc_gauss_weights(k, noise_size, _ptr(w, ctypes.c_float), ctypes.byref(norm))
    return w, np.float32(norm.value)


def bandpass(frame: np.ndarray, k: int, noise_size: float = 1.0, f64: bool = False) -> np.ndarray:
    H, W = frame.shape
    if f64:
        x = np.ascontiguousarray(frame, np.float64)
        out = np.zeros((H, W), np.float64)
        rc = lib().orc_bandpass_f64(_ptr(x, ctypes.c_double), H, W, k, noise_size,
                                    _ptr(out, ctypes.c_double))
    else:
        x = np.ascontiguousarray(frame, np.float32)
        out = np.zeros((H, W), np.float32)
        rc = lib().orc_bandpass_f32(_ptr(x, ctypes.c_float), H, W, k, noise_size,
                                    _ptr(out, ctypes.c_float))
    if rc != 0:
        raise ValueError(f"bandpass rejected input (rc={rc})")
    return out


def masses_centers(P: np.ndarray, diameter: int):
    H, W = P.shape
    x = np.ascontiguousarray(P, np.float32)
    m = np.zeros((H, W), np.float32)
    c = np.zeros((H, W, 2), np.float32)
    lib().orc_masses_centers_f32(_ptr(x, ctypes.c_float), H, W, diameter,
                                 _ptr(m, ctypes.c_float), _ptr(c, ctypes.c_float))
    return m, c


def is_max_map(P: np.ndarray, separation: int) -> np.ndarray:
    H, W = P.shape
    x = np.ascontiguousarray(P, np.float32)
    out = np.zeros((H, W), np.uint8)
    lib().orc_is_max_map(_ptr(x, ctypes.c_float), H, W, separation, _ptr(out, ctypes.c_uint8))
    return out.astype(bool)


def percentile_threshold(P: np.ndarray, percentile: float) -> float:
    x = np.ascontiguousarray(P, np.float32).ravel()
    return lib().orc_percentile_threshold(_ptr(x, ctypes.c_float), x.size, percentile)


def ncols(p: Params) -> int:
    cp = p.c()
    return lib().orc_ncols(ctypes.byref(cp))


def locate_frame(frame: np.ndarray, p: Params, frame_index: int = 0, f64: bool = False):
    """Rows (float64, column_names order) and candidate pixel index per row,
    in raster order of the candidate pixel; also the percentile threshold."""
    H, W = frame.shape
    x = np.ascontiguousarray(frame, np.float32)
    cp = p.c()
    n = ctypes.c_int64()
    thr = ctypes.c_double()
    L = lib()
    nc = L.orc_ncols(ctypes.byref(cp))
    cap = 4096
    while True:
        rows = np.zeros((cap, nc), np.float64)
        cand = np.zeros(cap, np.int64)
        rc = L.orc_locate_frame(_ptr(x, ctypes.c_float), H, W, ctypes.byref(cp), frame_index,
                                int(f64), _ptr(rows, ctypes.c_double),
                                _ptr(cand, ctypes.c_int64), cap, ctypes.byref(n),
                                ctypes.byref(thr))
        if rc == -3:
            cap = int(n.value)
            continue
        if rc != 0:
            raise ValueError(f"locate rejected input (rc={rc})")
        k = int(n.value)
        return rows[:k].copy(), cand[:k].copy(), thr.value


def candidates_frame(frame: np.ndarray, p: Params):
    H, W = frame.shape
    x = np.ascontiguousarray(frame, np.float32)
    cp = p.c()
    n = ctypes.c_int64()
    cap = 4096
    while True:
        cand = np.zeros(cap, np.int64)
        y = np.zeros(cap, np.float32)
        xx = np.zeros(cap, np.float32)
        m = np.zeros(cap, np.float32)
        rc = lib().orc_candidates_frame(_ptr(x, ctypes.c_float), H, W, ctypes.byref(cp),
                                        _ptr(cand, ctypes.c_int64), _ptr(y, ctypes.c_float),
                                        _ptr(xx, ctypes.c_float), _ptr(m, ctypes.c_float),
                                        cap, ctypes.byref(n))
        if rc == -3:
            cap = int(n.value)
            continue
        if rc != 0:
            raise ValueError(f"candidates rejected input (rc={rc})")
        k = int(n.value)
        return cand[:k], y[:k], xx[:k], m[:k]


def dedup_bruteforce(y, x, mass, idx, separation: int) -> np.ndarray:
    y = np.ascontiguousarray(y, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    mass = np.ascontiguousarray(mass, np.float32)
    idx = np.ascontiguousarray(idx, np.int64)
    keep = np.zeros(len(y), np.uint8)
    lib().orc_dedup_bruteforce(_ptr(y, ctypes.c_float), _ptr(x, ctypes.c_float),
                               _ptr(mass, ctypes.c_float), _ptr(idx, ctypes.c_int64),
                               len(y), separation, _ptr(keep, ctypes.c_uint8))
    return keep.astype(bool)


def adjoint_particle(frame: np.ndarray, p: Params, cand: int, f64_forward: bool = True):
    """Binary64 Jacobian of (y, x, mass) of the particle started at pixel
    `cand` w.r.t. the input pixels, on its S x S patch.  Returns (jac
    [3, S, S], origin (row, col), (y, x, mass), state (ue, ve, uf, vf))."""
    f = np.ascontiguousarray(frame, np.float32)
    H, W = f.shape
    g = geometry(p)
    S = 2 * (g.rint + g.R) + 1
    jac = np.zeros((3, S, S), np.float64)
    s_out = ctypes.c_int32()
    origin = np.zeros(2, np.int32)
    yxm = np.zeros(3, np.float64)
    state = np.zeros(4, np.int32)
    cp = p.c()
    rc = lib().orc_adjoint_particle(_ptr(f, ctypes.c_float), H, W, ctypes.byref(cp),
                                    int(f64_forward), int(cand), _ptr(jac, ctypes.c_double),
                                    ctypes.byref(s_out), _ptr(origin, ctypes.c_int32),
                                    _ptr(yxm, ctypes.c_double), _ptr(state, ctypes.c_int32))
    if rc != 0:
        raise RuntimeError(f"orc_adjoint_particle failed ({rc})")
    return jac, (int(origin[0]), int(origin[1])), yxm, state


def particle_f64(frame: np.ndarray, p: Params, cand: int):
    """Binary64 forward of one particle from a float64 frame: ((y, x, mass),
    state (ue, ve, uf, vf), number of disk pixels with P > 0)."""
    f = np.ascontiguousarray(frame, np.float64)
    H, W = f.shape
    yxm = np.zeros(3, np.float64)
    state = np.zeros(4, np.int32)
    na = ctypes.c_int32()
    cp = p.c()
    rc = lib().orc_particle_f64(_ptr(f, ctypes.c_double), H, W, ctypes.byref(cp), int(cand),
                                _ptr(yxm, ctypes.c_double), _ptr(state, ctypes.c_int32),
                                ctypes.byref(na))
    if rc != 0:
        raise RuntimeError(f"orc_particle_f64 failed ({rc})")
    return yxm, state, int(na.value)


# Planted mistakes of the characterisation (gt_oracle.h ORC_MUT_*), for the
# tests that show their checks catch each one.
MUTATIONS = {
    "gpu_radius_band": 1,
    "cpu_radius_rint": 2,
    "cpu_candidate_mass": 3,
    "theta_not_2theta": 4,
    "ecc_full_mass": 5,
    "signal_centre": 6,
    "r_not_r2": 7,
    "middle_most_off": 8,
    "raw_from_processed": 9,
}


class mutation:
    """Context manager: run the oracle with one planted mistake."""

    def __init__(self, name: str):
        self.m = MUTATIONS[name]

    def __enter__(self):
        self.old = lib().orc_set_mutation(self.m)
        return self

    def __exit__(self, *exc):
        lib().orc_set_mutation(self.old)
        return False


def char_tables(r: int):
    """(r2, sin2, cos2, middle index) of the disk of integer radius r, binary32."""
    cap = (2 * r + 1) ** 2
    r2 = np.zeros(cap, np.float32)
    s2 = np.zeros(cap, np.float32)
    c2 = np.zeros(cap, np.float32)
    mid = ctypes.c_int32()
    n = lib().orc_char_tables(r, _ptr(r2, ctypes.c_float), _ptr(s2, ctypes.c_float),
                              _ptr(c2, ctypes.c_float), cap, ctypes.byref(mid))
    if n < 0:
        raise ValueError("invalid radius")
    return r2[:n].copy(), s2[:n].copy(), c2[:n].copy(), int(mid.value)


def characterise_at(P: np.ndarray, raw: np.ndarray, p: Params, u: int, v: int,
                    walked_mass: float):
    """(Rg, ecc, signal) and (Rg_gpu, raw_mass_gpu, signal_gpu, ecc_gpu) of the
    candidate pixel (u, v) of the bandpassed frame P, binary32."""
    Pc = np.ascontiguousarray(P, np.float32)
    Rc = np.ascontiguousarray(raw, np.float32)
    if Pc.shape != Rc.shape:
        raise ValueError("P and raw differ in shape")
    H, W = Pc.shape
    cpu3 = np.zeros(3, np.float32)
    gpu4 = np.zeros(4, np.float32)
    cp = p.c()
    rc = lib().orc_characterise_at(_ptr(Pc, ctypes.c_float), _ptr(Rc, ctypes.c_float), H, W,
                                   ctypes.byref(cp), int(u), int(v), float(walked_mass),
                                   _ptr(cpu3, ctypes.c_float), _ptr(gpu4, ctypes.c_float))
    if rc != 0:
        raise ValueError(f"orc_characterise_at rejected input (rc={rc})")
    return cpu3, gpu4
