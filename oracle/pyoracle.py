"""ctypes front end of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker or the timed CPU baseline.  The product
library (nbfmm_amd) never loads it.

All arrays use the reference layouts: position float32[N, 2], weight
float32[N], index int32[N, 2], perm int32[N], head int32[B*B+1] and cell cubes
[levels][B][B] (see nbfmm_oracle.cpp).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None
_isa = None

# x86-64 ISA levels of the builds (oracle/Makefile) and the CPU flags each needs
_ISA_BUILDS = [
    ("x86-64-v4", "liboracle_v4.so", {"avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl"}),
    ("x86-64-v3", "liboracle_v3.so", {"avx2", "fma", "bmi2", "movbe"}),
    ("x86-64-v2", "liboracle.so", set()),
]


def _cpu_flags():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("flags"):
                return set(line.split(":", 1)[1].split())
    except OSError:
        pass
    return set()


def cpu_model() -> str:
    """The host CPU's model name (for the CPU baseline's record)."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _pick():
    flags = _cpu_flags()
    for isa, name, need in _ISA_BUILDS:
        path = os.path.join(_HERE, "_build", name)
        if need <= flags and os.path.exists(path):
            return isa, path
    return "x86-64-v2", _LIB_PATH


def isa() -> str:
    """ISA level of the loaded build."""
    lib()
    return _isa

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build() -> str:
    """Compile the oracle with its Makefile (g++ only; no GPU needed)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib, _isa
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _isa, path = _pick()
        L = ctypes.CDLL(path)
        c_int, c_float, c_double = ctypes.c_int, ctypes.c_float, ctypes.c_double
        L.orc_max_threads.restype = c_int
        L.orc_set_threads.argtypes = [c_int]
        L.orc_compute_index_f32.argtypes = [c_int, c_int, _f32p, _f32p, _i32p]
        L.orc_compute_index_f32.restype = c_int
        L.orc_predo_f32.argtypes = [c_int, c_int, _f32p, _f32p, _f32p, _f32p, _f32p, _i32p, _i32p, _i32p]
        L.orc_predo_f32.restype = c_int
        L.orc_p2p_f32.argtypes = [c_int, c_int, _f32p, _f32p, _i32p, _i32p, c_float, _f32p]
        L.orc_p2m_f32.argtypes = [c_int, c_int, _f32p, _f32p, _i32p, _f32p, _f32p]
        L.orc_m2m_f32.argtypes = [c_int, _f32p, _f32p]
        L.orc_m2l_f32.argtypes = [c_int, c_float, _f32p, _f32p, _f32p]
        L.orc_l2l_f32.argtypes = [c_int, _f32p]
        L.orc_l2p_f32.argtypes = [c_int, c_int, _i32p, _f32p, _f32p]
        L.orc_postdo_f32.argtypes = [c_int, _i32p, _f32p, _f32p]
        L.orc_solve_f32.argtypes = [c_int, c_int, _f32p, c_float, _f32p, _f32p, _f32p]
        L.orc_solve_f32.restype = c_int
        L.orc_solve_f64.argtypes = [c_int, c_int, _f64p, c_double, _f64p, _f64p, _f64p]
        L.orc_solve_f64.restype = c_int
        L.orc_update_f32.argtypes = [c_int, c_float, _f32p, _f32p, _f32p]
        L.orc_update_f64.argtypes = [c_int, c_double, _f64p, _f64p, _f64p]
        L.orc_run_f32.argtypes = [c_int, c_int, _f32p, c_float, c_float, c_int, _f32p, _f32p, _f32p, _f32p]
        L.orc_run_f32.restype = c_int
        L.orc_run_f64.argtypes = [c_int, c_int, _f64p, c_double, c_double, c_int, _f64p, _f64p, _f64p, _f64p]
        L.orc_run_f64.restype = c_int
        L.orc_run_f64_cells.argtypes = [c_int, c_int, _f64p, c_double, c_double, c_int, _f64p, _f64p, _f64p, _f64p,
                                        _i32p, c_int]
        L.orc_run_f64_cells.restype = c_int
        L.orc_direct_f64.argtypes = [c_int, _f32p, _f32p, c_double, c_int, ctypes.c_void_p, _f64p]
        L.orc_prune_f32.argtypes = [c_int, _f32p, _f32p, _f32p, _f32p]
        L.orc_prune_f32.restype = c_int
        c_uint = ctypes.c_uint
        L.orc_model_uniform.argtypes = [c_uint, c_int, c_int]
        L.orc_model_uniform.restype = c_float
        L.orc_model_disk.argtypes = [c_int, c_float, c_float, c_float, c_float, c_float, c_int, c_int, c_float,
                                     c_float, c_uint, _f32p, _f32p, _f32p]
        L.orc_model_circle.argtypes = [c_int, c_float, c_float, c_float, c_float, c_float, c_float, c_float, c_int,
                                       c_uint, _f32p, _f32p, _f32p]
        L.orc_model_rectangle.argtypes = [c_int, c_float, c_float, c_float, c_float, c_float, c_uint, _f32p, _f32p,
                                          _f32p]
        L.orc_double_disk_offsets.argtypes = [c_float] * 11 + [_f32p]
        L.orc_display.argtypes = [c_int, c_int, c_int, c_float, _f32p, _f32p, _f32p,
                                  np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")]
        L.orc_display.restype = None
        _lib = L
    return _lib


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def base_dim(L: int) -> int:
    return 1 << (L + 1)


def cube_levels(L: int) -> int:
    return max(L, 1)


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def max_threads() -> int:
    return int(lib().orc_max_threads())


def compute_index(L, pos, limits):
    pos = _f32(pos).reshape(-1, 2)
    index = np.zeros((pos.shape[0], 2), np.int32)
    clamped = lib().orc_compute_index_f32(L, pos.shape[0], pos, _f32(limits), index)
    return index, clamped


def predo(L, pos0, w0, limits):
    pos0 = _f32(pos0).reshape(-1, 2)
    w0 = _f32(w0)
    n = pos0.shape[0]
    B = base_dim(L)
    pos = np.zeros_like(pos0)
    w = np.zeros_like(w0)
    index = np.zeros((n, 2), np.int32)
    perm = np.zeros(n, np.int32)
    head = np.zeros(B * B + 1, np.int32)
    clamped = lib().orc_predo_f32(L, n, pos0, w0, _f32(limits), pos, w, index, perm, head)
    return dict(position=pos, weight=w, index=index, perm=perm, head=head, clamped=clamped)


def p2p(L, pos, w, index, head, eps):
    pos = _f32(pos).reshape(-1, 2)
    eff = np.zeros_like(pos)
    lib().orc_p2p_f32(L, pos.shape[0], pos, _f32(w), _i32(index), _i32(head), float(eps), eff)
    return eff


def p2m(L, pos, w, head):
    B = base_dim(L)
    cube = B * B * cube_levels(L)
    cpos = np.zeros((cube, 2), np.float32)
    cw = np.zeros(cube, np.float32)
    pos = _f32(pos).reshape(-1, 2)
    lib().orc_p2m_f32(L, pos.shape[0], pos, _f32(w), _i32(head), cpos, cw)
    return cpos, cw


def m2m(L, cell_pos, cell_w):
    cpos = _f32(cell_pos).copy()
    cw = _f32(cell_w).copy()
    lib().orc_m2m_f32(L, cpos, cw)
    return cpos, cw


def m2l(L, eps, cell_pos, cell_w):
    cpos = _f32(cell_pos)
    ceff = np.zeros_like(cpos)
    lib().orc_m2l_f32(L, float(eps), cpos, _f32(cell_w), ceff)
    return ceff


def l2l(L, cell_eff):
    ceff = _f32(cell_eff).copy()
    lib().orc_l2l_f32(L, ceff)
    return ceff


def l2p(L, index, cell_eff, eff):
    out = _f32(eff).copy()
    lib().orc_l2p_f32(L, out.shape[0], _i32(index), _f32(cell_eff), out)
    return out


def postdo(perm, eff):
    eff = _f32(eff).reshape(-1, 2)
    out = np.zeros_like(eff)
    lib().orc_postdo_f32(eff.shape[0], _i32(perm), eff, out)
    return out


def solve(L, pos0, w0, limits, eps):
    pos0 = _f32(pos0).reshape(-1, 2)
    eff = np.zeros_like(pos0)
    clamped = lib().orc_solve_f32(L, pos0.shape[0], _f32(limits), float(eps), pos0, _f32(w0), eff)
    return eff, clamped


def solve_f64(L, pos0, w0, limits, eps):
    pos0 = np.ascontiguousarray(pos0, np.float64).reshape(-1, 2)
    eff = np.zeros_like(pos0)
    lib().orc_solve_f64(L, pos0.shape[0], np.ascontiguousarray(limits, np.float64), float(eps), pos0,
                        np.ascontiguousarray(w0, np.float64), eff)
    return eff


def update(tick, xcur, xpre, acc):
    xcur = _f32(xcur).copy()
    xpre = _f32(xpre).copy()
    lib().orc_update_f32(xcur.shape[0], float(tick), xcur, xpre, _f32(acc))
    return xcur, xpre


def update_f64(tick, xcur, xpre, acc):
    xcur = np.ascontiguousarray(xcur, np.float64).copy()
    xpre = np.ascontiguousarray(xpre, np.float64).copy()
    lib().orc_update_f64(xcur.shape[0], float(tick), xcur, xpre, np.ascontiguousarray(acc, np.float64))
    return xcur, xpre


def run(L, limits, eps, tick, steps, xcur, xpre, w):
    """Stars::update without prune, `steps` times: returns (xcur, xpre, acc, clamped)."""
    xcur = _f32(xcur).copy()
    xpre = _f32(xpre).copy()
    acc = np.zeros_like(xcur)
    clamped = lib().orc_run_f32(L, xcur.shape[0], _f32(limits), float(eps), float(tick), int(steps), xcur, xpre,
                                _f32(w), acc)
    return xcur, xpre, acc, clamped


def run_f64(L, limits, eps, tick, steps, xcur, xpre, w):
    """fp64 Stars::update loop (solve + Verlet), for finite differences."""
    xcur = np.ascontiguousarray(xcur, np.float64).copy()
    xpre = np.ascontiguousarray(xpre, np.float64).copy()
    acc = np.zeros_like(xcur)
    lib().orc_run_f64(L, xcur.shape[0], np.ascontiguousarray(limits, np.float64), float(eps), float(tick),
                      int(steps), xcur, xpre, np.ascontiguousarray(w, np.float64), acc)
    return xcur, xpre


def run_f64_cells(L, limits, eps, tick, steps, xcur, xpre, w, cells=None):
    """run_f64 with every step's cell index pairs recorded (cells None:
    returned, shape [steps, n, 2]) or replayed from `cells` (the cells held
    fixed): returns (xcur, xpre, cells)."""
    xcur = np.ascontiguousarray(xcur, np.float64).copy()
    xpre = np.ascontiguousarray(xpre, np.float64).copy()
    n = xcur.shape[0]
    acc = np.zeros_like(xcur)
    record = cells is None
    c = np.zeros((steps, n, 2), np.int32) if record else np.ascontiguousarray(cells, np.int32).copy()
    assert c.shape == (steps, n, 2)
    lib().orc_run_f64_cells(L, n, np.ascontiguousarray(limits, np.float64), float(eps), float(tick), int(steps),
                            xcur, xpre, np.ascontiguousarray(w, np.float64), acc, c, int(record))
    return xcur, xpre, c


def direct_f64(pos, w, eps, targets=None):
    pos = _f32(pos).reshape(-1, 2)
    n = pos.shape[0]
    if targets is None:
        out = np.zeros((n, 2), np.float64)
        lib().orc_direct_f64(n, pos, _f32(w), float(eps), n, None, out)
    else:
        t = _i32(targets)
        out = np.zeros((t.shape[0], 2), np.float64)
        lib().orc_direct_f64(n, pos, _f32(w), float(eps), t.shape[0], t.ctypes.data_as(ctypes.c_void_p), out)
    return out


def prune(limits, xcur, xpre, w):
    """Stars::prune: (xcur, xpre, w) with the kept stars first (both groups in
    their previous order) and the number kept."""
    xcur = _f32(xcur).copy()
    xpre = _f32(xpre).copy()
    w = _f32(w).copy()
    kept = lib().orc_prune_f32(w.shape[0], _f32(limits), xcur, xpre, w)
    return xcur, xpre, w, kept


# ---- initial-condition models (source/nbfmm/model/*.cu) ----

def model_uniform(seed, i, k):
    return float(lib().orc_model_uniform(int(seed), int(i), int(k)))


def _model_out(n):
    return np.zeros((n, 2), np.float32), np.zeros((n, 2), np.float32), np.zeros(n, np.float32)


def model_disk(n, center, radius, weight, tick, eps=0.1, seed=0, center_weight=0.0, centered=False, rotate=True):
    """generateDisk / generateDiskCenter (centered) / generateDiskStatic (rotate=False)."""
    xc, xp, w = _model_out(n)
    lib().orc_model_disk(n, float(center[0]), float(center[1]), float(radius), float(weight), float(center_weight),
                         int(centered), int(rotate), float(tick), float(eps), int(seed), xc, xp, w)
    return xc, xp, w


def model_circle(n, center, radius, weight, center_weight, tick, eps=0.1, uniform=False, seed=0):
    """generateCircleCenter / generateCircleUniformCenter (uniform)."""
    xc, xp, w = _model_out(n)
    lib().orc_model_circle(n, float(center[0]), float(center[1]), float(radius), float(weight), float(center_weight),
                           float(tick), float(eps), int(uniform), int(seed), xc, xp, w)
    return xc, xp, w


def model_rectangle(n, center, width, height, max_weight, seed=0):
    xc, xp, w = _model_out(n)
    lib().orc_model_rectangle(n, float(center[0]), float(center[1]), float(width), float(height), float(max_weight),
                              int(seed), xc, xp, w)
    return xc, xp, w


def double_disk_offsets(c1, c2, m1, m2, q1, q2, ecc, tick, eps=0.1):
    o = np.zeros(4, np.float32)
    lib().orc_double_disk_offsets(float(c1[0]), float(c1[1]), float(c2[0]), float(c2[1]), float(m1), float(m2),
                                  float(q1), float(q2), float(ecc), float(tick), float(eps), o)
    return o[:2].copy(), o[2:].copy()


def model_double_disk(n1, n2, c1, c2, r1, r2, weight, ecc, tick, eps=0.1, seed=0, centered=False, cw1=0.0, cw2=0.0):
    """generateDoubleDisk / generateDoubleDiskCenter: disk 1 then disk 2."""
    a = model_disk(n1, c1, r1, weight, tick, eps, seed, cw1, centered)
    b = model_disk(n2, c2, r2, weight, tick, eps, seed, cw2, centered)
    f = np.float32
    m1 = f(weight) * f(n1 - 1) + f(cw1) if centered else f(weight) * f(n1)
    m2 = f(weight) * f(n2 - 1) + f(cw2) if centered else f(weight) * f(n2)
    q1, q2 = (m1, m2) if centered else (f(n1), f(n2))
    o1, o2 = double_disk_offsets(c1, c2, m1, m2, q1, q2, ecc, tick, eps)
    xp = np.concatenate([a[1] + o1, b[1] + o2]).astype(np.float32)
    return np.concatenate([a[0], b[0]]), xp, np.concatenate([a[2], b[2]])


# ---- video frames (source/nbfmm/display/stars/display.cu) ----

def display(width, height, size_scale, limits, pos, w):
    """Stars::display's frame: uint8[height * width * 3 // 2], planar YUV 4:2:0."""
    pos = _f32(pos).reshape(-1, 2)
    frame = np.empty(width * height * 3 // 2, np.uint8)
    lib().orc_display(pos.shape[0], int(width), int(height), float(size_scale), _f32(limits), pos, _f32(w), frame)
    return frame
