"""Initial conditions of the reference's star models (source/nbfmm/model/*.cu).

Each function returns (position_cur, position_pre, weight) as CUDA float32
tensors, generated on the device by nbfmm_amd's model kernels: the same
geometry, chord-angle previous positions and double-disk orbit offsets as the
reference, with uniform draws from a counter-based hash of (seed, particle,
draw) instead of cuRAND streams.  `tick` already includes sqrt(G) (Stars
multiplies it in, source/nbfmm/display/stars.cu:31); `eps` is the kernel
function's softening (0.1 in the reference).
"""
from __future__ import annotations

import numpy as np
import torch

from ._lib import check, lib


def _out(n):
    dev = torch.device("cuda")
    return (torch.empty((n, 2), dtype=torch.float32, device=dev), torch.empty((n, 2), dtype=torch.float32, device=dev),
            torch.empty(n, dtype=torch.float32, device=dev))


def _stream():
    return torch.cuda.current_stream().cuda_stream


def disk(n, center, radius, weight, tick, eps=0.1, seed=0):
    """generateDisk (model/disk.cu)."""
    xc, xp, w = _out(n)
    check(lib().nbfmm_model_disk(n, float(center[0]), float(center[1]), float(radius), float(weight), float(tick),
                                 float(eps), int(seed), xc.data_ptr(), xp.data_ptr(), w.data_ptr(), _stream()))
    return xc, xp, w


def disk_static(n, center, radius, weight, seed=0):
    """generateDiskStatic (model/disk_static.cu): at rest."""
    xc, xp, w = _out(n)
    check(lib().nbfmm_model_disk_static(n, float(center[0]), float(center[1]), float(radius), float(weight),
                                        int(seed), xc.data_ptr(), xp.data_ptr(), w.data_ptr(), _stream()))
    return xc, xp, w


def disk_center(n, center, radius, weight, center_weight, tick, eps=0.1, seed=0):
    """generateDiskCenter (model/disk_center.cu): star 0 is the heavy centre."""
    xc, xp, w = _out(n)
    check(lib().nbfmm_model_disk_center(n, float(center[0]), float(center[1]), float(radius), float(weight),
                                        float(center_weight), float(tick), float(eps), int(seed), xc.data_ptr(),
                                        xp.data_ptr(), w.data_ptr(), _stream()))
    return xc, xp, w


def circle_center(n, center, radius, weight, center_weight, tick, eps=0.1, uniform=False, seed=0):
    """generateCircleCenter (random angles) / generateCircleUniformCenter (uniform=True)."""
    xc, xp, w = _out(n)
    check(lib().nbfmm_model_circle_center(n, float(center[0]), float(center[1]), float(radius), float(weight),
                                          float(center_weight), float(tick), float(eps), int(bool(uniform)),
                                          int(seed), xc.data_ptr(), xp.data_ptr(), w.data_ptr(), _stream()))
    return xc, xp, w


def rectangle(n, center, width, height, max_weight, seed=0):
    """generateRectangle (model/rectangle.cu): uniform, at rest, weights in (0, max_weight]."""
    xc, xp, w = _out(n)
    check(lib().nbfmm_model_rectangle(n, float(center[0]), float(center[1]), float(width), float(height),
                                      float(max_weight), int(seed), xc.data_ptr(), xp.data_ptr(), w.data_ptr(),
                                      _stream()))
    return xc, xp, w


def double_disk(n1, n2, center1, center2, radius1, radius2, weight, eccentricity, tick, eps=0.1, seed=0,
                centered=False, center_weight1=0.0, center_weight2=0.0):
    """generateDoubleDisk / generateDoubleDiskCenter (centered=True): disk 1 in
    entries [0, n1), disk 2 in [n1, n1 + n2), on an orbit about each other."""
    xc, xp, w = _out(n1 + n2)
    c1 = np.ascontiguousarray(center1, np.float32)
    c2 = np.ascontiguousarray(center2, np.float32)
    check(lib().nbfmm_model_double_disk(n1, n2, c1.ctypes.data, c2.ctypes.data, float(radius1), float(radius2),
                                        float(weight), float(center_weight1), float(center_weight2),
                                        float(eccentricity), float(tick), float(eps), int(bool(centered)),
                                        int(seed), xc.data_ptr(), xp.data_ptr(), w.data_ptr(), _stream()))
    return xc, xp, w
