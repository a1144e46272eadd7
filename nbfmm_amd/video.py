"""Video frames of the stars and the YUV4MPEG2 file they go to
(source/nbfmm/display/stars/display.cu, main/main.cu:234-250).

A frame is a uint8 CUDA tensor of width * height * 3 // 2 bytes, planar YUV
4:2:0: luma rows first (row 0 at the top of the display box), then the two
chroma planes, which stay at 128.

    frame = display(x, w, width=1024, height=768, size_scale=1.0, display_limits=(-4, -3, 4, 3))
    with Y4MWriter("out.y4m", 1024, 768, fps=60) as video:
        video.write(frame)
    header, frames = read_y4m("out.y4m")
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from ._lib import check, lib


def _stream(stream=None) -> int:
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


def frame_bytes(width: int, height: int) -> int:
    return width * height * 3 // 2


def display(position: torch.Tensor, weight: torch.Tensor, width: int, height: int, size_scale: float,
            display_limits, num_star: int | None = None, out: torch.Tensor | None = None, stream=None) -> torch.Tensor:
    """Stars::display without its trailing prune (display.cu:219-226): the
    frame of the first num_star stars (all of them by default)."""
    n = position.shape[0] if num_star is None else int(num_star)
    if position.dtype != torch.float32 or weight.dtype != torch.float32:
        raise TypeError("position and weight must be float32")
    if not (position.is_cuda and weight.is_cuda and position.is_contiguous() and weight.is_contiguous()):
        raise ValueError("position and weight must be contiguous CUDA tensors")
    if n > position.shape[0] or n > weight.shape[0]:
        raise ValueError("num_star exceeds the arrays")
    if out is None:
        out = torch.empty(frame_bytes(width, height), dtype=torch.uint8, device=position.device)
    elif out.dtype != torch.uint8 or out.numel() < frame_bytes(width, height) or not out.is_contiguous():
        raise ValueError("out must be a contiguous uint8 tensor of width * height * 3 // 2 bytes")
    lim = (ctypes.c_float * 4)(*[float(v) for v in display_limits])
    check(lib().nbfmm_display(n, int(width), int(height), float(size_scale), lim, position.data_ptr(),
                              weight.data_ptr(), out.data_ptr(), _stream(stream)))
    return out


def luma(frame: torch.Tensor | np.ndarray, width: int, height: int):
    """The luma plane of a frame as [height, width]."""
    return frame[: width * height].reshape(height, width)


class Y4MWriter:
    """The YUV4MPEG2 stream of main/main.cu:234-250 through nbfmm_y4m_*."""

    def __init__(self, path: str, width: int, height: int, fps: int):
        self.width, self.height, self.fps = int(width), int(height), int(fps)
        h = ctypes.c_void_p()
        check(lib().nbfmm_y4m_open(str(path).encode(), self.width, self.height, self.fps, ctypes.byref(h)))
        self._h = h
        self.frames = 0

    def write(self, frame: torch.Tensor, stream=None) -> None:
        if self._h is None:
            raise ValueError("writer is closed")
        if frame.dtype != torch.uint8 or not frame.is_cuda or frame.numel() < frame_bytes(self.width, self.height):
            raise ValueError("frame must be a CUDA uint8 tensor of width * height * 3 // 2 bytes")
        check(lib().nbfmm_y4m_write_frame(self._h, frame.data_ptr(), _stream(stream)))

    def close(self) -> int:
        if self._h is not None:
            n = ctypes.c_long()
            h, self._h = self._h, None
            check(lib().nbfmm_y4m_close(h, ctypes.byref(n)))
            self.frames = n.value
        return self.frames

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def read_y4m(path: str):
    """Parses a 4:2:0 YUV4MPEG2 file: ({"W", "H", "F", ...}, uint8[frames, W * H * 3 // 2])."""
    with open(path, "rb") as f:
        data = f.read()
    end = data.index(b"\n")
    fields = data[:end].split(b" ")
    if fields[0] != b"YUV4MPEG2":
        raise ValueError("not a YUV4MPEG2 file")
    header = {p[:1].decode(): p[1:].decode() for p in fields[1:]}
    w, h = int(header["W"]), int(header["H"])
    size = frame_bytes(w, h)
    frames = []
    pos = end + 1
    while pos < len(data):
        tag_end = data.index(b"\n", pos)
        if not data[pos:tag_end].startswith(b"FRAME"):
            raise ValueError(f"expected FRAME at byte {pos}")
        pos = tag_end + 1
        if pos + size > len(data):
            raise ValueError("truncated frame")
        frames.append(np.frombuffer(data, np.uint8, size, pos))
        pos += size
    return header, (np.stack(frames) if frames else np.zeros((0, size), np.uint8))
