"""Frames (source/nbfmm/display/stars/display.cu) and the YUV4MPEG2 stream of
main/main.cu on the GPU: nbfmm_display against the oracle bit for bit, the
writer's file read back, and nbfmm_main (main.cu's arguments) against the
same loop driven through the library in this process."""
import os
import subprocess

import numpy as np
import pytest
import torch

from oracle import pyoracle as orc

pytestmark = pytest.mark.gpu


def stars(n, lim, seed):
    rng = np.random.default_rng(seed)
    span = np.array([lim[2] - lim[0], lim[3] - lim[1]])
    pos = (rng.uniform(-0.1, 1.1, (n, 2)) * span + np.array(lim[:2])).astype(np.float32)
    w = rng.exponential(4.0, n).astype(np.float32)
    if n >= 8:
        pos[0] = (np.nan, 0.0)
        pos[1] = (lim[0], lim[3])          # top-left corner: pixel (0, 0)
        pos[2] = (lim[2], lim[1])          # bottom-right edge: outside
        w[3] = np.nan
        w[4] = 1e6                         # largest stamp
        w[5] = 0.0
    return pos, w


@pytest.mark.parametrize("W,H,n", [(64, 48, 5000), (30, 22, 3000), (1024, 768, 200000), (2, 2, 10)])
def test_display_matches_oracle(W, H, n):
    from nbfmm_amd import video
    lim = (-4.0, -3.0, 4.0, 3.0)
    pos, w = stars(n, lim, W + n)
    x = torch.from_numpy(pos).cuda()
    m = torch.from_numpy(w).cuda()
    got = video.display(x, m, W, H, 1.0, lim)
    torch.cuda.synchronize()
    ref = orc.display(W, H, 1.0, lim, pos, w)
    assert np.array_equal(got.cpu().numpy(), ref)
    # an unaligned frame pointer and a first-n subset
    buf = torch.full((video.frame_bytes(W, H) + 3,), 7, dtype=torch.uint8, device="cuda")
    video.display(x, m, W, H, 0.5, lim, num_star=n // 2, out=buf[3:])
    torch.cuda.synchronize()
    b = buf.cpu().numpy()
    assert np.all(b[:3] == 7)
    assert np.array_equal(b[3:], orc.display(W, H, 0.5, lim, pos[:n // 2], w[:n // 2]))


def test_display_no_stars():
    from nbfmm_amd import video
    x = torch.zeros((1, 2), device="cuda")
    f = video.display(x, torch.ones(1, device="cuda"), 16, 8, 1.0, (0, 0, 1, 1), num_star=0)
    f = f.cpu().numpy()
    assert np.all(f[:128] == 0) and np.all(f[128:] == 128)


def test_y4m_roundtrip(tmp_path):
    from nbfmm_amd import video
    W, H, fps = 48, 32, 24
    path = str(tmp_path / "t.y4m")
    frames = []
    with video.Y4MWriter(path, W, H, fps) as out:
        buf = torch.empty(video.frame_bytes(W, H), dtype=torch.uint8, device="cuda")
        for k in range(5):
            pos, w = stars(100 * (k + 1), (0, 0, 1, 1), k)
            video.display(torch.from_numpy(pos).cuda(), torch.from_numpy(w).cuda(), W, H, 1.0, (0, 0, 1, 1),
                          out=buf)
            out.write(buf)   # the same device buffer is reused: the writer must copy in stream order
            frames.append(orc.display(W, H, 1.0, (0, 0, 1, 1), pos, w))
    assert out.frames == 5
    header, got = video.read_y4m(path)
    with open(path, "rb") as f:
        assert f.readline() == b"YUV4MPEG2 W48 H32 F24:1 Ip A1:1 C420\n"
    assert os.path.getsize(path) == len(b"YUV4MPEG2 W48 H32 F24:1 Ip A1:1 C420\n") + 5 * (6 + W * H * 3 // 2)
    assert header["W"] == "48" and header["H"] == "32" and header["C"] == "420"
    assert np.array_equal(got, np.stack(frames))


def run_main(args, tmp_path):
    from nbfmm_amd import build
    out = str(tmp_path / "main.y4m")
    print("nbfmm_main", *args, flush=True)
    r = subprocess.run([build.MAIN, out, *[str(a) for a in args]], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    from nbfmm_amd import video
    return video.read_y4m(out), r.stdout


def replay(n, level, lim, dlim, W, H, size_scale, tick, frames, x, xp, w):
    """main.cu's frame loop through the Python API: frame 0, prune, then step,
    prune and display per frame."""
    from nbfmm_amd import Solver, video
    s = Solver(level, n, lim, 0.1)
    acc = torch.zeros_like(x)
    out = [video.display(x, w, W, H, size_scale, dlim, num_star=n).cpu().numpy()]
    n = s.prune(n, x, xp, w)
    for _ in range(1, frames):
        if n > 0:
            s.step(n, tick, x, xp, w, acc)
            n = s.prune(n, x, xp, w)
        out.append(video.display(x, w, W, H, size_scale, dlim, num_star=n).cpu().numpy())
    s.close()
    return np.stack(out), n


def test_main_disk_scenario(tmp_path):
    """The disk launcher's argument layout (main/script/disk.py), 12 frames."""
    from nbfmm_amd import models
    W, H, fps, sec, level, n = 256, 192, 12, 1, 4, 4000
    pw, ph, scale, G, size_scale = 16.0, 12.0, 0.5, 1.0, 1.0
    radius, weight = min(pw, ph) * scale / 8, 1.0
    (header, frames), log = run_main([W, H, fps, sec, level, n, pw, ph, scale, G, size_scale, "disk", radius,
                                      weight], tmp_path)
    assert header["W"] == str(W) and header["F"] == f"{fps}:1"
    assert frames.shape == (fps * sec, W * H * 3 // 2)
    tick = np.float32(np.float32(0.05 / fps) * np.sqrt(np.float32(G)))
    lim = (-pw / 2, -ph / 2, pw / 2, ph / 2)
    dlim = tuple(np.float32(v) * np.float32(scale) for v in lim)
    x, xp, w = models.disk(n, (0.0, 0.0), radius, weight, float(tick), 0.1, seed=0)
    ref, left = replay(n, level, lim, dlim, W, H, size_scale, float(tick), fps * sec, x, xp, w)
    assert np.array_equal(frames, ref)
    assert f"{left} stars left" in log
    assert (frames[:, :W * H] == 255).any(axis=1).all()


def test_main_double_disk_center(tmp_path):
    """double-disk-center with its eight parameters; each disk keeps its own
    centre weight."""
    from nbfmm_amd import models
    W, H, fps, sec, level, n = 160, 120, 4, 1, 3, 3000
    pw, ph, scale = 16.0, 12.0, 0.8
    q1, q2, r1, r2, weight, cw1, cw2, ecc = 5.0, 3.0, 0.5, 0.3, 1.0, 40.0, 9.0, 0.5
    (header, frames), _ = run_main([W, H, fps, sec, level, n, pw, ph, scale, 1.0, 2.0, "double-disk-center", q1, q2,
                                    r1, r2, weight, cw1, cw2, ecc], tmp_path)
    assert frames.shape[0] == fps * sec
    f32 = np.float32
    p1, p2 = f32(q1) / f32(q1 + q2), f32(q2) / f32(q1 + q2)
    n1, n2 = int(f32(n) * p1), int(f32(n) * p2)
    lim = (-pw / 2, -ph / 2, pw / 2, ph / 2)
    d = [f32(v) * f32(scale) for v in lim]
    a1, b1 = p1, p1 + f32(2) * p2
    a2, b2 = p2, p2 + f32(2) * p1
    c1 = ((d[0] * a1 + d[2] * b1) / f32(2), (d[1] * a1 + d[3] * b1) / f32(2))
    c2 = ((d[2] * a2 + d[0] * b2) / f32(2), (d[3] * a2 + d[1] * b2) / f32(2))
    tick = float(f32(0.05 / fps))
    x, xp, w = models.double_disk(n1, n2, c1, c2, r1, r2, weight, ecc, tick, 0.1, seed=0, centered=True,
                                  center_weight1=cw1, center_weight2=cw2)
    wh = w.cpu().numpy()
    assert wh[0] == f32(cw1) and wh[n1] == f32(cw2)
    ref, _ = replay(n1 + n2, level, lim, tuple(d), W, H, 2.0, tick, fps * sec, x, xp, w)
    assert np.array_equal(frames, ref)


def test_main_rejects_unknown_model(tmp_path):
    from nbfmm_amd import build
    args = [64, 48, 10, 1, 3, 100, 16, 12, 0.5, 1, 1, "torus", 1]
    r = subprocess.run([build.MAIN, str(tmp_path / "o.y4m"), *[str(a) for a in args]], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 1
    assert "Unknown model torus." in r.stdout
