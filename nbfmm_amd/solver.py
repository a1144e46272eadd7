"""The nbfmm::Solver mirror over the C ABI (include/nbfmm_amd.h).

Reference interface: include/nbfmm/core/solver.hpp:22-209 and the Verlet
update of source/nbfmm/display/stars/update.cu:26-55.
"""
from __future__ import annotations

import ctypes

import torch

from ._lib import check, lib

FLAG_OUT_OF_BOX = 1
FLAG_P2P_WAIT = 2

# buffer name -> (id, dtype, components); names follow the Solver members
# gpuptr_position_, gpuptr_effect_, ... (solver.hpp:45-76)
BUFFERS = {
    "position": (0, torch.float32, 2),
    "effect": (1, torch.float32, 2),
    "weight": (2, torch.float32, 1),
    "index": (3, torch.int32, 2),
    "perm": (4, torch.int32, 1),
    "head": (5, torch.int32, 1),
    "cell_position": (6, torch.float32, 2),
    "cell_effect": (7, torch.float32, 2),
    "cell_weight": (8, torch.float32, 1),
}


def _stream(stream=None) -> int:
    if stream is None:
        stream = torch.cuda.current_stream()
    return int(stream.cuda_stream)


def _ptr(t: torch.Tensor, dtype=torch.float32, name="array") -> int:
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError(f"{name} must be a CUDA (HIP) tensor")
    if t.dtype != dtype:
        raise TypeError(f"{name} must be {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    return t.data_ptr()


def update(num_star: int, tick: float, position_cur, position_pre, acceleration, stream=None) -> None:
    """updateDevice (update.cu:26-42): x' = 2x - x_pre + a*tick^2, x_pre' = x."""
    check(lib().nbfmm_update(int(num_star), float(tick), _ptr(position_cur), _ptr(position_pre),
                             _ptr(acceleration), _stream(stream)))


class Solver:
    """nbfmm::Solver(num_level, max_num_particle, position_limits) with a
    runtime softening `epsilon` (the reference hard-codes 0.1,
    include/nbfmm/core/kernel_function.hpp:35)."""

    def __init__(self, num_level: int, max_num_particle: int, position_limits, epsilon: float = 0.1):
        if not torch.cuda.is_available():
            raise RuntimeError("nbfmm_amd needs a HIP GPU; none is available")
        L = lib()
        lim = (ctypes.c_float * 4)(*[float(x) for x in position_limits])
        h = ctypes.c_void_p()
        check(L.nbfmm_solver_create(int(num_level), int(max_num_particle), lim, float(epsilon), ctypes.byref(h)))
        self._h = h
        self.num_level = int(num_level)
        self.base_dim = 1 << (self.num_level + 1)
        self.max_num_particle = int(max_num_particle)
        self.position_limits = tuple(float(x) for x in position_limits)
        self.epsilon = float(epsilon)

    # -- lifetime ---------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            check(lib().nbfmm_solver_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- solve.cu:11-27 --------------------------------------------------
    def solve(self, num_particle: int, position_origin, weight_origin, effect_origin, stream=None) -> None:
        check(lib().nbfmm_solve(self._h, int(num_particle), _ptr(position_origin, name="position_origin"),
                                _ptr(weight_origin, name="weight_origin"),
                                _ptr(effect_origin, name="effect_origin"), _stream(stream)))

    # -- the eight stages (solver.hpp:111-207) ----------------------------
    def predo(self, num_particle, position_origin, weight_origin, stream=None):
        check(lib().nbfmm_predo(self._h, int(num_particle), _ptr(position_origin), _ptr(weight_origin),
                                _stream(stream)))

    def p2p(self, num_particle, stream=None):
        check(lib().nbfmm_p2p(self._h, int(num_particle), _stream(stream)))

    def p2m(self, num_particle, stream=None):
        check(lib().nbfmm_p2m(self._h, int(num_particle), _stream(stream)))

    def m2m(self, stream=None):
        check(lib().nbfmm_m2m(self._h, _stream(stream)))

    def m2l(self, stream=None):
        check(lib().nbfmm_m2l(self._h, _stream(stream)))

    def l2l(self, stream=None):
        check(lib().nbfmm_l2l(self._h, _stream(stream)))

    def l2p(self, num_particle, stream=None):
        check(lib().nbfmm_l2p(self._h, int(num_particle), _stream(stream)))

    def postdo(self, num_particle, effect_origin, stream=None):
        check(lib().nbfmm_postdo(self._h, int(num_particle), _ptr(effect_origin), _stream(stream)))

    # -- buffer access (the NBFMM_CHECK members) -------------------------
    def buffer_shape(self, name: str, count: int = 0):
        bid, dtype, comps = BUFFERS[name]
        nbytes = ctypes.c_size_t()
        check(lib().nbfmm_buffer_bytes(self._h, bid, int(count), ctypes.byref(nbytes)))
        elems = nbytes.value // 4 // comps
        return ((elems, comps) if comps > 1 else (elems,)), dtype

    def read_buffer(self, name: str, count: int = 0, stream=None) -> torch.Tensor:
        bid, dtype, _ = BUFFERS[name]
        shape, dtype = self.buffer_shape(name, count)
        out = torch.empty(shape, dtype=dtype, device="cuda")
        check(lib().nbfmm_buffer_read(self._h, bid, int(count), out.data_ptr(), _stream(stream)))
        return out

    def write_buffer(self, name: str, data: torch.Tensor, count: int = 0, stream=None) -> None:
        bid, dtype, _ = BUFFERS[name]
        shape, dtype = self.buffer_shape(name, count)
        data = data.to(device="cuda", dtype=dtype).contiguous()
        if data.numel() != int(torch.tensor(shape).prod()):
            raise ValueError(f"{name} needs {shape} elements, got {tuple(data.shape)}")
        check(lib().nbfmm_buffer_write(self._h, bid, int(count), data.data_ptr(), _stream(stream)))
        # keep the source alive until the copy has run
        torch.cuda.current_stream().synchronize()

    def flags(self, clear: bool = False) -> int:
        f = ctypes.c_uint()
        check(lib().nbfmm_solver_flags(self._h, ctypes.byref(f), int(clear)))
        return f.value

    # -- time stepping (Stars::update without prune) ---------------------
    def step(self, num_particle, tick, position_cur, position_pre, weight, acceleration, stream=None):
        check(lib().nbfmm_step(self._h, int(num_particle), float(tick), _ptr(position_cur), _ptr(position_pre),
                               _ptr(weight), _ptr(acceleration), _stream(stream)))

    def run(self, num_particle, tick, num_steps, position_cur, position_pre, weight, acceleration,
            use_graph=True, stream=None):
        check(lib().nbfmm_run(self._h, int(num_particle), float(tick), int(num_steps), _ptr(position_cur),
                              _ptr(position_pre), _ptr(weight), _ptr(acceleration), int(bool(use_graph)),
                              _stream(stream)))

    # -- adjoint ----------------------------------------------------------
    def solve_adjoint(self, num_particle, position_origin, weight_origin, grad_effect, grad_position, grad_weight,
                      stream=None):
        """dloss/dposition and dloss/dweight from dloss/deffect (cells held fixed)."""
        check(lib().nbfmm_solve_adjoint(self._h, int(num_particle), _ptr(position_origin), _ptr(weight_origin),
                                        _ptr(grad_effect), _ptr(grad_position), _ptr(grad_weight),
                                        _stream(stream)))

    def step_adjoint(self, num_particle, tick, position, weight, lam, pi, grad_weight, stream=None):
        """Adjoint of one Verlet step at positions x_k; lam, pi updated in place."""
        check(lib().nbfmm_step_adjoint(self._h, int(num_particle), float(tick), _ptr(position), _ptr(weight),
                                       _ptr(lam), _ptr(pi), _ptr(grad_weight), _stream(stream)))

    def trajectory_workspace(self, num_particle, num_steps, checkpoint_every):
        a = ctypes.c_size_t()
        b = ctypes.c_size_t()
        check(lib().nbfmm_trajectory_workspace(int(num_particle), int(num_steps), int(checkpoint_every),
                                               ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def trajectory_forward(self, num_particle, tick, num_steps, checkpoint_every, position_cur, position_pre, weight,
                           acceleration, checkpoints, obs_steps, obs_position, stream=None):
        obs = (ctypes.c_int * max(len(obs_steps), 1))(*[int(k) for k in obs_steps])
        check(lib().nbfmm_trajectory_forward(self._h, int(num_particle), float(tick), int(num_steps),
                                             int(checkpoint_every), _ptr(position_cur), _ptr(position_pre),
                                             _ptr(weight), _ptr(acceleration), _ptr(checkpoints), len(obs_steps),
                                             obs, _ptr(obs_position) if len(obs_steps) else None,
                                             _stream(stream)))

    def trajectory_backward(self, num_particle, tick, num_steps, checkpoint_every, weight, checkpoints, obs_steps,
                            obs_grad, grad_position0, grad_position_pre0, grad_weight, work, stream=None):
        obs = (ctypes.c_int * max(len(obs_steps), 1))(*[int(k) for k in obs_steps])
        check(lib().nbfmm_trajectory_backward(self._h, int(num_particle), float(tick), int(num_steps),
                                              int(checkpoint_every), _ptr(weight), _ptr(checkpoints),
                                              len(obs_steps), obs, _ptr(obs_grad) if len(obs_steps) else None,
                                              _ptr(grad_position0), _ptr(grad_position_pre0), _ptr(grad_weight),
                                              _ptr(work), _stream(stream)))

    def prune(self, num_star, position_cur, position_pre, weight, stream=None) -> int:
        """Stars::prune (prune.cu:28-62): stars outside the position limits move
        behind the others (both groups keep their order, arrays permuted in
        place); returns the number kept.  Synchronises."""
        kept = ctypes.c_int()
        check(lib().nbfmm_prune(self._h, int(num_star), _ptr(position_cur), _ptr(position_pre), _ptr(weight),
                                ctypes.byref(kept), _stream(stream)))
        return kept.value

    def update_stars(self, num_star, tick, position_cur, position_pre, weight, acceleration, stream=None) -> int:
        """Stars::update (update.cu:47-55): solve, Verlet update, prune;
        returns the new star count."""
        self.step(num_star, tick, position_cur, position_pre, weight, acceleration, stream=stream)
        return self.prune(num_star, position_cur, position_pre, weight, stream=stream)

    def warm_start(self, enable: bool = True) -> None:
        """Start each sort from the previous one of the same N (default on);
        results are identical either way."""
        check(lib().nbfmm_set_warm_start(self._h, int(bool(enable))))

    def small_front(self, enable: bool = True) -> None:
        """num_level 1..4 and N <= 4096: sort, P2M and M2M in one one-block
        kernel (default on); results are identical either way."""
        check(lib().nbfmm_set_small_front(self._h, int(bool(enable))))

    def set_p2p_wait_polls(self, polls: int) -> None:
        """Bound on every wait of a P2P wave for another wave of its block;
        past it the wave sums its run alone, segment after segment
        (FLAG_P2P_WAIT; the same bits).  Default 2^20; tests set 1 to force
        that path."""
        check(lib().nbfmm_set_p2p_wait_polls(self._h, int(polls)))

    # -- measurement ------------------------------------------------------
    def profile(self, enable: bool = True) -> None:
        check(lib().nbfmm_profile_enable(self._h, int(bool(enable))))

    def p2p_time(self, reset: bool = True):
        ms = ctypes.c_double()
        n = ctypes.c_int()
        check(lib().nbfmm_p2p_time(self._h, ctypes.byref(ms), ctypes.byref(n), int(bool(reset))))
        return ms.value, n.value

    def stage_times(self, reset: bool = True):
        """Summed milliseconds of the sort, upward (P2M+M2M), M2L and P2P
        stages of the calls timed since the last reset, and their number."""
        ms = (ctypes.c_double * 4)()
        n = ctypes.c_int()
        check(lib().nbfmm_stage_times(self._h, ms, ctypes.byref(n), int(bool(reset))))
        return list(ms), n.value

    def pair_count(self, num_particle):
        pairs = ctypes.c_double()
        occ = ctypes.c_int()
        check(lib().nbfmm_pair_count(self._h, int(num_particle), ctypes.byref(pairs), ctypes.byref(occ)))
        return pairs.value, occ.value
