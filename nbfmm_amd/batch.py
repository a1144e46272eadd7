"""Many independent systems advanced together (C ABI nbfmm_batch_*).

No reference counterpart: the reference runs one nbfmm::Solver per system
(include/nbfmm/core/solver.hpp:22).  A batch keeps each system's own particle
count, level, box and softening; every stage of a step is one launch over all
systems and `run` replays the step as one hipGraph.  System by system the
results are bit-identical to separate Solver objects.

Particle arrays are the systems' arrays concatenated in system order; use
`offsets` (S+1 entries) to slice them.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from ._lib import check, lib
from .solver import _ptr, _stream


class Batch:
    def __init__(self, num_particle, num_level, position_limits, epsilon, ticks=None):
        if not torch.cuda.is_available():
            raise RuntimeError("nbfmm_amd needs a HIP GPU; none is available")
        n = np.ascontiguousarray(num_particle, np.int32)
        L = np.ascontiguousarray(num_level, np.int32)
        lim = np.ascontiguousarray(position_limits, np.float32).reshape(-1, 4)
        eps = np.ascontiguousarray(np.broadcast_to(np.asarray(epsilon, np.float32), n.shape))
        if not (n.shape == L.shape == eps.shape and lim.shape[0] == n.shape[0]):
            raise ValueError("one num_particle, num_level, 4 limits and epsilon per system")
        h = ctypes.c_void_p()
        check(lib().nbfmm_batch_create(int(n.size), n.ctypes.data, L.ctypes.data, lim.ctypes.data,
                                       eps.ctypes.data, ctypes.byref(h)))
        self._h = h
        self.num_systems = int(n.size)
        self.num_particle = n.copy()
        self.num_level = L.copy()
        off = np.zeros(self.num_systems + 1, np.int32)
        check(lib().nbfmm_batch_offsets(self._h, off.ctypes.data))
        self.offsets = off
        self.total = int(off[-1])
        if ticks is not None:
            self.set_ticks(ticks)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            check(lib().nbfmm_batch_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_ticks(self, ticks):
        t = np.ascontiguousarray(np.broadcast_to(np.asarray(ticks, np.float32), (self.num_systems,)))
        check(lib().nbfmm_batch_set_ticks(self._h, t.ctypes.data))

    def set_p2p_config(self, variant, blocks_per_cu):
        """P2P kernel variant (see nbfmm_amd.autotune.variants) and grid of
        every launch; all configurations give identical results."""
        check(lib().nbfmm_batch_set_p2p_config(self._h, int(variant), int(blocks_per_cu)))

    def set_p2p_wait_polls(self, polls):
        """Solver.set_p2p_wait_polls for every system of the batch."""
        check(lib().nbfmm_batch_set_p2p_wait_polls(self._h, int(polls)))

    def warm_start(self, enable=True):
        """Start each sort from the previous one (on by default); results are
        identical either way."""
        check(lib().nbfmm_batch_set_warm_start(self._h, int(bool(enable))))

    def balance(self, position, weight, stream=None):
        """Share the P2P grids by each system's pair count at these positions
        (the batch's first call does it by itself); synchronises."""
        check(lib().nbfmm_batch_balance(self._h, _ptr(position), _ptr(weight), _stream(stream)))

    def p2p_config(self):
        v, b = ctypes.c_int(), ctypes.c_int()
        check(lib().nbfmm_batch_get_p2p_config(self._h, ctypes.byref(v), ctypes.byref(b)))
        return v.value, b.value

    def solve(self, position, weight, effect, stream=None):
        check(lib().nbfmm_batch_solve(self._h, _ptr(position), _ptr(weight), _ptr(effect), _stream(stream)))

    def step(self, position_cur, position_pre, weight, acceleration, stream=None):
        check(lib().nbfmm_batch_step(self._h, _ptr(position_cur), _ptr(position_pre), _ptr(weight),
                                     _ptr(acceleration), _stream(stream)))

    def run(self, num_steps, position_cur, position_pre, weight, acceleration, use_graph=True, stream=None):
        check(lib().nbfmm_batch_run(self._h, int(num_steps), _ptr(position_cur), _ptr(position_pre),
                                    _ptr(weight), _ptr(acceleration), int(bool(use_graph)), _stream(stream)))

    def solve_adjoint(self, position, weight, grad_effect, grad_position, grad_weight, stream=None):
        check(lib().nbfmm_batch_solve_adjoint(self._h, _ptr(position), _ptr(weight), _ptr(grad_effect),
                                              _ptr(grad_position), _ptr(grad_weight), _stream(stream)))

    def step_adjoint(self, position, weight, lam, pi, grad_weight, stream=None):
        """Adjoint of one Verlet step of every system (each with its own tick)."""
        check(lib().nbfmm_batch_step_adjoint(self._h, _ptr(position), _ptr(weight), _ptr(lam), _ptr(pi),
                                             _ptr(grad_weight), _stream(stream)))

    def keep_sort(self, slot, stream=None):
        """Keep the current sort and tree (of the last step's positions) in `slot`."""
        check(lib().nbfmm_batch_keep_sort(self._h, int(slot), _stream(stream)))

    def run_keep(self, num_steps, first_slot, position_cur, position_pre, weight, acceleration, stream=None):
        """num_steps eager steps, keeping the sort of step k (of the positions it
        starts from) in slot first_slot + k and that of the final positions in
        slot first_slot + num_steps: step + keep_sort, step by step, with the
        sorts chained as in a run (the same bits); acceleration is scratch."""
        check(lib().nbfmm_batch_run_keep(self._h, int(num_steps), int(first_slot), _ptr(position_cur),
                                         _ptr(position_pre), _ptr(weight), _ptr(acceleration), _stream(stream)))

    def step_adjoint_kept(self, slot, lam, pi, grad_weight, stream=None):
        """step_adjoint at the positions whose sort `slot` kept, without sorting again."""
        check(lib().nbfmm_batch_step_adjoint_kept(self._h, int(slot), _ptr(lam), _ptr(pi), _ptr(grad_weight),
                                                  _stream(stream)))

    def flags(self, clear=False):
        f = np.zeros(self.num_systems, np.uint32)
        check(lib().nbfmm_batch_flags(self._h, f.ctypes.data, int(bool(clear))))
        return f
