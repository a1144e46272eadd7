"""torch.autograd.Function wrappers over the HIP forward and adjoint kernels.

No reference counterpart: the reference only evaluates forces.  Gradients
hold each particle's cell fixed (the cell assignment is piecewise constant
in the positions), so they are the derivatives of the FMM as evaluated.

    acc = fmm_forces(solver, x, m)                      # [N,2] -> [N,2]
    obs = fmm_trajectory(solver, tick, steps, K, obs_steps, x0, x_pre0, m)
    obs = fmm_batch_trajectory(batch, steps, K, obs_steps, x0, x_pre0, m)

fmm_trajectory runs the Verlet loop (Stars::update without prune,
source/nbfmm/display/stars/update.cu:47-55) with (x, x_pre) checkpoints
every K steps and returns the positions at the observation steps; its
backward recomputes each segment from its checkpoint.  Tensors are float32
on the GPU.
"""
from __future__ import annotations

import torch


def _f32(t):
    return t.detach().to(torch.float32).contiguous()


class _Forces(torch.autograd.Function):
    @staticmethod
    def forward(ctx, solver, x, m):
        x, m = _f32(x), _f32(m)
        n = x.shape[0]
        acc = torch.empty_like(x)
        solver.solve(n, x, m, acc)
        ctx.solver = solver
        ctx.save_for_backward(x, m)
        return acc

    @staticmethod
    def backward(ctx, g):
        x, m = ctx.saved_tensors
        n = x.shape[0]
        gx = torch.empty_like(x)
        gm = torch.empty_like(m)
        ctx.solver.solve_adjoint(n, x, m, _f32(g), gx, gm)
        return None, gx, gm


def fmm_forces(solver, x, m):
    return _Forces.apply(solver, x, m)


class _Trajectory(torch.autograd.Function):
    @staticmethod
    def forward(ctx, solver, tick, steps, K, obs_steps, x0, xpre0, m):
        x, p, m = _f32(x0).clone(), _f32(xpre0).clone(), _f32(m)
        n = x.shape[0]
        ck_bytes, work_bytes = solver.trajectory_workspace(n, steps, K)
        ckpt = torch.empty(ck_bytes // 4, dtype=torch.float32, device=x.device)
        acc = torch.empty_like(x)
        obs = torch.empty((len(obs_steps), n, 2), dtype=torch.float32, device=x.device)
        solver.trajectory_forward(n, tick, steps, K, x, p, m, acc, ckpt, obs_steps, obs)
        ctx.solver, ctx.tick, ctx.steps, ctx.K, ctx.obs_steps = solver, tick, steps, K, list(obs_steps)
        ctx.work_bytes = work_bytes
        ctx.save_for_backward(m, ckpt)
        return obs

    @staticmethod
    def backward(ctx, g):
        m, ckpt = ctx.saved_tensors
        n = m.shape[0]
        gx = torch.empty((n, 2), dtype=torch.float32, device=m.device)
        gp = torch.empty_like(gx)
        gm = torch.empty_like(m)
        work = torch.empty(ctx.work_bytes // 4, dtype=torch.float32, device=m.device)
        ctx.solver.trajectory_backward(n, ctx.tick, ctx.steps, ctx.K, m, ckpt, ctx.obs_steps, _f32(g), gx, gp, gm,
                                       work)
        return None, None, None, None, None, gx, gp, gm


def fmm_trajectory(solver, tick, steps, checkpoint_every, obs_steps, x0, xpre0, m):
    return _Trajectory.apply(solver, float(tick), int(steps), int(checkpoint_every), list(obs_steps), x0, xpre0, m)


class _BatchTrajectory(torch.autograd.Function):
    """Same as _Trajectory for a Batch (every system steps together; the
    ticks are the batch's); the checkpoint loop runs here over the batched
    step and step-adjoint entry points."""

    @staticmethod
    def forward(ctx, batch, steps, K, obs_steps, x0, xpre0, m):
        x, p, m = _f32(x0).clone(), _f32(xpre0).clone(), _f32(m)
        acc = torch.empty_like(x)
        ckpts = []
        seen = {}
        k = 0
        while True:
            if k % K == 0:
                ckpts.append((x.clone(), p.clone()))
            if k in obs_steps:
                seen[k] = x.clone()
            if k == steps:
                break
            # the steps up to the next checkpoint or observation as one eager
            # run (the same bits as that many batch.step calls, with the sorts
            # chained; no graph, as x and p are new on every call)
            nxt = min([steps, (k // K + 1) * K] + [o for o in obs_steps if o > k])
            batch.run(nxt - k, x, p, m, acc, use_graph=False)
            k = nxt
        ctx.batch, ctx.steps, ctx.K, ctx.obs_steps = batch, steps, K, list(obs_steps)
        ctx.ckpts = ckpts
        ctx.save_for_backward(m)
        if not obs_steps:
            return torch.empty((0,) + tuple(x.shape), device=x.device)
        return torch.stack([seen[k] for k in obs_steps])

    @staticmethod
    def backward(ctx, g):
        (m,) = ctx.saved_tensors
        b, steps, K = ctx.batch, ctx.steps, ctx.K
        g = _f32(g)
        lam = torch.zeros((m.shape[0], 2), dtype=torch.float32, device=m.device)
        pi = torch.zeros_like(lam)
        gm = torch.zeros_like(m)

        def add_obs(k):
            for o, ks in enumerate(ctx.obs_steps):
                if ks == k:
                    lam.add_(g[o])

        add_obs(steps)
        # Segment sg is recomputed on a second stream into kept-state set
        # sg % 2 (slots (sg % 2) * K ..): steps k0 .. k1-2 again, each kept for
        # its adjoint, and the sort of x_{k1-1}; meanwhile this stream steps
        # back through segment sg + 1 from the other set.  The adjoint steps
        # read only kept states and their own scratch, the recompute only the
        # batch's own sort.
        main = torch.cuda.current_stream()
        rec = torch.cuda.Stream(device=m.device)
        rec.wait_stream(main)
        done = [torch.cuda.Event(), torch.cuda.Event()]
        kept = [torch.cuda.Event(), torch.cuda.Event()]
        for e in done:
            e.record(main)
        nseg = (steps + K - 1) // K
        for sg in range(nseg - 1, -1, -1):
            k0, k1 = sg * K, min(sg * K + K, steps)
            first = (sg % 2) * K
            with torch.cuda.stream(rec):
                rec.wait_event(done[sg % 2])  # the set's previous segment is stepped back
                x, p = (t.clone() for t in ctx.ckpts[sg])
                acc = torch.empty_like(x)
                b.run_keep(k1 - 1 - k0, first, x, p, m, acc, stream=rec)
                kept[sg % 2].record(rec)
            main.wait_event(kept[sg % 2])
            for k in range(k1 - 1, k0 - 1, -1):
                b.step_adjoint_kept(first + k - k0, lam, pi, gm, stream=main)
                add_obs(k)
            done[sg % 2].record(main)
        return None, None, None, None, lam, pi, gm


def fmm_batch_trajectory(batch, steps, checkpoint_every, obs_steps, x0, xpre0, m):
    return _BatchTrajectory.apply(batch, int(steps), int(checkpoint_every), list(obs_steps), x0, xpre0, m)
