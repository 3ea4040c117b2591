"""Differentiable rollouts of the periodic MHD step for PyTorch.

rollout(U0, grid, solver, dts, every) advances the (8, n) float64 CUDA
tensor U0 by len(dts) main-loop steps of the given lengths (a frozen time-step
schedule) on `grid` and returns the final state.  Its backward runs the
library's hand-written adjoint (pamhd_adjoint_step) step by step in reverse,
with checkpointing: the forward keeps the state every `every` steps, and the
backward recomputes the states of one segment at a time from its checkpoint.
All calls are ordered on torch's current stream.
"""
import torch

from .solver import solver_id


class MHDRollout(torch.autograd.Function):

    @staticmethod
    def forward(ctx, U0, grid, solver, dts, every):
        if U0.dtype != torch.float64 or not U0.is_cuda or tuple(U0.shape) != (8, grid.n):
            raise ValueError("U0 must be a float64 CUDA tensor of shape (8, %d)" % grid.n)
        stream = torch.cuda.current_stream(U0.device).cuda_stream
        sid = solver_id(solver)
        x = U0.detach().contiguous()
        grid.set_state_device(x.data_ptr(), stream)
        ckpts = [x.clone()]
        for s, dt in enumerate(dts):
            grid.step_fixed(sid, dt, stream)
            if (s + 1) % every == 0 and s + 1 < len(dts):
                c = torch.empty_like(x)
                grid.get_state_device(c.data_ptr(), stream)
                ckpts.append(c)
        out = torch.empty_like(x)
        grid.get_state_device(out.data_ptr(), stream)
        ctx.grid, ctx.sid, ctx.dts, ctx.every = grid, sid, tuple(dts), every
        ctx.ckpts = ckpts
        return out

    @staticmethod
    def backward(ctx, lam_out):
        grid, sid, dts, every = ctx.grid, ctx.sid, ctx.dts, ctx.every
        stream = torch.cuda.current_stream(lam_out.device).cuda_stream
        lam = lam_out.detach().to(torch.float64).contiguous().clone()
        nsteps = len(dts)
        for seg in reversed(range(len(ctx.ckpts))):
            start = seg * every
            end = min(start + every, nsteps)
            # states U^start .. U^(end-1) of the segment, recomputed
            states = [ctx.ckpts[seg]]
            grid.set_state_device(states[0].data_ptr(), stream)
            for s in range(start, end - 1):
                grid.step_fixed(sid, dts[s], stream)
                u = torch.empty_like(states[0])
                grid.get_state_device(u.data_ptr(), stream)
                states.append(u)
            for s in reversed(range(start, end)):
                grid.adjoint_step(sid, dts[s], states[s - start].data_ptr(), lam.data_ptr(), stream)
        ctx.ckpts = None
        return lam, None, None, None, None


def rollout(U0, grid, solver, dts, every=20):
    return MHDRollout.apply(U0, grid, solver, tuple(float(d) for d in dts), int(every))
