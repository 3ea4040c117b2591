"""Differentiable rollouts of the periodic MHD step for PyTorch.

rollout(U0, grid, solver, dts, every) advances the (8, n) float64 CUDA
tensor U0 by len(dts) main-loop steps of the given lengths (a frozen time-step
schedule) on `grid` and returns the final state.  Its backward runs the
library's hand-written adjoint (pamhd_adjoint_step) step by step in reverse,
with checkpointing: the forward keeps the state every `every` steps, and the
backward recomputes the states of one segment at a time from its checkpoint,
the previous segment's on a second stream while the current segment's
reverse steps run, and keeps each recomputed step's magnetic fluxes for its
reverse step.  The forward and the reverse steps are ordered on torch's
current stream.
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
        dev = lam_out.device
        main = torch.cuda.current_stream(dev)
        # the recomputation of the previous segment's states runs on a side
        # stream while the current segment's reverse steps run on the main
        # one: the two only share the checkpoints (read-only), and on a grid
        # that does not fill the GPU alone they overlap
        side = torch.cuda.Stream(device=dev)
        lam = lam_out.detach().to(torch.float64).contiguous().clone()
        nsteps = len(dts)
        nseg = len(ctx.ckpts)
        n = ctx.ckpts[0].shape[1]
        nb = max(every - 1, 0)
        bufs = [[torch.empty_like(ctx.ckpts[0]) for _ in range(nb)] for _ in range(2)]
        # the recomputed steps' magnetic fluxes, which the reverse steps then
        # need not recompute
        fbufs = [[torch.empty((6, n), dtype=torch.float64, device=dev) for _ in range(nb)]
                 for _ in range(2)]
        for b in bufs + fbufs:
            for t in b:
                t.record_stream(side)

        def recompute(seg, stream):
            """U^start .. U^(end-1) of segment seg (its checkpoint and the
            states after each of its steps but the last) and the magnetic
            fluxes of those steps"""
            start = seg * every
            end = min(start + every, nsteps)
            ckpt = ctx.ckpts[seg]
            buf, fbuf = bufs[seg % 2], fbufs[seg % 2]
            sp = stream.cuda_stream
            # each step from the previous step's buffer straight into the
            # next, the update pass writing the step's magnetic fluxes
            states = [ckpt] + buf[:end - 1 - start]
            for s in range(start, end - 1):
                grid.step_fixed_from(sid, dts[s], states[s - start].data_ptr(),
                                     states[s - start + 1].data_ptr(),
                                     fbuf[s - start].data_ptr(), sp)
            return states, fbuf[:end - 1 - start]

        side.wait_stream(main)
        states, fluxes = recompute(nseg - 1, main)
        for seg in reversed(range(nseg)):
            ready = None
            if seg > 0:
                # its buffers were last read by the reverse of segment seg + 1
                side.wait_stream(main)
                with torch.cuda.stream(side):
                    prev = recompute(seg - 1, side)
                    ready = torch.cuda.Event()
                    ready.record(side)
            start = seg * every
            end = min(start + every, nsteps)
            order = list(reversed(range(start, end)))
            # the segment's last step was not recomputed: its fluxes are not kept
            grid.adjoint_steps(sid, [dts[s] for s in order],
                               [states[s - start].data_ptr() for s in order],
                               [fluxes[s - start].data_ptr() if s - start < len(fluxes) else None
                                for s in order],
                               lam.data_ptr(), main.cuda_stream)
            if ready is not None:
                main.wait_event(ready)
                states, fluxes = prev
        ctx.ckpts = None
        return lam, None, None, None, None


def rollout(U0, grid, solver, dts, every=20):
    return MHDRollout.apply(U0, grid, solver, tuple(float(d) for d in dts), int(every))
