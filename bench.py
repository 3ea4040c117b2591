#!/usr/bin/env python3
"""Benchmark of the MI355X ideal-MHD step: cell-updates/s of the 256^3
periodic smooth-vortex workload (gamma 5/3, +-1% seeded density noise) with
the HLLD solver on one GPU, FP64 and mixed precision, with the fraction of
the HBM roofline and the oracle's CPU rate on the host beside it.

    python bench.py [--gpus 1] [--steps 100] [--warmup 5] [--size 256]
                    [--solver hlld-athena] [--cfl 0.3]

Prints one JSON line on stdout; progress goes to stderr.  A step is one
main-loop iteration (timestep + apply_boundaries, pamhd_step) over the whole
grid with the state resident in HBM.  The time-step factor defaults to 0.3:
with 0.5 per dimension the unsplit first-order scheme's 3-D Courant sum
reaches ~1.3 on this state and grid-scale noise grows.
"""
import argparse
import json
import math
import os
import platform
import subprocess
import sys
import time

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, AMD specification
BYTES_PER_CELL_UPDATE = 128    # 8 FP64 read + 8 FP64 written per cell-update


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return platform.processor() or "unknown"


def host_threads():
    """Threads for the CPU leg: the cores this process may run on, at most
    16 (a one-GPU box's share of its host)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(solver_name, cfl, seconds, threads, state, cells, length, gamma, mu0,
                 max_steps=1 << 30):
    """The FP64 oracle (a port of the reference's step: same per-face
    arithmetic, CT, refits and lagged CFL, OpenMP over cells) on the same
    workload as the GPU, after its final-init step, stepped until about
    `seconds` of wall time (at least one step, at most max_steps)."""
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests"))
    import oracle_lib as O
    sid = {"rusanov": O.RUSANOV, "hll-athena": O.HLL, "hlld-athena": O.HLLD,
           "roe-athena": O.ROE}[solver_name]
    used = O.lib().or_set_threads(threads)
    g = O.make_grid(cells, length, gamma, mu0)
    # final init (not timed), then timed steps
    st, new, sm, zc = O.step(g, sid, 0.0, state, refits=2)
    cur = new
    n = cells[0] * cells[1] * cells[2]
    done = 0
    t0 = time.perf_counter()
    while True:
        dt = O.cfl_dt(g, cfl, sm, zc)
        st, new, sm, zc = O.step(g, sid, dt, cur, refits=3)
        if st != 0:
            raise RuntimeError("oracle step failed with %d" % st)
        cur = new
        done += 1
        el = time.perf_counter() - t0
        if el >= seconds or done >= max_steps:
            break
    return {"value": n * done / el, "unit": "cell-updates/s", "cores": used, "kind": "port",
            "sample": "%d step(s) of the %dx%dx%d periodic vortex after its final init, %s, "
                      "FP64 oracle (oracle/pamhd_oracle.c, -O2, no FMA), %d thread(s), %s"
                      % (done, *cells, solver_name, used, cpu_model()),
            "steps": done, "seconds": round(el, 3)}


def flux_calibration(faces=200000, reps=5, seed=7):
    """Per-face time of the reference's own compiled Rusanov and HLL flux
    functions (oracle/_ref/libref_flux.so, built from the reference's
    headers) against the oracle's restatement of them, one thread, on the
    same random faces (best of `reps` runs each); the ratio says how the
    restated CPU baseline relates to the reference's arithmetic.  None when
    the reference build is absent."""
    import ctypes
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests"))
    import numpy as np
    import oracle_lib as O
    import states
    if not os.path.exists(O.REF_SO):
        return None
    R = ctypes.CDLL(O.REF_SO)
    dp = ctypes.POINTER(ctypes.c_double)
    f = R.ref_flux_batch
    f.argtypes = [ctypes.c_int, ctypes.c_size_t, dp, dp, dp, ctypes.c_double, ctypes.c_double,
                  dp, dp]
    f.restype = ctypes.c_longlong
    L, Rr, _ = states.random_faces(faces, seed, kind="generic")
    L, Rr = np.ascontiguousarray(L), np.ascontiguousarray(Rr)
    F, sp = np.zeros((8, faces)), np.zeros(faces)
    p = lambda a: a.ctypes.data_as(dp)
    out = {}
    for name, which, sid in (("rusanov", 0, O.RUSANOV), ("hll-athena", 2, O.HLL)):
        t_ref, t_or = [], []
        for _ in range(reps):
            t0 = time.perf_counter()
            f(which, faces, p(L), p(Rr), None, 5.0 / 3.0, 1.0, p(F), p(sp))
            t1 = time.perf_counter()
            Fo, so, st = O.flux_batch(sid, L, Rr)
            t2 = time.perf_counter()
            t_ref.append(t1 - t0)
            t_or.append(t2 - t1)
        same = bool(np.array_equal(F, Fo) and np.array_equal(sp, so))
        out[name] = {"reference_ns_per_face": round(min(t_ref) / faces * 1e9, 2),
                     "oracle_ns_per_face": round(min(t_or) / faces * 1e9, 2),
                     "reference_over_oracle": min(t_ref) / min(t_or), "identical": same}
    out["faces"] = faces
    out["note"] = ("reference flux headers compiled with clang++ -O2 -ffp-contract=off; "
                   "HLLD and Roe headers need gensimcell / muparserx and are not built")
    return out


def run_gpu(prec, state, cells, length, gamma, mu0, solver, cfl, steps, warmup):
    import torch
    from pamhd_amd.solver import Grid
    n = cells[0] * cells[1] * cells[2]
    final = None
    with Grid(cells, length, gamma, mu0, precision=prec) as g:
        g.upload(state)
        stream = torch.cuda.current_stream().cuda_stream
        g.apply_boundaries(0.0, stream=stream)
        g.timestep(solver, 0.0, cfl, sync=True, stream=stream)   # final init, dt = 0
        tuning = g.autotune(solver, stream=stream)
        log("  autotune: " + tuning)
        for _ in range(warmup):
            g.step(solver, 1e300, cfl, sync=False, stream=stream)
        torch.cuda.synchronize()
        g.check()
        start = torch.cuda.Event(enable_timing=True)
        stop = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        start.record()
        for k in range(steps):
            g.step(solver, 1e300, cfl, sync=False, stream=stream)
        stop.record()
        torch.cuda.synchronize()
        ms = start.elapsed_time(stop)
        g.check()
        speeds = g.max_signal_speeds()
        final = g.download()
        # per-kernel breakdown (kernels serialised while timed), after the
        # state used for the error figures has been read back
        g.enable_timing(True)
        for k in range(10):
            g.step(solver, 1e300, cfl, sync=False, stream=stream)
        kt = {k: g.kernel_time(k) for k in ("dt", "flux_pass", "update")}
        g.enable_timing(False)
    return {"ms": ms, "value": n * steps / (ms / 1e3), "kernels": kt, "speeds": speeds,
            "tuning": tuning, "final": final}


def substep_run(size=256, steps=10, solver="hlld-athena", cfl=0.3):
    """Per-cell 2^N substepping with the reference's default options
    (substep-min 0, substep-max 999) on a 'size'^3 periodic box at rest with
    a hot sphere (30x the background pressure): FP64 ms per main-loop step,
    the largest period and the cells' mean period, against global steps on
    the same state (a main-loop step covers `largest period` global steps at
    the smallest cells' time step)."""
    import torch
    from pamhd_amd import configs
    from pamhd_amd.solver import Grid
    cells = (size,) * 3
    state, length, gamma, mu0 = configs.hot_blob(cells, contrast=30.0, radius=0.25, seed=3)
    n = size ** 3
    with Grid(cells, length, gamma, mu0) as g:
        g.upload(state)
        g.set_substepping(0, 999)
        stream = torch.cuda.current_stream().cuda_stream
        g.apply_boundaries(0.0, stream=stream)
        g.timestep(solver, 0.0, cfl, stream=stream)
        for _ in range(2):
            g.step(solver, 1e300, cfl, stream=stream)
        torch.cuda.synchronize()
        start = torch.cuda.Event(enable_timing=True)
        stop = torch.cuda.Event(enable_timing=True)
        start.record()
        for _ in range(steps):
            g.step(solver, 1e300, cfl, stream=stream)
        stop.record()
        torch.cuda.synchronize()
        ms = start.elapsed_time(stop) / steps
        g.check()
        largest = g.last_max_substep()
        periods = g.last_substep_periods()
    with Grid(cells, length, gamma, mu0) as g:
        g.upload(state)
        g.apply_boundaries(0.0, stream=stream)
        g.timestep(solver, 0.0, cfl, stream=stream)
        for _ in range(2):
            g.step(solver, 1e300, cfl, stream=stream)
        torch.cuda.synchronize()
        start.record()
        for _ in range(steps):
            g.step(solver, 1e300, cfl, stream=stream)
        stop.record()
        torch.cuda.synchronize()
        global_ms = start.elapsed_time(stop) / steps
        g.check()
    del state
    import numpy as np
    vals, counts = np.unique(periods, return_counts=True)
    return {"cells": list(cells), "solver": solver, "cfl": cfl, "substep_min": 0,
            "substep_max": 999, "steps": steps, "ms_per_step": round(ms, 4),
            "cell_updates_per_s": n / (ms / 1e3), "largest_period": largest,
            "mean_period": float(periods.mean()),
            "period_share": {str(int(v)): float(c) / n for v, c in zip(vals, counts)},
            "global_step_ms": round(global_ms, 4),
            "global_steps_per_main_step": int(largest),
            "speedup_vs_global_steps": largest * global_ms / ms,
            "workload": "periodic box at rest, uniform field, hot sphere of 30x the pressure "
                        "(radius 0.25), +-1% density noise"}


def rel_l2(a, b):
    """Relative L2 error per conserved variable of a against b; momentum and
    face field as vectors (their z components are roundoff-sized in this
    quasi-2-D state)."""
    import numpy as np
    out = {}
    for name, rows in (("rho", [0]), ("m", [1, 2, 3]), ("E", [4]), ("B", [5, 6, 7])):
        out[name] = float(np.linalg.norm(a[rows] - b[rows]) / np.linalg.norm(b[rows]))
    return out


def divergence_cleaning(cells, cycles_max=50):
    """Times one multigrid divergence cleaning of the vortex state with a
    seeded Gaussian (sigma 1e-3) divergence-carrying face-field perturbation:
    V(2,2) cycles to a 1e-10 relative residual, projection and energy refit."""
    import torch
    from pamhd_amd import configs
    from pamhd_amd.solver import Grid
    state, length, gamma, mu0 = configs.orszag_tang_vortex(cells, noise=0.01, seed=2024,
                                                           divergence_sigma=1e-3)
    res = {}
    with Grid(cells, length, gamma, mu0) as g:
        for fused in (True, False):
            g.upload(state)
            g.set_poisson_smoother(fused)
            g.clean_divergence(tol=1e-10, max_cycles=cycles_max)     # warm-up, allocation
            g.upload(state)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            before, rel, cyc, st = g.clean_divergence(tol=1e-10, max_cycles=cycles_max)
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t0) * 1e3
            after = g.div_face2volume()
            res["fused" if fused else "per_colour"] = {
                "ms": round(ms, 3), "v_cycles": cyc, "rel_residual": rel,
                "div_before": before, "div_after": after}
    return res


def ensemble_run(nbox, steps, solver="hll-athena", cfl=0.3, seed=2024, imf_steps=0):
    """Config of the ensemble workload: `nbox` solar-wind boxes with seeded
    edges in {32, 48, 64, 96, 128}, HLL, `steps` main-loop steps after the
    final-init step, without the IMF (the boxes stop within 50-200 steps
    with it; see configs.solar_wind_ensemble).  Times the ensemble (one
    launch per stage for all boxes,
    steps replayed from a captured HIP graph, autotuned tile shape) against a
    loop over the boxes with one handle each (plain launches), on the same
    stream, and the same steps in one persistent launch; results are
    compared box by box."""
    import numpy as np
    import torch
    from pamhd_amd import configs
    from pamhd_amd.swbox import BoxGrid, Ensemble
    boxes = configs.solar_wind_ensemble(nbox=nbox, seed=seed, imf=False)
    stream = torch.cuda.current_stream().cuda_stream
    t0 = time.perf_counter()
    with Ensemble(boxes) as e:
        log("  ensemble of %d boxes, %d cells, set up in %.1f s"
            % (nbox, e.cells, time.perf_counter() - t0))
        e.timestep(solver, 0.0, cfl, stream=stream)
        tune = e.autotune(solver, cfl, stream=stream)
        log("  " + tune)
        e.run(solver, cfl, 2, use_graph=False, stream=stream)     # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e.run(solver, cfl, steps, use_graph=True, stream=stream)
        torch.cuda.synchronize()
        ens_s = time.perf_counter() - t0
        dts = e.check(stream=stream)
        probe = {q: e.download(q) for q in (0, nbox - 1)}
    normal = 0
    grids = []
    for sb in boxes:
        g = BoxGrid(sb)
        ct, _ = g.cell_types()
        normal += int((ct == 1).sum())
        grids.append(g)
    for g in grids:
        g.timestep(solver, 0.0, cfl, stream=stream)
        for _ in range(2):
            g.step(solver, 1e300, cfl, sync=False, stream=stream)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        for g in grids:
            g.step(solver, 1e300, cfl, sync=False, stream=stream)
    torch.cuda.synchronize()
    loop_s = time.perf_counter() - t0
    same = all(np.array_equal(grids[q].download(), probe[q]) for q in probe)
    for g in grids:
        g.close()
    imf = ensemble_imf_run(nbox, imf_steps, normal) if imf_steps > 0 else None
    # the same steps in one persistent launch (last: a cooperative launch is
    # not to precede the other timings) (stages separated by grid-wide
    # barriers), from the same start and tile shape
    tiles = {"64x4": 0, "32x8": 1, "16x16": 2}
    with Ensemble(boxes) as e2:
        e2.set_tile(tiles.get(tune.split("; chosen ")[-1].strip(), 1))
        e2.timestep(solver, 0.0, cfl, stream=stream)
        e2.run(solver, cfl, 2, use_graph=False, stream=stream)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e2.run(solver, cfl, steps, persistent=True, stream=stream)
        torch.cuda.synchronize()
        pers_s = time.perf_counter() - t0
        pers_same = all(np.array_equal(e2.download(q), probe[q]) for q in probe)
    return {"boxes": nbox, "cells": int(sum(b.n for b in boxes)), "normal_cells": normal,
            "persistent_s": round(pers_s, 4), "persistent_cell_updates_per_s": normal * steps / pers_s,
            "persistent_identical": bool(pers_same), "with_imf": imf,
            "steps": steps, "solver": solver, "cfl": cfl,
            "ensemble_s": round(ens_s, 4), "per_box_loop_s": round(loop_s, 4),
            "ensemble_cell_updates_per_s": normal * steps / ens_s,
            "per_box_loop_cell_updates_per_s": normal * steps / loop_s,
            "speedup_vs_per_box_loop": loop_s / ens_s,
            "tuning": tune.split("; chosen ")[-1], "graph": True, "imf": False,
            "results_identical": bool(same), "min_dt": float(dts.min())}


def ensemble_imf_run(nbox, max_steps, normal, chunk=10, solver="hll-athena", cfl=0.3, seed=2024):
    """The ensemble workload with the interplanetary field it specifies
    (By, Bz uniform in [-10, 10] nT): the boxes are advanced `chunk` steps at
    a time (graph-replayed), checked after each chunk, until a box reports a
    physics error or `max_steps` are done.  Returns the steps completed, the
    throughput over them (`normal` cells per step: the boxes' classification
    does not depend on the field) and the error that ended the run."""
    import torch
    from pamhd_amd import configs
    from pamhd_amd._lib import PamhdError
    from pamhd_amd.swbox import Ensemble
    boxes = configs.solar_wind_ensemble(nbox=nbox, seed=seed, imf=True)
    stream = torch.cuda.current_stream().cuda_stream
    done, secs, stop = 0, 0.0, None
    with Ensemble(boxes) as e:
        e.timestep(solver, 0.0, cfl, stream=stream)
        e.check(stream=stream)
        while done < max_steps:
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            e.run(solver, cfl, chunk, use_graph=True, stream=stream)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            try:
                e.check(stream=stream)
            except PamhdError as err:
                stop = str(err)
                break
            done += chunk
            secs += t1 - t0
    return {"boxes": nbox, "imf": True, "steps_completed": done, "chunk": chunk,
            "cell_updates_per_s": (normal * done / secs) if secs > 0 else None,
            "stopped_by": stop}


def adjoint_run(n=512, steps=200, every=20, directions=32, seed=5):
    """The adjoint workload: 2-D periodic n x n vortex with seeded
    random-Fourier perturbations (|k| <= 8, amplitudes 1e-2 on rho, v, p,
    A_z), HLL, `steps` steps on the frozen schedule of a CFL-0.3 run; loss =
    magnetic energy of a fixed 64 x 64 window at the end.  Times the forward
    rollout and the backward (hand-written adjoint, checkpoints every
    `every` steps; medians of five repetitions) and checks the gradient against central differences
    (h = 1e-6) along `directions` smooth random directions (random-Fourier
    fields of every state component) and for the 4 amplitudes."""
    import numpy as np
    import torch
    from pamhd_amd import configs
    from pamhd_amd.autograd import rollout
    from pamhd_amd.solver import Grid
    dev = torch.device("cuda", 0)
    w = configs.adjoint_vortex_2d(n=n, seed=seed)
    amps = torch.tensor([1e-2] * 4, dtype=torch.float64, device=dev)
    U0 = configs.adjoint_initial_state(w, amps).contiguous()
    loss = lambda U: configs.window_magnetic_energy(U, n, lo=n // 2 - 32, size=64)
    with Grid((n, n, 1), w["length"], w["gamma"], w["mu0"]) as g:
        g.upload(U0.cpu().numpy())
        g.timestep("hll-athena", 0.0, 0.3)
        dts = [g.step("hll-athena", 1e300, 0.3) for _ in range(steps)]
        # warm-up, then timed forward and backward
        x = U0.clone().requires_grad_(True)
        loss(rollout(x, g, "hll-athena", dts, every)).backward()
        torch.cuda.synchronize()
        # five timed repetitions (one ~25 ms backward is a noisy sample);
        # the medians are reported, with their ranges
        fwd, bwd = [], []
        for _ in range(5):
            a = amps.clone().requires_grad_(True)
            t0 = time.perf_counter()
            L = loss(rollout(configs.adjoint_initial_state(w, a), g, "hll-athena", dts, every))
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            L.backward()
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            fwd.append((t1 - t0) * 1e3)
            bwd.append((t2 - t1) * 1e3)
        fwd_ms, bwd_ms = float(np.median(fwd)), float(np.median(bwd))
        x = U0.clone().requires_grad_(True)
        loss(rollout(x, g, "hll-athena", dts, every)).backward()
        lam0 = x.grad.detach()
        rng = np.random.default_rng(17)
        scale = U0.abs().mean(1).clamp_min(1e-3).cpu().numpy()
        h = 1e-6
        errs = []
        with torch.no_grad():
            for _ in range(directions):
                d = torch.as_tensor(configs.smooth_direction(n, scale, rng), device=dev)
                lp = loss(rollout(U0 + h * d, g, "hll-athena", dts, every)).item()
                lm = loss(rollout(U0 - h * d, g, "hll-athena", dts, every)).item()
                fd, ad = (lp - lm) / (2 * h), float((lam0 * d).sum())
                errs.append(abs(fd - ad) / max(abs(fd), abs(ad), 1e-300))
            amp_errs = []
            ga = a.grad.detach().cpu().numpy()
            for q in range(4):
                e = torch.zeros(4, dtype=torch.float64, device=dev)
                e[q] = h
                lp = loss(rollout(configs.adjoint_initial_state(w, amps + e), g, "hll-athena", dts, every)).item()
                lm = loss(rollout(configs.adjoint_initial_state(w, amps - e), g, "hll-athena", dts, every)).item()
                fd = (lp - lm) / (2 * h)
                amp_errs.append(abs(fd - ga[q]) / max(abs(fd), abs(ga[q]), 1e-300))
    return {"cells": [n, n, 1], "steps": steps, "checkpoint_every": every, "solver": "hll-athena",
            "loss": float(L.item()), "forward_ms": round(fwd_ms, 3), "backward_ms": round(bwd_ms, 3),
            "forward_ms_range": [round(min(fwd), 3), round(max(fwd), 3)],
            "backward_ms_range": [round(min(bwd), 3), round(max(bwd), 3)], "repetitions": len(fwd),
            "backward_over_forward": bwd_ms / fwd_ms, "fd_directions": directions,
            "fd_max_rel_err": max(errs), "amplitude_grad": [float(v) for v in ga],
            "amplitude_fd_rel_err": amp_errs}


def precision_study(size=512, steps=160, clean_every=10, report_every=20, solver="hlld-athena",
                    cfl=0.3):
    """The precision workload: the vortex at size^3 with a seeded Gaussian
    (sigma 1e-3) divergence-carrying face-field perturbation, `steps` HLLD
    steps in FP64 and in mixed precision side by side on the GPU, each with
    divergence cleaning (multigrid to a 1e-10 relative residual) before the
    first step and every `clean_every` steps; the relative L2 error per
    variable of the mixed run against the FP64 run every `report_every`
    steps, and the time of the steps and of the cleanings of each run (each
    phase bracketed by device synchronisation).

    160 steps, not 200: the scheme itself (the FP64 step, bit-identical to
    the CPU oracle) loses positivity on this perturbed state at step 180
    (512^3; the oracle and the device both stop at step 184 at 256^3)."""
    import torch
    from pamhd_amd import configs
    from pamhd_amd.solver import Grid
    cells = (size,) * 3
    n = size ** 3
    log("  building %d^3 state" % size)
    state, length, gamma, mu0 = configs.orszag_tang_vortex(cells, noise=0.01, seed=2024,
                                                           divergence_sigma=1e-3)
    stream = torch.cuda.current_stream().cuda_stream
    precs = ("fp64", "mixed")
    grids, tuning = {}, {}
    for prec in precs:
        g = Grid(cells, length, gamma, mu0, precision=prec)
        g.upload(state)
        g.apply_boundaries(0.0, stream=stream)
        g.timestep(solver, 0.0, cfl, sync=True, stream=stream)
        tuning[prec] = g.autotune(solver, stream=stream).split("; chosen ")[-1]
        grids[prec] = g
    del state
    bufs = {p: torch.empty((8, n), dtype=torch.float64, device="cuda") for p in precs}
    step_ms = dict.fromkeys(precs, 0.0)
    clean_ms = dict.fromkeys(precs, 0.0)
    cycles = {p: [] for p in precs}
    errors = {}

    def timed(f):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = f()
        torch.cuda.synchronize()
        return r, (time.perf_counter() - t0) * 1e3

    try:
        for k in range(steps):
            for p in precs:
                g = grids[p]
                if k % clean_every == 0:
                    r, ms = timed(lambda: g.clean_divergence(tol=1e-10, max_cycles=50, stream=stream))
                    clean_ms[p] += ms
                    cycles[p].append(r[2])
                _, ms = timed(lambda: g.step(solver, 1e300, cfl, sync=False, stream=stream))
                step_ms[p] += ms
            if (k + 1) % report_every == 0 or k + 1 == steps:
                for p in precs:
                    grids[p].check()
                    grids[p].get_state_device(bufs[p].data_ptr(), stream)
                torch.cuda.synchronize()
                a, b = bufs["mixed"], bufs["fp64"]
                errors[k + 1] = {name: float(torch.linalg.vector_norm(a[r] - b[r]) /
                                             torch.linalg.vector_norm(b[r]))
                                 for name, r in (("rho", slice(0, 1)), ("m", slice(1, 4)),
                                                 ("E", slice(4, 5)), ("B", slice(5, 8)))}
                log("  step %d: mixed vs fp64 relative L2 %s" % (k + 1, errors[k + 1]))
        div_end = {p: grids[p].div_face2volume() for p in precs}
    finally:
        for g in grids.values():
            g.close()
    out = {p: {"step_ms": round(step_ms[p] / steps, 4), "clean_ms_total": round(clean_ms[p], 2),
               "cell_updates_per_s": n * steps / (step_ms[p] / 1e3), "v_cycles": cycles[p],
               "div_end": div_end[p], "tuning": tuning[p]} for p in precs}
    return {"cells": list(cells), "steps": steps, "clean_every": clean_every, "solver": solver,
            "cfl": cfl, "fp64": out["fp64"], "mixed": out["mixed"],
            "mixed_speedup_steps": step_ms["fp64"] / step_ms["mixed"],
            "mixed_speedup_with_cleaning": (step_ms["fp64"] + clean_ms["fp64"]) /
                                           (step_ms["mixed"] + clean_ms["mixed"]),
            "mixed_rel_l2_vs_fp64_by_step": errors}


# HBM traffic of one step (FETCH_SIZE x 2.00 + WRITE_SIZE x 1.00 summed over
# the step's launches) measured with rocprofv3 --pmc for the step structure
# and launch variants the autotuner picks (keyed by precision, size, solver
# and HLLD step structure): profiles/step_kernels_256cubed_hlld_counters.md
STEP_TRAFFIC_BYTES = {("fp64", 256, "hlld-athena", "accumulate"): 13.64e9,
                      ("mixed", 256, "hlld-athena", "accumulate"): 14.94e9,
                      ("fp64", 256, "hlld-athena", "passes"): 14.34e9}
# VALU wave instructions of one step's flux launches (SQ_INSTS_VALU of the
# same profiled runs; the flux launches are VALU-issue bound)
FLUX_VALU_WAVE_INSTR = {("fp64", 256, "hlld-athena", "accumulate"): 8.44e8,
                        ("mixed", 256, "hlld-athena", "accumulate"): 5.95e8,
                        ("fp64", 256, "hlld-athena", "passes"): 7.47e8}
STRUCTURES = {"0": "passes", "1": "fused", "2": "accumulate"}
# one wave64 VALU instruction per clock per CU: 256 CUs x 2.4 GHz (the
# 78.6 TFLOP/s FP64 vector peak is 128 FLOP, one 64-lane FMA, per clock per CU)
VALU_PEAK_WAVE_INSTR_PER_S = 256 * 2.4e9


def structure_of(tuning):
    """the HLLD step structure the autotuner chose ("passes", "fused",
    "accumulate"), from its report's "chosen fused=M..." entry"""
    chosen = tuning.split("; chosen ")[-1]
    for tok in chosen.split():
        if tok.startswith("fused="):
            return STRUCTURES.get(tok.split("=")[1].split("/")[0], "passes")
    return "passes"


def roofline(res, n, steps, key=None):
    """Step roofline: algorithmic 128 B per cell-update over the device time
    of one step (HIP events around the timed steps on the stream the kernels
    run on; with z-slab pipelining the flux pass and update overlap).  The
    per-kernel times come from a separate run with the kernels serialised."""
    kt = res["kernels"]
    flux_ms, flux_n = kt["flux_pass"]
    upd_ms, upd_n = kt["update"]
    step_ms = res["ms"] / steps
    achieved = BYTES_PER_CELL_UPDATE * n / (step_ms / 1e3) / 1e9
    valu = None
    if key in FLUX_VALU_WAVE_INSTR and flux_n:
        rate = FLUX_VALU_WAVE_INSTR[key] / (flux_ms / flux_n / 1e3)
        valu = {"bound": "valu", "kernel": "flux_pass (x, y, z faces)", "achieved": rate,
                "peak": VALU_PEAK_WAVE_INSTR_PER_S, "unit": "wave instructions/s",
                "frac": round(rate / VALU_PEAK_WAVE_INSTR_PER_S, 4),
                "instructions": FLUX_VALU_WAVE_INSTR[key]}
    kernel = {"passes": "step (flux passes + update)",
              "fused": "step (fused flux/update pass + face-field update)",
              "accumulate": "step (x, y, z accumulating marches + face-field update)"}
    return {"bound": "hbm", "kernel": kernel.get(key[3] if key else "passes"),
            "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": STEP_TRAFFIC_BYTES.get(key),
            "traffic_unit": "bytes per step (rocprofv3 FETCH_SIZE/WRITE_SIZE, copy-calibrated)",
            "step_ms": round(step_ms, 4),
            "tuning": res["tuning"].split("; chosen ")[-1],
            "kernel_ms_serialised": {"flux_pass": round(flux_ms / flux_n, 4),
                                     "update": round(upd_ms / upd_n, 4),
                                     "dt": round(kt["dt"][0] / max(kt["dt"][1], 1), 4)},
            "flux_valu": valu}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--solver", default="hlld-athena")
    ap.add_argument("--cfl", type=float, default=0.3)
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="wall time of the all-core CPU baseline (the 1-core figure is one step)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--adjoint-steps", type=int, default=200,
                    help="0 skips the adjoint workload")
    ap.add_argument("--other-solver-steps", type=int, default=20,
                    help="timed FP64 steps of Rusanov, HLL and Roe on the same workload (0: skip)")
    ap.add_argument("--substep-steps", type=int, default=10,
                    help="timed steps of the per-cell substepping record (0: skip)")
    ap.add_argument("--no-cleaning", action="store_true",
                    help="skip the divergence-cleaning record")
    ap.add_argument("--ensemble-boxes", type=int, default=64)
    ap.add_argument("--ensemble-imf-steps", type=int, default=300,
                    help="most steps of the ensemble with its IMF (it stops at the first physics error)")
    ap.add_argument("--ensemble-steps", type=int, default=500,
                    help="0 skips the ensemble workload")
    ap.add_argument("--precision-study", action="store_true",
                    help="only run the 512^3 mixed-vs-FP64 workload (160 HLLD steps) with "
                         "divergence cleaning every 10 steps and print its JSON")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device-to-device copy (scales the "
                         "FETCH_SIZE / WRITE_SIZE counters of a --pmc run)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(requested --gpus %d)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    from pamhd_amd import configs
    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    if args.copy_calibration:
        a = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        log("copied 3 x 1 GiB")
        return
    if args.precision_study:
        log("precision workload: 512^3, 160 HLLD steps, cleaning every 10")
        print(json.dumps({"metric": "precision workload", "precision_study": precision_study()}),
              flush=True)
        return
    cells = (args.size,) * 3
    n = args.size ** 3
    log("building %d^3 vortex state" % args.size)
    state, length, gamma, mu0 = configs.orszag_tang_vortex(cells, noise=0.01, seed=2024)

    results = {}
    for prec in ("fp64", "mixed"):
        log("gpu run: %s, %s, %d warmup + %d timed steps" % (prec, args.solver, args.warmup, args.steps))
        results[prec] = run_gpu(prec, state, cells, length, gamma, mu0, args.solver,
                                args.cfl, args.steps, args.warmup)
        log("  %s: %.3f ms/step, %.3e cell-updates/s" % (prec, results[prec]["ms"] / args.steps,
                                                          results[prec]["value"]))
    # the other solvers of the same workload (FP64, autotuned), as a record
    # beside the headline solver
    others = {}
    for other in ("rusanov", "hll-athena", "roe-athena"):
        if args.other_solver_steps <= 0 or other == args.solver:
            continue
        log("gpu run: fp64, %s, 2 warmup + %d timed steps" % (other, args.other_solver_steps))
        r = run_gpu("fp64", state, cells, length, gamma, mu0, other, args.cfl,
                    args.other_solver_steps, 2)
        r.pop("final")
        others[other] = {"ms_per_step": r["ms"] / args.other_solver_steps,
                         "cell_updates_per_s": r["value"], "tuning": r["tuning"].split("chosen ")[-1]}
        log("  %s: %.3f ms/step, %.3e cell-updates/s" % (other, others[other]["ms_per_step"],
                                                          r["value"]))
    del state
    err = rel_l2(results["mixed"].pop("final"), results["fp64"].pop("final"))
    log("  mixed vs fp64 relative L2 after %d steps: %s" % (args.warmup + args.steps, err))
    clean = None
    if not args.no_cleaning:
        log("divergence cleaning, %d^3" % args.size)
        clean = divergence_cleaning(cells)
        log("  %s" % clean)

    sub = None
    if args.substep_steps > 0:
        log("per-cell substepping, %d^3 hot sphere, options (0, 999)" % args.size)
        sub = substep_run(args.size, args.substep_steps, args.solver, args.cfl)
        log("  %.3f ms/step, largest period %d" % (sub["ms_per_step"], sub["largest_period"]))

    adj = None
    if args.adjoint_steps > 0:
        log("adjoint: 512^2 HLL, %d steps" % args.adjoint_steps)
        adj = adjoint_run(steps=args.adjoint_steps)
        log("  forward %.1f ms, backward %.1f ms, FD max rel err %.2e, amplitudes %s"
            % (adj["forward_ms"], adj["backward_ms"], adj["fd_max_rel_err"], adj["amplitude_fd_rel_err"]))

    # last of the GPU legs: its persistent run is a cooperative launch, which
    # is not to precede the other legs' timings
    ens = None
    if args.ensemble_steps > 0:
        log("ensemble: %d solar-wind boxes, %d steps" % (args.ensemble_boxes, args.ensemble_steps))
        ens = ensemble_run(args.ensemble_boxes, args.ensemble_steps, imf_steps=args.ensemble_imf_steps)
        log("  ensemble %.3e cell-updates/s, per-box loop %.3e, x%.2f, identical %s"
            % (ens["ensemble_cell_updates_per_s"], ens["per_box_loop_cell_updates_per_s"],
               ens["speedup_vs_per_box_loop"], ens["results_identical"]))
        log("  persistent launch %.3e cell-updates/s, identical %s"
            % (ens["persistent_cell_updates_per_s"], ens["persistent_identical"]))
        imf = ens["with_imf"]
        if imf:
            log("  with the IMF: %d steps completed (%s)" % (imf["steps_completed"], imf["stopped_by"]))

    base = None
    if not args.no_cpu_baseline:
        threads = host_threads()
        log("cpu baseline: oracle on the %d^3 workload, %d threads (nproc %s), ~%.0f s"
            % (args.size, threads, os.cpu_count(), args.cpu_seconds))
        state, length, gamma, mu0 = configs.orszag_tang_vortex(cells, noise=0.01, seed=2024)
        base = cpu_baseline(args.solver, args.cfl, args.cpu_seconds, threads, state, cells,
                            length, gamma, mu0)
        log("  %.3e cell-updates/s (%d steps)" % (base["value"], base["steps"]))
        log("cpu baseline: 1 thread, one step")
        one = cpu_baseline(args.solver, args.cfl, 0.0, 1, state, cells, length, gamma, mu0,
                           max_steps=1)
        del state
        log("  %.3e cell-updates/s" % one["value"])
        base["single_core"] = {"value": one["value"], "cores": 1, "steps": one["steps"],
                               "seconds": one["seconds"]}
        base["nproc"] = os.cpu_count()
        base["cpu_model"] = cpu_model()
        log("flux calibration: reference's compiled flux vs the oracle's, per face")
        base["calibration"] = flux_calibration()
        log("  %s" % base["calibration"])

    f64 = results["fp64"]
    mix = results["mixed"]
    out = {
        "metric": "MHD cell-updates/s at 256^3 HLLD (FP64 & mixed) on 1 GPU; % of HBM roofline",
        "value": f64["value"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": f64["ms"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "periodic Orszag-Tang smooth vortex with +-1%% density "
                               "noise, %d^3 cells, %s, time-step factor %.2f"
                               % (args.size, args.solver, args.cfl),
                   "cells": list(cells), "solver": args.solver, "cfl": args.cfl,
                   "gamma": gamma, "precision": "fp64"},
        "roofline": roofline(f64, n, args.steps, ("fp64", args.size, args.solver,
                                                  structure_of(f64["tuning"]))),
        "cpu_baseline": base,
        "mixed": {"value": mix["value"], "ms_per_step": mix["ms"] / args.steps,
                  "dtype": "f32 flux/wave speeds, f64 state",
                  "speedup_vs_fp64": mix["value"] / f64["value"],
                  "rel_l2_vs_fp64": err,
                  "roofline": roofline(mix, n, args.steps, ("mixed", args.size, args.solver,
                                                            structure_of(mix["tuning"])))},
        "other_solvers_fp64": others,
        "substepping": sub,
        "divergence_cleaning": clean,
        "ensemble": ens,
        "adjoint": adj,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
