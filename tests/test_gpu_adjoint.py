"""The hand-written adjoint of the periodic MHD step (HLL, frozen time-step
schedule) behind torch.autograd, against FP64 central finite differences of
the FP64 forward, on a small 2-D version of the adjoint workload, and
against torch.autograd of a PyTorch restatement of one step.

The forward of the rollout is the library's FP64 step, so it equals an
ordinary run with the same time steps bit for bit.  Gradients are checked
along random directions of the initial state and for the four perturbation
amplitudes (whose gradient flows through torch's autograd of the initial
state construction into the adjoint).  Tolerance: central differences with
h = 1e-6 have truncation error O(h^2) and roundoff O(eps L / h); 1e-5
relative leaves room for HLL's max/min kinks crossed by a perturbation.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N, STEPS, EVERY, LO, SIZE = 48, 30, 10, 16, 16


def _setup(gpu):
    import torch
    from pamhd_amd import configs
    from pamhd_amd.solver import Grid
    w = configs.adjoint_vortex_2d(n=N, seed=5)
    amps = torch.tensor([1e-2] * 4, dtype=torch.float64, device=gpu)
    U0 = configs.adjoint_initial_state(w, amps).contiguous()
    g = Grid((N, N, 1), w["length"], w["gamma"], w["mu0"])
    # time-step schedule of an ordinary CFL-limited run from U0
    g.upload(U0.cpu().numpy())
    g.timestep("hll-athena", 0.0, 0.3)
    dts = [g.step("hll-athena", 1e300, 0.3) for _ in range(STEPS)]
    return w, amps, U0, g, dts


def _loss(U):
    from pamhd_amd import configs
    return configs.window_magnetic_energy(U, N, lo=LO, size=SIZE)


def test_rollout_forward_is_the_fp64_step(gpu):
    import torch
    from pamhd_amd.autograd import rollout
    w, amps, U0, g, dts = _setup(gpu)
    with g:
        UT = rollout(U0, g, "hll-athena", dts, every=EVERY)
        g.upload(U0.cpu().numpy())
        for dt in dts:
            g.step_fixed("hll-athena", dt)
        ref = g.download()
    np.testing.assert_array_equal(UT.cpu().numpy(), ref)
    assert torch.isfinite(UT).all()


def test_adjoint_matches_finite_differences(gpu):
    """Central differences along smooth random directions of all 8 state
    components (random-Fourier fields, |k| <= 8)."""
    import torch
    from pamhd_amd import configs
    from pamhd_amd.autograd import rollout
    w, amps, U0, g, dts = _setup(gpu)
    with g:
        x = U0.clone().requires_grad_(True)
        L = _loss(rollout(x, g, "hll-athena", dts, every=EVERY))
        L.backward()
        lam0 = x.grad.detach()
        assert torch.isfinite(lam0).all()
        rng = np.random.default_rng(17)
        scale = U0.abs().mean(1).clamp_min(1e-3).cpu().numpy()
        h = 1e-6
        errs = []
        for q in range(8):
            d = torch.as_tensor(configs.smooth_direction(N, scale, rng), device=gpu)
            with torch.no_grad():
                lp = _loss(rollout(U0 + h * d, g, "hll-athena", dts, every=EVERY)).item()
                lm = _loss(rollout(U0 - h * d, g, "hll-athena", dts, every=EVERY)).item()
            fd = (lp - lm) / (2 * h)
            ad = float((lam0 * d).sum())
            err = abs(fd - ad) / max(abs(fd), abs(ad), 1e-300)
            print("direction", q, "fd", fd, "adjoint", ad, "rel err", err, flush=True)
            errs.append(err)
        assert max(errs) < 1e-5


def test_amplitude_gradients(gpu):
    import torch
    from pamhd_amd import configs
    from pamhd_amd.autograd import rollout
    w, amps, U0, g, dts = _setup(gpu)
    with g:
        a = amps.clone().requires_grad_(True)
        L = _loss(rollout(configs.adjoint_initial_state(w, a), g, "hll-athena", dts, every=EVERY))
        L.backward()
        grad = a.grad.detach().cpu().numpy()
        h = 1e-6
        for q in range(4):
            e = torch.zeros(4, dtype=torch.float64, device=gpu)
            e[q] = h
            with torch.no_grad():
                lp = _loss(rollout(configs.adjoint_initial_state(w, amps + e), g, "hll-athena", dts,
                                   every=EVERY)).item()
                lm = _loss(rollout(configs.adjoint_initial_state(w, amps - e), g, "hll-athena", dts,
                                   every=EVERY)).item()
            fd = (lp - lm) / (2 * h)
            err = abs(fd - grad[q]) / max(abs(fd), abs(grad[q]), 1e-300)
            print("amplitude", q, "fd", fd, "adjoint", grad[q], "rel err", err, flush=True)
            assert err < 1e-5


def test_checkpoint_interval_does_not_change_gradient(gpu):
    import torch
    from pamhd_amd.autograd import rollout
    w, amps, U0, g, dts = _setup(gpu)
    with g:
        grads = []
        for every in (1, 7, STEPS):
            x = U0.clone().requires_grad_(True)
            _loss(rollout(x, g, "hll-athena", dts, every=every)).backward()
            grads.append(x.grad.detach().cpu().numpy())
    np.testing.assert_array_equal(grads[0], grads[1])
    np.testing.assert_array_equal(grads[0], grads[2])


@pytest.mark.parametrize("reverse", [True, False])
@pytest.mark.parametrize("cells", [(8, 8, 1), (6, 5, 4)])
def test_adjoint_step_matches_autograd_of_restated_step(gpu, cells, reverse):
    """One reverse step against torch.autograd of the PyTorch restatement of
    the same step (tests/torch_step.py) on the CPU, component by component,
    for a random lambda, with the flux Jacobian by reverse accumulation and
    by dual numbers; the restatement's forward equals the device step to
    rounding."""
    import torch
    import torch_step
    from pamhd_amd import configs
    from pamhd_amd.solver import Grid
    state, length, gamma, mu0 = configs.orszag_tang_vortex(cells, noise=0.05, seed=3)
    rng = np.random.default_rng(1)
    state[1:4] *= 1 + 0.1 * rng.normal(size=state[1:4].shape)
    state[3] += 0.01 * rng.normal(size=state[3].shape)
    state[7] += 0.01 * rng.normal(size=state[7].shape)
    state[5:8] += 0.005 * rng.normal(size=state[5:8].shape)    # divergence-carrying
    n = state.shape[1]
    nx, ny, nz = cells
    dt = 0.2 * min(length) / 3
    lam1 = rng.normal(size=(8, n))
    with Grid(cells, length, gamma, mu0) as g:
        g.set_adjoint_jacobian(reverse)
        g.upload(state)
        g.step_fixed("hll-athena", dt)
        fwd = g.download()
        U = torch.from_numpy(state).to(gpu).contiguous()
        lam = torch.from_numpy(lam1).to(gpu).contiguous()
        g.adjoint_step("hll-athena", dt, U.data_ptr(), lam.data_ptr(),
                       torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        lam0 = lam.cpu().numpy()
    x = torch.from_numpy(state.reshape(8, nz, ny, nx).copy()).requires_grad_(True)
    y = torch_step.step(x, dt, length, gamma, mu0)
    np.testing.assert_allclose(y.detach().numpy().reshape(8, n), fwd, rtol=1e-11, atol=1e-13)
    (y * torch.from_numpy(lam1.reshape(8, nz, ny, nx))).sum().backward()
    ref = x.grad.numpy().reshape(8, n)
    for q in range(8):
        err = np.abs(lam0[q] - ref[q]).max() / max(np.abs(ref[q]).max(), 1e-300)
        print("component", q, "max rel err", err, flush=True)
    scale = np.abs(ref).max(1, keepdims=True) + 1e-300
    assert np.all(np.abs(lam0 - ref) <= 1e-9 * scale)


def test_reverse_and_dual_jacobians_agree(gpu):
    """the two forms of the flux Jacobian on a hot blob with a field-free
    region (both branches of the fast speed) and random lambda: the same
    reverse step to rounding"""
    import torch
    from pamhd_amd import configs
    from pamhd_amd.solver import Grid
    cells = (16, 12, 10)
    state, length, gamma, mu0 = configs.hot_blob(cells, contrast=30.0, radius=0.25, seed=3)
    state[5:8, : state.shape[1] // 3] = 0.0      # no field in a third of the cells
    n = state.shape[1]
    kin = 0.5 * (state[1:4] ** 2).sum(0) / state[0]
    vb = 0.5 * (state[5:8] + np.roll(state[5:8], -1, 1))
    state[4] = np.maximum(state[4], kin + 0.5 / mu0 * (vb ** 2).sum(0) + 0.1)
    rng = np.random.default_rng(7)
    lam1 = rng.normal(size=(8, n))
    outs = []
    for reverse in (True, False):
        with Grid(cells, length, gamma, mu0) as g:
            g.set_adjoint_jacobian(reverse)
            U = torch.from_numpy(state).to(gpu).contiguous()
            lam = torch.from_numpy(lam1).to(gpu).contiguous()
            g.adjoint_step("hll-athena", 0.01, U.data_ptr(), lam.data_ptr(),
                           torch.cuda.current_stream().cuda_stream)
            torch.cuda.synchronize()
            outs.append(lam.cpu().numpy())
    scale = np.abs(outs[1]).max(1, keepdims=True) + 1e-300
    err = np.abs(outs[0] - outs[1]) / scale
    print("max relative difference per component", err.max(1), flush=True)
    assert np.all(err <= 1e-12)


@pytest.mark.parametrize("cells", [(16, 12, 1), (10, 9, 8)])
def test_adjoint_with_the_forward_steps_fluxes(gpu, cells):
    """The reverse step fed with the forward step's own magnetic fluxes
    (pamhd_get_magnetic_fluxes after the step) gives the lambda of the one
    that recomputes them, bit for bit; the fluxes are refused after a step
    that kept no FP64 flux array."""
    import torch
    from pamhd_amd import _lib, configs
    from pamhd_amd.solver import Grid
    state, length, gamma, mu0 = configs.orszag_tang_vortex(cells, noise=0.05, seed=4)
    n = state.shape[1]
    dt = 0.2 * min(length) / 3
    rng = np.random.default_rng(2)
    lam1 = rng.normal(size=(8, n))
    stream = torch.cuda.current_stream().cuda_stream
    with Grid(cells, length, gamma, mu0) as g:
        U = torch.from_numpy(state).to(gpu).contiguous()
        g.set_state_device(U.data_ptr(), stream)
        g.step_fixed("hll-athena", dt, stream)
        pe = torch.empty((6, n), dtype=torch.float64, device=gpu)
        g.get_magnetic_fluxes(pe.data_ptr(), stream)
        a = torch.from_numpy(lam1).to(gpu).contiguous()
        b = a.clone()
        g.adjoint_step("hll-athena", dt, U.data_ptr(), a.data_ptr(), stream)
        g.adjoint_step_fluxes("hll-athena", dt, U.data_ptr(), pe.data_ptr(), b.data_ptr(), stream)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(a.cpu().numpy(), b.cpu().numpy())
        g.set_fused_step("accumulate")
        g.step_fixed("hlld-athena", dt, stream)
        with pytest.raises(_lib.PamhdError):
            g.get_magnetic_fluxes(pe.data_ptr(), stream)


@pytest.mark.parametrize("cells", [(70, 9, 1), (66, 5, 6)])
def test_adjoint_steps_match_single_steps(gpu, cells):
    """Four reverse steps in one call (the latest step's fluxes recomputed,
    the others given, so the final gather of each step also makes the next
    step's G and mu_flux) give the lambda of four pamhd_adjoint_step calls,
    bit for bit; grids wider than one 64-cell tile exercise the tiles'
    boundary terms."""
    import torch
    from pamhd_amd import configs
    from pamhd_amd.solver import Grid
    state, length, gamma, mu0 = configs.orszag_tang_vortex(cells, noise=0.05, seed=6)
    n = state.shape[1]
    dts = [0.15 * min(length) / 3, 0.1 * min(length) / 3, 0.12 * min(length) / 3, 0.08 * min(length) / 3]
    rng = np.random.default_rng(3)
    lam1 = rng.normal(size=(8, n))
    stream = torch.cuda.current_stream().cuda_stream
    with Grid(cells, length, gamma, mu0) as g:
        U = [torch.from_numpy(state).to(gpu).contiguous()]
        F = []
        g.set_state_device(U[0].data_ptr(), stream)
        for dt in dts:
            g.step_fixed("hll-athena", dt, stream)
            f = torch.empty((6, n), dtype=torch.float64, device=gpu)
            g.get_magnetic_fluxes(f.data_ptr(), stream)
            F.append(f)
            u = torch.empty_like(U[0])
            g.get_state_device(u.data_ptr(), stream)
            U.append(u)
        a = torch.from_numpy(lam1).to(gpu).contiguous()
        b = a.clone()
        for s in reversed(range(len(dts))):
            g.adjoint_step("hll-athena", dts[s], U[s].data_ptr(), a.data_ptr(), stream)
        order = list(reversed(range(len(dts))))
        g.adjoint_steps("hll-athena", [dts[s] for s in order], [U[s].data_ptr() for s in order],
                        [None if s == order[0] else F[s].data_ptr() for s in order],
                        b.data_ptr(), stream)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(a.cpu().numpy(), b.cpu().numpy())


@pytest.mark.parametrize("cells,update_kz", [((16, 12, 1), 0), ((10, 9, 8), 0), ((10, 9, 8), 4)])
def test_step_fixed_from_matches_step_fixed(gpu, cells, update_kz):
    """pamhd_step_fixed_from steps between two caller buffers: the same new
    state as pamhd_step_fixed and, written by the update pass (per-cell and
    column-marching kernels), the magnetic fluxes of
    pamhd_get_magnetic_fluxes, bit for bit; the grid's own state is left as
    it was, and overlapping buffers or fluxes from a step that keeps no flux
    array are refused."""
    import torch
    from pamhd_amd import _lib, configs
    from pamhd_amd.solver import Grid
    state, length, gamma, mu0 = configs.orszag_tang_vortex(cells, noise=0.05, seed=8)
    n = state.shape[1]
    dt = 0.15 * min(length) / 3
    stream = torch.cuda.current_stream().cuda_stream
    with Grid(cells, length, gamma, mu0) as g:
        g.set_launch(update_kz=update_kz)
        U0 = torch.from_numpy(state).to(gpu).contiguous()
        g.set_state_device(U0.data_ptr(), stream)
        g.step_fixed("hll-athena", dt, stream)
        want_f = torch.empty((6, n), dtype=torch.float64, device=gpu)
        g.get_magnetic_fluxes(want_f.data_ptr(), stream)
        want_u = torch.empty_like(U0)
        g.get_state_device(want_u.data_ptr(), stream)
        other = torch.from_numpy(state[::-1].copy()).to(gpu).contiguous()
        g.set_state_device(other.data_ptr(), stream)
        u1 = torch.full_like(U0, np.nan)
        f1 = torch.full_like(want_f, np.nan)
        g.step_fixed_from("hll-athena", dt, U0.data_ptr(), u1.data_ptr(), f1.data_ptr(), stream)
        u2 = torch.full_like(U0, np.nan)
        g.step_fixed_from("hll-athena", dt, U0.data_ptr(), u2.data_ptr(), None, stream)
        own = torch.empty_like(U0)
        g.get_state_device(own.data_ptr(), stream)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(u1.cpu().numpy(), want_u.cpu().numpy())
        np.testing.assert_array_equal(u2.cpu().numpy(), want_u.cpu().numpy())
        np.testing.assert_array_equal(f1.cpu().numpy(), want_f.cpu().numpy())
        np.testing.assert_array_equal(own.cpu().numpy(), other.cpu().numpy())
        with pytest.raises(_lib.PamhdError):
            g.step_fixed_from("hll-athena", dt, U0.data_ptr(), U0.data_ptr() + 8, None, stream)
        g.set_fused_step("accumulate")
        with pytest.raises(_lib.PamhdError):
            g.step_fixed_from("hlld-athena", dt, U0.data_ptr(), u1.data_ptr(), f1.data_ptr(), stream)
