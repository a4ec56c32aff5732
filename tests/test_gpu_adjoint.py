"""GPU adjoint (d3q19_adjoint_* through the C ABI) against the FP64 CPU
adjoint of the oracle and against central finite differences of the GPU
forward map (no reference counterpart: the reference has no adjoint; its
gradient-check pattern is common/common-warp/src/common_warp/autodiff/finite_diff.py:26-47).
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from conftest import field_scaled_err, print_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mods():
    import torch

    from lattice_boltzmann_d3q19_amd import autograd as ag

    return torch, ag


def problem(shape, seed=42, rms=0.02, famp=1e-4):
    rng = np.random.default_rng(seed)
    u0 = rms * rng.standard_normal((3,) + shape)
    F = famp * rng.standard_normal((3,) + shape)
    ut = rms * rng.standard_normal((3,) + shape)
    return u0, F, ut


def gpu_grad(torch, ag, shape, tau, T, K, storage, u0, F, ut):
    solver = ag.DifferentiableLBM(shape, tau, T, checkpoint_every=K, storage=storage)
    u0t = torch.tensor(u0, device="cuda", requires_grad=True)
    Ft = torch.tensor(F, device="cuda", requires_grad=True)
    utt = torch.tensor(ut, device="cuda")
    loss = ag.velocity_matching_loss(solver(u0t, Ft), utt)
    loss.backward()
    return float(loss.detach()), u0t.grad.cpu().numpy(), Ft.grad.cpu().numpy(), solver


# (5, 7, 70): rows of 70 cells in blocks of 128 z x 2 rows (58 padding lanes, a last block
# with one live row), as the recomputation's moments and the reverse step see them
@pytest.mark.parametrize("T,K,shape", [(7, 3, (6, 5, 4)), (6, 6, (6, 5, 4)), (5, 1, (6, 5, 4)), (7, 3, (5, 7, 70))])
def test_fp64_adjoint_matches_oracle_adjoint(mods, T, K, shape):
    torch, ag = mods
    print_case("adjoint-fp64-vs-oracle", T=T, K=K, shape=shape)
    u0, F, ut = problem(shape)
    L, gu, gF, _ = gpu_grad(torch, ag, shape, 0.6, T, K, "f64", u0, F, ut)
    Lr, gur, gFr = oracle.inverse_grad(u0, F, ut, 0.6, T)
    assert L == pytest.approx(Lr, rel=1e-12)
    assert field_scaled_err(gu, gur) < 1e-10
    assert field_scaled_err(gF, gFr) < 1e-10


def test_fp32_adjoint_matches_oracle_adjoint(mods):
    torch, ag = mods
    print_case("adjoint-fp32-vs-oracle")
    shape = (10, 8, 12)
    u0, F, ut = problem(shape, seed=7)
    L, gu, gF, _ = gpu_grad(torch, ag, shape, 0.6, 40, 10, "f32", u0, F, ut)
    Lr, gur, gFr = oracle.inverse_grad(u0, F, ut, 0.6, 40)
    assert L == pytest.approx(Lr, rel=1e-5)
    assert field_scaled_err(gu, gur) < 1e-4
    assert field_scaled_err(gF, gFr) < 1e-4


@pytest.mark.parametrize("storage,eps,tol", [("f64", 1e-5, 1e-7), ("f32", 1e-3, 1e-4)])
def test_adjoint_matches_central_differences(mods, storage, eps, tol):
    """<grad, d> vs (L(p + eps d) - L(p - eps d)) / 2 eps with an FP64 forward.

    FP32 adjoint at the specified eps = 1e-3 within 1e-4; the FP64 adjoint at
    eps = 1e-5, where the central-difference truncation error is ~1e-9."""
    torch, ag = mods
    print_case("adjoint-fd", storage=storage)
    shape = (16, 16, 16)
    T, K, tau = 60, 10, 0.6
    u0, F, ut = problem(shape, seed=3, famp=0.0)
    _, gu, gF, _ = gpu_grad(torch, ag, shape, tau, T, K, storage, u0, F, ut)
    fwd = ag.DifferentiableLBM(shape, tau, T, checkpoint_every=K, storage="f64")
    utt = torch.tensor(ut, device="cuda")

    def loss(a, b):
        u = fwd.forward_raw(torch.tensor(a, device="cuda"), torch.tensor(b, device="cuda"))
        return float(ag.velocity_matching_loss(u, utt))

    rng = np.random.default_rng(42)
    fds, ads = [], []
    for k in range(4):
        du = rng.standard_normal(u0.shape)
        dF = 1e-3 * rng.standard_normal(F.shape)  # forces live on a 1e-3 smaller scale
        fds.append((loss(u0 + eps * du, F + eps * dF) - loss(u0 - eps * du, F - eps * dF)) / (2 * eps))
        ads.append(float(np.sum(gu.astype(np.float64) * du) + np.sum(gF.astype(np.float64) * dF)))
    # scaled by the largest directional derivative, as the testkit scales by the field maximum
    scale = max(abs(v) for v in fds)
    for k, (ad, fd) in enumerate(zip(ads, fds)):
        assert abs(ad - fd) <= tol * scale, (k, ad, fd)


def test_checkpoint_interval_autotune_and_memory(mods):
    torch, ag = mods
    print_case("adjoint-autotune")
    s = ag.DifferentiableLBM((8, 8, 8), 0.6, 100, checkpoint_every=0)
    # the least memory: ceil(100 / K) checkpoints of 64 rows x 19 x 64 padded
    # FP32 values (311,296 B) against K moment slots of 4 x 512 FP32 (8,192 B)
    assert s.checkpoint_every == 50
    assert s.memory_bytes == 2 * 311296 + 50 * 8192 + 4 * 311296 + 4 * 3 * 512 * 4
    s.close()


def test_adjoint_errors(mods):
    torch, ag = mods
    from lattice_boltzmann_d3q19_amd import D3Q19Error

    print_case("adjoint-errors")
    s = ag.DifferentiableLBM((4, 4, 4), 0.6, 5)
    with pytest.raises(D3Q19Error, match="ESTATE"):
        s.backward_raw(torch.zeros((3, 4, 4, 4), device="cuda"))
    # a completed forward pass enables backward; a failed one disables it again
    u0 = torch.zeros((3, 4, 4, 4), device="cuda")
    s.forward_raw(u0, u0)
    s.backward_raw(u0)
    from lattice_boltzmann_d3q19_amd._device import ptr
    from lattice_boltzmann_d3q19_amd._lib import call

    with pytest.raises(D3Q19Error, match="EINVAL"):
        call("d3q19_adjoint_forward", s._h, ptr(u0), ptr(u0), None, ptr(u0), 99, 1)
    with pytest.raises(D3Q19Error, match="ESTATE"):
        s.backward_raw(u0)
    s.close()
    with pytest.raises(D3Q19Error, match="EINVAL"):
        ag.DifferentiableLBM((4, 4, 4), 0.6, 5, storage="f16")


def test_one_backward_per_forward(mods):
    """The recomputation in backward overwrites the buffer that holds H_T, so
    a second backward of one forward is refused (ESTATE), and so is the
    backward of a forward that a later forward replaced; the next forward
    gives the same gradients bit for bit (T = 10, K = 5)."""
    torch, ag = mods
    from lattice_boltzmann_d3q19_amd import D3Q19Error

    print_case("adjoint-one-backward-per-forward", T=10, K=5)
    shape = (6, 5, 4)
    u0, F, ut = problem(shape, seed=11)
    s = ag.DifferentiableLBM(shape, 0.6, 10, checkpoint_every=5, storage="f64")
    u0t, Ft, g = (torch.tensor(a, device="cuda") for a in (u0, F, ut))
    s.forward_raw(u0t, Ft)
    first = [t.clone() for t in s.backward_raw(g)[:2]]
    with pytest.raises(D3Q19Error, match="ESTATE"):
        s.backward_raw(g)
    s.forward_raw(u0t, Ft)
    again = s.backward_raw(g)[:2]
    assert all(torch.equal(a, b) for a, b in zip(first, again))
    # through autograd: a retained graph cannot run the adjoint twice
    a = u0t.clone().requires_grad_(True)
    loss = ag.outlet_velocity(s(a, Ft))
    loss.backward(retain_graph=True)
    with pytest.raises(RuntimeError):
        loss.backward()
    # the backward of a replaced forward raises instead of using the newer checkpoints
    a = u0t.clone().requires_grad_(True)
    old = ag.velocity_matching_loss(s(a, Ft), g)
    s(a.detach() * 2, Ft)
    with pytest.raises(RuntimeError, match="replaced"):
        old.backward()
    s.close()


def channel_problem(shape, seed=5):
    rng = np.random.default_rng(seed)
    u0 = 0.01 * rng.standard_normal((3,) + shape)
    u0[:, :, 0, :] = 0.0
    u0[:, :, -1, :] = 0.0
    F = np.zeros((3,) + shape)
    F[0] = 1e-4
    s = np.clip(rng.random(shape) * 1.4 - 0.6, 0.0, 1.0)
    return u0, F, s


@pytest.mark.parametrize("storage,tol", [("f64", 1e-6), ("f32", 2e-4)])
def test_halfway_brinkman_adjoint_matches_central_differences(mods, storage, tol):
    """Walls on -y/+y, Brinkman field, per-cell force: <grad, d> vs FP64 central differences."""
    torch, ag = mods
    print_case("adjoint-halfway-brinkman-fd", storage=storage)
    shape = (10, 9, 8)
    T, K, tau = 30, 7, 0.7
    u0, F, s = channel_problem(shape)
    kw = dict(boundary="halfway", wall_faces=("-y", "+y"), brinkman=True, checkpoint_every=K)
    solver = ag.DifferentiableLBM(shape, tau, T, storage=storage, **kw)
    a = torch.tensor(u0, device="cuda", requires_grad=True)
    b = torch.tensor(F, device="cuda", requires_grad=True)
    c = torch.tensor(s, device="cuda", requires_grad=True)
    J = ag.outlet_velocity(solver(a, b, c))
    J.backward()
    gu, gF, gs = (t.grad.double().cpu().numpy() for t in (a, b, c))
    fwd = ag.DifferentiableLBM(shape, tau, T, storage="f64", **kw)

    def Jat(x, y, z):
        return float(ag.outlet_velocity(fwd.forward_raw(*(torch.tensor(v, device="cuda") for v in (x, y, z)))))

    rng = np.random.default_rng(11)
    eps = 1e-5 if storage == "f64" else 1e-3
    fds, ads = [], []
    for _ in range(4):
        du = rng.standard_normal(u0.shape) * 0.1
        dF = rng.standard_normal(F.shape) * 1e-3
        ds = rng.standard_normal(s.shape)
        fds.append((Jat(u0 + eps * du, F + eps * dF, s + eps * ds) - Jat(u0 - eps * du, F - eps * dF, s - eps * ds)) / (2 * eps))
        ads.append(float(np.sum(gu * du) + np.sum(gF * dF) + np.sum(gs * ds)))
    scale = max(abs(v) for v in fds)
    for k, (ad, fd) in enumerate(zip(ads, fds)):
        assert abs(ad - fd) <= tol * scale, (k, ad, fd)


def test_fp16_checkpoints_close_to_fp32_checkpoints(mods):
    torch, ag = mods
    print_case("adjoint-fp16-checkpoints")
    shape = (16, 12, 10)
    u0, F, s = channel_problem(shape, seed=9)
    kw = dict(boundary="halfway", wall_faces=("-y", "+y"), brinkman=True, checkpoint_every=8)
    outs = {}
    for cks in (None, "f16"):
        solver = ag.DifferentiableLBM(shape, 0.7, 40, checkpoint_storage=cks, **kw)
        uT = solver.forward_raw(*(torch.tensor(v, device="cuda") for v in (u0, F, s)))
        g = torch.zeros_like(uT)
        g[0, -1] = 1.0 / (shape[1] * shape[2])
        outs[cks] = (uT.cpu().numpy(), [t.cpu().numpy() for t in solver.backward_raw(g)])
    # the forward never reads checkpoints; the gradients see the rounded states
    assert np.array_equal(outs[None][0], outs["f16"][0])
    for a, b in zip(outs[None][1], outs["f16"][1]):
        assert field_scaled_err(b, a) < 2e-3


def test_memory_budget_autotune(mods):
    torch, ag = mods
    from lattice_boltzmann_d3q19_amd import D3Q19Error

    print_case("adjoint-budget")
    shape = (16, 16, 16)
    s = ag.DifferentiableLBM(shape, 0.6, 400, checkpoint_every=0, checkpoint_storage="f16",
                             memory_budget=1 << 30)
    # FP16 checkpoints of 256 rows x 19 x 64 (622,592 B) against moment slots of 4 x 4096 FP32 (65,536 B)
    K = s.checkpoint_every
    assert K == 67 and s.memory_bytes <= 1 << 30
    s.close()
    with pytest.raises(D3Q19Error, match="ENOMEM"):
        ag.DifferentiableLBM(shape, 0.6, 400, checkpoint_every=0, memory_budget=1 << 20)


# Moving walls: the reference y-walls (both plates moving) and a half-way box
# with two moving faces.  The moving-wall terms are constants of the half-way
# map but enter the y-wall map through the wall density, so its transpose
# carries a rank-one term.
YWALLS = dict(boundary="ywalls", wall_velocity_top=(0.05, 0.0, 0.01), wall_velocity_bottom=(-0.02, 0.0, 0.0))
LID_FACES = {"+y": (0.05, 0.0, 0.02), "-x": (0.0, 0.01, 0.0)}
LID = dict(boundary="halfway", wall_faces=("-x", "+x", "-y", "+y", "-z", "+z"), face_velocity=LID_FACES)


def wall_problem(shape, seed):
    rng = np.random.default_rng(seed)
    u0 = 0.01 * rng.standard_normal((3,) + shape)
    F = 1e-5 * rng.standard_normal((3,) + shape)
    F[0] += 1e-4
    s = np.clip(rng.random(shape) * 1.4 - 0.6, 0.0, 1.0)
    ut = 0.01 * rng.standard_normal((3,) + shape)
    return u0, F, s, ut


@pytest.mark.parametrize("ny", [9, 2, 1])
def test_ywalls_adjoint_forward_is_the_reference_loop(mods, ny):
    """u^T of the y-wall chain = macroscopic_velocity(f^T, F) of the reference
    loop f <- apply_bounce_back_y_walls(bgk_step(f, tau, F)) from feq(1, u0)."""
    torch, ag = mods
    print_case("adjoint-ywalls-forward", ny=ny)
    shape = (6, ny, 5)
    T, tau = 13, 0.7
    u0, F, _, _ = wall_problem(shape, 3)
    solver = ag.DifferentiableLBM(shape, tau, T, checkpoint_every=4, storage="f64", **YWALLS)
    u = solver.forward_raw(torch.tensor(u0, device="cuda"), torch.tensor(F, device="cuda")).cpu().numpy()
    f = oracle.run_reference(oracle.feq_field(np.ones(shape), u0), tau, T, force_lattice=F, ywalls=True,
                             wall_velocity_top=YWALLS["wall_velocity_top"],
                             wall_velocity_bottom=YWALLS["wall_velocity_bottom"])
    assert field_scaled_err(u, oracle.macroscopic_velocity(f, F)) < 1e-11


def test_moving_lid_adjoint_forward_is_the_halfway_loop(mods):
    """u^T of the half-way chain with moving faces and Brinkman drag = the
    oracle's half-way trajectory from feq(1, u0)."""
    torch, ag = mods
    print_case("adjoint-lid-forward")
    shape = (7, 6, 5)
    T, tau = 11, 0.7
    u0, F, s, _ = wall_problem(shape, 4)
    solver = ag.DifferentiableLBM(shape, tau, T, checkpoint_every=3, storage="f64", brinkman=True, **LID)
    u = solver.forward_raw(*(torch.tensor(v, device="cuda") for v in (u0, F, s))).cpu().numpy()
    fu = np.zeros((6, 3))
    fu[3], fu[0] = LID_FACES["+y"], LID_FACES["-x"]
    f = oracle.run_halfway(oracle.feq_field(np.ones(shape), u0), tau, T, wall_mask=63, face_velocity=fu,
                           force_field=F, brinkman=s)
    Feff = F - s[None] * oracle.momentum_field(f)
    assert field_scaled_err(u, oracle.macroscopic_velocity(f, Feff)) < 1e-11


@pytest.mark.parametrize("storage,tol", [("f64", 1e-6), ("f32", 2e-4)])
@pytest.mark.parametrize("case", ["ywalls", "ywalls-ny1", "lid"])
def test_moving_wall_adjoint_matches_central_differences(mods, case, storage, tol):
    """<grad, d> in (u0, F, s) vs FP64 central differences of the same forward map."""
    torch, ag = mods
    print_case("adjoint-moving-wall-fd", case=case, storage=storage)
    shape = (7, 1, 6) if case == "ywalls-ny1" else (8, 7, 6)
    T, K, tau = 25, 6, 0.7
    u0, F, s, ut = wall_problem(shape, 5)
    kw = dict(LID if case == "lid" else YWALLS, brinkman=True, checkpoint_every=K)
    solver = ag.DifferentiableLBM(shape, tau, T, storage=storage, **kw)
    a, b, c = (torch.tensor(v, device="cuda", requires_grad=True) for v in (u0, F, s))
    utt = torch.tensor(ut, device="cuda")
    ag.velocity_matching_loss(solver(a, b, c), utt).backward()
    gu, gF, gs = (t.grad.double().cpu().numpy() for t in (a, b, c))
    fwd = ag.DifferentiableLBM(shape, tau, T, storage="f64", **kw)

    def Lat(x, y, z):
        u = fwd.forward_raw(*(torch.tensor(v, device="cuda") for v in (x, y, z)))
        return float(ag.velocity_matching_loss(u, utt))

    rng = np.random.default_rng(12)
    eps = 1e-5 if storage == "f64" else 1e-3
    fds, ads = [], []
    for _ in range(4):
        du = rng.standard_normal(u0.shape) * 0.1
        dF = rng.standard_normal(F.shape) * 1e-3
        ds = rng.standard_normal(s.shape)
        fds.append((Lat(u0 + eps * du, F + eps * dF, s + eps * ds) - Lat(u0 - eps * du, F - eps * dF, s - eps * ds)) / (2 * eps))
        ads.append(float(np.sum(gu * du) + np.sum(gF * dF) + np.sum(gs * ds)))
    scale = max(abs(v) for v in fds)
    for k, (ad, fd) in enumerate(zip(ads, fds)):
        assert abs(ad - fd) <= tol * scale, (k, ad, fd)


def test_ywalls_adjoint_with_fp16_checkpoints(mods):
    """The y-wall chain's f^0 slot also round-trips through FP16 checkpoints."""
    torch, ag = mods
    print_case("adjoint-ywalls-fp16-ckpt")
    shape = (8, 7, 6)
    u0, F, s, ut = wall_problem(shape, 6)
    grads = []
    for cks in (None, "f16"):
        solver = ag.DifferentiableLBM(shape, 0.7, 20, checkpoint_every=6, storage="f32", brinkman=True,
                                      checkpoint_storage=cks, **YWALLS)
        a, b, c = (torch.tensor(v, device="cuda", requires_grad=True) for v in (u0, F, s))
        ag.velocity_matching_loss(solver(a, b, c), torch.tensor(ut, device="cuda")).backward()
        grads.append([t.grad.double().cpu().numpy() for t in (a, b, c)])
    for g32, g16 in zip(*grads):
        assert field_scaled_err(g16, g32) < 1e-3
