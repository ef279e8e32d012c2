"""Differentiation: the hand-written adjoint of the deterministic Heun step against central finite
differences of the FP64 forward (relative 1e-4), the FP32 adjoint against the FP64 one, the forward
against the oracle, and the torch.autograd wrapper."""
import numpy as np
import pytest

from oracle import oracle as O
from vampire_amd import system as S
from vampire_amd.fit import Trajectory, torch_function

pytestmark = pytest.mark.gpu

J0, KU0, ALPHA0 = S.J_CO, S.KU_CO, 0.1
HZ = (0.0, 0.0, 1.0)


def tilted(n, seed, deg=30.0, jitter=8.0):
    rng = np.random.default_rng(seed)
    th = np.radians(deg + jitter * rng.standard_normal(n))
    ph = 2 * np.pi * rng.random(n) * 0.05
    return np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])


@pytest.fixture(scope="module")
def setup():
    s = S.ferromagnet("bcc", 5, alpha=ALPHA0)        # 250 atoms
    sp = tilted(s.n, 3)
    return s, sp


def loss_and_grad(traj, target, J, Ku, a, h=None):
    mz = traj.forward(J, Ku, a, h)
    r = mz - target
    L = float((r * r).mean())
    return L, 2.0 * r / r.size


def test_forward_matches_oracle(gpu, setup):
    s, sp = setup
    traj = Trajectory(s, sp, nsteps=200, record_every=10, checkpoint_every=50, H=HZ)
    mz = traj.forward(J0, KU0, ALPHA0)
    o = O.Oracle(s)
    o.set_conditions(0.0, HZ, thermal=False)
    cur, ref = sp, []
    for _ in range(20):
        cur = o.llg_heun(cur, 10)
        ref.append(cur[2].mean())
    assert np.abs(mz - np.array(ref)).max() < 1e-11
    traj.close()


def test_adjoint_matches_finite_differences(gpu, setup):
    s, sp = setup
    traj = Trajectory(s, sp, nsteps=300, record_every=10, checkpoint_every=40, H=HZ)
    rng = np.random.default_rng(5)
    h = 0.05 * rng.standard_normal((3, s.n))
    target = traj.forward(J0 * 1.1, KU0 * 0.85, ALPHA0 * 1.15, h)
    L0, dmz = loss_and_grad(traj, target, J0, KU0, ALPHA0, h)
    dJ, dKu, da, dh = traj.backward(dmz)
    for name, g, x, f in (
        ("J", dJ, J0, lambda v: loss_and_grad(traj, target, v, KU0, ALPHA0, h)[0]),
        ("Ku", dKu, KU0, lambda v: loss_and_grad(traj, target, J0, v, ALPHA0, h)[0]),
        ("alpha", da, ALPHA0, lambda v: loss_and_grad(traj, target, J0, KU0, v, h)[0]),
    ):
        eps = 1e-4 * abs(x)
        fd = (f(x + eps) - f(x - eps)) / (2 * eps)
        assert abs(g - fd) <= 1e-4 * abs(fd), (name, g, fd)
    for (k, i) in ((0, 3), (2, 17), (1, 101)):
        eps = 1e-4
        hp, hm = h.copy(), h.copy()
        hp[k, i] += eps
        hm[k, i] -= eps
        fd = (loss_and_grad(traj, target, J0, KU0, ALPHA0, hp)[0]
              - loss_and_grad(traj, target, J0, KU0, ALPHA0, hm)[0]) / (2 * eps)
        assert abs(dh[k, i] - fd) <= 1e-4 * abs(fd) + 1e-12, (k, i, dh[k, i], fd)
    traj.close()


def test_fp32_adjoint_close_to_fp64(gpu, setup):
    s, sp = setup
    out = {}
    for prec in ("fp64", "fp32"):
        traj = Trajectory(s, sp, nsteps=200, record_every=10, checkpoint_every=50, precision=prec, H=HZ)
        target = np.linspace(0.86, 0.9, 20)
        L, dmz = loss_and_grad(traj, target, J0, KU0, ALPHA0)
        out[prec] = np.array(traj.backward(dmz, want_h=False)[:3])
        traj.close()
    np.testing.assert_allclose(out["fp32"], out["fp64"], rtol=2e-3)


def test_torch_wrapper_gradients_and_descent(gpu, setup):
    torch = pytest.importorskip("torch")
    s, sp = setup
    traj = Trajectory(s, sp, nsteps=200, record_every=10, checkpoint_every=50, H=HZ)
    target = torch.as_tensor(traj.forward(J0 * 1.1, KU0 * 0.9, ALPHA0 * 1.2))
    F = torch_function()
    J = torch.tensor(J0 / 1e-21, dtype=torch.float64, requires_grad=True)       # optimise in 1e-21 J
    Ku = torch.tensor(KU0 / 1e-24, dtype=torch.float64, requires_grad=True)
    a = torch.tensor(ALPHA0, dtype=torch.float64, requires_grad=True)
    losses = []
    opt = torch.optim.Adam([J, Ku, a], lr=0.02)
    for it in range(15):
        opt.zero_grad()
        mz = F.apply(J * 1e-21, Ku * 1e-24, a, None, traj)
        loss = ((mz - target) ** 2).mean()
        loss.backward()
        if it == 0:
            dJ, dKu, da, _ = traj.backward((2 * (mz - target) / mz.numel()).detach().numpy(), want_h=False)
            assert abs(J.grad.item() - dJ * 1e-21) <= 1e-9 * abs(dJ * 1e-21) + 1e-300
        opt.step()
        losses.append(loss.item())
    assert losses[-1] < 0.5 * losses[0]
    traj.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_adjoint_graph_replay_equals_eager(gpu, setup, precision):
    """Backward calls with two or more full checkpoint segments replay one captured HIP graph per
    segment; a single segment runs the same kernels eagerly.  Checkpoint spacing does not change
    the recomputed states, so 250 steps with checkpoints every 100 (two replays and an eager
    remainder) and every 250 (one eager segment) give the same gradients bit for bit (each
    workgroup sums its own terms over the steps, and the workgroup sums are added in a fixed
    order), and a second backward call (a fresh capture) repeats them."""
    s, sp = setup
    rng = np.random.default_rng(9)
    h = 0.05 * rng.standard_normal((3, s.n))
    out = []
    for K in (100, 250):
        traj = Trajectory(s, sp, nsteps=250, record_every=10, checkpoint_every=K, H=HZ, precision=precision)
        mz = traj.forward(J0, KU0, ALPHA0, h)
        dmz = 2.0 * (mz - 0.5) / mz.size
        g1 = traj.backward(dmz)
        g2 = traj.backward(dmz)
        out += [g1, g2]
        traj.close()
    for g in out[1:]:
        assert np.array_equal(np.array(g[:3]), np.array(out[0][:3]))
        assert np.array_equal(g[3], out[0][3])
