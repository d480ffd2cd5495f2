"""torch.autograd through the GPU adjoint: the density-loss gradient that
autograd returns equals the engine's adjoint and the oracle's tangent at the
FP32 trajectory (1e-4)."""
import numpy as np
import pytest

from world_amd import workloads
from world_amd.adjoint import DensityWindow, density_window

pytestmark = pytest.mark.gpu


def test_autograd_density_loss_matches_adjoint_and_oracle(oracle_lib):
    import torch

    c = workloads.column(8, 8, 8, friction=0.3)
    win = DensityWindow(c, 60)  # checkpoints every ceil(sqrt(60)) = 8 steps
    try:
        assert win.every == 8
        x_init = win.x0.copy()
        with torch.no_grad():
            target = density_window(torch.tensor(x_init), torch.tensor(0.5, dtype=torch.float64),
                                    torch.tensor(150.0), win)[:, 0].clone()
        x0 = torch.tensor(x_init, requires_grad=True)
        mu = torch.tensor(0.3, dtype=torch.float64, requires_grad=True)
        k_n = torch.tensor(150.0, requires_grad=True)
        rho = density_window(x0, mu, k_n, win)
        loss = ((rho[:, 0] - target) ** 2).sum()
        loss.backward()
        assert x0.grad.shape == x0.shape and mu.grad.shape == mu.shape and k_n.grad.shape == k_n.shape

        G = np.zeros((1, 4, win.nodes), np.float32)
        G[:, 0] = (2.0 * (rho[:, 0] - target)).detach().numpy()
        fin = win.sim.state()
        gx, gmu, gk = win.backward(G)
        np.testing.assert_allclose(x0.grad.numpy(), gx, rtol=1e-6, atol=1e-12)
        assert float(mu.grad) == pytest.approx(float(gmu[0]), rel=1e-6)
        assert float(k_n.grad) == pytest.approx(gk, rel=1e-6)

        st0 = np.concatenate([x_init, win.v0], axis=1)
        R = oracle_lib.Recording(c, st0, np.ones(len(st0), np.int32), 60)
        np.testing.assert_array_equal(R.final_state(), fin)
        rng = np.random.default_rng(6)
        tx = rng.normal(size=x_init.shape)
        _, _, _, drho = R.replay(66, x_init, win.v0, float(np.float32(0.3)), 150.0, tx, None, 1.0, 10.0)
        jvp = G.reshape(-1).astype(np.float64) @ drho
        vjp = (x0.grad.numpy() * tx).sum() + float(mu.grad) + 10.0 * float(k_n.grad)
        assert vjp == pytest.approx(jvp, rel=1e-4)
    finally:
        win.close()


def test_two_forwards_then_one_backward_and_broadcast_friction():
    """L(f(x_a)) + L(f(x_b)) through one window: each term is differentiated
    along its own trajectory (the same gradients, bit for bit, as two
    separate forward/backward passes).  A scalar friction broadcast over a
    two-member batch gets the sum of the members' gradients."""
    import torch

    cs = [workloads.column(5, 5, 6, friction=0.3), workloads.column(4, 4, 7, friction=0.3, seed=9)]
    win = DensityWindow(cs, 30)
    try:
        xa = win.x0.copy()
        xb = xa + np.random.default_rng(1).normal(0, 2e-4, xa.shape).astype(np.float32)
        target = torch.zeros(2, win.nodes, dtype=torch.float32)

        def loss_of(x, mu):
            return ((density_window(x, mu, torch.tensor(150.0), win)[:, 0] - target) ** 2).sum()

        ta = torch.tensor(xa, requires_grad=True)
        tb = torch.tensor(xb, requires_grad=True)
        mu = torch.tensor(0.3, dtype=torch.float64, requires_grad=True)
        (loss_of(ta, mu) + loss_of(tb, mu)).backward()

        sep = []
        for x in (xa, xb):
            tx = torch.tensor(x, requires_grad=True)
            tmu = torch.tensor([0.3, 0.3], dtype=torch.float64, requires_grad=True)
            loss_of(tx, tmu).backward()
            sep.append((tx.grad.numpy().copy(), tmu.grad.numpy().copy()))
        np.testing.assert_array_equal(ta.grad.numpy(), sep[0][0])
        np.testing.assert_array_equal(tb.grad.numpy(), sep[1][0])
        assert not np.array_equal(sep[0][0], sep[1][0])
        assert mu.grad.shape == mu.shape
        assert float(mu.grad) == pytest.approx(float(sep[0][1].sum() + sep[1][1].sum()), rel=1e-12)
    finally:
        win.close()


DEVICE_SCRIPT = r"""
import numpy as np
import torch

torch.cuda.init()  # torch (imported first) holds the HIP runtime the library binds to
from world_amd import _capi, workloads
from world_amd.adjoint import DensityWindow, density_window

c = workloads.column(6, 6, 7, friction=0.3)
win = DensityWindow(c, 40)
x_init = win.x0.copy()
target = torch.zeros(1, win.nodes, dtype=torch.float32)
grads, grids = [], []
for dev in ("cpu", "cuda"):
    x0 = torch.tensor(x_init, device=dev, requires_grad=True)
    mu = torch.tensor(0.3, dtype=torch.float64, device=dev, requires_grad=True)
    k_n = torch.tensor(150.0, device=dev, requires_grad=True)
    rho = density_window(x0, mu, k_n, win)
    assert rho.device.type == dev
    ((rho[:, 0] - target.to(dev)) ** 2).sum().backward()
    assert x0.grad.device.type == dev
    grids.append(rho.detach().cpu().numpy())
    grads.append((x0.grad.cpu().numpy(), float(mu.grad), float(k_n.grad)))
np.testing.assert_array_equal(grids[0], grids[1])
np.testing.assert_array_equal(grads[0][0], grads[1][0])
assert grads[0][1:] == grads[1][1:]
assert np.count_nonzero(grads[0][0]) > 0
host = np.zeros((win.sim.n, 6), np.float32)
try:
    _capi.check(_capi.lib().wd_set_state_device(win.sim._h, _capi.ptr(host)))
    raise AssertionError("a host buffer was accepted")
except _capi.WdError as e:
    assert "not device memory" in str(e)
win.close()
print("device path: the host path's bits", flush=True)
"""


def test_device_tensors_give_the_host_path_bits():
    """Initial positions on the GPU: the grids and gradients come back on the
    GPU (device buffers across the C ABI) with the same bits as the host path;
    a host pointer given to a device entry point is refused.  It runs in a
    process of its own that imports torch before the library loads (the
    order the device path needs: see world_amd/adjoint.py)."""
    import subprocess
    import sys
    from pathlib import Path

    root = Path(__file__).resolve().parent.parent
    r = subprocess.run([sys.executable, "-c", DEVICE_SCRIPT], cwd=root, capture_output=True, text=True,
                       timeout=240)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    assert "the host path's bits" in r.stdout
