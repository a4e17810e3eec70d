"""The adjoint (gt_adj_forward / gt_adj_backward): gradients with respect to
the initial field of the run fft_dir+trunc -> nsteps x (step, forcing) ->
fft_inv.

* Dot-product test against the oracle's tangent-linear model (pinned on the
  CPU by finite differences): <gT, M' dw0> = <M'^T gT, dw0> to 1e-10.
* Central finite differences of the GPU forward model along 16 seeded
  directions on the adjoint workload (N = 512, alpha = 1, RK4, dt = 5e-4,
  200 steps, forcing off, checkpoints every 20 steps), eps = 1e-6: within
  1e-6.
* The torch.autograd wrapper gives the same gradient as the C ABI.
"""
import numpy as np
import pytest

from conftest import rel_l2
from helpers import params_for
from gturbo_amd import ic, workloads
from gturbo_amd.adjoint import Adjoint
from gturbo_amd.params import InitialSpectrum

pytestmark = pytest.mark.gpu


def _tlm(o, w0, dw0, nsteps, seed):
    """Oracle tangent of w0 -> wT along dw0 (forcing included in the state)."""
    n = o.n
    z = np.zeros((n, n + 2))
    z[:, :n] = w0
    dz = np.zeros((n, n + 2))
    dz[:, :n] = dw0
    for a in (z, dz):
        o.fft_dir(a)
        o.trunc(a)
    for _ in range(nsteps):
        o.step_tl(z, dz)
        seed = o.forcing(z, seed)
    o.fft_inv(z)
    o.fft_inv(dz)
    return z[:, :n], dz[:, :n]


@pytest.mark.parametrize("n,alpv,rko,nsteps,kchk", [
    (64, 2.0, 4, 7, 3), (64, 1.0, 4, 6, 6), (128, 0.5, 4, 5, 2), (64, 1.0, 2, 8, 3),
    (256, 1.0, 4, 4, 2), (4096, 2.0, 4, 2, 1)])
def test_adjoint_dot_product_with_oracle_tangent(oracle_mod, n, alpv, rko, nsteps, kchk):
    p = params_for(n, alpv=alpv, rko=rko, famp=1.0)
    w0 = ic.physical_field(p, InitialSpectrum(seed=11, p=4.0, k0=n / 16.0, energy=0.5))
    w0 = w0.astype(np.float64)
    rng = np.random.default_rng(n + nsteps)
    dw0 = rng.standard_normal((n, n))
    gT = rng.standard_normal((n, n))
    seed = 481929
    o = oracle_mod.Oracle(p)
    wT_o, dwT = _tlm(o, w0, dw0, nsteps, seed)
    with Adjoint(p, nsteps, kchk) as adj:
        wT, seed_out = adj.forward(w0, seed)
        g0 = adj.backward(gT)
    assert rel_l2(wT, wT_o) < 1e-11
    lhs = float(np.sum(gT * dwT))
    rhs = float(np.sum(g0 * dw0))
    print(n, alpv, rko, lhs, rhs, flush=True)
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)


def test_adjoint_gradient_matches_finite_differences():
    """The adjoint workload: loss 0.5 ||wT - obs||^2 on every 4th grid point
    in each direction; obs = the seed-200 truth run plus Gaussian noise of
    1e-3 x its rms; gradient at a first guess (truth + 5% of another field)."""
    w = workloads()["adjoint512"]
    p = w.params
    n, nsteps = p.n, p.npas
    truth = ic.physical_field(p, w.ic).astype(np.float64)
    other = ic.physical_field(p, InitialSpectrum(seed=201, p=2.0, k0=16.0, energy=0.5))
    guess = truth + 0.05 * other
    rng = np.random.default_rng(200)
    mask = np.zeros((n, n))
    mask[::4, ::4] = 1.0
    with Adjoint(p, nsteps, kchk=20) as adj:
        wt, _ = adj.forward(truth, w.forcing_seed)
        obs = wt + 1e-3 * np.sqrt(np.mean(wt ** 2)) * rng.standard_normal((n, n))

        def loss(w0):
            wT, _ = adj.forward(w0, w.forcing_seed)
            return 0.5 * float(np.sum(mask * (wT - obs) ** 2)), wT

        L0, wT = loss(guess)
        g0 = adj.backward(mask * (wT - obs))
        scale = np.sqrt(np.mean(guess ** 2))
        eps = 1e-6
        worst = 0.0
        for k in range(16):
            d = np.random.default_rng(1000 + k).standard_normal((n, n))
            d *= scale / np.sqrt(np.mean(d ** 2))
            fd = (loss(guess + eps * d)[0] - loss(guess - eps * d)[0]) / (2 * eps)
            ad = float(np.sum(g0 * d))
            err = abs(fd - ad) / abs(fd)
            worst = max(worst, err)
            print(k, fd, ad, err, flush=True)
            assert err < 1e-6
    print("worst relative difference", worst, "loss", L0)


def _torch_case():
    from helpers import params_for as pf
    p = pf(128, alpv=1.0, famp=1.0)
    w0 = ic.physical_field(p, InitialSpectrum(seed=3, p=2.0, k0=8.0, energy=0.5))
    obs = np.random.default_rng(5).standard_normal((128, 128)) * 0.01
    return p, w0.astype(np.float64), obs


def test_torch_autograd_cpu_tensors_give_the_library_gradient():
    torch = pytest.importorskip("torch")
    from gturbo_amd.adjoint import run
    p, w0, obs = _torch_case()
    with Adjoint(p, 9, kchk=4) as adj:
        wT_np, _ = adj.forward(w0, 17)
        g_np = adj.backward(wT_np - obs)
        wt = torch.tensor(w0, requires_grad=True)
        wT = run(wt, adj, seed=17)
        loss = 0.5 * ((wT - torch.tensor(obs)) ** 2).sum()
        loss.backward()
        assert np.array_equal(wT.detach().numpy(), wT_np)
        assert np.array_equal(wt.grad.numpy(), g_np)


_DEVICE_SCRIPT = """
import sys
import torch  # first: the library then binds to torch's HIP runtime
sys.path.insert(0, {root!r}); sys.path.insert(0, {tests!r})
import numpy as np
from test_gpu_adjoint import _torch_case
from gturbo_amd.adjoint import Adjoint, run
p, w0, obs = _torch_case()
with Adjoint(p, 9, kchk=4) as adj:
    wT_np, _ = adj.forward(w0, 17)
    g_np = adj.backward(wT_np - obs)
    wt = torch.tensor(w0, device="cuda:0", requires_grad=True)
    wT = run(wt, adj, seed=17)
    loss = 0.5 * ((wT - torch.tensor(obs, device="cuda:0")) ** 2).sum()
    loss.backward()
    torch.cuda.synchronize()
    assert np.array_equal(wT.detach().cpu().numpy(), wT_np)
    assert np.array_equal(wt.grad.cpu().numpy(), g_np)
print("device autograd ok")
"""


def test_torch_autograd_gpu_tensors_give_the_library_gradient():
    """GPU tensors go to the library in place; run in a fresh process that
    imports torch first (one HIP runtime shared with torch)."""
    pytest.importorskip("torch")
    import os
    import subprocess
    import sys
    here = os.path.dirname(os.path.abspath(__file__))
    code = _DEVICE_SCRIPT.format(root=os.path.dirname(here), tests=here)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                       timeout=240)
    print(r.stdout[-2000:], r.stderr[-3000:])
    assert r.returncode == 0 and "device autograd ok" in r.stdout


def test_gpu_tensor_after_library_load_is_refused():
    torch = pytest.importorskip("torch")
    from gturbo_amd import _lib
    from gturbo_amd.adjoint import run
    _lib.load()
    if _lib.TORCH_FIRST:
        pytest.skip("torch was imported before the library in this process")
    p, w0, _ = _torch_case()
    with Adjoint(p, 2, kchk=1) as adj:

        class FakeCuda:  # a GPU tensor stand-in: the check precedes any use
            shape = (128, 128)
            dtype = torch.float64
            is_cuda = True

        with pytest.raises(_lib.GturboError, match="import torch first"):
            run(FakeCuda(), adj)


def test_adjoint_errors():
    from gturbo_amd._lib import GturboError
    p = params_for(64)
    with Adjoint(p, 4, kchk=2) as adj:
        with pytest.raises(GturboError, match="forward first"):
            adj.backward(np.zeros((64, 64)))
    with pytest.raises(GturboError, match="kchk"):
        with Adjoint(p, 4, kchk=5) as adj:
            adj.forward(np.zeros((64, 64)))
    with pytest.raises(GturboError, match="2/3"):
        with Adjoint(params_for(64, truncate=0), 4, kchk=2) as adj:
            adj.forward(np.zeros((64, 64)))
