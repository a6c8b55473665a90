"""Differentiation on the GPU: ysim_hip_adj_forward / ysim_hip_adj_backward.

Parity bar:
* forward positions/velocities bit-exact vs the FP32 CPU rollout (and hence vs
  the scene oracle, which that rollout reproduces bit for bit);
* dL/dlog_k, dL/dkd, dL/dx_0, dL/dv_0 bit-exact vs the FP32 CPU adjoint (same
  operation order);
* against the FP64 adjoint on the same piece of the map (FP64 replays the
  FP32 run's contacts and gates): relative norm error <= 1e-3 and per-spring
  error <= 1e-3 * max |g|;
* the FP64 adjoint against central finite differences (eps 1e-3 on log k,
  1e-3 relative, plus the FD rounding floor) on 64 springs of the C4 drape
  (128 x 128, 60 substeps, checkpoints every 10).

The torch.autograd.Function below is test harness only: it routes a torch
loss through the two C-ABI entry points.
"""
import numpy as np
import pytest

from oracle import oracle as O
from ysim_amd import scenes
from ysim_amd._lib import YsimError
from ysim_amd.hip import HipScene

pytestmark = pytest.mark.gpu

FD_EPS = 1e-3
FD_REL = 1e-3


def _loss(x, xT):
    return 0.5 * float(((np.asarray(x, np.float64) - xT) ** 2).sum())


def _fd_atol(L, eps):
    return 64 * np.finfo(np.float64).eps * max(L, 1e-30) / eps


def _run(ctx, n, steps=60, K=10):
    sc = scenes.config_c4(n=n, steps=steps, checkpoint_every=K)
    tt, t0 = scenes.c4_theta(sc)
    kd = float(np.float32(sc.params["kd"]))
    V = n * n
    # target drape from the hidden stiffness, on the GPU
    hs_t = HipScene(ctx, sc)
    hs_t.adj_forward(tt, kd, steps, K)
    xT = hs_t.state()[0][:V].astype(np.float64)
    hs = HipScene(ctx, sc)
    hs.adj_forward(t0, kd, steps, K)
    return sc, hs, tt, t0, kd, xT


@pytest.mark.parametrize("n,steps,K", [(24, 60, 10), (32, 37, 8), (128, 60, 10)])
def test_adjoint_bit_exact_vs_fp32_oracle(ctx, oracle_threads, n, steps, K):
    print(f"adjoint parity n={n} steps={steps} K={K}", flush=True)
    sc, hs, tt, t0, kd, xT = _run(ctx, n, steps, K)
    V = n * n
    xg, vg, _ = hs.state()
    R = O.AdjointRollout(sc, 32)
    xo, vo = R.forward(t0.astype(np.float64), kd, steps, K)
    assert np.array_equal(xg[:V], xo.astype(np.float32))
    assert np.array_equal(vg[:V], vo.astype(np.float32))
    assert R.contacts()[0] > 0
    assert hs.flags() == 0
    seed = (xg[:V].astype(np.float64) - xT).astype(np.float32)
    vseed = np.random.default_rng(1).standard_normal((V, 3)).astype(np.float32) * 1e-3
    g, gkd, gx, gv = hs.adj_backward(seed, vseed)
    go, gkdo, gxo, gvo = R.backward(seed.astype(np.float64), vseed.astype(np.float64))
    assert np.array_equal(g, go.astype(np.float32))
    assert np.float32(gkd) == np.float32(gkdo)
    assert np.array_equal(gx, gxo.astype(np.float32))
    assert np.array_equal(gv, gvo.astype(np.float32))
    assert np.abs(g).max() > 0
    # the scene is left where the forward ended
    xa, va, _ = hs.state()
    assert np.array_equal(xa, xg) and np.array_equal(va, vg)


def test_c4_fp64_and_fd(ctx, oracle_threads):
    n, steps, K = 128, 60, 10
    print("C4 128x128: GPU adjoint vs FP64 adjoint vs frozen FD", flush=True)
    sc, hs, tt, t0, kd, xT = _run(ctx, n, steps, K)
    V = n * n
    xg = hs.state()[0][:V].astype(np.float64)
    g32, gkd32, gx32, _ = hs.adj_backward((xg - xT).astype(np.float32))
    # FP64 on the GPU run's piece: replay FP32's contacts and gates
    R32 = O.AdjointRollout(sc, 32)
    x32, _ = R32.forward(t0.astype(np.float64), kd, steps, K, mode="record")
    assert np.array_equal(x32.astype(np.float32), xg.astype(np.float32))
    R64 = O.AdjointRollout(sc, 64)
    R64.adopt_record(R32)
    x64, _ = R64.forward(t0.astype(np.float64), kd, steps, K, mode="frozen")
    pos_err = np.abs(x64 - xg).max()
    print(f"FP32 positions vs FP64 (same piece): max abs {pos_err:.3e}", flush=True)
    assert pos_err < 1e-5
    seed = (xg - xT).astype(np.float32).astype(np.float64)
    g64, gkd64, gx64, _ = R64.backward(seed)
    rel = np.linalg.norm(g32 - g64) / np.linalg.norm(g64)
    print(f"dL/dlog_k FP32 vs FP64: rel norm {rel:.3e}; kd {gkd32:.6e} vs {gkd64:.6e}", flush=True)
    assert rel <= 1e-3
    assert np.abs(g32 - g64).max() <= 1e-3 * np.abs(g64).max()
    assert abs(gkd32 - gkd64) <= 1e-3 * abs(gkd64)
    assert np.linalg.norm(gx32 - gx64) <= 1e-3 * np.linalg.norm(gx64)
    # central FD of the FP64 frozen map on 64 random springs (32 uniform, 32
    # with an endpoint that touched the sphere), against the FP64 adjoint of
    # that same loss (seeded with the FP64 residual)
    g64, gkd64, _, _ = R64.backward(x64 - xT)
    rng = np.random.default_rng(64)
    hist = R64.contact_hist()
    touched = np.nonzero(hist)[0]
    S = g64.size
    N = n
    # struct-x springs (r, c)-(r, c+1) of touched vertices: index r*(N-1)+c
    near = np.unique([(v // N) * (N - 1) + (v % N) for v in touched if v % N < N - 1])
    sel = np.unique(np.concatenate([rng.choice(S, 32, replace=False), rng.choice(near, 32, replace=False)]))
    L = _loss(x64, xT)
    atol = _fd_atol(L, FD_EPS)
    th = t0.astype(np.float64)
    bad = []
    strong = 0
    for s in sel:
        tp = th.copy(); tp[s] += FD_EPS
        tm = th.copy(); tm[s] -= FD_EPS
        xp, _ = R64.forward(tp, kd, steps, K, mode="frozen")
        xm, _ = R64.forward(tm, kd, steps, K, mode="frozen")
        fd = (_loss(xp, xT) - _loss(xm, xT)) / (2 * FD_EPS)
        if abs(fd - g64[s]) > FD_REL * abs(fd) + atol:
            bad.append((int(s), float(g64[s]), fd))
        strong += abs(fd) > 1e3 * atol
    print(f"FD springs checked {sel.size}, above noise {strong}, failures {len(bad)}", flush=True)
    assert not bad, bad
    assert strong >= 16


def _drape_function(hs, kd, steps, K, V):
    """torch.autograd.Function over the two C-ABI entry points (CPU tensors
    staged through device arrays).  adj_forward runs from the scene's current
    state, so every forward first restores the initial state."""
    import torch
    hs.snapshot()

    class Drape(torch.autograd.Function):
        @staticmethod
        def forward(fctx, log_k):
            hs.restore()
            hs.adj_forward(log_k.detach().numpy(), kd, steps, K)
            return torch.from_numpy(hs.state()[0][:V].copy())

        @staticmethod
        def backward(fctx, gx):
            g, *_ = hs.adj_backward(gx.detach().contiguous().numpy())
            return torch.from_numpy(g)

    return Drape


def test_autograd_function_wrapper(ctx, oracle_threads):
    import torch
    n, steps, K = 32, 40, 10
    sc = scenes.config_c4(n=n, steps=steps, checkpoint_every=K)
    tt, t0 = scenes.c4_theta(sc)
    kd = float(np.float32(sc.params["kd"]))
    V = n * n
    hs_t = HipScene(ctx, sc)
    hs_t.adj_forward(tt, kd, steps, K)
    xT = torch.from_numpy(hs_t.state()[0][:V].copy())
    hs = HipScene(ctx, sc)
    Drape = _drape_function(hs, kd, steps, K, V)
    theta = torch.from_numpy(t0.copy()).requires_grad_(True)
    x = Drape.apply(theta)
    loss = 0.5 * ((x - xT) ** 2).sum()
    loss.backward()
    # the FP32 CPU adjoint of the same loss, bit for bit
    R = O.AdjointRollout(sc, 32)
    xo, _ = R.forward(t0.astype(np.float64), kd, steps, K)
    assert np.array_equal(x.detach().numpy(), xo.astype(np.float32))
    go, *_ = R.backward((x.detach() - xT).numpy().astype(np.float64))
    assert np.array_equal(theta.grad.numpy(), go.astype(np.float32))
    assert torch.isfinite(theta.grad).all() and theta.grad.abs().max() > 0
    # a few steps of gradient descent on log k reduce the drape loss
    opt = torch.optim.SGD([theta], lr=float(0.5 / theta.grad.abs().max()))
    losses = [float(loss.detach())]
    for _ in range(3):
        opt.step()
        opt.zero_grad()
        x = Drape.apply(theta)
        loss = 0.5 * ((x - xT) ** 2).sum()
        loss.backward()
        losses.append(float(loss.detach()))
    print("drape loss under SGD on log k:", losses, flush=True)
    assert losses[-1] < losses[0]


_TORCH_DEVICE_SCRIPT = r"""
import sys, numpy as np
import torch                      # loads torch's HIP runtime first: one runtime per process
torch.cuda.init()
sys.path.insert(0, sys.argv[1])
from ysim_amd import scenes
from ysim_amd.hip import HipContext, HipScene
n, steps, K = 24, 30, 10
sc = scenes.config_c4(n=n, steps=steps, checkpoint_every=K)
tt, t0 = scenes.c4_theta(sc)
kd = float(np.float32(sc.params["kd"]))
ctx = HipContext(0)
hs = HipScene(ctx, sc)
theta = torch.from_numpy(t0).cuda()
torch.cuda.synchronize()
hs.adj_forward_ptr(theta.data_ptr(), kd, steps, K)
x = hs.state()[0][: n * n]
seed = torch.from_numpy((x - 0.5).astype(np.float32)).cuda()
g = torch.empty(hs.num_springs(), dtype=torch.float32, device="cuda")
gkd = torch.empty(1, dtype=torch.float32, device="cuda")
torch.cuda.synchronize()
hs.adj_backward_ptr(seed.data_ptr(), None, g.data_ptr(), gkd.data_ptr())
ctx.sync()
ref = HipScene(ctx, sc)
ref.adj_forward(t0, kd, steps, K)
gr, gkdr, *_ = ref.adj_backward(x - 0.5)
assert np.array_equal(g.cpu().numpy(), gr), "device-pointer gradient differs"
assert np.float32(gkd.cpu().numpy()[0]) == np.float32(gkdr)
print("torch device tensors OK", float(g.abs().max()))
"""


def test_torch_device_pointers_subprocess(tmp_path):
    """The *_ptr entry points on torch CUDA tensors, in a fresh process that
    imports torch before the library (so one HIP runtime is loaded)."""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    script = tmp_path / "torch_dev.py"
    script.write_text(_TORCH_DEVICE_SCRIPT)
    print("subprocess: torch device-pointer adjoint", flush=True)
    r = subprocess.run([sys.executable, str(script), root], capture_output=True, text=True, timeout=150)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    assert "torch device tensors OK" in r.stdout


def test_graph_after_adjoint_uses_new_stiffness(ctx, oracle_threads):
    """scene_step through the captured graph after adj_forward runs with the
    per-spring stiffness and damping adj_forward set (the graph is re-captured)."""
    n = 24
    sc = scenes.config_c4(n=n)
    tt, t0 = scenes.c4_theta(sc)
    kd = 0.05
    a, b = HipScene(ctx, sc), HipScene(ctx, sc)
    a.step(4, graph=True)   # capture with the family constants
    b.step(4, graph=False)
    a.adj_forward(tt, kd, 6, 3)
    b.adj_forward(tt, kd, 6, 3)
    a.step(10, graph=True)
    b.step(10, graph=False)
    assert np.array_equal(a.state()[0], b.state()[0])


def test_adjoint_rejects_self_collision(ctx):
    sc = scenes.config_c4(n=16)
    sc.params["self_collision"] = 1
    hs = HipScene(ctx, sc)
    with pytest.raises(YsimError):
        hs.num_springs()
    sc2 = scenes.config_c4(n=16)
    hs2 = HipScene(ctx, sc2)
    with pytest.raises(YsimError):
        hs2.adj_backward(np.zeros((256, 3), np.float32))   # no forward yet
