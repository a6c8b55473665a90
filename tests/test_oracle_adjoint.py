"""Pins of the CPU adjoint oracle (oracle/ysim_oracle_adjoint.cpp).

The reference has no differentiation, so there are no reference vectors for
this path; the oracle is pinned by mathematics and by the (already pinned)
forward oracle:

* FP64 adjoint vs central finite differences (eps 1e-3 on log k, 1e-3
  relative) of the same map with each step's contact set and gates frozen.
  The frozen map is the one the adjoint differentiates.  The unfrozen map has
  jumps: at resting contact the gate `distance < thickness` compares values
  that are equal up to rounding (a pushed-out vertex sits at exactly
  thickness), so its velocity projection flips on rounding noise and no FD step
  resolves a derivative across it.
* dL/dx_0 and dL/dv_0 by directional finite differences of the frozen map.
* The FP32 rollout (the device's operation order) reproduces the scene oracle's
  substep bit for bit when handed the same per-spring stiffness.
* FP32 vs FP64 on the same piece of the map (FP64 replays FP32's contacts).
"""
import dataclasses

import numpy as np
import pytest

from ysim_amd import scenes
from oracle import oracle as O

FD_EPS = 1e-3
FD_REL = 1e-3


def _fd_atol(L, eps):
    # rounding noise of a central difference of an FP64 loss of size L
    return 64 * np.finfo(np.float64).eps * max(L, 1e-30) / eps


def _loss(x, xT):
    return 0.5 * float(((x - xT) ** 2).sum())


def _setup(n, steps=60, K=10, precision=64, kd=None):
    sc = scenes.config_c4(n=n, steps=steps, checkpoint_every=K)
    tt, t0 = scenes.c4_theta(sc)
    kd = sc.params["kd"] if kd is None else kd
    R = O.AdjointRollout(sc, precision)
    xT, _ = R.forward(tt, kd, steps, K)
    x0, _ = R.forward(t0, kd, steps, K, mode="record")
    return sc, R, t0.astype(np.float64), kd, xT, x0


def _pick_springs(R, n, count, seed=0):
    """Half uniform, half among springs with an endpoint that touched the sphere
    (chosen from the forward's contacts, not from the gradient)."""
    rng = np.random.default_rng(seed)
    hist = R.contact_hist()
    N = n
    touched = set(np.nonzero(hist)[0].tolist())
    # spring endpoints, same numbering as the oracle's
    SY = N * (N - 1); SA = SY + (N - 1) * N; SB = SA + (N - 1) ** 2; BX = SB + (N - 1) ** 2; BY = BX + N * (N - 2)

    def ends(s):
        if s < SY: r, c = divmod(s, N - 1); return r * N + c, r * N + c + 1
        if s < SA: q = s - SY; r, c = divmod(q, N); return r * N + c, (r + 1) * N + c
        if s < SB: q = s - SA; r, c = divmod(q, N - 1); return r * N + c, (r + 1) * N + c + 1
        if s < BX: q = s - SB; r, c = divmod(q, N - 1); return r * N + c + 1, (r + 1) * N + c
        if s < BY: q = s - BX; r, c = divmod(q, N - 2); return r * N + c, r * N + c + 2
        q = s - BY; r, c = divmod(q, N); return r * N + c, (r + 2) * N + c

    S = R.S
    cand = [s for s in range(S) if ends(s)[0] in touched or ends(s)[1] in touched]
    half = count // 2
    a = rng.choice(S, half, replace=False)
    b = rng.choice(cand, count - half, replace=False) if len(cand) >= count - half else np.array(cand)
    return np.unique(np.concatenate([a, b]).astype(np.int64))


# kd = 20 makes the damping terms of the spring Jacobian large enough for the
# FD tolerance to see (at the scene's kd = 0.01 they are below 1e-3 relative)
@pytest.mark.parametrize("n,kd", [(16, None), (24, None), (16, 20.0)])
def test_fp64_adjoint_matches_frozen_fd(n, kd, oracle_threads):
    sc, R, t0, kd, xT, x0 = _setup(n, kd=kd)
    hist = R.contact_hist()
    assert hist.sum() > 0, "the drape must touch the sphere"
    g, gkd, _, _ = R.backward(x0 - xT)
    L = _loss(x0, xT)
    atol = _fd_atol(L, FD_EPS)
    sel = _pick_springs(R, n, 24)
    bad = []
    for s in sel:
        tp = t0.copy(); tp[s] += FD_EPS
        tm = t0.copy(); tm[s] -= FD_EPS
        xp, _ = R.forward(tp, kd, 60, 10, mode="frozen")
        xm, _ = R.forward(tm, kd, 60, 10, mode="frozen")
        fd = (_loss(xp, xT) - _loss(xm, xT)) / (2 * FD_EPS)
        if abs(fd - g[s]) > FD_REL * abs(fd) + atol:
            bad.append((int(s), g[s], fd))
    assert not bad, bad
    # at least a few of the checked springs carry a gradient well above the noise
    assert np.sum(np.abs(g[sel]) > 1e3 * atol) >= 4
    # damping: the step is scaled to kd's size
    e = 1e-3 * kd
    xp, _ = R.forward(t0, kd + e, 60, 10, mode="frozen")
    xm, _ = R.forward(t0, kd - e, 60, 10, mode="frozen")
    fd = (_loss(xp, xT) - _loss(xm, xT)) / (2 * e)
    assert abs(fd - gkd) <= FD_REL * abs(fd) + _fd_atol(L, e), (gkd, fd)


def test_fp64_adjoint_initial_state_directional_fd(oracle_threads):
    n = 16
    sc, R, t0, kd, xT, x0 = _setup(n)
    _, _, gx0, gv0 = R.backward(x0 - xT)
    base_x = sc.meshes[0].x.astype(np.float64)
    base_v = np.zeros_like(base_x)
    rng = np.random.default_rng(3)
    for which, grad in (("x", gx0), ("v", gv0)):
        d = rng.standard_normal(base_x.shape)
        d /= np.linalg.norm(d)
        e = 1e-5
        Ls = []
        for sgn in (1, -1):
            if which == "x":
                R.set_initial(x0=base_x + sgn * e * d, v0=base_v)
            else:
                R.set_initial(x0=base_x, v0=base_v + sgn * e * d)
            xs, _ = R.forward(t0, kd, 60, 10, mode="frozen")
            Ls.append(_loss(xs, xT))
        R.set_initial(x0=base_x, v0=base_v)
        fd = (Ls[0] - Ls[1]) / (2 * e)
        pred = float((grad * d).sum())
        assert abs(fd - pred) <= FD_REL * abs(fd) + _fd_atol(_loss(x0, xT), e), (which, fd, pred)
        assert abs(pred) > 1e3 * _fd_atol(_loss(x0, xT), e)


def test_fp32_rollout_matches_scene_oracle(oracle_threads):
    n = 24
    sc = scenes.config_c4(n=n)
    tt, _ = scenes.c4_theta(sc)
    kd = sc.params["kd"]
    R = O.AdjointRollout(sc, 32)
    x32, v32 = R.forward(tt.astype(np.float64), kd, 60, 10)
    k = np.exp(tt.astype(np.float64)).astype(np.float32)
    os_ = O.OracleScene(sc, kspring=k)
    os_.step(60)
    xs, vs, _ = os_.state()
    V = n * n
    assert np.array_equal(x32.astype(np.float32), xs[:V])
    assert np.array_equal(v32.astype(np.float32), vs[:V])
    assert R.contacts()[0] > 0


def test_fp32_vs_fp64_same_piece(oracle_threads):
    n = 32
    sc, R32, t0, kd, xT, x32 = _setup(n, precision=32)
    g32, gkd32, gx32, _ = R32.backward(x32 - xT)
    R64 = O.AdjointRollout(sc, 64)
    R64.adopt_record(R32)
    x64, _ = R64.forward(t0, kd, 60, 10, mode="frozen")
    assert np.abs(x32 - x64).max() < 1e-5
    g64, gkd64, gx64, _ = R64.backward(x32 - xT)
    assert np.linalg.norm(g32 - g64) <= 1e-3 * np.linalg.norm(g64)
    assert np.abs(g32 - g64).max() <= 1e-3 * np.abs(g64).max()
    assert abs(gkd32 - gkd64) <= 1e-3 * abs(gkd64)
    assert np.linalg.norm(gx32 - gx64) <= 1e-3 * np.linalg.norm(gx64)


@pytest.mark.parametrize("bits,what,kd", [(1, "contact push pull-back", None),
                                          (2, "velocity projection pull-back", None),
                                          (4, "spring damping Jacobian", 20.0),
                                          (8, "stiffness gradient sign", None)])
def test_planted_mistake_is_caught(bits, what, kd, oracle_threads):
    """Each term of the backward pass is pinned: a planted mistake in the
    oracle's backward (dropping one term) fails the frozen-FD check."""
    n = 16
    sc, R, t0, kd, xT, x0 = _setup(n, kd=kd)
    R.plant(bits)
    g, gkd, gx0, _ = R.backward(x0 - xT)
    R.plant(0)
    L = _loss(x0, xT)
    sel = _pick_springs(R, n, 24)
    worst = 0.0
    for s in sel:
        tp = t0.copy(); tp[s] += FD_EPS
        tm = t0.copy(); tm[s] -= FD_EPS
        xp, _ = R.forward(tp, kd, 60, 10, mode="frozen")
        xm, _ = R.forward(tm, kd, 60, 10, mode="frozen")
        fd = (_loss(xp, xT) - _loss(xm, xT)) / (2 * FD_EPS)
        worst = max(worst, abs(fd - g[s]) / (FD_REL * abs(fd) + _fd_atol(L, FD_EPS)))
    e = 1e-3 * kd
    xp, _ = R.forward(t0, kd + e, 60, 10, mode="frozen")
    xm, _ = R.forward(t0, kd - e, 60, 10, mode="frozen")
    fdk = (_loss(xp, xT) - _loss(xm, xT)) / (2 * e)
    worst = max(worst, abs(fdk - gkd) / (FD_REL * abs(fdk) + _fd_atol(L, e)))
    # directional x_0 check
    rng = np.random.default_rng(5)
    base_x = sc.meshes[0].x.astype(np.float64)
    d = rng.standard_normal(base_x.shape)
    d /= np.linalg.norm(d)
    Ls = []
    for sgn in (1, -1):
        R.set_initial(x0=base_x + sgn * 1e-5 * d)
        xs, _ = R.forward(t0, kd, 60, 10, mode="frozen")
        Ls.append(_loss(xs, xT))
    R.set_initial(x0=base_x)
    fdx = (Ls[0] - Ls[1]) / 2e-5
    pred = float((gx0 * d).sum())
    worst = max(worst, abs(fdx - pred) / (FD_REL * abs(fdx) + _fd_atol(L, 1e-5)))
    assert worst > 1.0, f"planted mistake ({what}) not detected"


@pytest.mark.parametrize("n", [16, 24])
def test_unfrozen_map_jumps_where_the_gradient_lives(n, oracle_threads):
    """Why the adjoint differentiates the frozen map (each step's contacts and
    gates held fixed), measured: on the C4 drape every spring among those
    touching the sphere whose gradient exceeds 1e-9 changes the contact
    signature (step, vertex, facet, projection gate) of the ordinary,
    detecting forward under a log-stiffness perturbation of only +-1e-9, so
    the unfrozen map jumps at theta_0 there and no finite difference resolves
    its derivative; springs that keep the signature under +-1e-3 are where
    both maps coincide, and there their central differences agree."""
    sc = scenes.config_c4(n=n, steps=60, checkpoint_every=10)
    tt, t0 = scenes.c4_theta(sc)
    t0 = t0.astype(np.float64)
    kd = sc.params["kd"]
    R, Rd = O.AdjointRollout(sc, 64), O.AdjointRollout(sc, 64)
    xT, _ = R.forward(tt, kd, 60, 10)
    x0, _ = R.forward(t0, kd, 60, 10, mode="record")
    c0, s0 = R.contacts()
    assert c0 > 0
    g, _, _, _ = R.backward(x0 - xT)
    springs = _pick_springs(R, n, 48, seed=n)
    relevant = [j for j in springs if abs(g[j]) > 1e-9]
    assert len(relevant) >= 8
    for j in relevant:
        tp, tm = t0.copy(), t0.copy()
        tp[j] += 1e-9
        tm[j] -= 1e-9
        Rd.forward(tp, kd, 60, 10, mode="detect")
        sp = Rd.contacts()
        Rd.forward(tm, kd, 60, 10, mode="detect")
        sm = Rd.contacts()
        assert sp != (c0, s0) or sm != (c0, s0), j
    # where the signature holds (20 steps: the first contacts), the maps agree
    steps, e = 20, FD_EPS
    xT, _ = R.forward(tt, kd, steps, 10)
    x0, _ = R.forward(t0, kd, steps, 10, mode="record")
    c0, s0 = R.contacts()
    L = _loss(x0, xT)
    same = 0
    for j in springs:
        tp, tm = t0.copy(), t0.copy()
        tp[j] += e
        tm[j] -= e
        xdp, _ = Rd.forward(tp, kd, steps, 10, mode="detect")
        sp = Rd.contacts()
        xdm, _ = Rd.forward(tm, kd, steps, 10, mode="detect")
        sm = Rd.contacts()
        if sp != (c0, s0) or sm != (c0, s0):
            continue
        same += 1
        xfp, _ = R.forward(tp, kd, steps, 10, mode="frozen")
        xfm, _ = R.forward(tm, kd, steps, 10, mode="frozen")
        fd_unfrozen = (_loss(xdp, xT) - _loss(xdm, xT)) / (2 * e)
        fd_frozen = (_loss(xfp, xT) - _loss(xfm, xT)) / (2 * e)
        assert abs(fd_unfrozen - fd_frozen) <= 2 * _fd_atol(L, e), (j, fd_unfrozen, fd_frozen)
    assert same >= 4, (same, len(springs))
