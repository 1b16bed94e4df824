"""Learned warm start on the GPU: the BF16 matrix-core network against a NumPy
model of the same arithmetic (operands rounded to BF16 at every product,
accumulation in higher precision), the training loop, the per-vertex
features against a NumPy restatement, capture targets, and whole steps with
the warm start against plain steps (same converged positions)."""
import numpy as np
import pytest

from fastipc_amd import _lib
from fastipc_amd import scene as S
from fastipc_amd.newton import NewtonIPC
from fastipc_amd.warmstart import (DEPTH, N_IN, N_OUT, WIDTH, Capture, LearnedWarmStart, WarmStartMLP, bf16_round,
                                   scale_for)

pytestmark = pytest.mark.gpu


def _layers(w):
    dims = [N_IN] + [WIDTH] * DEPTH + [N_OUT]
    out, o = [], 0
    for fi, fo in zip(dims[:-1], dims[1:]):
        W = w[o:o + fi * fo].reshape(fi, fo).astype(np.float64)
        o += fi * fo
        b = w[o:o + fo].astype(np.float64)
        o += fo
        out.append((W, b))
    assert o == len(w)
    return out


def _bf(a):
    return bf16_round(np.asarray(a, np.float32)).astype(np.float64)


def _silu(z):
    return z / (1 + np.exp(-z))


def _ref_forward(w, X):
    H = X.astype(np.float64)
    Zs, Hs = [], [H]
    L = _layers(w)
    for l, (W, b) in enumerate(L):
        Z = _bf(H) @ _bf(W) + b
        Z = Z.astype(np.float32).astype(np.float64)
        Zs.append(Z)
        H = _silu(Z) if l < len(L) - 1 else Z
        H = H.astype(np.float32).astype(np.float64)
        Hs.append(H)
    return H, Zs, Hs


def _ref_grad(w, X, T):
    Y, Zs, Hs = _ref_forward(w, X)
    L = _layers(w)
    n = Y.size
    dZ = 2 * (Y - T) / n
    grads = []
    for l in range(len(L) - 1, -1, -1):
        W, _ = L[l]
        grads.append((_bf(Hs[l]).T @ _bf(dZ), dZ.sum(0)))
        if l > 0:
            z = Zs[l - 1]
            s = 1 / (1 + np.exp(-z))
            dZ = (_bf(dZ) @ _bf(W).T) * (s * (1 + z * (1 - s)))
    grads.reverse()
    return float(np.mean((Y - T) ** 2)), np.concatenate([np.concatenate([g.ravel(), b]) for g, b in grads])


def test_forward_matches_bf16_model():
    rng = np.random.default_rng(0)
    m = WarmStartMLP(seed=3)
    w = m.params()
    # larger output layer than the initial 0.1 scale so errors are visible
    w = (w * rng.uniform(0.5, 2.0, size=w.size)).astype(np.float32)
    m.set_params(w)
    for n in (1, 37, 64, 1000):
        X = rng.normal(size=(n, N_IN)).astype(np.float32)
        Y = m.forward(X)
        R, _, _ = _ref_forward(w, X)
        err = np.abs(Y - R).max() / max(np.abs(R).max(), 1e-6)
        assert err < 2e-3, (n, err)
    m.close()


def test_gradient_matches_bf16_model():
    rng = np.random.default_rng(1)
    m = WarmStartMLP(seed=4)
    w = m.params()
    X = rng.normal(size=(777, N_IN)).astype(np.float32)
    T = rng.normal(size=(777, N_OUT)).astype(np.float32)
    loss, g = m.grad(X, T)
    lr, gr = _ref_grad(w, X, T)
    assert loss == pytest.approx(lr, rel=1e-3)
    # per layer, relative to that layer's largest entry
    o = 0
    dims = [N_IN] + [WIDTH] * DEPTH + [N_OUT]
    for fi, fo in zip(dims[:-1], dims[1:]):
        for sz in (fi * fo, fo):
            a, b = g[o:o + sz], gr[o:o + sz]
            assert np.abs(a - b).max() <= 1e-2 * np.abs(b).max() + 1e-9, (fi, fo, sz)
            o += sz
    m.close()


def test_training_reduces_loss():
    rng = np.random.default_rng(2)
    m = WarmStartMLP(seed=5)
    X = rng.normal(size=(20000, N_IN)).astype(np.float32)
    T = np.stack([np.sin(X[:, 0]) + 0.5 * X[:, 1], X[:, 2] * X[:, 3], np.tanh(X[:, 4] - X[:, 5])], 1).astype(np.float32)
    m.dataset_add(X, T)
    assert m.dataset_size() == 20000
    Xb, Tb = m.dataset()
    assert np.array_equal(Xb, X) and np.array_equal(Tb, T)
    l0 = float(np.mean((m.forward(X) - T) ** 2))
    losses = m.train(400, batch=4096, lr=1e-3, seed=1)
    l1 = float(np.mean((m.forward(X) - T) ** 2))
    assert np.all(np.isfinite(losses))
    assert l1 < 0.1 * l0, (l0, l1, losses[:3], losses[-3:])
    m.close()


def test_save_load_round_trip(tmp_path):
    m = WarmStartMLP(seed=6)
    p = str(tmp_path / "w.npz")
    m.save(p, {"k": 1})
    m2 = WarmStartMLP.load(p)
    assert np.array_equal(m2.params(), bf16_round(m.params()))
    assert m2.meta == {"k": 1}
    m.close()
    m2.close()


def _features_ref(sim, sc, scale):
    """NumPy restatement of the 16 per-vertex features."""
    x, v = sim.x, sim.v
    dt, g = sc.dt, sc.gravity
    F = np.zeros((len(x), 16))
    F[:, 0:3] = v * dt / scale
    F[:, 7] = g * dt * dt / scale
    F[:, 9] = 1.0
    F[:, 13] = sc.dirichlet_mask
    return F


def test_features_and_capture_targets():
    sc = S.tiny_pair()
    sim = NewtonIPC(sc)
    for _ in range(2):
        sim.step()
    scale = scale_for(sc)
    sim.compute_xn_and_xTilde(sim.frame * sc.dt)
    sim.find_constraints()
    f = np.zeros((sim.nv, 16), np.float32)
    _lib.call("fipc_ws_features", sim.ctx, scale, _lib.fptr(f))
    R = _features_ref(sim, sc, scale)
    cols = [0, 1, 2, 6, 7, 8, 13]
    assert np.allclose(f[:, cols], R[:, cols], rtol=1e-5, atol=1e-6)
    # previous correction: after a step, x - xTilde of that step (zero on Dirichlet vertices)
    assert np.all(f[sc.dirichlet_mask, 3:6] == 0)
    assert np.all((f[:, 9] > 0) & (f[:, 9] <= 1))
    nrm = np.linalg.norm(f[:, 10:13], axis=1)
    assert np.all((np.abs(nrm - 1) < 1e-4) | (nrm == 0))
    assert np.all((f[:, 9] < 1) == (nrm > 0))
    assert np.all(f[f[:, 9] == 1, 14:16] == 0) and np.all(f[:, 15] >= 0)
    # capture: targets are the change of the correction,
    # ((x_{n+1} - xTilde) - (x_n - xTilde_{n-1})) / scale, of the captured vertices
    m = WarmStartMLP(seed=7)
    sim2 = NewtonIPC(sc)
    sim2.compute_xn_and_xTilde(0.0)
    xt1 = sim2.vector(3)
    sim2.step()
    dprev = sim2.x - xt1
    dprev[sc.dirichlet_mask] = 0.0
    cap = Capture(m, scale, nsample=100, seed=3)
    sim2.compute_xn_and_xTilde(sim2.frame * sc.dt)
    xt2 = sim2.vector(3)
    st = sim2.step(capture=cap)
    assert st.converged
    X, T = m.dataset()
    assert len(X) == 100
    d = (sim2.x - xt2 - dprev) / scale
    free = d[~sc.dirichlet_mask]
    assert np.abs(T).max() > 0
    for t in T[:20]:
        # every captured target is the value of one of the free vertices
        assert np.abs(free - t).max(axis=1).min() < 1e-5 * max(1.0, np.abs(t).max())
    m.close()


def _deformed(sc, amp, seed):
    rng = np.random.default_rng(seed)
    x = sc.x.copy()
    k = rng.normal(size=(3, 3)) * 3.0
    ph = rng.uniform(0, 2 * np.pi, size=3)
    free = ~sc.dirichlet_mask
    x[free] += amp * np.sin(x[free] @ k + ph)
    return x


@pytest.mark.parametrize("scene_fn,steps,amp", [(S.tiny_pair, 5, 0.001), (lambda: S.two_cubes(4), 10, 0.004)])
def test_warm_started_steps_match_plain(scene_fn, steps, amp):
    """With the Newton tolerance tightened, a warm-started step converges to
    the same positions as the plain step (the warm start only changes the
    starting point); every step stays intersection-free.  Both runs start from
    the same pre-strained state, away from F = I where the reference's 2x2
    make_pd jumps."""
    sc = scene_fn()
    sc.tol_rel *= 1e-3
    x0 = _deformed(sc, amp, 2)
    m = WarmStartMLP(seed=8)
    # train briefly on the plain trajectory so the prediction is not trivial
    cap_sim = NewtonIPC(sc, pcg_tol=1e-12)
    cap_sim.set_state(x0)
    cap = Capture(m, scale_for(sc), nsample=512, seed=1)
    for _ in range(steps):
        cap_sim.step(capture=cap)
    m.train(100, batch=1024, lr=1e-3, seed=2)
    a = NewtonIPC(sc, pcg_tol=1e-12)
    b = NewtonIPC(sc, pcg_tol=1e-12)
    a.set_state(x0)
    b.set_state(x0)
    hook = LearnedWarmStart(m, scale_for(sc))
    accepted = 0
    for k in range(steps):
        sa = a.step()
        sb = b.step(warm_start=hook)
        accepted += int(sb.warm_accepted)
        assert sa.converged and sb.converged
        assert b.check_collision() == 0
        err = np.abs(a.x - b.x).max()
        assert err <= 1e-6 * sc.bbox_diag, (k, err, sa.newton_iters, sb.newton_iters)
    print(f"warm start accepted in {accepted} of {steps} steps")
    m.close()


def test_capture_near_contacts():
    """With near_frac = 1 every captured row comes from a vertex within d-hat
    of a contact primitive (feature 9 below 1), when the step has any."""
    sc = S.two_cubes(4)
    sim = NewtonIPC(sc)
    for _ in range(3):
        sim.step()
    m = WarmStartMLP(seed=9)
    sim.compute_xn_and_xTilde(sim.frame * sc.dt)
    sim.find_constraints()
    f = np.zeros((sim.nv, 16), np.float32)
    _lib.call("fipc_ws_features", sim.ctx, scale_for(sc), _lib.fptr(f))
    near = (f[:, 9] < 1) & ~sc.dirichlet_mask
    cap = Capture(m, scale_for(sc), nsample=256, seed=4, near_frac=1.0)
    sim.step(capture=cap)
    X, _ = m.dataset()
    assert len(X) == 256
    if near.any():
        assert np.all(X[:, 9] < 1), (near.sum(), X[:, 9].max())
    print(f"{near.sum()} vertices near contacts")
    m.close()


@pytest.mark.parametrize("family,steps", [("slab", 30), ("cubes", 60)])
def test_shipped_weights_step_their_family(family, steps):
    """The shipped weights (fastipc_amd/weights/ws_<family>.npz) load, carry
    their provenance, and warm-start a held-out member of their scene family
    (seed 1000, the small size): every step ends intersection-free and
    converged (or stalled, see NewtonIPC.step), and the prediction is kept in
    some steps."""
    import os
    from fastipc_amd import warmstart as W
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    net = W.WarmStartMLP.load(os.path.join(root, "fastipc_amd", "weights", f"ws_{family}.npz"))
    assert net.meta.get("family") == family
    sc = W.family_scene(family, 1000, "small")
    sim = NewtonIPC(sc)
    hook = W.LearnedWarmStart(net, W.scale_for(sc))
    kept = 0
    for k in range(steps):
        st = sim.step(warm_start=hook)
        assert st.converged or st.stalled, k
        assert sim.check_collision() == 0, k
        kept += int(st.warm_accepted)
    print(f"{family}: warm start kept in {kept} of {steps} steps")
    assert kept > 0
    net.close()
