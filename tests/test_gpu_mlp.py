"""The bucket-width MLP: CPU double oracle pinned by finite differences, GPU
forward/backward kernels checked against it, GPU training, and the learned
policy inside the solver (scheduling only: answers stay bit-exact)."""
import numpy as np
import pytest

from oracle import mlp_oracle as MO
from oracle import oracle as O
from vine_amd import mlp as M
from vine_amd import solver as S
from vine_amd import workload as W


def _data(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 1, (n, M.NIN)).astype(np.float32) * 3 + 2
    t = np.stack([np.sin(x[:, 0]) + 0.3 * x[:, 1], 0.1 * x[:, 2] * x[:, 3]], 1).astype(np.float32)
    m = (rng.random((n, M.NOUT)) > 0.2).astype(np.float32)
    return x, t, m


def test_oracle_gradient_matches_finite_differences():
    """Pins the CPU oracle: analytic gradient vs central differences in double."""
    x, t, m = _data(40, 1)
    p = M.init_params(3, xmu=x.mean(0), xsd=x.std(0)).astype(np.float64)
    _, g = MO.gradient(p, x, t, m)
    rng = np.random.default_rng(0)
    idx = rng.choice(M.NTRAIN, 60, replace=False)
    eps = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (MO.loss(pp, x, t, m) - MO.loss(pm, x, t, m)) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-5 * max(1.0, abs(fd)), (i, fd, g[i])
    # a planted sign error in the oracle would be caught
    assert not np.allclose(-g[idx], g[idx], atol=1e-9)


@pytest.mark.gpu
def test_gpu_forward_and_gradient_match_oracle(gpu):
    x, t, m = _data(1000, 2)
    p = M.init_params(5, xmu=x.mean(0), xsd=x.std(0))
    net = M.MLP(p)
    y = net.forward(x)
    np.testing.assert_allclose(y, MO.forward(p, x), rtol=1e-4, atol=1e-4)
    g, loss = net.gradient(x, t, m)
    l_ref, g_ref = MO.gradient(p.astype(np.float64), x, t, m)
    assert abs(loss - l_ref) <= 1e-4 * max(1.0, abs(l_ref))
    # fp32 accumulation over 1,000 samples: relative 1e-3 of the gradient scale
    scale = np.abs(g_ref).max()
    np.testing.assert_allclose(g, g_ref, rtol=1e-3, atol=1e-4 * scale)


@pytest.mark.gpu
def test_gpu_training_reduces_loss(gpu):
    x, t, m = _data(20000, 4)
    net = M.MLP(M.init_params(7, xmu=x.mean(0), xsd=x.std(0)))
    _, l0 = net.gradient(x, t, m)
    curve = net.train(x, t, m, epochs=15, batch=256, lr=2e-3, seed=1)
    _, l1 = net.gradient(x, t, m)
    assert l1 < 0.5 * l0, (l0, l1, curve)
    assert np.all(np.isfinite(curve))


@pytest.mark.gpu
def test_set_params_rejects_bad_values(gpu):
    net = M.MLP()
    p = M.init_params(1)
    p[3] = np.nan
    with pytest.raises(RuntimeError):
        net.set_params(p)
    p = M.init_params(1)
    p[-1] = 0.0  # a zero standard deviation
    with pytest.raises(RuntimeError):
        net.set_params(p)


@pytest.mark.gpu
@pytest.mark.parametrize("engine", [S.ENGINE_ROUNDS, S.ENGINE_WORKGROUP])
def test_learned_policy_keeps_answers_exact(gpu, engine):
    """Whatever the network says (random weights here), windows only reorder
    work: answers equal the oracle's."""
    g, q = W.preset("grid64")
    ref = [O.cspp(g, q.src[i], q.dst[i], q.bound[i]) for i in range(len(q))]
    print("grid64 learned policy engine=%d" % engine, flush=True)
    for seed in (0, 1):
        p = M.init_params(seed)
        with S.Solver(g, max_batch=50, prune=S.PRUNE_BOUNDS, delta_policy=S.DELTA_MLP, delta_fixed=50.0,
                      record_trace=1, engine=engine) as s:
            s.set_mlp(p)
            a = s.solve(q.src, q.dst, q.bound)
            tr = s.trace()
        for i in range(len(q)):
            assert bool(a.feasible[i]) == ref[i].feasible
            if ref[i].feasible:
                assert (a.cost[i], a.delay[i]) == (np.float32(ref[i].cost), np.float32(ref[i].delay))
        assert tr.shape[1] == 10 and tr.shape[0] > 0 and np.all(np.isfinite(tr))


@pytest.mark.gpu
def test_mlp_policy_requires_params(gpu):
    g, q = W.preset("grid64")
    with S.Solver(g, max_batch=4, delta_policy=S.DELTA_MLP) as s:
        with pytest.raises(S.VineError):
            s.solve(q.src[:4], q.dst[:4], q.bound[:4])
