"""The learned list-length predictor (gt_group_config.predict = 1): after a
few evaluations gather samples, the MLP is trained on the GPU and sizes every
group's slots.  The lists (and so the forces) must not depend on how the
slots were sized: predicted slots, uniform slots and the plain two-pass solver
give bit-identical states, eager or from graphs; the predictor's error and
overflow rate are reported."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def run(hip, pos, vel, mass, steps, use_graph=False, **cfg):
    hip.state_set(pos, vel, mass)
    hip.prime(**cfg)
    hip.steps(steps, 0.01, use_graph=use_graph, **cfg)
    return hip.state_get()


def moving_system(oracle, n, seed):
    pos, mass = oracle.test_system(1, n, seed)
    rng = np.random.default_rng(seed)
    return pos, 0.5 * rng.standard_normal(pos.shape), mass


@pytest.mark.parametrize("exact", [True, False])
def test_predicted_slots_change_nothing(hip, oracle, exact):
    pos, vel, mass = moving_system(oracle, 30000, 71)
    cfg = dict(theta=0.6, eps=0.01, group_size=32, exact=exact)
    want = run(hip, pos, vel, mass, 12, predict=0, **cfg)
    two_pass = run(hip, pos, vel, mass, 12, predict=2, **cfg)
    hip.predictor_reset(4)  # learn from the first 4 evaluations
    got = run(hip, pos, vel, mass, 12, predict=1, **cfg)
    st = hip.predictor_stats()
    print(st)
    assert st["trained"] == 1 and st["samples"] > 0 and st["predicted"] >= 8
    assert st["mean_abs_log_error"] < 0.5
    assert st["list_entries"] <= st["slot_entries"] or st["overflows"] > 0
    for a, b, c in zip(got, want, two_pass):
        assert np.array_equal(bits(a), bits(b)) and np.array_equal(bits(c), bits(b))


def test_predicted_slots_from_graphs(hip, oracle):
    pos, vel, mass = moving_system(oracle, 20000, 72)
    cfg = dict(theta=0.7, eps=0.01, group_size=32, predict=1)
    hip.predictor_reset(3)
    eager = run(hip, pos, vel, mass, 8, **cfg)
    hip.predictor_reset(3)
    graph = run(hip, pos, vel, mass, 8, use_graph=True, **cfg)
    st = hip.predictor_stats()
    print(st)
    assert st["trained"] == 1
    for a, b in zip(graph, eager):
        assert np.array_equal(bits(a), bits(b))


def test_predictor_is_deterministic(hip, oracle):
    pos, vel, mass = moving_system(oracle, 20000, 73)
    cfg = dict(theta=0.7, eps=0.01, group_size=32, predict=1)
    stats = []
    for _ in range(2):
        hip.predictor_reset(3)
        run(hip, pos, vel, mass, 5, **cfg)
        stats.append(hip.predictor_stats())
    assert stats[0] == stats[1]


# ---- the network's kernels against the CPU restatement (oracle/mlp_oracle.py) ----

def mlp_case(seed, n=5000):
    import mlp_oracle as M
    rng = np.random.default_rng(seed)
    p = np.concatenate([rng.normal(0, 1 / np.sqrt(M.F), M.H * M.F), rng.normal(0, 0.1, M.H),
                        rng.normal(0, 1 / np.sqrt(M.H), M.H * M.H), rng.normal(0, 0.1, M.H),
                        rng.normal(0, 1 / np.sqrt(M.H), M.O * M.H), rng.normal(0, 0.1, M.O)])
    p = p.astype(np.float32)
    # raw features on their own scales and offsets, targets log(1 + length)
    feat = rng.normal(size=(n, M.F)) * rng.uniform(0.5, 4, M.F) + rng.uniform(-3, 3, M.F)
    targ = rng.uniform(3.0, 8.0, (n, M.O))
    samples = np.concatenate([feat, targ], 1).astype(np.float32)
    s64 = samples.astype(np.float64)
    stats = np.concatenate([s64.mean(0), s64.std(0)]).astype(np.float32)
    return p, samples, stats


def test_forward_kernel_equals_the_oracle(hip):
    import mlp_oracle as M
    p, samples, stats = mlp_case(11)
    x, _ = M.standardise(samples, stats)
    got = hip.predictor_forward(x.astype(np.float32), p)
    want = M.forward(p.astype(np.float64), x.astype(np.float32).astype(np.float64))[0]
    err = np.abs(got - want).max() / np.abs(want).max()
    print(f"forward: max difference {err:.2e} of the largest output", flush=True)
    assert err <= 1e-5


@pytest.mark.parametrize("iteration", [0, 5])
def test_gradient_and_adam_kernels_equal_the_oracle(hip, iteration):
    """mlp_gradients: each workgroup's gradient of its 64 samples' share of the
    minibatch loss against the oracle's analytic backprop (itself checked
    against finite differences on the CPU), relative 1e-4; the squared-error
    sums; and adam_step's update from the gradients added in workgroup order."""
    import mlp_oracle as M
    p, samples, stats = mlp_case(12 + iteration)
    rng = np.random.default_rng(iteration)
    m = (rng.normal(size=M.NPARAMS) * 1e-3 * (iteration > 0)).astype(np.float32)
    v = (rng.uniform(0, 1e-5, M.NPARAMS) * (iteration > 0)).astype(np.float32)
    blocks, lr = 8, 3e-3
    p1, m1, v1, grad, loss = hip.predictor_train_step(samples, stats, p, m, v, blocks,
                                                      iteration, lr)
    idx = M.minibatch(len(samples), blocks, iteration)
    x, t = M.standardise(samples, stats)
    batch = blocks * M.BLOCK
    worst = 0.0
    for b in range(blocks):
        rows = idx[b]
        want = M.gradient(p.astype(np.float64), x[rows], t[rows], batch)
        rel = np.linalg.norm(grad[b] - want) / np.linalg.norm(want)
        worst = max(worst, rel)
        assert rel <= 1e-4
        assert np.all(np.abs(grad[b] - want) <= 1e-4 * np.abs(want).max() + 1e-4 * np.abs(want))
        y = M.forward(p.astype(np.float64), x[rows])[0]
        assert abs(loss[b] - np.sum((y - t[rows]) ** 2)) <= 1e-5 * np.sum((y - t[rows]) ** 2)
    print(f"iteration {iteration}: worst workgroup gradient relative difference {worst:.2e}",
          flush=True)
    g = np.zeros(M.NPARAMS, np.float32)
    for b in range(blocks):  # added in workgroup order, as adam_step does
        g += grad[b]
    pw, mw, vw = M.adam(p, m, v, g.astype(np.float64), iteration, lr)
    # FP32 rounding of each term: bounded by the operands' size (0.9 m and
    # 0.1 g may nearly cancel)
    g64 = g.astype(np.float64)
    assert np.all(np.abs(m1 - mw) <= 1e-6 * (0.9 * np.abs(m) + 0.1 * np.abs(g64)) + 1e-15)
    assert np.all(np.abs(v1 - vw) <= 1e-6 * (0.999 * np.abs(v) + 0.001 * g64 * g64) + 1e-20)
    assert np.abs(p1 - pw).max() <= 1e-3 * lr + 1e-6 * np.abs(pw).max()
