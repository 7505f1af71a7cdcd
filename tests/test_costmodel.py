"""Cost-model MLP: the numpy restatement pinned by finite differences (CPU),
the GPU forward/backward kernels against it, and training behaviour (GPU)."""
import numpy as np
import pytest

from hisvsim_amd.circuit import random_layered
from hisvsim_amd.costmodel import N_FEATURES, part_features, candidate_plans, fallback_plan
from hisvsim_amd.partition import check_parts, greedy_parts
from oracle import mlp_ref

IN, H = 6, 16


def rand_params(rng, n_in=IN, h=H):
    P = n_in * h + h + h * h + h + h + 1
    return rng.normal(0, 0.4, P)


def test_reference_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = rand_params(rng)
    X = rng.normal(size=(20, IN))
    y = rng.normal(size=20)
    _, g = mlp_ref.loss_and_grad(p, X, y, IN, H)
    eps = 1e-6
    for i in rng.choice(len(p), 40, replace=False):
        d = np.zeros_like(p)
        d[i] = eps
        lp, _ = mlp_ref.loss_and_grad(p + d, X, y, IN, H)
        lm, _ = mlp_ref.loss_and_grad(p - d, X, y, IN, H)
        assert abs((lp - lm) / (2 * eps) - g[i]) <= 1e-6 * max(1.0, abs(g[i]))


def test_features_and_candidates():
    n = 24
    gates = random_layered(n, 6, seed=2)
    parts = greedy_parts(gates, 8)
    f = part_features(parts[0], n, "part")
    assert f.shape == (N_FEATURES,) and np.all(np.isfinite(f))
    plans = candidate_plans(gates, n)
    assert {p["mode"] for p in plans} == {"part", "fused64", "fusedbf16", "chain"}
    assert {p["partitioner"] for p in plans} == {"greedy", "nat", "dfs", "closure", "chain"}
    for p in plans:
        check_parts(gates, p["parts"])
        if p["mode"] == "chain":
            f = part_features(p["parts"][1], n, "chain", p["lows"][1])
            assert f[1] * 10 in (12, 13) and f[-1] > 0   # tile width, fused steps
    check_parts(gates, fallback_plan(gates, n)["parts"], 10)


@pytest.mark.gpu
def test_gpu_forward_and_gradient_match_reference():
    from hisvsim_amd.costmodel import MLP
    rng = np.random.default_rng(1)
    m = MLP(IN, H, seed=3)
    p = m.params().astype(np.float64)
    X = rng.normal(size=(300, IN)).astype(np.float32)
    y = rng.normal(size=300).astype(np.float32)
    want, _ = mlp_ref.forward(p, X, IN, H)
    assert np.abs(m.predict(X) - want).max() <= 1e-4 * max(1.0, np.abs(want).max())
    loss, g = m.gradient(X, y)
    rl, rg = mlp_ref.loss_and_grad(p, X, y, IN, H)
    assert abs(loss - rl) <= 1e-4 * rl
    assert np.abs(g - rg).max() <= 1e-4 * max(1.0, np.abs(rg).max())
    m.close()


@pytest.mark.gpu
def test_gpu_training_fits_and_is_deterministic():
    from hisvsim_amd.costmodel import MLP
    rng = np.random.default_rng(2)
    X = rng.uniform(-1, 1, size=(4096, N_FEATURES)).astype(np.float32)
    y = (np.sin(2 * X[:, 0]) + X[:, 1] * X[:, 2] + 0.5 * X[:, 3]).astype(np.float32)
    a, b = MLP(seed=5), MLP(seed=5)
    l0, _ = a.gradient(X, y)
    la = a.train(X, y, epochs=60, batch=256, lr=3e-3)
    lb = b.train(X, y, epochs=60, batch=256, lr=3e-3)
    assert la < 0.1 * l0
    assert la == lb and np.array_equal(a.params(), b.params())
    a.close()
    b.close()


def test_committed_model_matches_features():
    """The shipped parameters are for this feature vector and network."""
    import json
    from hisvsim_amd.costmodel import PARAMS_FILE
    d = json.load(open(PARAMS_FILE))
    assert d["in"] == N_FEATURES
    h = d["hidden"]
    assert len(d["params"]) == N_FEATURES * h + h + h * h + h + h + 1


@pytest.mark.gpu
def test_trained_model_beats_plain_partitioner():
    """What the learned part buys: on a 28-qubit random layered circuit the
    plan the committed model picks runs faster than the plain partitioner
    (natural order, k = 10), and the two plans give the same state to the
    picked plan's precision."""
    from hisvsim_amd import capi
    from hisvsim_amd.costmodel import MLP, choose_plan, fallback_plan, run_plan
    n = 28
    gates = random_layered(n, 30, seed=11)
    model = MLP.load()
    best, plans = choose_plan(model, gates, n)
    fb = fallback_plan(gates, n)
    with capi.State(n) as a, capi.State(n) as b:
        run_plan(a, best)      # warm-up (code objects, scratch)
        a.reset()
        t_model = float(run_plan(a, best).sum())
        t_fb = float(run_plan(b, fb).sum())
        infid = 1 - abs(a.inner(b)) ** 2
    model.close()
    assert t_model < 0.8 * t_fb, (best["mode"], best["k"], best["partitioner"], t_model, t_fb)
    assert infid <= (1e-12 if best["mode"] not in ("fusedbf16", "chain") else 1e-6)


@pytest.mark.gpu
@pytest.mark.parametrize("family", ["rl32", "qft30"])
def test_cost_model_choice_measured(family):
    """The committed model (trained on 2000 random layered circuits and 400
    controlled-phase circuits of 20-30 qubits) on circuits outside its
    training range: a 32-qubit random layered circuit of depth 30 and a
    30-qubit QFT. Its pick, its next two candidates, the fixed default
    (chain, k = 13) and the plain partitioner (natural order, k = 10) are
    each run on a dense state and timed. On the random layered circuit (the
    family it was trained on, at a larger n) the pick must be within 10% of
    the fastest of them (regret), no slower than the default by more than
    5%, and at least 3x faster than the plain partitioner; its plan-level
    prediction within 35% of its measured time (the absolute level moves
    from one GPU to another: the pick of this circuit measured 1.00 s on
    two GPUs and 0.86 s on a third, against 1.11 s predicted), and the
    ratio of predicted to measured time the same within 10% over the
    split-precision plans it ran, which is what the ranking rests on. The
    QFT (which the training covers only through QFT prefixes and random
    controlled-phase layers) holds looser bars: regret within 25%, within
    15% of the default, 1.5x over the plain partitioner, and no bar on the
    prediction error."""
    from hisvsim_amd import capi
    from hisvsim_amd.circuit import qft
    from hisvsim_amd.costmodel import MLP, compare_choice
    from hisvsim_amd.train_costmodel import dense_start
    if family == "rl32":
        n, gates = 32, random_layered(32, 30, seed=21)
        regret, plan_err, vs_default, plain_gain = 1.10, 0.35, 0.95, 3.0
    else:
        n, gates = 30, qft(30)
        regret, plan_err, vs_default, plain_gain = 1.25, None, 0.85, 1.5
    model = MLP.load()
    with capi.State(n) as s:
        dense_start(s)
        res = compare_choice(model, gates, n, s, top=3)
    model.close()
    print(f"{family}: pick {res['pick']}; measured {res['measured_ms']}; predicted {res['predicted_ms']}", flush=True)
    assert res["regret"] <= regret, res
    assert plan_err is None or res["plan_rel_err"] <= plan_err, res
    if plan_err is not None:
        ratios = [res["predicted_ms"][k] / res["measured_ms"][k] for k in res["measured_ms"] if not k.startswith("part ")]
        assert len(ratios) >= 2 and max(ratios) / min(ratios) <= 1.10, res
    assert res["gain_over_default"] >= vs_default, res
    assert res["gain_over_plain"] >= plain_gain, res


@pytest.mark.gpu
@pytest.mark.parametrize("family", ["bv28", "ghz30", "rl28_depth3"])
def test_cost_model_beats_the_default_where_it_differs(family):
    """Circuits the default plan (the moving-layout chain at k = 13) is not
    built for: a Bernstein-Vazirani circuit, a GHZ chain and a 3-layer random
    layered circuit. The committed model picks another plan there (fixed-
    layout wide parts), that plan runs at least 10% faster than the default,
    and nothing it ranked next or the plain partitioner beats it by more than
    5% (regret)."""
    from hisvsim_amd import capi
    from hisvsim_amd.costmodel import MLP, compare_choice, plan_name, default_plan
    from hisvsim_amd.train_costmodel import dense_start, family_circuits
    name, n, gates = next(f for f in family_circuits() if f[0] == family)
    model = MLP.load()
    with capi.State(n) as s:
        dense_start(s)
        res = compare_choice(model, gates, n, s)
    model.close()
    print(family, res["pick"], res["regret"], res["gain_over_default"], flush=True)
    assert res["pick"] != plan_name(default_plan(gates, n))
    assert res["gain_over_default"] >= 1.10, res
    assert res["regret"] <= 1.05, res
