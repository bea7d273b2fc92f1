"""Learned warm start: the MFMA forward / backward kernels against the FP64
restatement (oracle/mlp_oracle.py), Adam and the batch sampler against numpy,
the feature kernel against a numpy restatement, the residual-gated fallback
against the plain solver (bit for bit), and what a trained network gains."""
import numpy as np
import pytest

from oracle import mlp_oracle as O
from libatssim_amd import GpuVBD, Params, configs
from libatssim_amd import _lib
from libatssim_amd.warm_start import (NUM_PARAMS, OFF_B1, OFF_B4, OFF_W1, OFF_W4, WarmStartMLP, evaluate,
                                      train_recipe, unpack_params)

pytestmark = pytest.mark.gpu

MASK = (1 << 64) - 1


def _case(name):
    print(f"[case] {name}", flush=True)


def _mix64(z):
    z = (z + 0x9E3779B97F4A7C15) & MASK
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
    return z ^ (z >> 31)


def _live_params(seed):
    """Initial parameters with the output layer scaled up to O(1) outputs."""
    rng = np.random.default_rng(seed)
    m = WarmStartMLP(seed)
    p = m.params()
    W4 = p[OFF_W4:OFF_B4].reshape(128, 16)
    W4[:, :3] *= 50.0
    p[OFF_B1:OFF_B1 + 128] = rng.standard_normal(128) * 0.1
    p[OFF_B4:OFF_B4 + 3] = rng.standard_normal(3) * 0.1
    m.set_params(p)
    return m, m.params()


@pytest.mark.parametrize("rows", [1, 77, 200, 1000])
def test_forward_matches_oracle(gpu_available, rows):
    _case(f"mlp forward, {rows} rows")
    m, p = _live_params(3)
    X = np.random.default_rng(rows).standard_normal((rows, 48)).astype(np.float32)
    y = m.forward(X)
    yo, _ = O.forward(p, X, bf16_round=True)
    scale = np.sqrt(np.mean(yo ** 2))
    assert scale > 0.05
    # FP32 accumulation order and the rare BF16 rounding flip of a hidden
    # activation (one BF16 ulp, 2^-8 relative) separate the two; a wrong
    # operand layout or a missing term is O(scale)
    err = np.abs(y - yo)
    assert err.max() <= 1e-2 * scale, err.max() / scale
    assert err.mean() <= 5e-4 * scale, err.mean() / scale


def test_padding_stays_zero_and_init_is_small(gpu_available):
    _case("mlp init layout")
    p = WarmStartMLP(5).params()
    W1, _ = unpack_params(p)[0]
    W4, b4 = unpack_params(p)[3]
    assert np.all(W1[48:] == 0)
    assert np.all(W4[:, 3:] == 0) and np.all(b4 == 0)
    assert 0 < np.max(np.abs(W4[:, :3])) < 0.01


@pytest.mark.parametrize("rows", [64, 333])
def test_gradient_matches_oracle(gpu_available, rows):
    _case(f"mlp gradient, {rows} rows")
    m, p = _live_params(4)
    rng = np.random.default_rng(100 + rows)
    X = rng.standard_normal((rows, 48)).astype(np.float32)
    T = rng.standard_normal((rows, 3)).astype(np.float32)
    g, loss = m.gradient(X, T)
    go, lo = O.gradient(p, X, T, bf16_round=True)
    assert abs(loss - lo) <= 1e-4 * lo
    for k, ((Wg, bg), (Wo, bo)) in enumerate(zip(unpack_params(g), unpack_params(go))):
        for a, b, name in ((Wg, Wo, "W"), (bg, bo, "b")):
            nb = np.linalg.norm(b)
            if nb == 0:
                assert np.all(a == 0), (k, name)
                continue
            rel = np.linalg.norm(a - b) / nb
            assert rel < 5e-3, (k, name, rel)


def test_adam_step_and_sampler_match_numpy(gpu_available):
    _case("mlp adam step + sampler")
    m, p = _live_params(6)
    rng = np.random.default_rng(9)
    n = 500
    X = rng.standard_normal((n, 48)).astype(np.float32)
    T = rng.standard_normal((n, 3)).astype(np.float32)
    m.append_rows(X, T)
    Xs, Ts = m.dataset(0, n)
    assert np.array_equal(Xs, O.bf16(X).astype(np.float32))
    assert np.array_equal(Ts, T)
    batch, seed, lr = 256, 42, 1e-3
    loss = m.train(1, batch, lr, seed)
    rs = _mix64((seed + 0) & MASK) | 1
    rows = np.array([_mix64((rs + r) & MASK) % n for r in range(batch)])
    g, lo = O.gradient(p, X[rows], T[rows], bf16_round=True)
    assert abs(loss - lo) <= 1e-4 * lo
    # first Adam step: m = 0.1 g, v = 0.001 g^2, bias corrections 0.1 / 0.001
    g32 = g.astype(np.float32)
    mhat = (np.float32(0.1) * g32) / np.float32(0.1)
    vhat = (np.float32(0.001) * g32 * g32) / np.float32(0.001)
    want = p - np.float32(lr) * mhat / (np.sqrt(vhat) + np.float32(1e-8))
    got = m.params()
    moved = np.abs(g) > 1e-2 * np.max(np.abs(g))
    # the step is ~lr * sign(g) where the gradient is not tiny
    assert np.max(np.abs(got - want)[moved]) <= 2e-6
    assert np.all(np.abs(got - p) <= 1.01 * lr + 1e-9)


def test_training_reduces_loss_and_is_deterministic(gpu_available):
    _case("mlp training on synthetic rows")
    rng = np.random.default_rng(11)
    X = rng.standard_normal((8192, 48)).astype(np.float32)
    T = (0.5 * np.tanh(X[:, :3] + 0.5 * X[:, 3:6] * X[:, 6:9])).astype(np.float32)
    runs = []
    for _ in range(2):
        m = WarmStartMLP(12)
        m.append_rows(X, T)
        l0 = m.train(10, 1024, 3e-3, 1)
        m.train(400, 1024, 3e-3, 2)
        l1 = m.train(10, 1024, 3e-3, 3)
        runs.append((l0, l1, m.params()))
    l0, l1, p = runs[0]
    assert l1 < 0.2 * l0, (l0, l1)
    assert np.array_equal(p, runs[1][2])


def _small_scene(n=24, iters=30, seed=5):
    cloth = configs.grid_cloth(n, "top", seed)
    wind = configs.make_wind(8, cloth.num_verts, 2.0, seed + 1)
    return cloth, Params(constraint_iter_count=iters), wind


def test_fallback_reproduces_plain_solver(gpu_available):
    _case("warm start fallback == plain solver")
    cloth, params, wind = _small_scene()
    m = WarmStartMLP(1)
    p = m.params()
    W4 = p[OFF_W4:OFF_B4].reshape(128, 16)
    W4[:, :3] = 1e3 * np.sign(W4[:, :3])       # garbage guesses, far off
    m.set_params(p)
    run = GpuVBD(cloth, params, _lib.LATS_STRATEGY_GRAPH)
    plain = GpuVBD(cloth, params, _lib.LATS_STRATEGY_GRAPH)
    m.bind(run)
    for f in range(4):
        run.set_external_force(wind[f])
        plain.set_external_force(wind[f])
        r = m.warm_start_frame(run, 0.0, params.constraint_iter_count, 4, 2, True)
        plain.physics_step()
        assert not r["kept"] and r["e_learned"] > r["e_inertial"]
        assert r["sweeps"] == params.constraint_iter_count
        xr, vr = run.state()
        xp, vp = plain.state()
        assert np.array_equal(xr, xp) and np.array_equal(vr, vp), f


def test_unbound_context_is_rejected(gpu_available):
    _case("warm start needs the bound context")
    cloth, params, _ = _small_scene()
    m = WarmStartMLP(1)
    a = GpuVBD(cloth, params)
    b = GpuVBD(cloth, params)
    m.bind(a)
    with pytest.raises(_lib.LatsError):
        m.record_frame(b)
    a.close()
    with pytest.raises(_lib.LatsError):
        m.warm_start_frame(b)


def _numpy_features(cloth, params, x0, v, corr1, corr2, wind_force):
    """FP64 restatement of the 48 feature columns (see warm_start.py)."""
    t = cloth.topology
    h = np.float64(np.float32(params.implicit_dt))
    g = np.array(params.gravity, np.float64)
    m = np.float64(cloth.mass[0])
    pinned = cloth.fixed.astype(bool)
    xp = x0 + h * (v + h * g)
    xp[pinned] = x0[pinned]
    V = len(x0)
    E = t.edges
    L = t.rest_length.astype(np.float64)
    s0 = np.float64(np.float32(L.mean()))
    ks = np.float64(np.float32(params.stiffness_stretch_spring))
    kb = np.float64(np.float32(params.stiffness_quadratic_bending))
    deg = np.zeros(V)
    rc1 = np.zeros((V, 3)); rc2 = np.zeros((V, 3)); rv = np.zeros((V, 3)); rx = np.zeros((V, 3))
    fs = np.zeros((V, 3)); strain = np.zeros(V); smax = np.zeros(V); tr = np.zeros(V)
    for e, (a, b) in enumerate(E):
        for s, o in ((a, b), (b, a)):
            deg[s] += 1
            rc1[s] += corr1[o]; rc2[s] += corr2[o]; rv[s] += v[o]; rx[s] += xp[o]
            d = xp[o] - xp[s]
            ln = np.linalg.norm(d)
            st = (ln - L[e]) / L[e]
            strain[s] += st
            smax[s] = max(smax[s], abs(st))
            fs[s] += ks * (ln - L[e]) * d / ln
            tr[s] += ks + 2 * ks * max(1 - L[e] / ln, 0.0)
    fb = np.zeros((V, 3)); hb = np.zeros(V); bdeg = np.zeros(V)
    Q = t.bending_Q.astype(np.float64)
    for b, ids in enumerate(t.bending):
        P = xp[ids]
        for o in range(4):
            fb[ids[o]] += -kb * (Q[b][o] @ P)
            hb[ids[o]] += kb * Q[b][o][o]
            bdeg[ids[o]] += 1
    colour = np.zeros(V, int)
    for c, vs in enumerate(t.clusters()):
        colour[vs] = c
    h2m = h * h / m
    F = np.zeros((V, 48))
    inv = 1.0 / deg
    F[:, 0:3] = v * h / s0
    F[:, 3:6] = corr1 / s0
    F[:, 6:9] = corr2 / s0
    F[:, 9:12] = rc1 * inv[:, None] / s0
    F[:, 12:15] = rc2 * inv[:, None] / s0
    F[:, 15:18] = rv * inv[:, None] * h / s0
    F[:, 18:21] = (rx * inv[:, None] - xp) / s0
    F[:, 21:24] = fs * h2m / s0
    F[:, 24:27] = fb * h2m / s0
    F[:, 27:30] = (m * g + wind_force) * h2m / s0
    F[:, 30] = pinned
    F[np.arange(V), 31 + colour] = 1.0
    F[:, 39] = strain * inv
    F[:, 40] = smax
    F[:, 41] = deg / 8
    F[:, 42] = bdeg / 16
    F[:, 43] = tr * h2m
    F[:, 44] = hb * h2m
    # 45-47: obstacle contacts, zero without an obstacle
    return F, xp


def test_features_match_numpy(gpu_available):
    _case("warm start features, two frames")
    cloth, params, wind = _small_scene(n=12, iters=20)
    m = WarmStartMLP(1)
    g = GpuVBD(cloth, params, _lib.LATS_STRATEGY_GRAPH)
    m.bind(g)
    V = cloth.num_verts
    perm = g.permutation()             # recorded rows are in slot order: slot s holds vertex perm[s]
    x0, v0 = g.state()
    corr1 = np.zeros((V, 3)); corr2 = np.zeros((V, 3))
    for f in range(3):
        g.set_external_force(wind[f])
        m.record_frame(g)
        X, _ = m.dataset(f * V, V)
        F, xp = _numpy_features(cloth, params, x0.astype(np.float64), v0.astype(np.float64), corr1, corr2,
                                wind[f].astype(np.float64))
        F = F[perm]
        rms = np.sqrt(np.mean(F ** 2, axis=0))
        err = np.abs(X - F)
        bad = err > 1e-2 * rms[None, :] + 1e-4 + 2 ** -7 * np.abs(F)
        assert not bad.any(), (f, np.argwhere(bad)[:5], X[bad][:5], F[bad][:5])
        x1, v1 = g.state()
        corr2 = corr1
        corr1 = x1.astype(np.float64) - xp       # vertex order
        x0, v0 = x1, v1


def test_trained_warm_start_needs_fewer_sweeps(gpu_available):
    """What the trained network gains: sweeps to the plain solver's
    100-iteration energy on frames of a scenario it was not trained on."""
    _case("trained warm start gain (48x48, 4 scenarios x 40 frames, 2 guided rounds)")
    n, iters = 48, 100
    params = Params(constraint_iter_count=iters)
    scen = [configs.warm_c3_scenario(k, n=n, frames=40) for k in range(4)]
    m = WarmStartMLP(2024)
    train_recipe(m, scen, params, 40, plain_steps=1000, guided_rounds=2, guided_steps=1000, batch_rows=4096)
    cloth, wind = configs.warm_c3_scenario(9, n=n, frames=30)
    learned = evaluate(m, cloth, params, wind, 30, use_model=True)
    plain = evaluate(m, cloth, params, wind, 30, use_model=False)
    sl = np.array([r["sweeps"] for r in learned[5:]])
    sp = np.array([r["sweeps"] for r in plain[5:]])
    print("sweeps learned", sl.tolist(), "plain", sp.tolist(), flush=True)
    assert all(r["e_final"] <= r["target"] for r in learned)
    assert all(r["e_final"] <= r["target"] for r in plain)
    assert sl.mean() <= 0.6 * sp.mean(), (sl.mean(), sp.mean())
    # the same frames on the patch strategy: identical sweeps give identical
    # energies, hence the same decisions and sweep counts, frame for frame

    def patch(g):
        g.set_patches(16, 4, 512)
        g.set_strategy(_lib.LATS_STRATEGY_PATCH)

    learned_p = evaluate(m, cloth, params, wind, 30, use_model=True, setup=patch)
    for a, b in zip(learned, learned_p):
        assert (a["sweeps"], a["kept"], a["e_final"]) == (b["sweeps"], b["kept"], b["e_final"])


def test_guided_recording_is_deterministic(gpu_available):
    """Regression: the guided rows' targets come from a second context whose
    frame-start state is copied in; that copy must be ordered with the second
    context's stream.  Two recordings with the same network give the same
    sweeps, energies and rows."""
    from libatssim_amd.warm_start import record_guided, record_scenarios, _train_chunks
    _case("guided recording determinism (128x128, 8 frames, twice)")
    params = Params(constraint_iter_count=100)
    scen = [configs.warm_c3_scenario(0, n=128, frames=8)]
    m = WarmStartMLP(2024)
    record_scenarios(m, scen, params, 8)
    m.fit_feature_scale()
    _train_chunks(m, 100, 8192, 1e-3, 0)
    n0 = m.dataset_rows()
    runs = []
    for _ in range(2):
        start = m.dataset_rows()
        stats = record_guided(m, scen, params, 8)
        X, T = m.dataset(start, m.dataset_rows() - start)
        runs.append((stats, X, T))
    (s0, X0, T0), (s1, X1, T1) = runs
    assert m.dataset_rows() == n0 + 2 * 8 * 128 * 128
    assert [(r["sweeps"], r["kept"], r["e_final"], r["target"]) for r in s0] == \
           [(r["sweeps"], r["kept"], r["e_final"], r["target"]) for r in s1]
    assert np.array_equal(X0, X1) and np.array_equal(T0, T1)


def test_stop_rule_fallback_is_the_plain_frame(gpu_available):
    """With the target-free rule on, a rejected guess runs the plain frame."""
    _case("stop rule: rejected guesses run the plain frame")
    cloth, params, wind = _small_scene()
    m = WarmStartMLP(1)
    p = m.params()
    W4 = p[OFF_W4:OFF_B4].reshape(128, 16)
    W4[:, :3] = 1e3 * np.sign(W4[:, :3])
    m.set_params(p)
    m.set_stop_rule(1e-5)
    run = GpuVBD(cloth, params, _lib.LATS_STRATEGY_GRAPH)
    plain = GpuVBD(cloth, params, _lib.LATS_STRATEGY_GRAPH)
    m.bind(run)
    for f in range(3):
        run.set_external_force(wind[f])
        plain.set_external_force(wind[f])
        r = m.warm_start_frame(run, 0.0, 3 * params.constraint_iter_count, 4, 2, True)
        plain.physics_step()
        assert not r["kept"] and r["sweeps"] == params.constraint_iter_count
        assert np.array_equal(run.state()[0], plain.state()[0]), f


def test_stop_rule_needs_no_target(gpu_available):
    """A trained network with the rule on: frames given no target stop
    between the first check and the cap (rejected guesses: the plain count),
    and from the same frame start a looser tolerance stops no later (its
    condition holds wherever the tighter one does)."""
    _case("stop rule without a target (24x24, 3 scenarios x 30 frames)")
    from libatssim_amd.warm_start import evaluate_rule
    n, iters = 24, 60
    params = Params(constraint_iter_count=iters)
    scen = [configs.warm_c3_scenario(k, n=n, frames=30) for k in range(3)]
    m = WarmStartMLP(2024)
    train_recipe(m, scen, params, 30, plain_steps=800, guided_rounds=1, guided_steps=800, batch_rows=4096)
    cloth, wind = configs.warm_c3_scenario(7, n=n, frames=20)
    loose = evaluate_rule(m, cloth, params, wind, 20, 1e-3, max_sweeps=150)
    tight = evaluate_rule(m, cloth, params, wind, 20, 1e-7, max_sweeps=150)
    for res in (loose, tight):
        for r in res:
            assert r["sweeps"] <= 150
            if r["kept"]:
                assert r["sweeps"] >= 10
            else:
                assert r["sweeps"] == iters
    assert sum(r["kept"] for r in loose) > 0
    assert np.mean([r["sweeps"] for r in loose]) < 150
    assert loose[0]["sweeps"] <= tight[0]["sweeps"]
