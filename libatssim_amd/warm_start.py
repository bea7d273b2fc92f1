"""Learned warm start of the VBD frame solve (lats_mlp_* of libatssim_hip).

The reference starts every frame's VBD iterations from the inertial guess
(Core::predict_position, src/Solver/xpbd_constraints.h:321-384).  Here a
per-vertex MLP (48 features -> 128 -> 128 -> 128 -> 3, SiLU, BF16 weights on
the matrix cores, FP32 accumulation and FP32 master weights) predicts the
frame's correction x_end - x_pred of every vertex from its 1-ring state; the
prediction is kept only when it lowers the frame's VBD energy, otherwise the
frame starts from the inertial guess as in the reference.  The network is
trained with Adam by the library's own forward / backward kernels on frames
that the plain solver produces (record_frame).

Feature columns (displacements in units of the mean rest length s0, forces
as the displacement h^2/m * f they cause over one step):
  0-2 velocity * h; 3-5 / 6-8 own correction of the last / second-last
  frame; 9-11 / 12-14 1-ring mean of those corrections; 15-17 1-ring mean
  velocity * h; 18-20 umbrella Laplacian of x_pred; 21-23 spring force,
  24-26 bending force, 27-29 gravity + external force at x_pred; 30 pinned;
  31-38 colour one-hot; 39 mean spring strain; 40 max |strain|;
  41 spring degree / 8; 42 bending degree / 16; 43 spring stiffness trace,
  44 bending stiffness, both * h^2/m; 45-47 obstacle contacts: the sum over
  the vertex's contacts of (thickness - (x_pred - x_s).n) n where positive
  (zero without an obstacle).  Each column is divided by its feature scale.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check, lib

NUM_FEATURES = 48
NUM_OUTPUTS = 3
HIDDEN = 128
# Sweeps a kept learned guess always gets, and sweeps between energy checks
# after that.  An energy check costs about a third of two sweeps on a 256x256
# cloth (one host round trip); checking first after 10 sweeps (the frames
# need 14 on average) beat checking after 4, 6, 8 or 16.
FIRST_CHECK = 10
CHECK_EVERY = 2

# offsets of the parameter blocks in the flat float[43408] vector
OFF_W1 = 0
OFF_B1 = OFF_W1 + 64 * HIDDEN
OFF_W2 = OFF_B1 + HIDDEN
OFF_B2 = OFF_W2 + HIDDEN * HIDDEN
OFF_W3 = OFF_B2 + HIDDEN
OFF_B3 = OFF_W3 + HIDDEN * HIDDEN
OFF_W4 = OFF_B3 + HIDDEN
OFF_B4 = OFF_W4 + HIDDEN * 16
NUM_PARAMS = OFF_B4 + 16
# multiply-adds x 2 of one row's forward pass (unpadded layer sizes)
FLOPS_PER_ROW = 2 * (NUM_FEATURES * HIDDEN + 2 * HIDDEN * HIDDEN + HIDDEN * NUM_OUTPUTS)


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def unpack_params(p):
    """Flat parameter vector -> [(W [in][out], b)] of the 4 layers, padding included."""
    p = np.asarray(p)
    return [
        (p[OFF_W1:OFF_B1].reshape(64, HIDDEN), p[OFF_B1:OFF_W2]),
        (p[OFF_W2:OFF_B2].reshape(HIDDEN, HIDDEN), p[OFF_B2:OFF_W3]),
        (p[OFF_W3:OFF_B3].reshape(HIDDEN, HIDDEN), p[OFF_B3:OFF_W4]),
        (p[OFF_W4:OFF_B4].reshape(HIDDEN, 16), p[OFF_B4:NUM_PARAMS]),
    ]


class WarmStartMLP:
    """The per-vertex warm-start network and its training rows (on the GPU)."""

    _close_rank = 0   # closed before the contexts it may be bound to

    def __init__(self, seed=0):
        n = ctypes.c_uint32()
        check(lib().lats_mlp_num_params(ctypes.byref(n)))
        if n.value != NUM_PARAMS:
            raise RuntimeError(f"library has {n.value} MLP parameters, expected {NUM_PARAMS}")
        h = ctypes.c_void_p()
        check(lib().lats_mlp_create(int(seed), ctypes.byref(h)))
        self._h = h
        _lib.track(self)
        self._bound = None

    def close(self):
        if getattr(self, "_h", None):
            lib().lats_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    # ---- parameters ------------------------------------------------------
    def params(self):
        p = np.zeros(NUM_PARAMS, np.float32)
        check(lib().lats_mlp_get_params(self._h, _p(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        if p.size != NUM_PARAMS:
            raise ValueError(f"expected {NUM_PARAMS} parameters")
        check(lib().lats_mlp_set_params(self._h, _p(p)))

    def feature_scale(self):
        s = np.zeros(NUM_FEATURES, np.float32)
        check(lib().lats_mlp_get_feature_scale(self._h, _p(s)))
        return s

    def set_feature_scale(self, s):
        s = np.ascontiguousarray(s, np.float32)
        if s.size != NUM_FEATURES:
            raise ValueError(f"expected {NUM_FEATURES} scales")
        check(lib().lats_mlp_set_feature_scale(self._h, _p(s)))

    # ---- checks on caller rows --------------------------------------------
    def forward(self, X):
        X = np.ascontiguousarray(X, np.float32).reshape(-1, NUM_FEATURES)
        Y = np.zeros((len(X), NUM_OUTPUTS), np.float32)
        check(lib().lats_mlp_forward_host(self._h, _p(X), len(X), _p(Y)))
        return Y

    def gradient(self, X, T):
        """(flat gradient of mean((Y - T)^2), loss) on caller rows, no update."""
        X = np.ascontiguousarray(X, np.float32).reshape(-1, NUM_FEATURES)
        T = np.ascontiguousarray(T, np.float32).reshape(-1, NUM_OUTPUTS)
        if len(T) != len(X):
            raise ValueError("X and T need the same number of rows")
        g = np.zeros(NUM_PARAMS, np.float32)
        loss = ctypes.c_double()
        check(lib().lats_mlp_gradient_host(self._h, _p(X), _p(T), len(X), _p(g), ctypes.byref(loss)))
        return g, loss.value

    # ---- solver side -------------------------------------------------------
    def bind(self, gpu):
        """Attach a GpuVBD (clears the correction history)."""
        check(lib().lats_mlp_bind(self._h, gpu._h))
        self._bound = gpu

    def set_guess_options(self, smooth_passes=0, half_step=False):
        """1-ring smoothing passes of the prediction; half-length retry before the fallback."""
        check(lib().lats_mlp_set_guess_options(self._h, int(smooth_passes), int(bool(half_step))))

    def set_stop_rule(self, rel_decrease):
        """Target-free stopping rule (0 turns it off): a kept guess sweeps
        until the energy decrease per sweep falls to rel_decrease * |energy|."""
        check(lib().lats_mlp_set_stop_rule(self._h, float(rel_decrease)))

    def reset_history(self):
        check(lib().lats_mlp_reset_history(self._h))

    def reserve_rows(self, rows):
        check(lib().lats_mlp_reserve_rows(self._h, int(rows)))

    def record_frame(self, gpu):
        """One plain frame of `gpu`, recorded as training rows."""
        check(lib().lats_mlp_record_frame(self._h, gpu._h))

    def append_rows(self, X, T):
        """Adds caller rows (features already divided by the feature scales)."""
        X = np.ascontiguousarray(X, np.float32).reshape(-1, NUM_FEATURES)
        T = np.ascontiguousarray(T, np.float32).reshape(-1, NUM_OUTPUTS)
        if len(T) != len(X):
            raise ValueError("X and T need the same number of rows")
        check(lib().lats_mlp_dataset_append(self._h, _p(X), _p(T), len(X)))

    def dataset_rows(self):
        n = ctypes.c_uint64()
        check(lib().lats_mlp_dataset_rows(self._h, ctypes.byref(n)))
        return n.value

    def dataset(self, first, count):
        X = np.zeros((count, NUM_FEATURES), np.float32)
        T = np.zeros((count, NUM_OUTPUTS), np.float32)
        check(lib().lats_mlp_dataset_read(self._h, int(first), int(count), _p(X), _p(T)))
        return X, T

    def clear_dataset(self):
        check(lib().lats_mlp_dataset_clear(self._h))

    def fit_feature_scale(self):
        s = np.zeros(NUM_FEATURES, np.float32)
        check(lib().lats_mlp_fit_feature_scale(self._h, _p(s)))
        return s

    def time_kernels(self, rows, iters=20):
        """(ms of one inference forward over `rows` recorded rows, ms of one training step on them)."""
        f = ctypes.c_float()
        t = ctypes.c_float()
        check(lib().lats_mlp_time_kernels(self._h, int(rows), int(iters), ctypes.byref(f), ctypes.byref(t)))
        return f.value, t.value

    def train(self, steps, batch_rows=8192, lr=1e-3, seed=0):
        loss = ctypes.c_double()
        check(lib().lats_mlp_train(self._h, int(steps), int(batch_rows), float(lr), int(seed), ctypes.byref(loss)))
        return loss.value

    def warm_start_frame(self, gpu, target=0.0, max_sweeps=100, first_check=FIRST_CHECK, check_every=CHECK_EVERY,
                         use_model=True):
        """One frame of `gpu` from the learned guess (if it lowers the energy)
        or the inertial guess, sweeping until the energy reaches `target`
        (a kept guess gets at least first_check sweeps)."""
        out = np.zeros(6, np.float32)
        check(lib().lats_mlp_warm_start_frame(self._h, gpu._h, float(target), int(max_sweeps), int(first_check),
                                              int(check_every), int(bool(use_model)), _p(out)))
        return _frame_result(out)

    def record_guided_frame(self, gpu, ref, max_sweeps=200, first_check=FIRST_CHECK, check_every=CHECK_EVERY,
                            use_model=True):
        """One warm-started frame of `gpu` whose target comes from the plain
        solve of `ref` (same mesh) from the same frame start; the frame is
        recorded as training rows (features of gpu, correction of ref)."""
        out = np.zeros(6, np.float32)
        check(lib().lats_mlp_record_guided_frame(self._h, gpu._h, ref._h, int(max_sweeps), int(first_check),
                                                 int(check_every), int(bool(use_model)), _p(out)))
        return _frame_result(out)


def _frame_result(out):
    return {"sweeps": int(out[0]), "kept": bool(out[1]), "e_inertial": float(out[2]),
            "e_learned": float(out[3]), "e_final": float(out[4]), "target": float(out[5])}


def record_scenarios(mlp, scenarios, params, frames, lanes=8, record_every=1, setup=None):
    """Plain-solver trajectories -> training rows (every record_every-th
    frame is recorded, the others are plain frames).  scenarios: [(cloth,
    wind)], wind None for no external force; setup(gpu) prepares each new
    context (an obstacle, say)."""
    from .vbd import GpuVBD
    from . import _lib
    total = sum(c.num_verts for c, _ in scenarios) * ((frames + record_every - 1) // record_every)
    mlp.reserve_rows(mlp.dataset_rows() + total)
    for cloth, wind in scenarios:
        g = GpuVBD(cloth, params, _lib.LATS_STRATEGY_GRAPH)
        g.set_lanes_per_vertex(lanes)
        if setup:
            setup(g)
        mlp.bind(g)
        for f in range(frames):
            if wind is not None:
                g.set_external_force(wind[f % len(wind)])
            if f % record_every == 0:
                mlp.record_frame(g)
            else:
                mlp.warm_start_frame(g, 0.0, params.constraint_iter_count, use_model=False)
        g.synchronize()
        g.close()
    return mlp.dataset_rows()


def record_guided(mlp, scenarios, params, frames, lanes=8, first_check=FIRST_CHECK, check_every=CHECK_EVERY,
                  use_model=True, record_every=1, setup=None):
    """Rows from the states the warm-started solver visits: each scenario is
    advanced by warm-started frames, each recorded frame's target is the
    plain solve from the same frame start (a second context)."""
    from .vbd import GpuVBD
    from . import _lib
    total = sum(c.num_verts for c, _ in scenarios) * ((frames + record_every - 1) // record_every)
    mlp.reserve_rows(mlp.dataset_rows() + total)
    stats = []
    for cloth, wind in scenarios:
        run = GpuVBD(cloth, params, _lib.LATS_STRATEGY_GRAPH)
        ref = GpuVBD(cloth, params, _lib.LATS_STRATEGY_GRAPH)
        run.set_lanes_per_vertex(lanes)
        ref.set_lanes_per_vertex(lanes)
        if setup:
            setup(run)
            setup(ref)
        mlp.bind(run)
        for f in range(frames):
            if wind is not None:
                run.set_external_force(wind[f % len(wind)])
            if f % record_every == 0:
                stats.append(mlp.record_guided_frame(run, ref, 2 * params.constraint_iter_count, first_check,
                                                     check_every, use_model))
            else:
                # target of an unrecorded frame: the plain solve on ref, as when recording
                x, v = run.state()
                ref.set_state(x, v)
                if wind is not None:
                    ref.set_external_force(wind[f % len(wind)])
                ref.frame_begin()
                ref.iterate(params.constraint_iter_count)
                stats.append(mlp.warm_start_frame(run, ref.energy_now(), 2 * params.constraint_iter_count,
                                                  first_check, check_every, use_model))
        run.close()
        ref.close()
    return stats


def train_recipe(mlp, scenarios, params, frames, plain_steps=2000, guided_rounds=3, guided_steps=2000,
                 batch_rows=16384, lr=1e-3, seed=0, smooth_passes=1, half_step=True, record_every=1, setup=None,
                 log=None):
    """Plain trajectories -> feature scales -> Adam; then rounds of rows from
    warm-started trajectories (targets from the plain solve) -> Adam on all
    rows (the states a warm-started solver visits differ from the plain
    solver's, so a network trained on plain rows alone degrades over a run).
    Returns the mean loss of the last chunk of steps."""
    import time
    t0 = time.perf_counter()
    mlp.set_guess_options(smooth_passes, half_step)
    record_scenarios(mlp, scenarios, params, frames, record_every=record_every, setup=setup)
    mlp.fit_feature_scale()
    loss = _train_chunks(mlp, plain_steps, batch_rows, lr, seed)
    if log:
        log(f"plain rows {mlp.dataset_rows()} loss {loss:.3e} ({time.perf_counter() - t0:.1f} s)")
    for r in range(guided_rounds):
        stats = record_guided(mlp, scenarios, params, frames, record_every=record_every, setup=setup)
        loss = _train_chunks(mlp, guided_steps, batch_rows, lr, seed + 1000 * (r + 1))
        if log:
            sw = [s["sweeps"] for s in stats]
            log(f"guided round {r}: rows {mlp.dataset_rows()} mean sweeps while recording {sum(sw) / len(sw):.1f} "
                f"loss {loss:.3e} ({time.perf_counter() - t0:.1f} s)")
    return loss


def _train_chunks(mlp, steps, batch_rows, lr, seed, chunk=250):
    """Adam steps in chunks; the last 30 % of the steps at lr / 5."""
    loss = 0.0
    done = 0
    while done < steps:
        k = min(chunk, steps - done)
        loss = mlp.train(k, batch_rows, lr if done < 0.7 * steps else lr * 0.2, seed + done)
        done += k
    return loss


def evaluate(mlp, cloth, params, wind, frames, first_check=FIRST_CHECK, check_every=CHECK_EVERY, use_model=True,
             lanes=8, max_sweeps=None, setup=None, regions=None, stop_rule=None, colour_block=None):
    """Frames of `cloth` solved from the learned guess until each frame's
    energy reaches the plain solver's energy after its full iteration count
    from the same frame start (computed on a second context).

    stop_rule: tolerance of the target-free stopping rule
    (lats_mlp_set_stop_rule).  The frames then never see the plain target:
    they stop by the rule, and the target is only reported beside the result
    (computed before the frame, outside its timing) to check it.

    lanes / regions / colour_block: colour-pass configuration of the
    warm-started context (results are identical for every choice; None keeps
    the default).

    Returns per-frame dicts (sweeps, kept, energies, target, wall_ms)."""
    import time
    from .vbd import GpuVBD
    from . import _lib
    iters = params.constraint_iter_count
    max_sweeps = max_sweeps or 2 * iters
    run = GpuVBD(cloth, params, _lib.LATS_STRATEGY_GRAPH)
    ref = GpuVBD(cloth, params, _lib.LATS_STRATEGY_GRAPH)
    run.set_lanes_per_vertex(lanes)
    ref.set_lanes_per_vertex(lanes)
    if regions:
        run.set_regions(regions)
        ref.set_regions(regions)
    if colour_block:
        run.set_colour_block(colour_block)
        ref.set_colour_block(colour_block)
    if setup:
        setup(run)
        setup(ref)
    for g in (run, ref):
        if g.strategy()[0] == _lib.LATS_STRATEGY_PATCH:
            g.patch_info()   # the patch layout is built here, not inside the first timed frame
    mlp.bind(run)
    out = []
    for f in range(frames):
        x, v = run.state()
        ref.set_state(x, v)
        if wind is not None:
            ref.set_external_force(wind[f % len(wind)])
            run.set_external_force(wind[f % len(wind)])
        ref.frame_begin()
        ref.iterate(iters)
        target = ref.energy_now()
        run.synchronize()
        t0 = time.perf_counter()
        r = mlp.warm_start_frame(run, 0.0 if stop_rule else target, max_sweeps, first_check, check_every,
                                 use_model)
        r["wall_ms"] = (time.perf_counter() - t0) * 1e3
        if stop_rule:
            r["rule_target"] = r["target"]     # the frame's own estimate
        r["target"] = target
        out.append(r)
    run.close()
    ref.close()
    return out


def evaluate_rule(mlp, cloth, params, wind, frames, rel_tol, **kw):
    """evaluate() with the target-free stopping rule at rel_tol."""
    mlp.set_stop_rule(rel_tol)
    try:
        return evaluate(mlp, cloth, params, wind, frames, stop_rule=rel_tol, **kw)
    finally:
        mlp.set_stop_rule(0.0)


STOP_RULE_CANDIDATES = (3e-4, 1e-4, 3e-5, 1e-5, 3e-6, 1e-6)
# (first check, sweeps between checks) of a kept guess
STOP_SCHEDULE_CANDIDATES = ((10, 2), (10, 3), (12, 2), (8, 3), (14, 2), (10, 4))


def calibrate_stop_rule(mlp, scenarios, params, frames, candidates=STOP_RULE_CANDIDATES, **kw):
    """Tolerance of the target-free stopping rule, chosen on training
    `scenarios` ((cloth, wind) pairs) only: the loosest candidate whose share
    of frames reaching their plain target is within one percentage point of
    the best candidate's.  Returns (tolerance, per-candidate summaries)."""
    import numpy as np
    tried = []
    for tol in sorted(candidates, reverse=True):
        res = []
        for cloth, wind in scenarios:
            res += evaluate_rule(mlp, cloth, params, wind, frames, tol, **kw)
        tried.append({"tol": tol, "reached": float(np.mean([r["e_final"] <= r["target"] for r in res])),
                      "mean_sweeps": float(np.mean([r["sweeps"] for r in res]))})
    best = max(t["reached"] for t in tried)
    return next(t["tol"] for t in tried if t["reached"] >= best - 0.01), tried


def calibrate_stop_schedule(mlp, scenarios, params, frames, candidates=STOP_RULE_CANDIDATES,
                            schedules=STOP_SCHEDULE_CANDIDATES, **kw):
    """Tolerance and check schedule (first check, sweeps between checks) of
    the target-free stopping rule, chosen on training `scenarios` only: of
    every (schedule, tolerance) pair, the fastest (mean wall time per frame)
    whose share of frames reaching their plain target is within one
    percentage point of the best pair's.  The plateau test is noisy over few
    sweeps (the energy of a warm-started frame is not monotone), so the
    schedule matters as much as the tolerance.  Returns a dict with tol,
    first_check, check_every and the per-pair summaries."""
    import numpy as np
    tried = []
    for fc, ce in schedules:
        for tol in candidates:
            res = []
            for cloth, wind in scenarios:
                res += evaluate_rule(mlp, cloth, params, wind, frames, tol, first_check=fc, check_every=ce, **kw)
            tried.append({"first_check": fc, "check_every": ce, "tol": tol,
                          "reached": float(np.mean([r["e_final"] <= r["target"] for r in res])),
                          "mean_sweeps": float(np.mean([r["sweeps"] for r in res])),
                          "ms_per_frame": float(np.mean([r["wall_ms"] for r in res]))})
    best = max(t["reached"] for t in tried)
    pick = min((t for t in tried if t["reached"] >= best - 0.01), key=lambda t: t["ms_per_frame"])
    return {"tol": pick["tol"], "first_check": pick["first_check"], "check_every": pick["check_every"],
            "tried": tried}
