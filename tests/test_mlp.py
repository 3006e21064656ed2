"""Learned warm start: the FP64 MLP oracle, the GPU forward / backward kernels
against it, the feature definition, training and the solve it warm-starts."""
import os

import numpy as np
import pytest

import cases
from oracle import mlp as omlp
from openmps_amd import scenes

FEAT, HID = 132, 256


def random_params(seed, scale=1.0):
    rng = np.random.Generator(np.random.PCG64(seed))
    p = {"w1": rng.uniform(-1, 1, (HID, FEAT)) / np.sqrt(FEAT),
         "b1": rng.uniform(-1, 1, HID) / np.sqrt(FEAT),
         "w2": rng.uniform(-1, 1, (HID, HID)) / np.sqrt(HID),
         "b2": rng.uniform(-1, 1, HID) / np.sqrt(HID),
         "w3": rng.uniform(-1, 1, HID) / np.sqrt(HID) * scale,
         "b3": rng.uniform(-1, 1, 1) / np.sqrt(HID)}
    return {k: v.astype(np.float32).astype(np.float64) for k, v in p.items()}


def test_oracle_gradients_match_finite_differences():
    """Pins the oracle's analytic backward pass (FP64) entry by entry."""
    p = random_params(1)
    rng = np.random.Generator(np.random.PCG64(2))
    x = rng.normal(size=(7, FEAT))
    t = rng.normal(size=7)
    _, g = omlp.loss_and_grads(p, x, t)
    picks = {"w1": [(0, 0), (17, 100), (255, 131)], "b1": [(3,), (200,)],
             "w2": [(5, 7), (250, 1)], "b2": [(9,)], "w3": [(0,), (128,)], "b3": [(0,)]}
    for name, idxs in picks.items():
        for idx in idxs:
            num = omlp.numeric_grad(p, x, t, name, idx)
            assert abs(num - g[name][idx]) <= 1e-6 * max(1.0, abs(num)), (name, idx)


def test_oracle_gelu_derivative():
    z = np.linspace(-6, 6, 1001)
    h = 1e-6
    num = (omlp.gelu(z + h) - omlp.gelu(z - h)) / (2 * h)
    assert np.allclose(omlp.gelu_grad(z), num, atol=1e-8)
    assert omlp.gelu(0.0) == 0.0 and abs(omlp.gelu(10.0) - 10.0) < 1e-12


def test_bf16_rounding():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.0e-39, 65504.0])
    got = omlp.to_bf16(vals)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: rounds to even (1.0);
    # 1 + 3 * 2^-9 rounds up to 1 + 2^-7
    assert got.tolist()[:4] == [1.0, 1.0, 1.0 + 2 ** -7, -2.5]
    assert got[5] == 65536.0


def emulate_gpu_forward(p, x):
    """The GPU's rounding points: BF16 x, w1, w2 and h1; FP32 elsewhere."""
    xb = omlp.to_bf16(x)
    z1 = xb @ omlp.to_bf16(p["w1"]).T + p["b1"]
    h1 = omlp.to_bf16(omlp.gelu(z1))
    z2 = h1 @ omlp.to_bf16(p["w2"]).T + p["b2"]
    return omlp.gelu(z2) @ p["w3"] + p["b3"][0]


@pytest.mark.gpu
def test_gpu_forward_matches_oracle():
    from openmps_amd.learned import Model
    p = random_params(3)
    m = Model(0)
    m.set_params(p)
    back = m.params()
    for k in p:
        assert np.array_equal(back[k].astype(np.float64), p[k]), k
    rng = np.random.Generator(np.random.PCG64(4))
    x = rng.normal(size=(1000, FEAT)).astype(np.float32).astype(np.float64)
    y = m.forward(x).astype(np.float64)
    ye = emulate_gpu_forward(p, x)
    assert np.abs(y - ye).max() <= 2e-3 * (np.abs(ye).max() + 1e-3)
    y64, _ = omlp.forward(p, x)  # unrounded FP64 reference: BF16-level agreement
    assert np.linalg.norm(y - y64) <= 3e-2 * np.linalg.norm(y64)


@pytest.mark.gpu
def test_gpu_forward_row_counts():
    """The inference kernel walks 64-row batches, several per block when the
    rows outnumber 64 x the compute units: tails, a single row and a
    multi-pass count all match the rounding-point emulation."""
    from openmps_amd.learned import Model
    p = random_params(11)
    m = Model(0)
    m.set_params(p)
    rng = np.random.Generator(np.random.PCG64(12))
    for rows in (1, 63, 65, 40_003):
        x = rng.normal(size=(rows, FEAT)).astype(np.float32).astype(np.float64)
        y = m.forward(x).astype(np.float64)
        ye = emulate_gpu_forward(p, x)
        assert y.shape == (rows,)
        assert np.abs(y - ye).max() <= 2e-3 * (np.abs(ye).max() + 1e-3), rows


@pytest.mark.gpu
def test_gpu_gradients_match_oracle():
    from openmps_amd.learned import Model
    p = random_params(5)
    pb = dict(p, w1=omlp.to_bf16(p["w1"]), w2=omlp.to_bf16(p["w2"]))
    m = Model(0)
    m.set_params(p)
    rng = np.random.Generator(np.random.PCG64(6))
    x = omlp.to_bf16(rng.normal(size=(4096, FEAT)))
    t = rng.normal(size=4096)
    loss, g = m.gradients(x, t)
    lo, go = omlp.loss_and_grads(pb, x, t)
    assert abs(loss - lo) <= 1e-2 * lo
    for k in ("w1", "b1", "w2", "b2", "w3", "b3"):
        err = np.linalg.norm(g[k] - go[k]) / np.linalg.norm(go[k])
        assert err < 3e-2, (k, err)


def _small_scene(seed=5):
    parts, dom = scenes.dam_break(0.1, 0.06, 0.025, 0.05, 0.001, jitter=0.02, seed=seed)
    env = dict(max_dt=1e-4, courant=0.1, g=9.8, rho=998.2, nu=1e-6, r_e_by_l0=2.1, l0=0.001,
               r_lap_by_l0=3.1, **dom)
    return parts, env


@pytest.mark.gpu
def test_features_definition():
    """The feature rows equal the definition evaluated in numpy from the
    particle state, the assembled system and r0 = b - A p_prev."""
    parts, env = _small_scene()
    c = cases.make("gpu", env, 1e-10)
    c.add_particles(parts)
    for _ in range(3):
        c.forward_time(1e-4)
    c.set_time(4e-4, 1e-4)
    for st in ("search", "density", "error_correction", "explicit", "density", "save_x",
               "ppe_assemble"):
        c.stage(st)
    _check_features(c)


@pytest.mark.gpu
def test_features_keep_the_32_nearest():
    """A lattice 1.2 times denser than l0 gives about 43 neighbours within
    r_lap: the rows keep the 32 nearest (the second walk of the feature
    kernel), and their count feature exceeds 1."""
    parts, env = _small_scene()
    env = dict(env, l0=0.0012)
    c = cases.make("gpu", env, 1e-10)
    c.add_particles(parts)
    c.set_time(1e-4, 1e-4)
    for st in ("search", "density", "error_correction", "explicit", "density", "save_x",
               "ppe_assemble"):
        c.stage(st)
    F = _check_features(c)
    assert (F[:, 0] > 1.0).sum() >= 500  # rows with more than 32 accepted neighbours


@pytest.mark.gpu
def test_features_wide_stencil():
    """r_lap = 4.5 l0: rows of about 60 columns, list bounds past 256 (16-bit
    column positions in the feature kernel), the 32 nearest kept."""
    parts, env = _small_scene()
    env = dict(env, r_lap_by_l0=4.5)
    c = cases.make("gpu", env, 1e-10)
    c.add_particles(parts)
    assert c.grid()["kmax"] > 256  # list bound: 16-bit column positions
    c.set_time(1e-4, 1e-4)
    for st in ("search", "density", "error_correction", "explicit", "density", "save_x",
               "ppe_assemble"):
        c.stage(st)
    F = _check_features(c)
    assert (F[:, 0] > 1.5).sum() >= 500


def _check_features(c):
    import scipy.sparse
    from openmps_amd import learned
    F = learned.features(c)
    rp, col, val, b, _ = c.ppe()
    P = c.particles()
    n = len(b)
    A = scipy.sparse.csr_matrix((val, col.astype(np.int64), rp.astype(np.int64)), shape=(n, n))
    r0 = b - A @ P["p"]
    alive = (P["type"] != 2) & (P["type"] != 3)
    s_r = np.sqrt((r0[alive] ** 2).sum() / alive.sum())
    s_p = np.sqrt((P["p"][alive] ** 2).sum() / alive.sum())
    isr = 1.0 / s_r if s_r > 0 else 0.0  # a zero scale zeroes its features
    isp = 1.0 / s_p if s_p > 0 else 0.0
    cnt, ids = c.neighbors()
    off = np.concatenate([[0], np.cumsum(cnt.astype(np.int64))])
    e = c.env()
    rl2 = e["r_lap"] ** 2
    bad = 0
    rng = np.random.Generator(np.random.PCG64(9))
    for i in rng.choice(np.nonzero(alive)[0], 300, replace=False):
        js = ids[off[i]:off[i + 1]]
        dx, dz = P["x"][js] - P["x"][i], P["z"][js] - P["z"][i]
        r2 = dx * dx + dz * dz
        keep = (P["type"][js] != 2) & (r2 < rl2)
        js, dx, dz, r2 = js[keep], dx[keep], dz[keep], r2[keep]
        order = np.lexsort((js, r2.astype(np.float32)))[:32]  # float r^2, then index
        want = np.zeros(FEAT)
        want[0] = len(js) / 32.0
        want[1] = (P["n"][i] - e["n0"]) / e["n0"]
        want[2] = P["p"][i] * isp
        want[3] = r0[i] * isr
        for k, q in enumerate(order):
            want[4 + 4 * k:8 + 4 * k] = [dx[q] / e["l0"], dz[q] / e["l0"], P["p"][js[q]] * isp,
                                         r0[js[q]] * isr]
        if not np.allclose(F[i], omlp.to_bf16(want), rtol=1e-2, atol=2e-2):
            bad += 1
    assert bad <= 3  # bf16 rounding of a feature on either side of a rounding boundary
    assert np.all(F[~alive] == 0)
    return F


@pytest.mark.gpu
def test_snapshots_and_training_reduce_loss():
    from openmps_amd import learned
    from openmps_amd import _lib as L
    parts, env = _small_scene(7)
    ds = learned.Dataset(200_000)
    c = cases.make("gpu", env, 1e-10)
    c.set_solver(kind=L.SOLVER_CG_FP64, stop_rule=L.STOP_RHS_NORM, tol=1e-10)
    c.add_particles(parts)
    learned.attach_dataset(c, ds, 0.25, 11)
    for _ in range(20):
        c.forward_time(1e-4)
    n = len(ds)
    assert n > 5000
    x, t = ds.rows(0, 10)
    assert np.all(np.isfinite(x)) and np.all(np.isfinite(t))
    m = learned.Model(3)
    losses = m.train(ds, 6, 1024, 2e-3, 5)
    assert losses[-1] < 0.7 * losses[0], losses


@pytest.mark.gpu
def test_learned_warm_start_is_judged_by_the_solver():
    """With any model (here untrained) the solve still reaches the stop rule:
    a worse guess is discarded for the previous-step guess."""
    from openmps_amd import learned
    from openmps_amd import _lib as L
    parts, env = _small_scene(8)
    c = cases.make("gpu", env, 1e-8)
    c.set_solver(kind=L.SOLVER_MIXED_IR, stop_rule=L.STOP_RHS_NORM, warm_start=L.WARM_LEARNED,
                 jacobi=True, tol=1e-8)
    m = learned.Model(1)
    learned.attach_model(c, m)
    c.add_particles(parts)
    for _ in range(5):
        c.forward_time(1e-4)
        s = c.solve_stats()
        assert s["rel_residual"] <= 1e-8 * 1.5
        if not s["learned_used"]:
            # the previous guess's residual was evaluated with A p_prev in FP32
            assert s["initial_rel_residual"] == pytest.approx(s["prev_rel_residual"], rel=1e-5)


@pytest.mark.gpu
def test_learned_start_with_the_multilevel_solver():
    """The learned start in front of the multilevel mixed solver (whose set-up
    then runs on a second stream beside it): every solve meets the stop rule,
    the trained model's guess is kept, and the pressures match the same
    solver started from the previous step's pressure to the solve tolerance."""
    from openmps_amd import learned
    from openmps_amd import _lib as L
    parts, env = _small_scene(12)
    runs = []
    for use_model in (True, False):
        c = cases.make("gpu", env, 1e-10)
        c.set_solver(kind=L.SOLVER_MIXED_IR, stop_rule=L.STOP_RHS_NORM, precond=L.PRECOND_MULTILEVEL,
                     warm_start=L.WARM_LEARNED if use_model else L.WARM_PREVIOUS, tol=1e-10)
        if use_model:
            learned.attach_model(c, learned.Model.load())
        c.add_particles(parts)
        used = 0
        for _ in range(6):
            c.forward_time(1e-4)
            s = c.solve_stats()
            assert s["rel_residual"] <= 1.5e-10
            used += s["learned_used"]
        if use_model:
            assert used >= 1
        runs.append(c.particles()["p"])
    scale = np.abs(runs[1]).max()
    assert np.abs(runs[0] - runs[1]).max() <= 1e-6 * scale


@pytest.mark.gpu
def test_learned_start_needs_a_particle_system():
    """A system installed with set_ppe over the particles (rows that are not
    the particles' stencils): the solve starts from the given guess with the
    model attached, and the feature query fails cleanly."""
    from openmps_amd import learned
    from openmps_amd import _lib as L
    from openmps_amd.computer import MpsError
    parts, env = _small_scene(3)
    c = cases.make("gpu", env, 1e-10)
    c.set_solver(kind=L.SOLVER_MIXED_IR, stop_rule=L.STOP_RHS_NORM, warm_start=L.WARM_LEARNED, tol=1e-10)
    learned.attach_model(c, learned.Model(1))
    c.add_particles(parts)
    n = c.size()
    # tridiagonal, diagonally dominant, with a known solution
    import scipy.sparse
    T = scipy.sparse.diags([-np.ones(n - 1), 4.0 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1], format="csr")
    T.sort_indices()
    xs = np.linspace(-1.0, 1.0, n)
    c.set_ppe(T.indptr.astype(np.uint64), T.indices.astype(np.uint32), T.data.astype(np.float64), T @ xs,
              np.zeros(n))
    c.stage("ppe_solve")
    s = c.solve_stats()
    assert s["learned_used"] == 0 and s["rel_residual"] <= 1e-10
    assert np.allclose(c.aux("ppe_x"), xs, atol=1e-8)
    with pytest.raises(MpsError):
        learned.features(c)


@pytest.mark.gpu
def test_trained_model_gains_over_previous_guess():
    """What the trained network gains: a model trained on the FP64 solver's own
    snapshots of one scene, applied to another seed, against the previous-step
    guess on the same assembled systems (initial residual and FP32 iterations
    of the mixed solver to |b - A x| <= 1e-8 |b|)."""
    from openmps_amd import learned
    from openmps_amd import _lib as L
    parts, env = _small_scene(21)
    ds = learned.Dataset(400_000)
    c = cases.make("gpu", env, 1e-10)
    c.set_solver(kind=L.SOLVER_CG_FP64, stop_rule=L.STOP_RHS_NORM, tol=1e-10)
    c.add_particles(parts)
    learned.attach_dataset(c, ds, 1.0, 5)
    for _ in range(60):
        c.forward_time(1e-4)
    m = learned.Model(4)
    losses = m.train(ds, 40, 512, 2e-3, 9)
    print(f"  {len(ds)} rows, loss {losses[0]:.3f} -> {losses[-1]:.3f}", flush=True)
    assert losses[-1] < 0.3 * losses[0]

    parts2, env2 = _small_scene(22)
    g = cases.make("gpu", env2, 1e-8)
    g.set_solver(kind=L.SOLVER_MIXED_IR, stop_rule=L.STOP_RHS_NORM, precond=L.PRECOND_JACOBI, tol=1e-8)
    g.add_particles(parts2)
    for _ in range(20):
        g.forward_time(1e-4)
    learned.attach_model(g, m)
    ratios, gains = [], []
    for k in range(6):
        g.set_time((21 + k) * 1e-4, 1e-4)
        for st in ("search", "density", "error_correction", "explicit", "density", "save_x",
                   "ppe_assemble"):
            g.stage(st)
        x_prev = g.aux("ppe_x")
        g.set_solver(kind=L.SOLVER_MIXED_IR, stop_rule=L.STOP_RHS_NORM, precond=L.PRECOND_JACOBI,
                     tol=1e-8, warm_start=L.WARM_LEARNED)
        g.stage("ppe_solve")
        s_l = g.solve_stats()
        x_l = g.aux("ppe_x")
        g.set_aux("ppe_x", x_prev)
        g.set_solver(kind=L.SOLVER_MIXED_IR, stop_rule=L.STOP_RHS_NORM, precond=L.PRECOND_JACOBI,
                     tol=1e-8, warm_start=L.WARM_PREVIOUS)
        g.stage("ppe_solve")
        s_p = g.solve_stats()
        assert s_l["rel_residual"] <= 1e-8 and s_p["rel_residual"] <= 1e-8
        ratios.append(s_l["initial_rel_residual"] / s_l["prev_rel_residual"])
        gains.append(s_p["iterations"] - s_l["iterations"])
        print(f"  step {k}: initial |r|/|b| learned {s_l['initial_rel_residual']:.3e} vs previous "
              f"{s_l['prev_rel_residual']:.3e}; iterations {s_l['iterations']} vs {s_p['iterations']}",
              flush=True)
        g.set_aux("ppe_x", x_l)
        for st in ("pressure_writeback", "pressure_gradient", "dynamic_stabilize"):
            g.stage(st)
    # the learned guess lowers the initial residual on every step and never costs iterations on
    # average
    assert max(ratios) < 0.8
    assert np.mean(gains) >= 0
