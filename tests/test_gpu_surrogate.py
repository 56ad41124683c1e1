"""GPU checks of the forecast surrogate (csrc/surrogate.hip) through the C-ABI.

  * forward output (FP32 and BF16 operands) against the FP64 network reference
    (tests/oracle_fno.py) on the same parameters and inputs;
  * every parameter gradient of the training loss against the reference's
    autograd gradients (themselves pinned by finite differences);
  * the state <-> grid transforms and the forecast update against the
    reference's band transforms, and the forecast noise variance;
  * training: Adam on pairs collected from the engine lowers the held-out loss
    far below persistence (predicting no change), i.e. the network has learned
    the flow map;
  * the surrogate-forecast particle filter tracks a twin truth within 5% of
    the plain-solver filter's analysis RMSE, and with the solver as live
    judge it falls back and reproduces the plain filter exactly.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle_fno as F

pytestmark = pytest.mark.gpu

SMALL = dict(state_n=64, grid=32, width=16, layers=2, modes=4)
TRAIN = dict(state_n=64, grid=32, width=32, layers=4, modes=8)
FULL = dict(state_n=128, grid=64, width=32, layers=4, modes=16)


def shape_of(cfg):
    return F.Shape(cfg["state_n"], cfg["grid"], cfg["width"], cfg["layers"], cfg["modes"])


def grf_fields(orc, n, count, base):
    return np.stack([orc.grf(n, orc.stream_seed(base, i)) for i in range(count)])


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("precision,tol", [("fp32", 2e-4), ("bf16", 4e-2)])
@pytest.mark.parametrize("cfg", [SMALL, FULL], ids=["small", "full"])
def test_forward_matches_reference(pkg, orc, cfg, precision, tol):
    s = shape_of(cfg)
    f = grf_fields(orc, s.n, 3, 11)
    nu = np.array([5e-4, 1e-3, 2e-3])
    amp = np.array([0.02, 0.1, 0.5])  # the forcing channel follows each sample's amplitude
    scale = 1.0 / np.sqrt(np.mean(F.sample_grid(s, f).numpy() ** 2))
    with pkg.Surrogate(**cfg, seed=7) as sur:
        sur.set_scales(scale, 1.0)
        got = sur.forward(f, nu, precision, amp=amp)
        default = sur.forward(f, nu, precision)
        p = torch.as_tensor(sur.get_params().astype(np.float64))
    want = F.forward(s, p, f, nu, scale, amp).numpy()
    assert rel(default, F.forward(s, p, f, nu, scale).numpy()) < tol  # amplitude 0.1
    print(cfg, precision, rel(got, want))
    assert rel(got, want) < tol


def test_fused_inference_matches_batched_products(pkg, orc):
    """The fused BF16 layer passes (csrc/fno_fused.hpp: one forward- and one
    inverse-transform pass per layer) against the training path's batched
    products on the same parameters and inputs: both round the same
    intermediates to BF16, so they agree to FP32 summation order and rare
    BF16 rounding flips; a forecast replayed from its captured graph equals
    the eager fused forecast bit for bit."""
    s = shape_of(FULL)
    S = 8
    f = grf_fields(orc, s.n, S, 31)
    nu = np.linspace(5e-4, 2e-3, S)
    amp = np.linspace(0.05, 0.4, S)
    scale = 1.0 / np.sqrt(np.mean(F.sample_grid(s, f).numpy() ** 2))
    with pkg.Surrogate(**FULL, seed=5) as sur:
        sur.set_scales(scale, 2.0)
        fused = sur.forward(f, nu, "bf16", amp=amp)
        sur.set_fused(False)
        batched = sur.forward(f, nu, "bf16", amp=amp)
        fp32 = sur.forward(f, nu, "fp32", amp=amp)
        sur.set_fused(True)
        p = torch.as_tensor(sur.get_params().astype(np.float64))
        want = F.forward(s, p, f, nu, scale, amp).numpy()
        print("fused vs batched", rel(fused, batched), "fused vs fp32", rel(fused, fp32),
              "fused vs fp64", rel(fused, want), "batched vs fp64", rel(batched, want))
        assert rel(fused, batched) < 5e-3
        assert rel(fused, want) < 4e-2
        # forecasts: captured and replayed (calls 2-4) vs a graph-free twin
        with pkg.Surrogate(**FULL, seed=5) as eager, \
                pkg.Ensemble(s.n, S, "fp32", sigma=0.2, seed=3) as a, \
                pkg.Ensemble(s.n, S, "fp32", sigma=0.2, seed=3) as b:
            eager.set_scales(scale, 2.0)
            eager.set_graph(False)
            a.set_state(f)
            b.set_state(f)
            for k in range(4):
                sur.forecast(a, 1e-3, 50, "bf16")
                eager.forecast(b, 1e-3, 50, "bf16")
                a.synchronize()
                b.synchronize()
                np.testing.assert_array_equal(a.get_state(), b.get_state())


@pytest.mark.parametrize("cfg,S", [(SMALL, 2), (FULL, 1)], ids=["small", "full"])
def test_gradients_match_autograd(pkg, orc, cfg, S):
    s = shape_of(cfg)
    f = grf_fields(orc, s.n, S, 21)
    nxt = f + 0.2 * grf_fields(orc, s.n, S, 22)
    nu = np.linspace(6e-4, 1.8e-3, S)
    amp = np.linspace(0.05, 0.3, S)
    ins, outs = 2.0, 3.0
    with pkg.Surrogate(**cfg, seed=8) as sur:
        sur.set_scales(ins, outs)
        loss, grad = sur.loss_grad(f, nxt, nu, amp=amp)
        p = sur.get_params().astype(np.float64)
    want_loss, want = F.loss_and_grad(s, p, f, nxt, nu, ins, outs, amp)
    print(cfg, loss, want_loss, rel(grad, want))
    assert abs(loss - want_loss) < 1e-5 * want_loss
    assert rel(grad, want) < 1e-4
    # every parameter group separately (small groups would hide in the total)
    o = 0
    for name, sh in s.sizes():
        k = int(np.prod(sh))
        g, w = grad[o:o + k], want[o:o + k]
        assert np.linalg.norm(g - w) <= 1e-3 * np.linalg.norm(w) + 1e-9, name
        o += k


def test_forecast_applies_band_increment_and_decay(pkg, orc):
    cfg = SMALL
    s = shape_of(cfg)
    B, dt, steps, nu, alpha = 8, 2e-3, 25, 1e-3, 0.01
    f = grf_fields(orc, s.n, B, 31)
    with pkg.Surrogate(**cfg, seed=9) as sur, \
            pkg.Ensemble(s.n, B, "fp32", nu=nu, alpha=alpha, sigma=0.0, amp=0.1, seed=4) as ens:
        sur.set_scales(1.5, 4.0)
        ens.set_state(f)
        f32 = ens.get_state()  # FP32-rounded retained states
        out = sur.forward(f32, np.full(B, nu), "fp32")
        sur.forecast(ens, dt, steps, "fp32")
        got = ens.get_state()
        assert ens.step_index == steps
    h = steps * dt
    want = f32 + F.band_increment(s, out / 4.0)
    k = np.fft.fftfreq(s.n, 1.0 / s.n)
    kx, ky = np.abs(k)[None, :], np.abs(k)[:, None]
    outside = (kx > s.KS) | (ky > s.KS)
    decay = np.exp(-(nu * (kx ** 2 + ky ** 2) + alpha) * h)
    want = np.where(outside[None], f32 * decay[None], want)
    assert rel(got, want) < 1e-5


def test_forecast_noise_has_the_forcing_variance(pkg, orc):
    cfg = SMALL
    s = shape_of(cfg)
    B, dt, steps, sigma = 512, 2e-3, 25, 0.5
    f = grf_fields(orc, s.n, 1, 41)
    with pkg.Surrogate(**cfg, seed=9) as sur:
        sur.set_scales(1.0, 1.0)
        runs = []
        for sg in (0.0, sigma):
            with pkg.Ensemble(s.n, B, "fp32", nu=1e-3, sigma=sg, kmin=1, kmax=4, seed=5) as ens:
                ens.set_state(np.repeat(f, B, axis=0))
                sur.forecast(ens, dt, steps, "fp32")
                runs.append(ens.get_state())
    d = runs[1] - runs[0]
    k = np.fft.fftfreq(s.n, 1.0 / s.n)
    k2 = k[None, :] ** 2 + k[:, None] ** 2
    shell = (k2 >= 1) & (k2 <= 16) & (k[None, :] > 0)
    var = np.mean(np.concatenate([d[:, shell].real.ravel(), d[:, shell].imag.ravel()]) ** 2)
    want = sigma ** 2 * steps * dt / 2
    print(var, want)
    assert abs(var / want - 1) < 0.1


def collect_and_train(pkg, cfg, epochs=20, windows=6, B=256, seed=3):
    """Pairs from a forced 64^2 ensemble with log-uniform viscosities; targets
    advanced without the shell noise (the deterministic propagator)."""
    n, dt, steps, sigma = cfg["state_n"], 4e-3, 25, 0.02
    rng = np.random.default_rng(seed)
    ens = pkg.Ensemble(n, B, "fp32", nu=1e-3, sigma=sigma, amp=0.1, seed=seed)
    for m in range(B):
        ens.set_member_parameters(m, float(np.exp(rng.uniform(np.log(5e-4), np.log(2e-3)))),
                                  0.0, sigma, 0.1)
    ens.init_grf()
    ens.step(dt, 50)
    sur = pkg.Surrogate(**cfg, seed=seed)
    sur.collect(ens, dt, steps, windows, noise_free_targets=True)
    assert sur.dataset_size == windows * B
    train = (windows - 1) * B
    losses = sur.fit(0, train, epochs, 32, 3e-3, seed=seed)
    return ens, sur, losses, train, B, dt, steps


def test_training_learns_the_flow_map(pkg):
    """What the network gains: on held-out pairs its squared error is a small
    fraction of persistence's (forecasting no change)."""
    ens, sur, losses, train, B, dt, steps = collect_and_train(pkg, TRAIN)
    held = sur.evaluate(train, B, "fp32")
    held_bf16 = sur.evaluate(train, B, "bf16")
    # persistence on the same held-out pairs: the all-zero network
    trained = sur.get_params()
    sur.set_params(np.zeros_like(trained))
    persistence = sur.evaluate(train, B, "fp32")
    sur.set_params(trained)
    print("epoch losses", losses, "held-out", held, held_bf16, "persistence", persistence)
    assert np.all(np.isfinite(losses)) and losses[-1] < 0.1 * losses[0]
    assert held < 0.1 * persistence
    assert held_bf16 < 1.5 * held + 0.005
    ens.close()
    sur.close()


def twin_64(pkg, dt, steps, cycles, nobs, sd, seed):
    """Hidden truth at 64^2 (one FP32 member from a GRF, nu = 1e-3, shell
    noise 0.02, Kolmogorov 0.1) observed every `steps` steps at nobs random
    grid points with noise sd.  The particles are a prior the truth could
    have been drawn from: background + independent GRFs kept on the forced
    large scales |k| <= 4, where the background is the truth plus one more
    such draw (so the free ensemble mean is off by that draw)."""
    n = TRAIN["state_n"]
    rng = np.random.default_rng(seed)
    with pkg.Ensemble(n, 1, "fp32", nu=1e-3, sigma=0.02, amp=0.1, seed=seed + 1) as truth:
        truth.init_grf()
        truth0 = truth.get_state()
        truths, rows = [], []
        for c in range(cycles):
            truth.step(dt, steps)
            w = truth.get_state()[0]
            truths.append(w)
            ph = (np.fft.ifft2(w) * n * n).real
            xs, ys = rng.integers(0, n, nobs), rng.integers(0, n, nobs)
            rows += [((c + 1) * steps, x, y, ph[y, x] + sd * rng.standard_normal(), sd)
                     for x, y in zip(xs, ys)]
    k = np.abs(np.fft.fftfreq(n, 1.0 / n))
    low = (k[None, :] ** 2 + k[:, None] ** 2) <= 16
    with pkg.Ensemble(n, 257, "fp32", seed=seed + 2) as pert:
        pert.init_grf()
        d = pert.get_state() * low[None]
    init = truth0 + d[0] + d[1:]
    return np.array(rows), truths, init


def analysis_rmse(history, truths):
    """Mean over groups of the grid RMS of (weighted mean - truth) in physical
    space: sqrt(sum |dw_k|^2) by Parseval, the fields being fft2(phys) / n^2."""
    return float(np.mean([np.sqrt(np.sum(np.abs(h["mean_field"] - t) ** 2))
                          for h, t in zip(history, truths)]))


def test_surrogate_filter_runs_and_tracks(pkg, orc):
    """Twin experiment at 64^2, 256 particles, 10 groups of 25 steps with 32
    observations (sd 0.1), shell noise 0.02: averaged over twelve filter
    seeds, the surrogate-forecast PF's analysis RMSE against the truth is
    within 5% of the plain-solver PF's from the same particles, and both have
    skill (below half the truth's own RMS).  The filters' weights are nearly
    degenerate here (ESS ~ 8 of 256), so one seed's RMSE varies by ~10%:
    the mean over twelve seeds has a standard error of ~1.5%.  With the solver
    as live judge and a threshold no network error can meet, the filter falls
    back to the solver in every group and reproduces the plain PF exactly."""
    ens, sur, _, _, _, dt, steps = collect_and_train(pkg, TRAIN, epochs=8)
    ens.close()
    n = TRAIN["state_n"]
    rows, truths, init = twin_64(pkg, dt, steps, 10, 32, 0.1, 4321)
    truth_rms = float(np.mean([np.sqrt(np.sum(np.abs(t) ** 2)) for t in truths]))
    errs = {"plain": [], "surrogate": []}
    for fseed in range(5, 17):
        with pkg.Ensemble(n, 256, "fp32", nu=1e-3, sigma=0.02, amp=0.1, seed=12) as pf:
            pf.set_state(init)
            plain = pkg.bootstrap_filter(pf, rows, dt, fseed)
        with pkg.Ensemble(n, 256, "fp32", nu=1e-3, sigma=0.02, amp=0.1, seed=12) as pf:
            pf.set_state(init)
            res = pkg.surrogate_filter(pf, sur, rows, dt, fseed, precision="bf16")
        assert len(plain["history"]) == len(res["history"]) == 10
        assert np.isfinite(res["log_likelihood"])
        errs["plain"].append(analysis_rmse(plain["history"], truths))
        errs["surrogate"].append(analysis_rmse(res["history"], truths))
    e = {k: float(np.mean(v)) for k, v in errs.items()}
    print("truth RMS", truth_rms, "analysis RMSE", e, errs)
    assert e["plain"] < 0.5 * truth_rms and e["surrogate"] < 0.5 * truth_rms
    assert e["surrogate"] <= 1.05 * e["plain"]
    # live judge + unreachable threshold: the solver forecasts every group
    with pkg.Ensemble(n, 256, "fp32", nu=1e-3, sigma=0.02, amp=0.1, seed=12) as pf, \
            pkg.Ensemble(n, 8, "fp32", nu=1e-3, sigma=0.0, amp=0.1, seed=13) as judge:
        pf.set_state(init)
        fell = pkg.surrogate_filter(pf, sur, rows, dt, 5, precision="bf16",
                                    fallback_threshold=1e-9, judge=judge, judge_every=3)
    # the same with the judges after the first one running beside the cycles
    with pkg.Ensemble(n, 256, "fp32", nu=1e-3, sigma=0.02, amp=0.1, seed=12) as pf, \
            pkg.Ensemble(n, 8, "fp32", nu=1e-3, sigma=0.0, amp=0.1, seed=13) as judge:
        pf.set_state(init)
        fell_async = pkg.surrogate_filter(pf, sur, rows, dt, 5, precision="bf16",
                                          fallback_threshold=1e-9, judge=judge,
                                          judge_every=3, judge_async=True)
    with pkg.Ensemble(n, 256, "fp32", nu=1e-3, sigma=0.02, amp=0.1, seed=12) as pf:
        pf.set_state(init)
        plain = pkg.bootstrap_filter(pf, rows, dt, 5)
    sur.close()
    # synchronous judges at groups 0, 3, 6, 9; enqueued ones started at 3 and 6
    # are folded in by group 6 and 9 at the latest (the one started at 9 after
    # the last group)
    assert [h["judged"] for h in fell["history"]] == [g % 3 == 0 for g in range(10)]
    assert fell_async["history"][0]["judged"]
    assert sum(h["judged"] for h in fell_async["history"]) == 3
    for run in (fell, fell_async):
        assert all(h["solver"] for h in run["history"])
        assert all(h["error_estimate"] > 0 for h in run["history"])
        assert run["log_likelihood"] == plain["log_likelihood"]
        for a, b in zip(run["history"], plain["history"]):
            assert a["increment"] == b["increment"] and a["resampled"] == b["resampled"]
            np.testing.assert_array_equal(a["mean_field"], b["mean_field"])


def test_spread_and_solver_fallback(pkg, orc):
    """The spread is the grid RMS deviation from the ensemble mean; the
    fallback runs the plain solver (bit-identical to stepping the engine)
    when the error estimate exceeds threshold x spread, else the network."""
    cfg = SMALL
    s = shape_of(cfg)
    B, dt, steps = 16, 2e-3, 25
    f = grf_fields(orc, s.n, B, 51)
    with pkg.Surrogate(**cfg, seed=9) as sur, \
            pkg.Ensemble(s.n, B, "fp32", sigma=0.05, seed=6) as a, \
            pkg.Ensemble(s.n, B, "fp32", sigma=0.05, seed=6) as b:
        sur.set_scales(1.0, 1.0)
        a.set_state(f)
        b.set_state(f)
        x = a.get_state()
        phys = np.stack([orc.fft2d(m, inverse=True).real for m in x])
        want = np.sqrt(np.mean((phys - phys.mean(0)) ** 2) * 1.0)
        assert abs(sur.spread(a) - want) < 1e-5 * want
        with pytest.raises(pkg.HzlInvalidArgument):
            sur.forecast_or_solve(a, dt, steps, 0.5)  # no error estimate yet
        sur.error_estimate = 1e3
        used, sp = sur.forecast_or_solve(a, dt, steps, 0.5)
        b.step(dt, steps)
        assert used and sp > 0
        np.testing.assert_array_equal(a.get_state(), b.get_state())
        assert a.step_index == steps
        sur.error_estimate = 0.0
        used, sp = sur.forecast_or_solve(a, dt, steps, 0.5, precision="fp32")
        assert not used
        assert abs(sp - sur.spread(a)) < 1e-12 * sp
        assert a.step_index == 2 * steps


def test_async_judge_rejects_unusable_judge_ensembles(pkg, orc):
    """judge_start checks before it enqueues anything: a judge ensemble with
    more members than the ensemble, the ensemble itself, an FP64 one and one
    running the OU forcing process are refused, and nothing is left in flight."""
    cfg = SMALL
    s = shape_of(cfg)
    with pkg.Surrogate(**cfg, seed=9) as sur, \
            pkg.Ensemble(s.n, 4, "fp32", sigma=0.05, seed=6) as ens, \
            pkg.Ensemble(s.n, 8, "fp32", sigma=0.0, seed=7) as big, \
            pkg.Ensemble(s.n, 2, "fp64", sigma=0.0, seed=8) as f64, \
            pkg.Ensemble(s.n, 2, "fp32", sigma=0.0, seed=9) as ou:
        sur.set_scales(1.0, 1.0)
        ens.init_grf()
        ou.set_forcing_process("ou")
        for judge in (big, ens, f64, ou):
            with pytest.raises(pkg.HzlInvalidArgument):
                sur.judge_start(ens, judge, 2e-3, 5, "fp32")
        assert sur.judge_poll(wait=True) == (False, -1.0)


def test_async_judge_outlives_closed_handles(pkg, orc):
    """A judge in flight when its judge ensemble is closed still folds in (the
    ensemble's close waits for its stream; the sums live in the surrogate's
    host memory), and closing the surrogate with a judge in flight waits for
    it: the ensembles stay usable and step on."""
    cfg = SMALL
    s = shape_of(cfg)
    B, J, dt, steps = 16, 6, 2e-3, 25
    f = grf_fields(orc, s.n, B, 75)
    with pkg.Ensemble(s.n, B, "fp32", sigma=0.05, amp=0.1, seed=6) as ens:
        ens.set_state(f)
        sur = pkg.Surrogate(**cfg, seed=9)
        sur.set_scales(1.5, 4.0)
        judge = pkg.Ensemble(s.n, J, "fp32", sigma=0.0, amp=0.1, seed=7)
        sur.judge_start(ens, judge, dt, steps, "fp32")
        judge.close()
        finished, v = sur.judge_poll(wait=True)
        assert finished and np.isfinite(v) and v > 0
        with pkg.Ensemble(s.n, J, "fp32", sigma=0.0, amp=0.1, seed=8) as judge2:
            sur.judge_start(ens, judge2, dt, steps, "fp32")
            sur.close()
            judge2.step(dt, 2)
            judge2.synchronize()
        ens.step(dt, 2)
        ens.synchronize()
        assert np.all(np.isfinite(ens.get_state()))


@pytest.mark.parametrize("B", [1, 100, 1024])
def test_spread_over_member_chunks(pkg, orc, B):
    """The spread sums members in chunks of 32 (a ragged last chunk at 100)
    with deviations from member 0: on members that share a mean field 1e3
    times their spread it still equals the FP64 grid RMS deviation of the
    same FP32 states to 1e-6 (raw moments would spend ~6 of FP64's 16 digits
    on the cancellation here; the shift by member 0 spends none)."""
    cfg = SMALL
    s = shape_of(cfg)
    base = grf_fields(orc, s.n, 1, 90)[0]
    pert = grf_fields(orc, s.n, B, 91)
    with pkg.Surrogate(**cfg, seed=9) as sur, \
            pkg.Ensemble(s.n, B, "fp32", sigma=0.05, seed=6) as a:
        sur.set_scales(1.0, 1.0)
        a.set_state(1e3 * base[None] + pert)
        x = a.get_state()
        got = sur.spread(a)
    phys = np.stack([orc.fft2d(m, inverse=True).real for m in x])
    dev = phys - phys.mean(0)
    want = np.sqrt(np.mean(dev ** 2))
    print("spread", B, got, want)
    if B == 1:
        assert got == 0.0
    else:
        assert abs(got - want) < 1e-6 * want


def test_graph_replays_match_eager_forecasts(pkg, orc):
    """Forecasts replayed from the captured graph equal kernel-by-kernel
    forecasts bit for bit, over several replays with eager engine work and
    synchronisations between them (and a new noise counter each time).  Call
    1 runs eagerly, call 2 captures, calls 3-5 replay."""
    cfg = SMALL
    s = shape_of(cfg)
    B, dt, steps = 16, 2e-3, 25
    f = grf_fields(orc, s.n, B, 61)
    with pkg.Surrogate(**cfg, seed=9) as g, pkg.Surrogate(**cfg, seed=9) as e, \
            pkg.Ensemble(s.n, B, "fp32", sigma=0.3, seed=6) as a, \
            pkg.Ensemble(s.n, B, "fp32", sigma=0.3, seed=6) as b:
        g.set_scales(1.5, 4.0)
        e.set_scales(1.5, 4.0)
        e.set_graph(False)
        a.set_state(f)
        b.set_state(f)
        for k in range(5):
            g.forecast(a, dt, steps, "bf16")
            e.forecast(b, dt, steps, "bf16")
            a.step(dt, 2)
            b.step(dt, 2)
            a.synchronize()
            b.synchronize()
            np.testing.assert_array_equal(a.get_state(), b.get_state())


def test_graph_cache_replays_alternating_ensembles(pkg, orc):
    """One surrogate forecasting two ensembles in turn (a filter's particles
    and a judge's scratch ensemble) keeps a captured forecast for each; every
    call equals a graph-free surrogate's on twin ensembles bit for bit."""
    cfg = SMALL
    s = shape_of(cfg)
    dt, steps = 2e-3, 25
    fa, fc = grf_fields(orc, s.n, 16, 81), grf_fields(orc, s.n, 4, 82)
    with pkg.Surrogate(**cfg, seed=9) as g, pkg.Surrogate(**cfg, seed=9) as e, \
            pkg.Ensemble(s.n, 16, "fp32", sigma=0.3, seed=6) as a, \
            pkg.Ensemble(s.n, 16, "fp32", sigma=0.3, seed=6) as b, \
            pkg.Ensemble(s.n, 4, "fp32", sigma=0.0, seed=7) as c, \
            pkg.Ensemble(s.n, 4, "fp32", sigma=0.0, seed=7) as d:
        for x in (g, e):
            x.set_scales(1.5, 4.0)
        e.set_graph(False)
        a.set_state(fa)
        b.set_state(fa)
        c.set_state(fc)
        d.set_state(fc)
        for k in range(6):
            g.forecast(a, dt, steps, "bf16")
            e.forecast(b, dt, steps, "bf16")
            if k % 2 == 0:
                g.forecast(c, dt, steps, "bf16")
                e.forecast(d, dt, steps, "bf16")
            for x in (a, b, c, d):
                x.synchronize()
            np.testing.assert_array_equal(a.get_state(), b.get_state())
            np.testing.assert_array_equal(c.get_state(), d.get_state())


def test_judge_measures_the_network_error_against_the_solver(pkg, orc):
    """hzl_surrogate_judge: the first J particles forecast by the network
    (no shell noise) and by the plain solver from the same states; the RMS
    of their difference equals the same comparison made by hand with
    separate ensembles, becomes the error estimate, and leaves the filtered
    ensemble untouched (states and step index)."""
    cfg = SMALL
    s = shape_of(cfg)
    B, J, dt, steps = 16, 6, 2e-3, 25
    f = grf_fields(orc, s.n, B, 71)
    nus = np.linspace(6e-4, 1.8e-3, B)
    with pkg.Surrogate(**cfg, seed=9) as sur, \
            pkg.Ensemble(s.n, B, "fp32", sigma=0.05, amp=0.1, seed=6) as ens, \
            pkg.Ensemble(s.n, J, "fp32", sigma=0.0, amp=0.1, seed=7) as judge, \
            pkg.Ensemble(s.n, J, "fp32", sigma=0.0, amp=0.1, seed=8) as net, \
            pkg.Ensemble(s.n, J, "fp32", sigma=0.0, amp=0.1, seed=8) as sol:
        sur.set_scales(1.5, 4.0)
        for m in range(B):
            ens.set_member_parameters(m, nus[m], 0.0, 0.05, 0.1 + 0.01 * m)
        for e in (net, sol):
            for m in range(J):
                e.set_member_parameters(m, nus[m], 0.0, 0.0, 0.1 + 0.01 * m)
        ens.set_state(f)
        ens.step(dt, 3)
        before = ens.get_state()
        rms = sur.judge(ens, judge, dt, steps, "fp32")
        assert np.array_equal(ens.get_state(), before) and ens.step_index == 3
        assert sur.error_estimate == rms
        net.set_state(before[:J])
        sol.set_state(before[:J])
        sur.forecast(net, dt, steps, "fp32")
        sol.step(dt, steps)
        d = net.get_state() - sol.get_state()
        want = np.sqrt(np.sum(np.abs(d) ** 2) / J)
    print("judge", rms, want)
    assert rms > 0 and abs(rms - want) <= 1e-5 * want


def test_async_judge_matches_the_synchronous_judge(pkg, orc):
    """hzl_surrogate_judge_start / _poll: the enqueued judge folds in the same
    RMS as hzl_surrogate_judge on the same states, although the ensemble
    steps on right after the start (its stream waits only for the copy of
    the judged states), and leaves the ensemble exactly where a twin without
    any judge ends; a poll with nothing in flight reports nothing new."""
    cfg = SMALL
    s = shape_of(cfg)
    B, J, dt, steps = 16, 6, 2e-3, 25
    f = grf_fields(orc, s.n, B, 73)
    nus = np.linspace(6e-4, 1.8e-3, B)
    with pkg.Surrogate(**cfg, seed=9) as sur, \
            pkg.Ensemble(s.n, B, "fp32", sigma=0.05, amp=0.1, seed=6) as ens, \
            pkg.Ensemble(s.n, B, "fp32", sigma=0.05, amp=0.1, seed=6) as twin, \
            pkg.Ensemble(s.n, J, "fp32", sigma=0.0, amp=0.1, seed=7) as judge:
        sur.set_scales(1.5, 4.0)
        for e in (ens, twin):
            for m in range(B):
                e.set_member_parameters(m, nus[m], 0.0, 0.05, 0.1 + 0.01 * m)
            e.set_state(f)
            e.step(dt, 3)
        assert sur.judge_poll() == (False, -1.0)
        for k in range(3):
            # the synchronous judge of the twin's (identical) states
            want = sur.judge(twin, judge, dt, steps, "fp32")
            assert sur.judge_poll(wait=True) == (False, want)
            sur.error_estimate = 123.0
            sur.judge_start(ens, judge, dt, steps, "fp32")
            ens.step(dt, 2)  # queued behind the copy of the judged states
            finished, v = sur.judge_poll(wait=True)
            print("async judge", k, v, "sync", want)
            assert finished and v == want and sur.error_estimate == want
            assert sur.judge_poll() == (False, want)
            twin.step(dt, 2)
            np.testing.assert_array_equal(ens.get_state(), twin.get_state())
        assert ens.step_index == twin.step_index == 9
