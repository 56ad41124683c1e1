"""FP64 oracle drivers for the filter loops, built from the pinned oracle
primitives (oracle/hzl_oracle.c).  TEST ONLY.

oracle_bootstrap_filter restates reference bootstrap_filter
(src/filter.cpp:38-126) with either of the engine's forcing processes -- white
shell noise keyed per particle by Philox, or the reference's own OU process
(mt19937_64 NormalStream per particle, children keep the ancestor's forcing and
restart their stream, src/filter.cpp:112-115) -- plus Kolmogorov forcing, so
that the GPU filter can be compared group by group: same particle keys, same
resampling uniforms (mt19937_64(stream_seed(seed, UINT64_MAX))), same
reseeding of children.
"""
from __future__ import annotations

import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

# Particles advance independently; the oracle's C calls release the GIL, so a
# thread pool runs them in parallel (at most 16 workers).
_POOL = ThreadPoolExecutor(max_workers=max(1, min(16, os.cpu_count() or 1)))


def _each(fn, count):
    return list(_POOL.map(fn, range(count)))

UINT64_MAX = (1 << 64) - 1


def full_from_half(h, n):
    """Reference Field (n x n) from a retained half plane [kx][ky + K]."""
    K = (n - 1) // 3
    out = np.zeros((n, n), dtype=np.complex128)
    for c in range(K + 1):
        for j, ky in enumerate(range(-K, K + 1)):
            out[ky % n, c] = h[c, j]
            if c > 0:
                out[(-ky) % n, n - c] = np.conj(h[c, j])
    return out


def physical(orc, field):
    return orc.fft2d(field, inverse=True).real


def oracle_bootstrap_filter(orc, initial, obs, dt, seed, nu, alpha=0.0, amp=0.1, kf=4,
                            sigma=0.05, kmin=1, kmax=4, forcing="white", tau=0.5):
    """initial: (P, n, n) spectral particle states.  obs rows (step, x, y, value, sd).
    forcing: "white" or "ou" (OU processes started from the particle keys)."""
    states = np.array(initial, dtype=np.complex128)
    P, n = states.shape[0], states.shape[1]
    keys = [orc.stream_seed(seed, i) for i in range(P)]
    ou = [orc.ou_process(n, sigma, tau, kmin, kmax, k) for k in keys] if forcing == "ou" else None
    kolmogorov = orc.step_forcing(n, amp, kf, 0.0, kmin, kmax, dt, 0, 0)
    weights = np.full(P, 1.0 / P)
    rng = orc.mt(orc.stream_seed(seed, UINT64_MAX))
    steps = obs[:, 0].astype(np.int64)
    groups = np.unique(steps)
    current, total, events, history = 0, 0.0, 0, []
    for g, target in enumerate(groups):
        sel = obs[steps == target]

        def advance(i):
            if ou is None:
                return orc.trajectory(n, nu, alpha, dt, int(target - current), keys[i], amp=amp,
                                      kf=kf, sigma=sigma, kmin=kmin, kmax=kmax, step0=current,
                                      initial=states[i])
            w = states[i]
            for _ in range(int(target - current)):
                ou[i].transition(dt / 2)
                w = orc.rk4(w, nu, alpha, dt, 1, ou[i].F + kolmogorov)
                ou[i].transition(dt / 2)
            return w

        states = np.stack(_each(advance, P))
        logs = np.zeros(P)
        for i in range(P):
            ph = physical(orc, states[i])
            pred = ph[sel[:, 2].astype(int), sel[:, 1].astype(int)]
            logs[i] = orc.pf_logweight(weights[i], pred, sel[:, 3], sel[:, 4])
        inc, w, ess, mw = orc.pf_normalize(logs)
        total += inc
        mean = np.tensordot(w, states, axes=1)
        resampled = False
        if ess < P / 2 and g + 1 < len(groups):
            u = orc.mt_uniform(rng)
            anc = orc.systematic_resample(w, u)
            states = states[anc].copy()
            keys = [orc.stream_seed(seed, (g + 1) * P + i) for i in range(P)]
            if ou is not None:
                kept = [ou[a].F.copy() for a in anc]
                ou = [orc.ou_process(n, sigma, tau, kmin, kmax, k) for k in keys]
                for proc, F, k in zip(ou, kept, keys):
                    proc.reseed(k)  # reseed_future: the stream restarts, no draws
                    proc.F = F      # the ancestor's forcing
            w = np.full(P, 1.0 / P)
            events += 1
            resampled = True
        weights = w
        history.append(dict(step=int(target), increment=inc, ess=ess, max_weight=mw,
                            resampled=resampled, mean_field=mean))
        current = int(target)
    return dict(log_likelihood=total, resampling_events=events, history=history)


def symmetrise(field):
    """Reference update_latent_state pair averaging (src/solver.cpp:414-421)."""
    n = field.shape[0]
    out = field.copy()
    for y in range(n):
        for x in range(n):
            i, j = y * n + x, ((n - y) % n) * n + (n - x) % n
            if i < j:
                a = 0.5 * (out.flat[i] + np.conj(out.flat[j]))
                out.flat[i], out.flat[j] = a, np.conj(a)
    return out


def gaspari_cohn(d, radius):
    """Gaspari & Cohn (1999) eq. 4.10, half support `radius` (src/filter.cpp:25-36)."""
    x = np.asarray(d, dtype=np.float64) / radius
    a = 1 - 5 / 3 * x**2 + 5 / 8 * x**3 + 0.5 * x**4 - 0.25 * x**5
    with np.errstate(divide="ignore", invalid="ignore"):
        b = 4 - 5 * x + 5 / 3 * x**2 + 5 / 8 * x**3 - 0.5 * x**4 + x**5 / 12 - 2 / (3 * x)
    return np.where(x <= 1, a, np.where(x < 2, b, 0.0))


def taper_field(n, xo, yo, radius):
    g = np.arange(n)
    dx = np.minimum(np.abs(g - xo), n - np.abs(g - xo))
    dy = np.minimum(np.abs(g - yo), n - np.abs(g - yo))
    return gaspari_cohn(np.hypot(dx[None, :], dy[:, None]), radius)


def localized_scalar_update(orc, states, h, y, sd, eps, xo, yo, radius):
    """Reference scalar update with the covariance tapered in physical space
    (src/filter.cpp:276-319)."""
    P, n = states.shape[0], states.shape[1]
    X = states.reshape(P, -1)
    d = h - h.mean()
    var = np.sum(d * d) / (P - 1)
    cov = np.sum((X - X.mean(0)) * (d / (P - 1))[:, None], axis=0).reshape(n, n)
    phys = orc.fft2d(cov, inverse=True) * taper_field(n, xo, yo, radius)
    cov = orc.fft2d(phys) * orc.active_mask(n)
    g = (y + sd * eps - h) / (var + sd * sd)
    return (X + g[:, None] * cov.ravel()[None, :]).reshape(P, n, n)


def project_forcing(orc, F, kmin, kmax):
    """update_latent_state's forcing part (src/solver.cpp:412-416): keep the
    shell pairs, each the mean of itself and its partner's conjugate."""
    n = F.shape[0]
    out = np.zeros_like(F)
    flat, res = F.ravel(), out.reshape(-1)
    for i in orc.ou_pairs(n, kmin, kmax):
        y, x = divmod(int(i), n)
        j = ((n - y) % n) * n + (n - x) % n
        res[i] = 0.5 * (flat[i] + np.conj(flat[j]))
        res[j] = np.conj(res[i])
    return out


def oracle_enkf(orc, initial, obs, dt, ens_seed, filter_seed, nu, alpha=0.0, amp=0.1, kf=4,
                sigma=0.05, kmin=1, kmax=4, inflation=1.0, mode="serial", radius=0.0,
                forcing="white", tau=0.5):
    """Perturbed-observation EnKF in FP64 on full spectra, restating reference
    ensemble_kalman_filter (src/filter.cpp:204-356) with either forcing process:
    "white" (the engine's, no forcing state) or "ou" (the reference's: each
    member's OU forcing is inflated and updated with its own covariance,
    tapered like the state's, then projected on the shell pairs); mode "batch"
    applies the dense all-observation analysis instead of the scalar sequence,
    with the same perturbations (white forcing only)."""
    states = np.array(initial, dtype=np.complex128)
    P, n = states.shape[0], states.shape[1]
    keys = [orc.stream_seed(ens_seed, i) for i in range(P)]
    ou = [orc.ou_process(n, sigma, tau, kmin, kmax, k) for k in keys] if forcing == "ou" else None
    if ou is not None and mode != "serial":
        raise ValueError("the OU forcing state is updated by the serial analysis only")
    kolmogorov = orc.step_forcing(n, amp, kf, 0.0, kmin, kmax, dt, 0, 0)
    normals = orc.mt(orc.stream_seed(filter_seed, P))
    steps = obs[:, 0].astype(np.int64)
    current, history = 0, []
    for target in np.unique(steps):
        sel = obs[steps == target]
        for i in range(P):
            if ou is None:
                states[i] = orc.trajectory(n, nu, alpha, dt, int(target - current), keys[i],
                                           amp=amp, kf=kf, sigma=sigma, kmin=kmin, kmax=kmax,
                                           step0=current, initial=states[i])
                continue
            for _ in range(int(target - current)):
                ou[i].transition(dt / 2)
                states[i] = orc.rk4(states[i], nu, alpha, dt, 1, ou[i].F + kolmogorov)
                ou[i].transition(dt / 2)
        if inflation > 1:
            mu = states.mean(axis=0)
            states = mu + inflation * (states - mu)
            if ou is not None:
                Fs = np.stack([p.F for p in ou])
                muF = Fs.mean(axis=0)
                for p, F in zip(ou, muF + inflation * (Fs - muF)):
                    p.F = project_forcing(orc, F, kmin, kmax)
        fc = physical(orc, states.mean(axis=0))
        xs, ys = sel[:, 1].astype(int), sel[:, 2].astype(int)
        rmse = float(np.sqrt(np.mean((fc[ys, xs] - sel[:, 3]) ** 2)))
        eps = np.array([[orc.normal(normals) for _ in range(P)] for _ in range(len(sel))])
        if mode == "serial":
            for o, row in enumerate(sel):
                h = np.array([physical(orc, s)[int(row[2]), int(row[1])] for s in states])
                if ou is not None:
                    # the forcing state's update (src/filter.cpp:273-319)
                    Fs = np.stack([p.F for p in ou])
                    d = h - h.mean()
                    var = np.sum(d * d) / (P - 1)
                    fcov = np.sum((Fs - Fs.mean(0)) * (d / (P - 1))[:, None, None], axis=0)
                    if radius > 0:
                        tap = taper_field(n, int(row[1]), int(row[2]), radius)
                        fcov = orc.fft2d(orc.fft2d(fcov, inverse=True) * tap)
                    g = (row[3] + row[4] * eps[o] - h) / (var + row[4] * row[4])
                    for p, gi in zip(ou, g):
                        p.F = project_forcing(orc, p.F + fcov * gi, kmin, kmax)
                if radius > 0:
                    upd = localized_scalar_update(orc, states, h, row[3], row[4], eps[o],
                                                  int(row[1]), int(row[2]), radius)
                else:
                    upd = orc.enkf_scalar_update(states.reshape(P, -1), h, row[3], row[4],
                                                 eps[o]).reshape(P, n, n)
                states = np.stack([symmetrise(f) for f in upd])
        else:
            X = np.concatenate([states.reshape(P, -1).real, states.reshape(P, -1).imag], 1)
            HX = np.stack([physical(orc, s)[ys, xs] for s in states])  # P x m
            A, HA = X - X.mean(0), HX - HX.mean(0)
            Pxy = A.T @ HA / (P - 1)
            S = HA.T @ HA / (P - 1) + np.diag(sel[:, 4] ** 2)
            D = sel[:, 3][:, None] + sel[:, 4][:, None] * eps - HX.T  # m x P
            Xa = X + (Pxy @ np.linalg.solve(S, D)).T
            d = n * n
            states = (Xa[:, :d] + 1j * Xa[:, d:]).reshape(P, n, n)
            states = np.stack([symmetrise(f) for f in states])
        mu = states.mean(axis=0)
        spread = float(np.sqrt(np.sum(np.abs(states - mu) ** 2) / (P - 1)))
        history.append(dict(step=int(target), forecast_sensor_rmse=rmse, posterior_spread=spread,
                            mean_field=mu, states=states.copy()))
        current = int(target)
    return history


STOCHASTIC_PRIOR = ((np.log(0.03), -10.0, 0.0), (np.log(0.1), -10.0, 0.0),
                    (np.log(0.2), -10.0, 2.0), (np.log(0.5), -10.0, 2.0))


def stochastic_log_prior(q):
    """Reference stochastic_log_prior (src/filter.cpp:127-140): Gaussian of
    width 1.5 around (log 0.03, log 0.1, log 0.2, log 0.5), -inf outside
    [-10, 0] (first two) and [-10, 2] (last two)."""
    total = 0.0
    for qi, (c, lo, hi) in zip(q, STOCHASTIC_PRIOR):
        if not (np.isfinite(qi) and lo <= qi <= hi):
            return -np.inf
        total -= 0.5 * ((qi - c) / 1.5) ** 2
    return total


def oracle_pmcmc(orc, initial, obs, dt, q0, samples, scale, seed, lower, upper, particles,
                 alpha=0.0, sigma=0.05, kf=4, kmin=1, kmax=4, forcing="white", tau=0.5,
                 prior="log_uniform", chain_seeds=None):
    """Reference particle_metropolis (src/filter.cpp:166-203) per chain.

    prior "log_uniform": q = (log nu, log amplitude), uniform in [lower, upper],
    alpha / sigma / tau fixed, Kolmogorov amplitude exp(q1) (the viscosity / amplitude chains);
    prior "stochastic": q = (log nu, log alpha, log sigma, log tau) with the
    reference's stochastic_log_prior and no Kolmogorov term (the reference's own
    sampler; use forcing="ou").  Chain c's stream seed s is chain_seeds[c] when
    given, else stream_seed(seed, c): NormalStream(s) for proposals,
    mt19937_64(stream_seed(s, 1)) for likelihood seeds and acceptance uniforms.
    Rows per sample: (q, log-likelihood, log-density, accepted, likelihood seed)."""
    n = initial.shape[0]
    K = (n - 1) // 3
    init = np.stack([initial] * particles)
    D = 4 if prior == "stochastic" else 2

    def log_prior(q):
        if prior == "stochastic":
            return stochastic_log_prior(q)
        ok = all(np.isfinite(q[i]) and lower[i] <= q[i] <= upper[i] for i in range(2))
        return 0.0 if ok else -np.inf

    def like(q, lseed):
        a = np.exp(q[1]) if prior == "stochastic" else alpha
        if log_prior(q) == -np.inf or dt * (2 * np.exp(q[0]) * K * K + a) > 2.5:
            return -np.inf
        if prior == "stochastic":
            kw = dict(alpha=a, amp=0.0, sigma=np.exp(q[2]), tau=np.exp(q[3]))
        else:
            kw = dict(alpha=alpha, amp=np.exp(q[1]), sigma=sigma, tau=tau)
        return oracle_bootstrap_filter(orc, init, obs, dt, lseed, np.exp(q[0]), kf=kf, kmin=kmin,
                                       kmax=kmax, forcing=forcing, **kw)["log_likelihood"]

    out = []
    for c, qc in enumerate(np.asarray(q0, dtype=np.float64).reshape(-1, D)):
        s = int(chain_seeds[c]) if chain_seeds is not None else orc.stream_seed(seed, c)
        normal, random = orc.mt(s), orc.mt(orc.stream_seed(s, 1))
        cur = qc.copy()
        lseed = orc.mt_next(random)
        ll = like(cur, lseed)
        dens = ll + log_prior(cur)
        if not np.isfinite(dens):
            raise ValueError("initial particle posterior estimate is not finite")
        rows = []
        for _ in range(samples):
            prop = cur + scale * np.array([orc.normal(normal) for _ in range(D)])
            pseed = orc.mt_next(random)
            pll = like(prop, pseed)
            pd = pll + log_prior(prop)
            u = orc.mt_uniform(random)
            acc = bool(np.isfinite(pd) and np.log(max(u, np.finfo(float).tiny)) < min(0.0, pd - dens))
            if acc:
                cur, ll, dens, lseed = prop, pll, pd, pseed
            rows.append((cur.copy(), ll, dens, acc, lseed))
        out.append(rows)
    return out


def change_grid(field, n):
    """Retained modes of `field` on an n x n grid: modes both grids retain are
    copied, the others are zero (truncation / zero padding)."""
    ns = field.shape[0]
    K = min((n - 1) // 3, (ns - 1) // 3)
    out = np.zeros((n, n), dtype=np.complex128)
    for ky in range(-K, K + 1):
        for kx in range(-K, K + 1):
            if kx or ky:
                out[ky % n, kx % n] = field[ky % ns, kx % ns]
    return out


def oracle_joint_filter(orc, initial, obs, dt, seed, nu, alpha=0.0, amp=0.1, kf=4, sigma=0.05,
                        kmin=1, kmax=4):
    """Bootstrap filter over particles of several grid sizes; `initial` is a
    list of (P_e, n_e, n_e) arrays, observation indices on the coarsest grid."""
    states = [s.copy() for part in initial for s in np.asarray(part, dtype=np.complex128)]
    N = len(states)
    coarse = min(s.shape[0] for s in states)
    grid = [s.shape[0] for s in states]
    keys = [orc.stream_seed(seed, i) for i in range(N)]
    weights = np.full(N, 1.0 / N)
    rng = orc.mt(orc.stream_seed(seed, UINT64_MAX))
    steps = obs[:, 0].astype(np.int64)
    groups = np.unique(steps)
    current, total, events, history = 0, 0.0, 0, []
    for g, target in enumerate(groups):
        sel = obs[steps == target]
        logs = np.zeros(N)
        states = _each(lambda i: orc.trajectory(grid[i], nu, alpha, dt, int(target - current),
                                                keys[i], amp=amp, kf=kf, sigma=sigma, kmin=kmin,
                                                kmax=kmax, step0=current, initial=states[i]), N)
        for i in range(N):
            r = grid[i] // coarse
            ph = physical(orc, states[i])
            pred = ph[(sel[:, 2] * r).astype(int), (sel[:, 1] * r).astype(int)]
            logs[i] = orc.pf_logweight(weights[i], pred, sel[:, 3], sel[:, 4])
        inc, w, ess, mw = orc.pf_normalize(logs)
        total += inc
        resampled = False
        if ess < N / 2 and g + 1 < len(groups):
            anc = orc.systematic_resample(w, orc.mt_uniform(rng))
            states = [change_grid(states[a], grid[i]) for i, a in enumerate(anc)]
            keys = [orc.stream_seed(seed, (g + 1) * N + i) for i in range(N)]
            w = np.full(N, 1.0 / N)
            events += 1
            resampled = True
        weights = w
        history.append(dict(step=int(target), increment=inc, ess=ess, max_weight=mw,
                            resampled=resampled))
        current = int(target)
    return dict(log_likelihood=total, resampling_events=events, history=history, states=states)


def oracle_likelihood_adjoint(orc, initial, obs, dt, q, stride):
    """likelihood_adjoint (src/adjoint.cpp:43-161) from the oracle's pinned
    primitives: forward RK4 with checkpoints every `stride` steps, Gaussian
    residuals at the observations, then per block a replay and the reverse
    sweep of the RK4 stages through adjoint_drift, with the parameter terms
    -nu k^2 <K, S> and -alpha <K, S>.  Returns (value, gradient[2])."""
    n = initial.shape[0]
    q0, q1 = float(q[0]), float(q[1])
    if not all(np.isfinite([q0, q1])) or not (-12 <= q0 <= 0 and -12 <= q1 <= 0):
        return np.inf, np.zeros(2)
    nu, alpha = np.exp(q0), np.exp(q1)
    K = (n - 1) // 3
    if dt * (2 * nu * K * K + alpha) > 2.5:
        return np.inf, np.zeros(2)
    k = np.fft.fftfreq(n, 1.0 / n)
    kx, ky = np.meshgrid(k, k)
    ksq = kx * kx + ky * ky
    mask = orc.active_mask(n).astype(bool)
    steps = obs[:, 0].astype(np.int64)
    last = int(steps[-1])
    w = initial.copy()
    checkpoints, cstep = [w.copy()], [0]
    residual = np.zeros(len(obs))
    value, o = 0.0, 0
    for step in range(last + 1):
        if step > 0:
            w = orc.rk4(w, nu, alpha, dt, 1)
            if step % stride == 0 or step == last:
                checkpoints.append(w.copy())
                cstep.append(step)
        if o < len(obs) and steps[o] == step:
            field = physical(orc, w)
            while o < len(obs) and steps[o] == step:
                _, x, y, v, sd = obs[o]
                r = (field[int(y), int(x)] - v) / sd
                value += 0.5 * r * r
                residual[o] = r / sd
                o += 1
    drift = lambda s: orc.rhs(s, nu, alpha)
    grad = np.zeros(2)
    cot = np.zeros((n, n), dtype=np.complex128)
    o = len(obs)

    def inject(step):
        nonlocal o, cot
        if o == 0 or steps[o - 1] != step:
            return
        phys = np.zeros((n, n), dtype=np.complex128)
        while o > 0 and steps[o - 1] == step:
            o -= 1
            phys[int(obs[o, 2]), int(obs[o, 1])] += residual[o] * n * n
        cot = cot + np.where(mask, orc.fft2d(phys), 0)

    def contribution(state, weight):
        inner = np.real(np.conj(weight) * state)
        grad[0] -= nu * np.sum(ksq * inner)
        grad[1] -= alpha * np.sum(inner)

    for b in range(len(checkpoints) - 1, 0, -1):
        begin, end = cstep[b - 1], cstep[b]
        states = [checkpoints[b - 1]]
        for _ in range(begin, end):
            s = states[-1]
            k1 = drift(s)
            k2 = drift(s + dt / 2 * k1)
            k3 = drift(s + dt / 2 * k2)
            k4 = drift(s + dt * k3)
            states.append(s + dt / 6 * (k1 + 2.0 * k2 + 2.0 * k3 + k4))
        for step in range(end, begin, -1):
            inject(step)
            s1 = states[step - begin - 1]
            s2 = s1 + dt / 2 * drift(s1)
            s3 = s1 + dt / 2 * drift(s2)
            s4 = s1 + dt * drift(s3)
            prev = cot.copy()
            K1, K2, K3, K4 = cot * (dt / 6), cot * (dt / 3), cot * (dt / 3), cot * (dt / 6)
            contribution(s4, K4)
            b4 = orc.adjoint_drift(s4, K4, nu, alpha)
            prev += b4
            K3 = K3 + dt * b4
            contribution(s3, K3)
            b3 = orc.adjoint_drift(s3, K3, nu, alpha)
            prev += b3
            K2 = K2 + dt / 2 * b3
            contribution(s2, K2)
            b2 = orc.adjoint_drift(s2, K2, nu, alpha)
            prev += b2
            K1 = K1 + dt / 2 * b2
            contribution(s1, K1)
            prev += orc.adjoint_drift(s1, K1, nu, alpha)
            cot = prev
    return value, grad


def oracle_step_tangent(orc, w, tangents, nu, alpha, dt):
    """One Solver::step_tangent step (src/solver.cpp:324-380) from the pinned
    drift and tangent drift: RK4 of the state and of t' = L_X t - c_j X with
    c_0 = nu |k|^2, c_1 = alpha.  Returns (state, [t0, t1])."""
    n = w.shape[0]
    k = np.fft.fftfreq(n, 1.0 / n)
    kx, ky = np.meshgrid(k, k)
    cj = (nu * (kx * kx + ky * ky), alpha)
    drift = lambda s: orc.rhs(s, nu, alpha)

    def slope(X, T):
        return [orc.linearized_drift(X, T[j], nu, alpha) - cj[j] * X for j in range(2)]

    T = [np.asarray(t, dtype=np.complex128) for t in tangents]
    k1 = drift(w)
    t1 = slope(w, T)
    X = w + dt / 2 * k1
    k2 = drift(X)
    t2 = slope(X, [T[j] + dt / 2 * t1[j] for j in range(2)])
    X = w + dt / 2 * k2
    k3 = drift(X)
    t3 = slope(X, [T[j] + dt / 2 * t2[j] for j in range(2)])
    X = w + dt * k3
    k4 = drift(X)
    t4 = slope(X, [T[j] + dt * t3[j] for j in range(2)])
    w_new = w + dt / 6 * (k1 + 2.0 * k2 + 2.0 * k3 + k4)
    T_new = [T[j] + dt / 6 * (t1[j] + 2.0 * t2[j] + 2.0 * t3[j] + t4[j]) for j in range(2)]
    return w_new, T_new


def oracle_likelihood_gradient(orc, initial, obs, dt, q):
    """likelihood_gradient (src/inference.cpp:340-382): the objective of
    likelihood_adjoint with its gradient by forward sensitivities."""
    n = initial.shape[0]
    q0, q1 = float(q[0]), float(q[1])
    if not all(np.isfinite([q0, q1])) or not (-12 <= q0 <= 0 and -12 <= q1 <= 0):
        return np.inf, np.zeros(2)
    nu, alpha = np.exp(q0), np.exp(q1)
    K = (n - 1) // 3
    if dt * (2 * nu * K * K + alpha) > 2.5:
        return np.inf, np.zeros(2)
    w = initial.copy()
    T = [np.zeros_like(w), np.zeros_like(w)]
    step, cached = 0, -1
    value, grad = 0.0, np.zeros(2)
    for s_, x, y, v, sd in obs:
        while step < int(s_):
            w, T = oracle_step_tangent(orc, w, T, nu, alpha, dt)
            step += 1
        if cached != step:
            field = physical(orc, w)
            pt = [physical(orc, t) for t in T]
            cached = step
        r = (field[int(y), int(x)] - v) / sd
        value += 0.5 * r * r
        for j in range(2):
            grad[j] += r * pt[j][int(y), int(x)] / sd
    return value, grad
