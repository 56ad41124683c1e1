"""FP64 oracle drivers for the filter loops, built from the pinned oracle
primitives (oracle/hzl_oracle.c).  TEST ONLY.

oracle_bootstrap_filter restates reference bootstrap_filter
(src/filter.cpp:38-126) with the engine's forcing model (white shell noise keyed
per particle by Philox, Kolmogorov forcing) so that the GPU filter can be
compared group by group: same particle keys, same resampling uniforms
(mt19937_64(stream_seed(seed, UINT64_MAX))), same reseeding of children.
"""
from __future__ import annotations

import numpy as np

UINT64_MAX = (1 << 64) - 1


def physical(orc, field):
    return orc.fft2d(field, inverse=True).real


def oracle_bootstrap_filter(orc, initial, obs, dt, seed, nu, alpha=0.0, amp=0.1, kf=4,
                            sigma=0.05, kmin=1, kmax=4):
    """initial: (P, n, n) spectral particle states.  obs rows (step, x, y, value, sd)."""
    states = np.array(initial, dtype=np.complex128)
    P, n = states.shape[0], states.shape[1]
    keys = [orc.stream_seed(seed, i) for i in range(P)]
    weights = np.full(P, 1.0 / P)
    rng = orc.mt(orc.stream_seed(seed, UINT64_MAX))
    steps = obs[:, 0].astype(np.int64)
    groups = np.unique(steps)
    current, total, events, history = 0, 0.0, 0, []
    for g, target in enumerate(groups):
        sel = obs[steps == target]
        for i in range(P):
            states[i] = orc.trajectory(n, nu, alpha, dt, int(target - current), keys[i],
                                       amp=amp, kf=kf, sigma=sigma, kmin=kmin, kmax=kmax,
                                       step0=current, initial=states[i])
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
            w = np.full(P, 1.0 / P)
            events += 1
            resampled = True
        weights = w
        history.append(dict(step=int(target), increment=inc, ess=ess, max_weight=mw,
                            resampled=resampled, mean_field=mean))
        current = int(target)
    return dict(log_likelihood=total, resampling_events=events, history=history)
