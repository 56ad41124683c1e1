"""Write the golden vectors of tests/golden/ from the UNMODIFIED reference CPU
library (oracle/_ref/libhzlref.so, built by `make -C oracle ref` from the
reference's own sources).  Run in the container that holds the reference:

    python tests/golden/make_golden.py

Each fixture stores inputs and reference outputs only (numpy .npz, no pickles)
plus a JSON sidecar with the recipe and sha256.  The engine's physics is driven
through the reference's public API (frozen OU forcing + update_latent_state
before every step; see oracle/ref_shim.cpp).

Fixtures
  trajectory_64.npz 64x64, nu 1e-3, dt 1e-3, alpha 0,
                   0.1 sin(4y) + white noise sigma 0.05 on 1<=|k|<=4, GRF start
                   (member 0 of seed 20261203); reference states at steps
                   0 / 100 / 1000 as retained half planes [kx 0..K][ky -K..K],
                   reference diagnostics at the same steps.
  drift_multimode_32.npz  reference Solver::drift of the multimode state, 32x32.
  pf_deterministic_16.npz reference bootstrap_filter / negative_log_likelihood /
                   ensemble_kalman_filter on the deterministic dataset of
                   tests/verification.cpp:184-210 (16x16).
  checkpoint_v1_16.txt  a reference Solver::save file (16x16, OU forcing on,
                   7 steps from multimode); checkpoint_v1_16.npz holds the
                   reference's Solver::load fields of it and the JSON sidecar its
                   diagnostics.

  adjoint_16.json  the reference's likelihood_adjoint (value, gradient) on the
                   pf_deterministic_16 dataset at four parameter vectors, strides 7, 4,
                   and its likelihood_gradient (forward sensitivities) at the same vectors.

Usage: python tests/golden/make_golden.py [trajectory64 drift32 pf16 checkpoint16 adjoint16]
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from oracle import Oracle, Reference  # noqa: E402


def half(field: np.ndarray) -> np.ndarray:
    n = field.shape[0]
    K = (n - 1) // 3
    ky = np.arange(-K, K + 1) % n
    return np.ascontiguousarray(field[ky][:, : K + 1].T)  # [kx][ky + K]


def save(name: str, recipe: dict, **arrays) -> None:
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **arrays)
    digest = hashlib.sha256(open(path, "rb").read()).hexdigest()
    meta = dict(recipe, file=name + ".npz", sha256=digest,
                arrays={k: [list(v.shape), str(v.dtype)] for k, v in arrays.items()})
    with open(os.path.join(HERE, name + ".json"), "w") as f:
        json.dump(meta, f, indent=2, sort_keys=True)
        f.write("\n")
    print("wrote", path, digest[:16])


def trajectory64(orc: Oracle, ref: Reference) -> None:
    n, nu, dt, seed = 64, 1e-3, 1e-3, 20261203
    key = orc.stream_seed(seed, 0)
    w0 = orc.grf(n, key)
    F = np.stack([orc.step_forcing(n, 0.1, 4, 0.05, 1, 4, dt, key, k) for k in range(1000)])
    w100 = ref.run(w0, nu, 0.0, dt, 100, F[:100], per_step=True)
    w1000 = ref.run(w100, nu, 0.0, dt, 900, F[100:], per_step=True)
    d = [ref.diagnostics(w, nu, 0.0) for w in (w0, w100, w1000)]
    diag = np.array([[x["energy"], x["enstrophy"], x["dissipation"], x["cfl_rate"]] for x in d])
    save("trajectory_64",
         dict(recipe="reference Solver, frozen OU + update_latent_state(F_step) per step",
              n=n, nu=nu, alpha=0.0, dt=dt, seed=seed, member=0, amp=0.1, kf=4, sigma=0.05,
              kmin=1, kmax=4, steps=[0, 100, 1000],
              diag_columns=["energy", "enstrophy", "dissipation", "cfl_rate"]),
         w0=half(w0), w100=half(w100), w1000=half(w1000), diag=diag)


def drift32(ref: Reference) -> None:
    n = 32
    w = ref.initialize(n, "multimode")
    d = ref.drift(w, 0.01, 0.1)
    save("drift_multimode_32", dict(recipe="reference Solver::drift of initialize('multimode')",
                                    n=n, nu=0.01, alpha=0.1),
         state=half(w), drift=half(d))


def pf16(ref: Reference) -> None:
    nu, alpha = 0.03, 0.1
    h, rows = ref.synthetic(16, 0.01, 30, 10, nu, alpha, 0.0, 0.5, 2, 3, 0.2, 987)
    pf = ref.bootstrap_filter(h, nu, alpha, 0.0, 0.5, 8, 555, means=True, n=16)
    nll = ref.negative_log_likelihood(h, np.log(nu), np.log(alpha))
    enkf = ref.enkf(h, nu, alpha, 0.0, 0.5, 8, 555, n=16)
    init = ref.initialize(16, "multimode")
    # the same dataset written by the reference's Dataset::save
    ref.dataset_save(h, os.path.join(HERE, "observations_v2_16.csv"))
    ref.dataset_free(h)
    save("pf_deterministic_16",
         dict(recipe="synthetic_observations(16, 0.01, 30, 10, p, 0.2, 987) with nu 0.03, "
                     "alpha 0.1, sigma 0; bootstrap_filter(data, p, 8, 555); "
                     "ensemble_kalman_filter(data, p, 8, 555)",
              n=16, dt=0.01, nu=nu, alpha=alpha, particles=8, seed=555,
              log_likelihood=pf["log_likelihood"], resampling_events=pf["resampling_events"],
              negative_log_likelihood=nll),
         rows=rows, records=pf["records"], means=pf["means"], initial=half(init),
         enkf_records=enkf["records"], enkf_means=enkf["means"])


def ou16(ref: Reference) -> None:
    """The reference's OU shell forcing: a Solver trajectory with sigma > 0 and
    a bootstrap filter over a stochastic synthetic dataset."""
    n, nu, alpha, sigma, tau, dt = 16, 0.03, 0.1, 0.2, 0.5, 0.01
    init = ref.initialize(n, "multimode")
    runs = {}
    for seed, kmin, kmax, steps in ((77, 2, 3, 25), (123456789, 1, 4, 40)):
        w, F = ref.run_ou(init, nu, alpha, sigma, tau, kmin, kmax, seed, dt, steps)
        runs[f"run_{seed}"] = (w, F, dict(seed=seed, kmin=kmin, kmax=kmax, steps=steps))
    h, rows = ref.synthetic(n, dt, 60, 10, nu, alpha, sigma, tau, 2, 3, 0.05, 987)
    pf = ref.bootstrap_filter(h, nu, alpha, sigma, tau, 8, 555, means=True, n=n)
    enkf = ref.enkf(h, nu, alpha, sigma, tau, 8, 555, n=n)
    enkf_loc = ref.enkf(h, nu, alpha, sigma, tau, 8, 555, inflation=1.05, radius=3.0, n=n)
    q0 = np.log([nu, alpha, sigma, tau])
    chain = ref.particle_metropolis(h, q0, 0.1, 6, 8, 123)
    # the same stochastic dataset as written by the reference's Dataset::save
    ref.dataset_save(h, os.path.join(HERE, "ou_observations_16.csv"))
    ref.dataset_free(h)
    energies = ref.ensemble(n, nu, alpha, sigma, tau, 2, 3, dt, 20, 4, 2, 42)
    arrays = dict(initial=half(init), rows=rows, records=pf["records"], means=pf["means"],
                  chain_q=np.array([c["q"] for c in chain]),
                  chain_values=np.array([[c["log_likelihood"], c["log_posterior"]]
                                         for c in chain]),
                  chain_seeds=np.array([c["likelihood_seed"] for c in chain], dtype=np.uint64),
                  chain_accepted=np.array([c["accepted"] for c in chain]),
                  ensemble_energy=energies, enkf_records=enkf["records"],
                  enkf_means=enkf["means"], enkf_loc_records=enkf_loc["records"],
                  enkf_loc_means=enkf_loc["means"])
    for k, (w, F, _) in runs.items():
        arrays[k + "_state"] = half(w)
        arrays[k + "_forcing"] = half(F)
    # a stochastic run checkpointed mid-way (Solver::save), and 15 steps later
    ref.checkpoint_ou(n, nu, alpha, sigma, tau, 2, 3, 31337, dt, 10, 15,
                      os.path.join(HERE, "checkpoint_ou_before_16.txt"),
                      os.path.join(HERE, "checkpoint_ou_after_16.txt"))
    save("ou_16",
         dict(recipe="Solver(16, p, seed) with OU forcing from initialize('multimode'), "
                     "`steps` x Solver::step(dt) (ref_run_ou); synthetic_observations(16, "
                     "0.01, 60, 10, p, 0.05, 987) with sigma 0.2, tau 0.5 and "
                     "bootstrap_filter(data, p, 8, 555); ensemble_kalman_filter(data, p, "
                     "8, 555) plain and with inflation 1.05, localisation radius 3; "
                     "particle_metropolis(data, "
                     "log(nu, alpha, sigma, tau), 0.1, 6, 8, 123); ensemble(16, p, 0.01, 20, "
                     "4, 2, 42) energies; Solver(16, p, 31337) from multimode checkpointed "
                     "after 10 steps of 0.01 and again after 15 more (checkpoint_ou_*_16.txt)",
              n=n, nu=nu, alpha=alpha, sigma=sigma, tau=tau, dt=dt, particles=8, seed=555,
              runs=[v[2] for v in runs.values()], log_likelihood=pf["log_likelihood"],
              resampling_events=pf["resampling_events"]),
         **arrays)


# The reference CLI's flows (src/main.cpp:384-526), run by the reference's own
# executable (oracle/_ref/hzl_ref_cli, make -C oracle ref-cli) in a scratch
# directory; every file it writes is kept under cli/<case>/.
CLI_CASES = {
    "filter_particle": ["filter", "--data", "{golden}/ou_observations_16.csv", "--particles", "8",
                        "--seed", "555", "--nu", "0.03", "--alpha", "0.1", "--sigma", "0.2",
                        "--tau", "0.5"],
    "filter_enkf_ou": ["filter", "--method", "enkf", "--data", "{golden}/ou_observations_16.csv",
                       "--particles", "8", "--seed", "555", "--nu", "0.03", "--alpha", "0.1",
                       "--sigma", "0.2", "--tau", "0.5", "--inflation", "1.05",
                       "--localization-radius", "3", "--forecast-steps", "3"],
    "filter_enkf": ["filter", "--method", "enkf", "--data", "{golden}/observations_v2_16.csv",
                    "--particles", "8", "--seed", "555", "--nu", "0.03", "--alpha", "0.1",
                    "--sigma", "0", "--inflation", "1.02", "--forecast-steps", "5",
                    "--forecast-every", "2"],
    "particle_sample": ["particle-sample", "--data", "{golden}/ou_observations_16.csv",
                        "--samples", "4", "--particles", "8", "--proposal", "0.1", "--seed",
                        "123"],
    "ensemble": ["ensemble", "--n", "16", "--steps", "20", "--count", "4", "--workers", "2",
                 "--dt", "0.01", "--seed", "42", "--nu", "0.03", "--alpha", "0.1", "--sigma",
                 "0.2", "--tau", "0.5"],
}


def cli16() -> None:
    import shutil
    import subprocess
    import tempfile
    exe = os.path.join(ROOT, "oracle", "_ref", "hzl_ref_cli")
    if not os.path.exists(exe):
        raise SystemExit("build the reference CLI first: make -C oracle ref-cli")
    for case, argv in CLI_CASES.items():
        args = [a.replace("{golden}", HERE) for a in argv]
        with tempfile.TemporaryDirectory() as tmp:
            p = subprocess.run([exe, *args, "--output", "out"], cwd=tmp, capture_output=True,
                               text=True, check=True)
            dst = os.path.join(HERE, "cli", case)
            shutil.rmtree(dst, ignore_errors=True)
            shutil.copytree(os.path.join(tmp, "out"), dst)
            with open(os.path.join(dst, "stdout.txt"), "w") as f:
                f.write(p.stdout)
        print("wrote", dst)
    with open(os.path.join(HERE, "cli", "cases.json"), "w") as f:
        json.dump(CLI_CASES, f, indent=2, sort_keys=True)
        f.write("\n")


def checkpoint16(ref: Reference) -> None:
    """A reference Solver::save file (OU forcing active, random stream
    advanced) and the reference's own Solver::load diagnostics of it."""
    path = os.path.join(HERE, "checkpoint_v1_16.txt")
    recipe = dict(n=16, nu=0.02, alpha=0.1, sigma=0.3, tau=0.5, kmin=2, kmax=3, seed=4242,
                  initial="multimode", dt=0.01, steps=7)
    ref.checkpoint_write(path, recipe["n"], recipe["nu"], recipe["alpha"], recipe["sigma"],
                         recipe["tau"], recipe["kmin"], recipe["kmax"], recipe["seed"],
                         recipe["initial"], recipe["dt"], recipe["steps"])
    digest = hashlib.sha256(open(path, "rb").read()).hexdigest()
    w, F = ref.checkpoint_fields(path, recipe["n"])
    save("checkpoint_v1_16",
         dict(recipe, recipe_text="hzl::Solver(n, p, seed); initialize(initial); steps x "
                                  "step(dt); save(path); arrays: Solver::load(path)."
                                  "coefficients() / forcing()",
              checkpoint_file="checkpoint_v1_16.txt", checkpoint_sha256=digest,
              reference_load_diagnostics=ref.checkpoint_diagnostics(path)),
         vorticity=w, forcing=F)


def adjoint16(ref: Reference) -> None:
    """The reference's likelihood_adjoint on the deterministic 16x16 dataset
    of pf_deterministic_16 at a few (log nu, log alpha), two strides."""
    h, rows = ref.synthetic(16, 0.01, 30, 10, 0.03, 0.1, 0.0, 0.5, 2, 3, 0.2, 987)
    qs = [(float(np.log(0.03)), float(np.log(0.1))), (float(np.log(0.02)), float(np.log(0.15))),
          (float(np.log(0.05)), float(np.log(0.05))), (-13.0, -1.0)]
    out = []
    for stride in (7, 4):
        for q in qs:
            r = ref.likelihood_adjoint(h, q[0], q[1], stride)
            out.append(dict(q=list(q), stride=stride, value=r["value"],
                            gradient=[float(x) for x in r["gradient"]],
                            checkpoint_fields=r["checkpoint_fields"],
                            replayed_steps=r["replayed_steps"]))
    tangent = []
    for q in qs:
        v, g = ref.likelihood_gradient(h, q[0], q[1])
        tangent.append(dict(q=list(q), value=v, gradient=[float(x) for x in g]))
    ref.dataset_free(h)
    meta = dict(recipe="likelihood_adjoint(synthetic_observations(16, 0.01, 30, 10, p, 0.2, "
                       "987), q, stride) with nu 0.03, alpha 0.1, sigma 0; initial state "
                       "initialize('multimode'); observations = pf_deterministic_16 rows",
                n=16, dt=0.01, evaluations=out, likelihood_gradient=tangent)
    with open(os.path.join(HERE, "adjoint_16.json"), "w") as f:
        json.dump(meta, f, indent=2, sort_keys=True)
        f.write("\n")
    print("wrote adjoint_16.json")


def main() -> None:
    orc, ref = Oracle(), Reference()
    which = sys.argv[1:] or ["trajectory64", "drift32", "pf16", "checkpoint16", "adjoint16", "ou16",
                             "cli16"]
    if "trajectory64" in which:
        trajectory64(orc, ref)
    if "drift32" in which:
        drift32(ref)
    if "pf16" in which:
        pf16(ref)
    if "checkpoint16" in which:
        checkpoint16(ref)
    if "adjoint16" in which:
        adjoint16(ref)
    if "ou16" in which:
        ou16(ref)
    if "cli16" in which:
        cli16()


if __name__ == "__main__":
    main()
