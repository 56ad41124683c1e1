"""Step chains (hzl_ensemble_set_chains): member ranges stepped as chains of
launches on their own streams give the same states, bit for bit, as one chain
over all members, launched eagerly and replayed from a captured graph."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def run(pkg, n, members, precision, chains, graphs, tuning=None, steps=(3, 5, 5)):
    with pkg.Ensemble(n, members, precision, nu=1e-3, sigma=0.05, amp=0.1, seed=20261204) as e:
        e.set_graph(graphs)
        if tuning is not None:
            e.tuning = tuning
        e.chains = chains
        e.init_grf()
        for s in steps:  # the repeated size is captured, then replayed
            e.step(1e-3, s)
        e.synchronize()
        return e.get_state(), e.diagnostics()


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
@pytest.mark.parametrize("members", [6, 7])
def test_chains_match_one_chain(pkg, precision, members):
    want, dw = run(pkg, 64, members, precision, 1, True)
    for chains in (2, 3, 4):
        for graphs in (True, False):
            got, dg = run(pkg, 64, members, precision, chains, graphs)
            print(f"chains {chains} graphs {graphs}: max |diff| "
                  f"{np.max(np.abs(got - want)):.3e}", flush=True)
            assert np.array_equal(got, want)
            np.testing.assert_array_equal(dg, dw)


def test_chains_tuned_shapes_256(pkg):
    want, _ = run(pkg, 256, 4, "fp32", 1, True, tuning=(10, 10), steps=(2, 2))
    got, _ = run(pkg, 256, 4, "fp32", 2, True, tuning=(10, 10), steps=(2, 2))
    assert np.array_equal(got, want)


def test_autotune_picks_a_schedule_and_keeps_the_state(pkg):
    with pkg.Ensemble(128, 256, "fp32", nu=1e-3, sigma=0.05, amp=0.1, seed=3) as e:
        e.init_grf()
        before = e.get_state()
        e.autotune()
        assert e.chains in (1, 2, 4)
        assert np.array_equal(e.get_state(), before)


def test_chains_reject_bad_schedules(pkg):
    with pkg.Ensemble(32, 2, "fp32") as e:
        with pytest.raises(ValueError):
            e.chains = 3  # more chains than members
        with pytest.raises(ValueError):
            e.chains = 0
        with pytest.raises(ValueError):
            e.chains = 5
        assert e.chains == 1
