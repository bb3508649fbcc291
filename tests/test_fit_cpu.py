"""The genetic algorithm's host logic on the CPU (no GPU): offspring count,
gene deviation, roulette index, and the whole loop on an analytic decay
model that stands in for the GPU's batched decays."""
import numpy as np

from randomwalk_nmr_amd.fit import ga_offspring_count, genetic_fit, genotype_deviation, roulette_index


class AnalyticModel:
    """S_p(t) = exp(-t / T2b_p) sum_l w_l exp(-rho_pl c_l t): a stand-in with
    the interface genetic_fit uses (times, decays)."""

    def __init__(self, echoes=400):
        self.times = np.linspace(0.0, 800.0, echoes + 1)
        self.w = np.linspace(1.0, 2.0, 8) / np.linspace(1.0, 2.0, 8).sum()
        self.c = np.linspace(2e-5, 6e-5, 8)
        self.calls = 0

    def decays(self, rho, t2b):
        self.calls += 1
        rho = np.atleast_2d(rho)
        t2b = np.atleast_1d(t2b)
        e = np.exp(-rho[:, :, None] * self.c[None, :, None] * self.times[None, None, :])
        return np.exp(-self.times[None, :] / t2b[:, None]) * np.einsum("l,plt->pt", self.w, e)


def test_offspring_count_matches_reference_formula():
    # (proportion * (P / 2)) * 2 with integer P / 2 (B/GA/ga_core.cpp:80)
    assert ga_offspring_count(24, 0.5) == 12
    assert ga_offspring_count(64, 0.5) == 32
    assert ga_offspring_count(25, 0.5) == 12
    assert ga_offspring_count(64, 0.25) == 16


def test_genotype_deviation_by_hand():
    pop = np.array([[1.0, 2.0], [3.0, 2.0]])
    # gene 0: mean 2, population std 1 -> 0.5; gene 1: 0 -> mean 0.25
    assert genotype_deviation(pop) == 0.25
    rng = np.random.default_rng(0)
    pop = rng.uniform(1.0, 5.0, (50, 4))
    want = np.mean(pop.std(axis=0, ddof=0) / pop.mean(axis=0))
    assert abs(genotype_deviation(pop) - want) <= 1e-15 * want


def test_roulette_index_is_first_reaching_cumulative():
    cum = np.cumsum([0.1, 0.2, 0.3, 0.4])
    assert roulette_index(cum, 0.0) == 0
    assert roulette_index(cum, 0.1) == 0
    assert roulette_index(cum, 0.1000001) == 1
    assert roulette_index(cum, 0.65) == 3
    assert roulette_index(cum, 1.5) == 3      # rounding past the last sum stays in the population


def test_genetic_fit_converges_on_analytic_model():
    m = AnalyticModel()
    truth_g = np.array([12.0, 30.0, 9.0, 23.0, 6.0, 16.0, 32.0, 10.0, 2800.0])
    truth = m.decays(truth_g[None, :-1], truth_g[-1:])[0]
    lo = np.r_[np.full(8, 1.0), 500.0]
    hi = np.r_[np.full(8, 60.0), 5000.0]
    r = genetic_fit(m, truth, bounds_lo=lo, bounds_hi=hi, truth=truth, tol=1e-4, population=32, max_generations=100)
    assert r["converged"] and r["generations"] <= 100
    S = m.decays(r["genotype"][None, :-1], r["genotype"][-1:])[0]
    assert np.sqrt(np.mean((S - truth) ** 2)) <= 1e-4
    # every batch is one call: the initial population plus at most one offspring batch and one
    # diversification batch per generation
    assert m.calls <= 1 + 2 * r["generations"]


def test_genetic_fit_diversification_resets_and_bounds():
    m = AnalyticModel()
    truth_g = np.array([12.0, 30.0, 9.0, 23.0, 6.0, 16.0, 32.0, 10.0, 2800.0])
    truth = m.decays(truth_g[None, :-1], truth_g[-1:])[0]
    lo = np.r_[np.full(8, 1.0), 500.0]
    hi = np.r_[np.full(8, 60.0), 5000.0]
    r = genetic_fit(m, truth, bounds_lo=lo, bounds_hi=hi, truth=truth, tol=1e-7, population=24, max_generations=120,
                    seed=3)
    assert not r["converged"]
    # the population collapses and is reset after every 5 imposed mutations
    assert r["resets"] >= 1
    assert np.all(r["genotype"] >= lo) and np.all(r["genotype"] <= hi)
    # the fittest individual never gets worse from one generation to the next (elitist selection;
    # mutations and resets spare the top individual)
    errs = [h["rms_truth"] for h in r["history"]]
    assert all(b <= a for a, b in zip(errs, errs[1:]))
