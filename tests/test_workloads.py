"""Synthetic inputs: seeded, deterministic, with the configured shapes and ranges."""
import numpy as np

from mpgos import workloads as W


def test_splitmix_stream_is_deterministic_and_uniform():
    a = W.Stream(W.SEED, 1).uniform(100000)
    b = W.Stream(W.SEED, 1).uniform(100000)
    c = W.Stream(W.SEED, 2).uniform(100000)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert 0.0 <= a.min() and a.max() < 1.0 and abs(a.mean() - 0.5) < 0.005
    # SplitMix64 known answer (Steele-Lea-Flood mixer): mix(0x9E3779B97F4A7C15) of state 0
    z = W._mix(np.array([0x9E3779B97F4A7C15], np.uint64))[0]
    assert int(z) == 0xE220A8397B1DCDAF


def test_normal_moments():
    z = W.Stream(W.SEED, 5).normal(200001, 0.1)
    assert abs(z.mean()) < 1e-3 and abs(z.std() - 0.1) < 1e-3


def test_linspace_matches_tutorial_definition():
    x = W.linspace(0.3, 10.0, 5)
    assert x[0] == 0.3 and x[-1] == 10.0 and x[1] == 0.3 + 1 * ((10.0 - 0.3) / 4)


def test_configs():
    c0 = W.config0_duffing_rk4()
    assert c0["nt"] == 16384 and c0["TD"][1, 0] == 64 * 2 * W.PI and c0["options"]["dt0"] == 2 * W.PI / 1000
    c1 = W.config1_lorenz(4096)
    assert 0 <= c1["cPAR"].min() and c1["cPAR"].max() <= 21 and abs(c1["X"].mean() - 10) < 0.01
    c2 = W.config2_duffing_uneven(4096)
    eps = c2["cPAR"][2]
    assert 1e-3 <= eps.min() and eps.max() <= 1.0
    periods = c2["TD"][1] * eps / (2 * W.PI)
    assert np.allclose(periods, np.round(periods)) and periods.min() >= 16 and periods.max() <= 4096
    c3 = W.config3_duffing_adjoint(1024)
    assert np.all(np.abs(c3["cPAR"][0] / 0.2 - 1) <= 0.05 + 1e-12)
