"""CPU checks of the synthetic-input generators against the oracle's."""
import numpy as np

from neb_amd import synthetic as syn
from oracle import neb_oracle as o


def test_blob_domain_matches_oracle_sampler():
    a, ca, vol = syn.blob_domain(5000, seed=0, batch=1 << 14)
    b, cb = o.blob_domain_points(5000, seed=0, batch=1 << 14)
    assert np.array_equal(a, b) and np.array_equal(ca, cb)
    assert 0.0 < vol < 8.0


def test_curl_targets_match_oracle_and_are_divergence_free():
    p = np.random.default_rng(1).uniform(-1, 1, (7, 3))
    assert np.allclose(syn.curl_targets(p, 3, seed=2), o.fourier_potential_targets(p, 3, seed=2), rtol=0, atol=0)


def test_periodic_grid():
    g = syn.periodic_grid(8)
    assert np.array_equal(g, o.periodic_grid(8))
    assert g.shape == (64, 2) and g[1, 0] > 0 and g[1, 1] == 0  # x fastest
