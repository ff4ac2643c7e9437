"""Host-side workload builders (no GPU): seeded, in the stated ranges."""
import numpy as np

from uncc_sar_focusing_amd import workloads
from uncc_sar_focusing_amd.scenes import CONFIGS, normal, uniform


def test_uneven_scene_specs_ranges_and_determinism():
    a = workloads.uneven_scene_specs(64, seed=4321)
    b = workloads.uneven_scene_specs(64, seed=4321)
    assert len(a) == 64
    for (Np, K, side, ntg, nfft, c), (Np2, K2, side2, ntg2, nfft2, c2) in zip(a, b):
        assert (Np, K, side, ntg, nfft) == (Np2, K2, side2, ntg2, nfft2)
        assert np.array_equal(c, c2)
        assert 256 <= Np <= 4096 and 256 <= K <= 2048 and 128 <= side <= 1024
        assert 10 <= ntg <= 50 and nfft >= 4 * K // 2 and nfft <= 16384
        assert len(c) == 9
    sides = [s[2] for s in a]
    assert max(sides) - min(sides) > 500  # actually uneven


def test_smoothed_random_walk_statistics():
    w = workloads.smoothed_random_walk(1024, 0.05, seed=3)
    assert w.shape == (1024, 3)
    assert np.allclose(w.std(axis=0), 0.05)
    assert np.allclose(w.mean(axis=0), 0.0, atol=1e-12)
    # smooth: neighbouring pulses differ far less than the spread
    assert np.abs(np.diff(w, axis=0)).max() < 0.01


def test_rng_streams_are_reproducible_and_distinct():
    assert np.array_equal(uniform(1, 2, 10), uniform(1, 2, 10))
    assert not np.array_equal(uniform(1, 2, 10), uniform(1, 3, 10))
    z = normal(5, 1, 200000)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1) < 0.01


def test_configs_match_the_workload_sizes():
    c0, c1, c2 = CONFIGS["C0"], CONFIGS["C1"], CONFIGS["C2"]
    assert (c0.Np, c0.K, c0.nfft, c0.n_pix, c0.spacing, c0.n_targets) == (512, 424, 4096, 256, 0.1, 5)
    assert (c1.Np, c1.K, c1.nfft, c1.n_pix, c1.spacing, c1.n_targets) == (4096, 2048, 16384, 2048,
                                                                           0.05, 100)
    assert (c2.Np, c2.K, c2.n_pix, c2.n_targets) == (1024, 1024, 512, 20)
    g = c1.geometry()
    assert abs(np.linalg.norm(g.ant[0]) - 12500.0) < 1e-6
    span = np.degrees(np.arctan2(g.ant[-1, 1], g.ant[-1, 0]) - np.arctan2(g.ant[0, 1], g.ant[0, 0]))
    assert abs(abs(span) - 4.0) < 1e-9
