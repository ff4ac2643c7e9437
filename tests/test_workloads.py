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


def test_reference_trajectory_oracle_closed_forms():
    """The oracle's restatement of the reference's candidate model
    (gridSearchErrorFunctions_dft.cuh:279-291) against what the formulas fix:
    pulse 0 sits at the constant terms; D = 7 moves by the constant step
    (p1, p3, p5) * p6 per pulse; D = 10 has constant second differences
    2 (p2, p5, p8) and first step (p1, p4, p7) * p9 + (p2, p5, p8)."""
    import oracle

    rng = np.random.default_rng(5)
    p7 = rng.standard_normal(7) * 100.0
    a = oracle.reference_trajectory_f64(p7, 7, 50)
    assert np.allclose(a[0], p7[[0, 2, 4]])
    assert np.allclose(np.diff(a, axis=0), (p7[[1, 3, 5]] * p7[6])[None, :])
    p10 = rng.standard_normal(10) * 10.0
    b = oracle.reference_trajectory_f64(p10, 10, 50)
    assert np.allclose(b[0], p10[[0, 3, 6]])
    assert np.allclose(np.diff(b, 2, axis=0), 2.0 * p10[[2, 5, 8]][None, :])
    assert np.allclose(b[1] - b[0], p10[[1, 4, 7]] * p10[9] + p10[[2, 5, 8]])


def test_reference_trajectory_candidates_reproduce_the_model():
    """The per-basis linear form that batch_entropy takes reproduces the
    reference's D = 7 and D = 10 candidate positions (FP64)."""
    import oracle
    from uncc_sar_focusing_amd.autofocus import reference_trajectory_candidates

    rng = np.random.default_rng(9)
    for D in (7, 10):
        params = rng.standard_normal((5, D)) * 50.0
        basis, direction, coefs = reference_trajectory_candidates(params, D, 300)
        assert basis.shape[0] == direction.shape[0] == coefs.shape[1] == (6 if D == 7 else 9)
        for j in range(5):
            pos = np.einsum("k,kn,knx->nx", coefs[j], basis, direction)
            ref = oracle.reference_trajectory_f64(params[j], D, 300)
            assert np.max(np.abs(pos - ref)) <= 1e-12 * np.max(np.abs(ref))
