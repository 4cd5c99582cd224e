"""The package's benchmark generator (fmm_amd.points) against the oracle's restatement of
fmm-validate's SplitMix64 and distributions, and the GPU direct sum against the oracle's
compensated direct sum."""
import numpy as np
import pytest

from fmm_amd import points as P
from oracle import oracle as O


def test_splitmix_and_cube_bit_exact():
    a, b = P.SplitMix64(P.BENCH_SEED), O.Rng(O.BENCH_SEED)
    assert np.array_equal(a.cube(5000), b.cube(5000))
    assert np.array_equal(a.cube(7, (1.0, 2.0, 3.0), 0.25), b.cube(7, (1.0, 2.0, 3.0), 0.25))


def test_normals_and_sample():
    a, b = P.SplitMix64(99), O.Rng(99)
    # numpy's log and cos may differ from the C library's in the last bit
    assert np.allclose(a.normals(4001), b.normals(4001), rtol=4e-16, atol=0)
    assert np.array_equal(a.sample(100000, 4096), b.sample(100000, 4096))
    assert np.array_equal(a.cube(10), b.cube(10))  # the streams stayed in step


def test_bench_problem_matches_oracle():
    pts, q, sample = P.bench_cube(200000, 1000)
    rp, rq, rs = O.bench_problem(200000, 1000, charges="normal")
    assert np.array_equal(pts, rp)
    assert np.allclose(q, rq, rtol=4e-16, atol=0)
    assert np.array_equal(sample, rs)
    pts, q, sample = P.bench_cube(100000, 1000, charges="uniform")
    rp, rq, rs = O.bench_problem(100000, 1000)
    assert np.array_equal(pts, rp) and np.array_equal(q, rq) and np.array_equal(sample, rs)


def test_gaussian_blob_matches_oracle_clusters():
    a, b = P.SplitMix64(7), O.Rng(7)
    x = a.gaussian_blob(3000, (0.5, 0.4, 0.6), 0.1)
    y = b.gaussian_clusters(3000, [(0.5, 0.4, 0.6)], 0.1)
    assert np.allclose(x, y, rtol=1e-15, atol=1e-16)
    assert np.array_equal(a.cube(5), b.cube(5))


@pytest.mark.gpu
def test_gpu_direct_sum_matches_oracle():
    import fmm_amd
    if fmm_amd.device_count() < 1:
        pytest.skip("no GPU")
    rng = O.Rng(5)
    src = np.vstack([rng.cube(20000), rng.cube(50)])
    src[-50:] = src[:50]  # coincident pairs are skipped
    q = rng.normals(len(src))
    tgt = np.vstack([src[:300], rng.cube(200)])
    phi, grad = fmm_amd.direct_sum(src, q, tgt)
    rphi, rgrad = O.direct_sum(src, q, tgt)
    assert np.allclose(phi, rphi, rtol=1e-14, atol=1e-12)
    assert np.allclose(grad, rgrad, rtol=1e-13, atol=1e-10)
    # the known answer of fmm-ref/src/p2p.rs:51-60
    p2, g2 = fmm_amd.direct_sum(np.array([[0.0, 0, 0], [3, 4, 0]]), np.array([2.0, -1.0]),
                                np.array([[0.0, 0, 0], [3, 4, 0]]))
    assert np.allclose(p2, [-1 / 5, 2 / 5], rtol=1e-15)
    assert np.allclose(g2[1], [-6 / 125, -8 / 125, 0], rtol=1e-15, atol=1e-18)


def test_sphere_and_plummer_match_the_oracle_bit_for_bit():
    """points::sphere and points::plummer (fmm-validate/src/points.rs:48-100): the package's
    vectorised parse of the stream against the oracle's point-by-point draws, and the
    stream left at the same state."""
    from fmm_amd import points as P

    for seed in (1, 99, 0x00BE9C06):
        r1, r2 = P.SplitMix64(seed), O.Rng(seed)
        assert np.array_equal(P.sphere(r1, 3000, (0.5, 0.5, 0.5), 0.5), r2.sphere(3000, (0.5, 0.5, 0.5), 0.5))
        assert np.array_equal(P.plummer(r1, 4000, (0.5, 0.5, 0.5), 0.1), r2.plummer(4000, (0.5, 0.5, 0.5), 0.1))
        assert np.array_equal(r1.uniform(5), np.array([r2.uniform() for _ in range(5)]))
