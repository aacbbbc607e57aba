"""cudaBVH_buildTLASFast on the GPU (lbvh.hip): the LBVH of the reference's
TLAS fast path (layer/cuda_bvh_backend.cu:257-389, :1620-1698), nodes (the
32-byte stackless DFS records) and leaf order byte-identical to the oracle's
host restatement, including tied Morton codes (stable order), boxes whose
centroids coincide on an axis (zero extent), signed zeros and sizes around
the radix sort's 2048-key tiles."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _boxes(rng, n, spread=10.0):
    lo = rng.uniform(-spread, spread, (n, 3)).astype(np.float32)
    return np.concatenate([lo, lo + rng.uniform(0.1, 3, (n, 3)).astype(np.float32)], axis=1)


@pytest.mark.parametrize("n", [1, 2, 3, 64, 777, 2047, 2048, 2049, 100_000])
def test_lbvh_matches_oracle(pkg, oracle, dev, n):
    aabbs = _boxes(np.random.default_rng(21 + n), n)
    nodes, order = pkg.build_tlas_fast(aabbs)
    onodes, oorder = oracle.lbvh_tlas(aabbs)
    assert nodes.shape == (2 * n - 1, 8)
    assert np.array_equal(order, oorder)
    assert np.array_equal(nodes, onodes)


def test_lbvh_ties_zero_extent_and_signed_zeros(pkg, oracle, dev):
    rng = np.random.default_rng(5)
    # many boxes on a coarse grid: equal Morton codes keep input order
    g = rng.integers(0, 4, (3000, 3)).astype(np.float32)
    a = np.concatenate([g, g + 1], axis=1)
    for aabbs in (a, a[:, [0, 1, 2, 3, 4, 5]] * np.float32(0)):        # second: every centroid equal
        nodes, order = pkg.build_tlas_fast(aabbs)
        onodes, oorder = oracle.lbvh_tlas(aabbs)
        assert np.array_equal(order, oorder) and np.array_equal(nodes, onodes)
    # signed zeros on box faces: merges follow the host's fminf / fmaxf
    z = _boxes(rng, 500, 1.0)
    z[rng.random((500, 6)) < 0.3] = np.float32(0.0)
    z[rng.random((500, 6)) < 0.3] = np.float32(-0.0)
    nodes, order = pkg.build_tlas_fast(z)
    onodes, oorder = oracle.lbvh_tlas(z)
    assert np.array_equal(order, oorder) and np.array_equal(nodes, onodes)


def test_lbvh_bad_arguments(pkg, dev):
    with pytest.raises(pkg.CudaBVHError):
        pkg.build_tlas_fast(np.zeros((0, 6), np.float32))
