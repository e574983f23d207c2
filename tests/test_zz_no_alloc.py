"""mist_plic_dist_batch_noalloc (the batch staged in the library's per-GPU device array, no
allocation) against mist_plic_dist_batch (which allocates its staging) through the C ABI:
the same kernel on the same inputs, so bit-exact; two calls in a row, the second smaller, so
the array's leftovers from the first call must not leak into the second."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _batch(rng, m, N):
    ijk = rng.integers(1, N + 1, size=(m, 3)).astype(np.int64)
    n = rng.standard_normal((m, 3))
    n /= np.sqrt((n ** 2).sum(1))[:, None]
    vf = rng.random(m)
    vf[::97] = 0.0   # empty and full cells take computeDist's early return
    vf[1::97] = 1.0
    return ijk, n, vf


def test_noalloc_dist_batch_matches_allocating_batch():
    from ns_multiphase_amd import solver
    from ns_multiphase_amd.workloads import deformation3d
    ctx = solver.Context(deformation3d(16)[0])
    rng = np.random.default_rng(5)
    for m in (65536, 1000):
        ijk, n, vf = _batch(rng, m, 16)
        ref = ctx.plic_dist_batch(ijk, n, vf)
        got = ctx.plic_dist_batch_noalloc(ijk, n, vf)
        assert np.array_equal(got, ref), m
        assert np.isfinite(got).all()
    # a second context on the same GPU shares the array
    ctx2 = solver.Context(deformation3d(12)[0])
    ijk, n, vf = _batch(rng, 300, 12)
    assert np.array_equal(ctx2.plic_dist_batch_noalloc(ijk, n, vf), ctx2.plic_dist_batch(ijk, n, vf))


def test_noalloc_dist_batch_rejects_oversized_batch():
    from ns_multiphase_amd import solver
    from ns_multiphase_amd.workloads import deformation3d
    ctx = solver.Context(deformation3d(8)[0])
    ijk, n, vf = _batch(np.random.default_rng(6), 65537, 8)
    with pytest.raises(solver.MistError, match="exceeds the limit of 65536"):
        ctx.plic_dist_batch_noalloc(ijk, n, vf)
