"""Multi-block Poisson: the reference's multi-block test app
(apps/c/poisson/poisson.cpp) on the GPU, one block and partitions into
blocks joined by ops_halo_transfer.

Jacobi updates are pointwise, so a partition must give the same field bits
as one block (halo transfers carry exactly the neighbours' values); the
total error is a sum, within 1e-12.  Against the reference app's own output
(its known answer and fixtures at other sizes) the tolerance is 1e-12
relative: the device sin/cos may differ from glibc in the last bit."""
import numpy as np
import pytest

from oracle import oracle_py as O

pytestmark = pytest.mark.gpu


def test_known_answer_and_reference_runs(gpu_ops):
    from ops_amd import apps
    g, m = O.load_golden()
    c = m["mb_poisson"]
    err, _ = apps.mb_poisson(20, 20, 1, 1, 10)
    assert abs(err - c["expected_app"]) <= 1e-12 * c["expected_app"]
    for (nx, ny, it), ref in zip(c["configs"], g["mb_poisson_err"]):
        err, u = apps.mb_poisson(nx, ny, 1, 1, it)
        assert abs(err - ref) <= 1e-12 * ref, (nx, ny, it)
        e_o, u_o = O.mb_poisson(nx, ny, it)
        np.testing.assert_allclose(u, u_o, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("parts", [(2, 1), (1, 3), (2, 2), (3, 2), (4, 3)])
def test_partition_bit_identical(gpu_ops, parts):
    from ops_amd import apps
    nx, ny, it = 97, 61, 40
    e1, u1 = apps.mb_poisson(nx, ny, 1, 1, it)
    ep, up = apps.mb_poisson(nx, ny, parts[0], parts[1], it)
    assert np.array_equal(u1, up), parts
    assert abs(ep - e1) <= 1e-12 * e1
