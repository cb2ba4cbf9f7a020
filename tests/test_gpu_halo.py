"""ops_halo_transfer on the GPU against the reference library's own
transfers (fixtures written by oracle/refdrv/ref_halo.cpp through libops_seq,
ops/c/src/sequential/ops_host_singlenode.cpp:172-446): the multi-block app's
five halo groups (apps/c/mblock/mblock.cpp:60-157), 3D permuted and flipped
axes with unequal halo depths, an overlapping self-transfer, two components
per point, float and double dats meeting, and x-padded rows.  Every dat's
whole array (halos included) must match bit for bit after every group."""
import numpy as np
import pytest

from oracle import oracle_py as O

pytestmark = pytest.mark.gpu

CASES = ["mblock", "perm3d", "multi", "mixed", "wide"]


def _as_whole(a, ndim, dim):
    # fixture [z, y, x, component] -> get_whole's [(z,) y, x(, component)]
    if ndim == 2:
        a = a[0]
    return a[..., 0] if dim == 1 else a


@pytest.mark.parametrize("name", CASES)
def test_halo_transfer_bit_exact(gpu_ops, name):
    ops = gpu_ops
    g, m = O.load_golden()
    case = [c for c in m["halo"] if c["name"] == name][0]
    fx = O.halo_fixture(g, case)
    nd = case["ndim"]
    dats, dims = {}, {}
    for dname, typ, dim, size, d_m, d_p in case["dats"]:
        blk = ops.ops_decl_block(nd, f"{name}_{dname}")
        dt = np.float32 if typ == "float" else np.float64
        d = ops.ops_decl_dat(blk, size, d_m, d_p, dtype=dt, name=dname, dim=dim)
        ops.set_whole(d, _as_whole(fx[dname][0], nd, dim))
        dats[dname], dims[dname] = d, dim
    for dname, d in dats.items():
        assert np.array_equal(ops.get_whole(d), _as_whole(fx[dname][0], nd, dims[dname]))
    for k, grp in enumerate(case["groups"], start=1):
        hs = [ops.ops_decl_halo(dats[f], dats[t], it, fb, tb, fd, td) for f, t, it, fb, tb, fd, td in grp]
        ops.ops_halo_transfer(ops.ops_decl_halo_group(hs))
        for dname, d in dats.items():
            want = _as_whole(fx[dname][k], nd, dims[dname])
            got = ops.get_whole(d)
            assert np.array_equal(got, want), (name, k, dname, np.argwhere(got != want)[:5].tolist())
    for d in dats.values():
        ops.ops_free_dat(d)


def test_halo_box_outside_dat_is_an_error(gpu_ops):
    ops = gpu_ops
    blk = ops.ops_decl_block(2, "hb_err")
    a = ops.ops_decl_dat(blk, [6, 5], name="a")
    b = ops.ops_decl_dat(blk, [6, 5], name="b")
    # three columns from x = 5 run past a's one-deep halo
    h = ops.ops_decl_halo(a, b, [3, 5], [5, 0], [-1, 0], [1, 2], [1, 2])
    with pytest.raises(ops.OPSError, match="leaves dat"):
        ops.ops_halo_transfer(ops.ops_decl_halo_group([h]))
    # a direction naming a dimension the block does not have
    h = ops.ops_decl_halo(a, b, [1, 5], [5, 0], [-1, 0], [1, 3], [1, 2])
    with pytest.raises(ops.OPSError, match="bad direction"):
        ops.ops_halo_transfer(ops.ops_decl_halo_group([h]))


def test_halo_transfer_soa_equals_aos(gpu_ops):
    """Two-component dats in SoA layout (ops_set_soa) exchange the same
    values as in AoS: a flipped, permuted 3D transfer between two dats."""
    ops = gpu_ops
    n = (9, 7, 5)
    out = []
    for soa in (0, 1):
        blk = ops.ops_decl_block(3, f"hsoa{soa}")
        ops.ops_set_soa(soa)
        try:
            a = ops.ops_decl_dat(blk, list(n), d_m=[-2, -1, -1], d_p=[2, 1, 1], name=f"ha{soa}", dim=2)
            b = ops.ops_decl_dat(blk, list(n), d_m=[-2, -1, -1], d_p=[2, 1, 1], name=f"hb{soa}", dim=2)
        finally:
            ops.ops_set_soa(0)
        whole = (n[2] + 2, n[1] + 2, n[0] + 4, 2)
        ops.set_whole(a, np.arange(np.prod(whole), dtype=np.float64).reshape(whole))
        ops.set_whole(b, -np.ones(whole))
        h = ops.ops_decl_halo(a, b, [3, 4, 2], [1, 0, 2], [-2, 1, 0], [2, -1, 3], [1, 2, -3])
        ops.ops_halo_transfer(ops.ops_decl_halo_group([h]))
        out.append(ops.get_whole(b))
    assert np.array_equal(out[0], out[1])
    assert np.any(out[0] != -1)
