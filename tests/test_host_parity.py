"""CPU: the product's host algorithms (DOF numbering, CSR pattern, gather map,
AMG coarsening -- built without a device) against the oracle, bit for bit."""
import numpy as np
import pytest

import afepack_amd as afe
from oracle import oracle_py as orc

CASES = [(1, 32, 0.0, 4), (1, 16, 0.2, 4), (2, 12, 0.2, 6), (2, 8, 0.0, 6), (3, 5, 0.0, 2)]


def _mesh(kind, n, jit, seed=7):
    return afe.mesh_cube(n, jit, seed) if kind == 3 else afe.mesh_square(n, jit, seed)


@pytest.mark.parametrize("kind,n,jit,acc", CASES)
def test_numbering_and_pattern(kind, n, jit, acc):
    c, ev, vm = _mesh(kind, n, jit)
    hs = afe.Space(None, kind, c, ev, vm)
    os_ = orc.Space(kind, c, ev, vm)
    ed, dm, dp = hs.dofs()
    oed, odm, odp = os_.dofs()
    assert np.array_equal(ed, oed) and np.array_equal(dm, odm)
    assert np.abs(dp - odp).max() <= 1e-15
    rp, col = hs.pattern()
    orp, ocol, _ = os_.build(orc.OP_STIFF, acc).csr()
    assert np.array_equal(rp, orp) and np.array_equal(col, ocol)


@pytest.mark.parametrize("lazy", [0, 1])
@pytest.mark.parametrize("kind,n,jit,acc", CASES)
def test_hierarchy_bit_identical(lazy, kind, n, jit, acc):
    c, ev, vm = _mesh(kind, n, jit)
    os_ = orc.Space(kind, c, ev, vm)
    A = os_.build(orc.OP_STIFF, acc)
    f = os_.rhs(acc, orc.F_POISSON)
    os_.apply_dirichlet(A, f)
    rp, col, val = A.csr()
    H = afe.HostAmg(rp, col, val, afe.AmgParams(lazy=lazy, build_fp32=0))
    g = orc.Amg(A, lazy=bool(lazy))
    info = H.info()
    assert info["levels"] == g.n_project + 1 and info["coarse_full"] == g.full
    for l in range(info["levels"]):
        _, _, r1, c1, v1 = H.level_csr(l, 0)
        r2, c2, v2 = g.A(l).csr()
        assert np.array_equal(r1, r2) and np.array_equal(c1, c2) and np.array_equal(v1, v2)
        if l < info["levels"] - 1:
            _, _, r1, c1, v1 = H.level_csr(l, 1)
            r2, c2, v2 = g.P(l).csr()
            assert np.array_equal(r1, r2) and np.array_equal(c1, c2) and np.array_equal(v1, v2)
    if g.full:
        assert np.array_equal(H.coarse_inverse(), g.coarse_inverse())


@pytest.mark.parametrize("nmpo,nmpsd", [(50, 0.382), (200, 0.0), (10, 0.9)])
def test_hierarchy_parameters(nmpo, nmpsd):
    c, ev, vm = _mesh(1, 24, 0.0)
    os_ = orc.Space(1, c, ev, vm)
    A = os_.build(orc.OP_STIFF, 4)
    rp, col, val = A.csr()
    H = afe.HostAmg(rp, col, val, afe.AmgParams(nmpo=nmpo, nmpsd=nmpsd, build_fp32=0))
    g = orc.Amg(A, nmpo=nmpo, nmpsd=nmpsd)
    assert H.info()["n"] == [g.A(l).m for l in range(g.n_project + 1)]


def test_lazy_reinit_host_matches_oracle():
    c, ev, vm = _mesh(2, 10, 0.2)
    os_ = orc.Space(2, c, ev, vm)
    A = os_.build(orc.OP_STIFF, 6)
    rp, col, val = A.csr()
    H = afe.HostAmg(rp, col, val, afe.AmgParams(lazy=1, build_fp32=0))
    g = orc.Amg(A, lazy=True)
    # moving-mesh style update: new values on the same pattern
    c2 = afe.mesh_displace(2, c, vm, 0.1 / 10, 2)
    os2 = orc.Space(2, c2, ev, vm)
    val2 = os2.build(orc.OP_STIFF, 6).csr()[2]
    H.lazy_reinit(val2)
    B = A.share_pattern(val2)
    assert g.lazy_reinit(B) == 0
    for l in range(H.info()["levels"]):
        assert np.array_equal(H.level_csr(l, 0)[4], g.A(l).csr()[2])


def test_mesh_generators():
    c, ev, vm = afe.mesh_square(8, 0.2, 3)
    p = c.reshape(-1, 2)
    t = ev.reshape(-1, 3)
    a = (p[t[:, 1], 0] - p[t[:, 0], 0]) * (p[t[:, 2], 1] - p[t[:, 0], 1]) - \
        (p[t[:, 1], 1] - p[t[:, 0], 1]) * (p[t[:, 2], 0] - p[t[:, 0], 0])
    assert np.all(a > 0) and abs(a.sum() / 2 - 1.0) < 1e-12
    assert vm.sum() == 4 * 8
    c3, ev3, vm3 = afe.mesh_cube(4)
    p3 = c3.reshape(-1, 3)
    t3 = ev3.reshape(-1, 4)
    d = np.linalg.det(np.stack([p3[t3[:, 1]] - p3[t3[:, 0]], p3[t3[:, 2]] - p3[t3[:, 0]],
                                p3[t3[:, 3]] - p3[t3[:, 0]]], axis=1))
    assert np.all(d > 0) and abs(d.sum() / 6 - 1.0) < 1e-12
