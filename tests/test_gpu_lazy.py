"""GPU tests of AMGSolver::lazyReinit on the device (afe_amg_lazy_reinit).

A lazy hierarchy refreshed with new values on the same pattern (moving-mesh
update, new coefficient) must equal, bit for bit, a hierarchy built from
scratch by the host coarsening on those values: every level matrix, the
coarsest inverse and the PCG solution.
"""
import numpy as np
import pytest

from tests.conftest import make_mesh

pytestmark = pytest.mark.gpu


def _levels(g):
    info = g.info()
    return [g.level_csr(l, 0) for l in range(info["levels"])]


@pytest.mark.parametrize("kind,n,jitter,op", [(2, 16, 0.2, 0), (1, 48, 0.1, 2), (3, 6, 0.05, 2)])
def test_device_lazy_reinit_equals_fresh_build(ctx, kind, n, jitter, op):
    import afepack_amd as afe
    print(f"case lazy kind={kind} n={n} op={op}", flush=True)
    c, ev, vm = make_mesh(kind, n, jitter, 17)
    sp = afe.Space(ctx, kind, c, ev, vm)
    dim = 3 if kind == 3 else 2
    f1 = afe.field_random(8, dim, 0.5, 1) if op == 2 else None
    f2 = afe.field_random(8, dim, 0.5, 2) if op == 2 else None
    A = afe.Matrix(sp)
    A.assemble(op, 4, f1)
    b = sp.rhs(4, afe.F_POISSON)
    A.apply_dirichlet(b)
    g = afe.Amg(A, afe.AmgParams(lazy=1, smooth=1, build_fp32=1))
    # moving-mesh update of the same space and a new coefficient
    c2 = afe.mesh_displace(dim, c, vm, 0.1 / n, 2)
    sp.update_coords(c2)
    A.assemble(op, 4, f2)
    b2 = sp.rhs(4, afe.F_POISSON)
    A.apply_dirichlet(b2)
    g.lazy_reinit(A)
    fresh = afe.Amg(A, afe.AmgParams(lazy=1, smooth=1, build_fp32=1))
    L1, L2 = _levels(g), _levels(fresh)
    assert len(L1) == len(L2)
    for (_, _, rp1, c1, v1), (_, _, rp2, c2_, v2) in zip(L1, L2):
        assert np.array_equal(rp1, rp2) and np.array_equal(c1, c2_)
        assert np.array_equal(v1, v2)
    if g.info()["coarse_full"]:
        assert np.array_equal(g.coarse_inverse(), fresh.coarse_inverse())
    x1 = afe.DeviceVector(ctx, sp.n_dof)
    x2 = afe.DeviceVector(ctx, sp.n_dof)
    s1 = g.pcg(x1, b2, rtol=1e-10, max_it=300, prec=afe.PREC_FP32)
    s2 = fresh.pcg(x2, b2, rtol=1e-10, max_it=300, prec=afe.PREC_FP32)
    assert s1.iterations == s2.iterations
    assert np.array_equal(x1.numpy(), x2.numpy())
