"""GPU parity of the drop-in boundary's remaining entry points.

  SparseMatrix::vmult / Tvmult   afe_matrix_vmult / afe_matrix_tvmult against
                                 the oracle's restatement: bit-exact
  AMGPreconditioner::vmult       afe_amg_precondition (iterate_step + 1
                                 stationary cycles from the given x) against a
                                 dense model of the same multicolour cycle
  isSolveMostProjectExactly()    afe_amg_set_coarse_exact(0): the cycle against
                                 the dense model with Gauss-Seidel on the
                                 coarsest level, and the converged solution
                                 against the oracle's AMGSolver with the same
                                 setting (1e-9 relative)
  AMGSolver::lazyReinit          the device refresh against the oracle's
                                 lazyReinit of the same values: every level
                                 bit-exact
  BoundaryConditionAdmin::apply  afe_apply_dirichlet_bc with and without
                                 preserve_symmetry, nonzero data on part of
                                 the boundary: matrix bit-exact, right-hand
                                 side within 1e-13 of its scale
"""
import numpy as np
import pytest

from tests.conftest import make_mesh
from tests.test_gpu_parity import _dense_levels, _device_colours, _system, _zeros

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind,n,jitter,acc", [(2, 16, 0.2, 6), (3, 6, 0.05, 2)])
def test_matrix_vmult_and_tvmult_match_oracle(ctx, kind, n, jitter, acc):
    import afepack_amd as afe
    from oracle import oracle_py as orc
    print(f"case vmult kind={kind} n={n}", flush=True)
    c, ev, vm = make_mesh(kind, n, jitter, 23)
    sp = afe.Space(ctx, kind, c, ev, vm)
    A = afe.Matrix(sp)
    A.assemble(afe.OP_STIFF, acc)
    b = sp.rhs(acc, afe.F_POISSON)
    A.apply_dirichlet(b)  # eliminated rows: the matrix is no longer symmetric
    rp, col = sp.pattern()
    OA = orc.Mat.from_csr(sp.n_dof, sp.n_dof, rp, col, A.values())
    x = np.random.default_rng(3).uniform(-1, 1, sp.n_dof)
    dx = afe.DeviceVector.from_numpy(ctx, x)
    dy = _zeros(ctx, sp.n_dof)
    A.vmult(dx, dy)
    assert np.array_equal(dy.numpy(), OA.vmult(x))
    A.tvmult(dx, dy)
    ty = OA.tvmult(x)
    assert np.array_equal(dy.numpy(), ty)
    assert not np.array_equal(ty, OA.vmult(x))
    # a matrix given as CSR (the path of host-assembled matrices)
    B = afe.Matrix(ctx=ctx, csr=(rp, col, A.values()))
    B.tvmult(dx, dy)
    assert np.array_equal(dy.numpy(), ty)


def _model_stationary(As, Ps, colours, nu, exact_coarse):
    """x <- one stationary V(nu, nu) cycle of AMGSolver::solve
    (src/AMGSolver.cpp:579-613) from the given x: multicolour Gauss-Seidel in
    increasing colour order before and after the coarse correction, coarse
    levels from zero, and on the coarsest level either the exact solve or
    nu sweeps from zero (isSolveMostProjectExactly() = false)."""
    L = len(As)
    fwd = [np.concatenate([np.nonzero(c == k)[0] for k in range(c.max() + 1)]) for c in colours]

    def sweep(l, x, b):
        Am = As[l]
        for i in fwd[l]:
            x[i] = (b[i] - Am[i] @ x + Am[i, i] * x[i]) / Am[i, i]

    def vc(l, x, b):
        if l == L - 1:
            if exact_coarse:
                return np.linalg.solve(As[l], b)
            for _ in range(nu):
                sweep(l, x, b)
            return x
        for _ in range(nu):
            sweep(l, x, b)
        x += Ps[l].T @ vc(l + 1, np.zeros(Ps[l].shape[0]), Ps[l] @ (b - As[l] @ x))
        for _ in range(nu):
            sweep(l, x, b)
        return x

    return vc


@pytest.mark.parametrize("exact", [True, False])
@pytest.mark.parametrize("tail", [4096, 0])
@pytest.mark.parametrize("kind,n,jitter,acc", [(1, 24, 0.1, 4), (3, 6, 0.05, 2)])
def test_precondition_matches_stationary_model(ctx, kind, n, jitter, acc, tail, exact):
    """AMGPreconditioner::vmult = solve(x, b, 0, iterate_step, 1): with
    iterate_step = 1 exactly two stationary cycles from the caller's x."""
    import afepack_amd as afe
    print(f"case precondition kind={kind} tail={tail} exact={exact}", flush=True)
    sp, osp, A, OA, f, of = _system(ctx, kind, n, jitter, acc)
    g = afe.Amg(A, afe.AmgParams(lazy=1, smooth=2))
    g.set_tail_rows(tail)
    g.set_coarse_exact(exact)
    As, Ps, Gs = _dense_levels(g)
    cols = _device_colours(g, Gs)
    x0 = np.random.default_rng(8).uniform(-1, 1, sp.n_dof)
    dx = afe.DeviceVector.from_numpy(ctx, x0)
    g.precondition(dx, f, iterate_step=1)
    cyc = _model_stationary(As, Ps, cols, 2, exact)
    b = f.numpy()
    ref = cyc(0, cyc(0, x0.copy(), b), b)
    assert np.linalg.norm(dx.numpy() - ref) <= 1e-12 * np.linalg.norm(ref)


@pytest.mark.parametrize("kind,n,jitter,acc", [(1, 32, 0.0, 4), (2, 10, 0.2, 6)])
def test_inexact_coarse_solution_matches_oracle(ctx, kind, n, jitter, acc):
    """isSolveMostProjectExactly() = false on both sides: the device
    stationary iteration and FP64 PCG, and the oracle's AMGSolver with
    Gauss-Seidel on its coarsest level, all converge to one solution."""
    import afepack_amd as afe
    from oracle import oracle_py as orc
    print(f"case inexact coarse kind={kind} n={n}", flush=True)
    sp, osp, A, OA, f, of = _system(ctx, kind, n, jitter, acc)
    og = orc.Amg(OA)
    ox = np.zeros(sp.n_dof)
    h0 = None
    for _ in range(40):
        h = og.solve(ox, of, smooth=3, step=49, exactly=False)
        h0 = h[0] if h0 is None else h0
        if h[-1] <= 1e-13 * h0:
            break
    g = afe.Amg(A)
    g.set_coarse_exact(False)
    x = _zeros(ctx, sp.n_dof)
    st = g.solve(x, f, step=5000, rtol=1e-12)
    assert st.rel_res_l2 <= 1e-12
    assert np.linalg.norm(x.numpy() - ox) <= 1e-9 * np.linalg.norm(ox)
    x2 = _zeros(ctx, sp.n_dof)
    st2 = g.pcg(x2, f, rtol=1e-12, max_it=500, prec=afe.PREC_FP64)
    assert st2.converged == 1
    assert np.linalg.norm(x2.numpy() - ox) <= 1e-9 * np.linalg.norm(ox)
    # switching back restores the exact coarsest solve
    g.set_coarse_exact(True)
    x3 = _zeros(ctx, sp.n_dof)
    st3 = g.pcg(x3, f, rtol=1e-12, max_it=500, prec=afe.PREC_FP64)
    assert st3.converged == 1 and st3.iterations <= st2.iterations


@pytest.mark.parametrize("kind,n,jitter,op", [(2, 16, 0.2, 0), (1, 48, 0.1, 2), (3, 6, 0.05, 2)])
def test_device_lazy_reinit_matches_oracle(ctx, kind, n, jitter, op):
    """afe_amg_lazy_reinit against or_amg_lazy_reinit (src/AMGSolver.cpp:
    505-545) on the same two value sets of one pattern: every level's
    pattern and values bit-exact."""
    import afepack_amd as afe
    from oracle import oracle_py as orc
    print(f"case lazy vs oracle kind={kind} n={n} op={op}", flush=True)
    c, ev, vm = make_mesh(kind, n, jitter, 29)
    sp = afe.Space(ctx, kind, c, ev, vm)
    dim = 3 if kind == 3 else 2
    A = afe.Matrix(sp)
    A.assemble(op, 4, afe.field_random(8, dim, 0.5, 1) if op == 2 else None)
    b = sp.rhs(4, afe.F_POISSON)
    A.apply_dirichlet(b)
    rp, col = sp.pattern()
    O1 = orc.Mat.from_csr(sp.n_dof, sp.n_dof, rp, col, A.values())
    g = afe.Amg(A, afe.AmgParams(lazy=1, smooth=1, build_fp32=1))
    og = orc.Amg(O1, lazy=True)
    sp.update_coords(afe.mesh_displace(dim, c, vm, 0.1 / n, 3))
    A.assemble(op, 4, afe.field_random(8, dim, 0.5, 2) if op == 2 else None)
    O2 = O1.share_pattern(A.values())
    g.lazy_reinit(A)
    assert og.lazy_reinit(O2) == 0
    assert g.info()["levels"] == og.n_project + 1
    for l in range(og.n_project + 1):
        _, _, lrp, lcol, lval = g.level_csr(l, 0)
        orp, ocol, oval = og.A(l).csr()
        assert np.array_equal(lrp, orp) and np.array_equal(lcol, ocol)
        assert np.array_equal(lval, oval), f"level {l}"


@pytest.mark.parametrize("kind,n,jitter,acc", [(1, 20, 0.1, 4), (2, 10, 0.2, 6), (3, 5, 0.05, 2)])
@pytest.mark.parametrize("sym", [True, False])
def test_dirichlet_with_and_without_symmetry_matches_oracle(ctx, kind, n, jitter, acc, sym):
    import afepack_amd as afe
    from oracle import oracle_py as orc
    print(f"case dirichlet kind={kind} sym={sym}", flush=True)
    c, ev, vm = make_mesh(kind, n, jitter, 17)
    dim = 3 if kind == 3 else 2
    cc = c.reshape(-1, dim)
    vm = vm.copy()
    vm[(vm != 0) & (cc[:, 0] == 0.0)] = 2  # data on the face x = 0 only
    sp = afe.Space(ctx, kind, c, ev, vm)
    osp = orc.Space(kind, c, ev, vm)
    A = afe.Matrix(sp)
    A.assemble(afe.OP_STIFF, acc)
    OA = osp.build(orc.OP_STIFF, acc)
    of = osp.rhs(acc, orc.F_POISSON)
    P = sp.dofs()[2].reshape(-1, dim)
    g = 2.0 + np.sin(3.0 * P[:, 1])
    df = afe.DeviceVector.from_numpy(ctx, of)
    A.apply_dirichlet_values([2], g, df, preserve_symmetry=sym)
    osp.apply_dirichlet_values(OA, [2], g, of, preserve_symmetry=sym)
    assert np.array_equal(A.values(), OA.csr()[2])
    assert np.abs(df.numpy() - of).max() <= 1e-13 * np.abs(of).max()
