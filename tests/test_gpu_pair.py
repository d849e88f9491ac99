"""GPU parity of BilinearOperator over two FE spaces on one mesh (the
same-mesh branch of BilinearOperator.templates.h:64-255) and L2InnerProduct
(:300-326) through the C ABI, against the oracle's two-space build.

Bars: pattern and nMaxCouplingDof bit-exact; L2InnerProduct values
bit-exact (the device follows the reference's arithmetic order: Jxw *
phi0_j * phi1_k per quadrature point, elements summed in ascending order);
caller element matrices gathered bit-exactly (ascending element order);
vmult / Tvmult bit-exact against the oracle's SparseMatrix restatement.
"""
import numpy as np
import pytest

from tests.conftest import make_mesh

pytestmark = pytest.mark.gpu

PAIRS = [
    # kind0, kind1, n, jitter, acc
    (1, 2, 12, 0.2, 4),
    (2, 1, 12, 0.2, 6),
    (1, 1, 10, 0.1, 3),   # two P1 spaces: square pattern, diagonal first
    (2, 2, 6, 0.2, 5),
    (3, 3, 4, 0.05, 2),
]


def _pair(ctx, k0, k1, n, jitter, seed=7):
    import afepack_amd as afe
    from oracle import oracle_py as orc
    c, ev, vm = make_mesh(k0, n, jitter, seed)
    s0, s1 = afe.Space(ctx, k0, c, ev, vm), afe.Space(ctx, k1, c, ev, vm)
    o0, o1 = orc.Space(k0, c, ev, vm), orc.Space(k1, c, ev, vm)
    return s0, s1, o0, o1


@pytest.mark.parametrize("k0,k1,n,jitter,acc", PAIRS)
def test_l2_inner_product_of_two_spaces_matches_oracle(ctx, k0, k1, n, jitter, acc):
    import afepack_amd as afe
    from oracle import oracle_py as orc
    print(f"case pair k0={k0} k1={k1} n={n} acc={acc}", flush=True)
    s0, s1, o0, o1 = _pair(ctx, k0, k1, n, jitter)
    A = afe.Matrix(s0, space1=s1)
    OA, omc = o0.build_pair(o1, orc.OP_MASS, acc)
    rp, col, oval = OA.csr()
    prp, pcol = A.pattern()
    assert (A.m, A.n) == (s0.n_dof, s1.n_dof)
    assert np.array_equal(prp, rp) and np.array_equal(pcol, col)
    assert A.max_coupling == omc
    A.assemble(afe.OP_MASS, acc)
    assert np.array_equal(A.values(), oval)
    # SparseMatrix::vmult / Tvmult on the rectangular matrix
    rng = np.random.default_rng(k0 * 10 + k1)
    x1 = rng.uniform(-1, 1, A.n)
    x0 = rng.uniform(-1, 1, A.m)
    y0 = afe.DeviceVector(ctx, A.m)
    y1 = afe.DeviceVector(ctx, A.n)
    A.vmult(afe.DeviceVector.from_numpy(ctx, x1), y0)
    A.tvmult(afe.DeviceVector.from_numpy(ctx, x0), y1)
    assert np.array_equal(y0.numpy(), OA.vmult(x1))
    assert np.array_equal(y1.numpy(), OA.tvmult(x0))


def test_two_space_caller_element_matrices(ctx):
    """afe_assemble_elements on a P2 x P1 matrix: n_loc0 x n_loc1 element
    matrices summed into the pattern in ascending element order"""
    import afepack_amd as afe
    s0, s1, _, _ = _pair(ctx, 2, 1, 5, 0.1)
    A = afe.Matrix(s0, space1=s1)
    ne = s0.n_elem
    rng = np.random.default_rng(3)
    K = rng.uniform(-1, 1, (ne, 6, 3))
    A.assemble_elements(K)
    rp, col = A.pattern()
    e0 = s0.dofs()[0].reshape(ne, 6)
    e1 = s1.dofs()[0].reshape(ne, 3)
    want = np.zeros(A.nnz)
    for e in range(ne):
        for a in range(6):
            i = e0[e, a]
            r = col[rp[i]:rp[i + 1]]
            for b in range(3):
                want[rp[i] + int(np.nonzero(r == e1[e, b])[0][0])] += K[e, a, b]
    assert np.array_equal(A.values(), want)


def test_two_space_matrix_rejections(ctx):
    import afepack_amd as afe
    s0, s1, _, _ = _pair(ctx, 1, 2, 4, 0.0)
    A = afe.Matrix(s0, space1=s1)
    with pytest.raises(afe.AfeError):
        A.assemble(afe.OP_STIFF, 2)  # the reference defines only L2InnerProduct on two spaces
    with pytest.raises(afe.AfeError):
        afe.Amg(A)
    f = afe.DeviceVector(ctx, A.m)
    with pytest.raises(afe.AfeError):
        A.apply_dirichlet(f)
    c, ev, vm = make_mesh(2, 5, 0.0, 0)  # another mesh
    other = afe.Space(ctx, 2, c, ev, vm)
    with pytest.raises(afe.AfeError):
        afe.Matrix(s0, space1=other)


def test_pair_transpose_and_row_sums_at_full_size(ctx):
    """size-independent checks on a 512 x 512 mesh (0.79M P1 x 1.05M P2):
    the P1 x P2 and P2 x P1 matrices are each other's transposes (within
    rounding: the products Jxw phi0 phi1 and Jxw phi1 phi0 round apart), the
    P1 x P2 row sums are the P1 mass row sums (partition of unity of P2)"""
    import afepack_amd as afe
    import scipy.sparse as sps
    print("case pair full size n=512", flush=True)
    c, ev, vm = make_mesh(1, 512, 0.2, 20261)
    s1, s2 = afe.Space(ctx, 1, c, ev, vm), afe.Space(ctx, 2, c, ev, vm)
    A12, A21, M1 = afe.Matrix(s1, space1=s2), afe.Matrix(s2, space1=s1), afe.Matrix(s1)
    A12.assemble(afe.OP_MASS, 4)
    A21.assemble(afe.OP_MASS, 4)
    M1.assemble(afe.OP_MASS, 4)
    rp, col = A12.pattern()
    X = sps.csr_matrix((A12.values(), col, rp), shape=(A12.m, A12.n))
    rp, col = A21.pattern()
    Y = sps.csr_matrix((A21.values(), col, rp), shape=(A21.m, A21.n))
    D = (X - Y.T).tocsr()
    assert abs(D).max() <= 1e-15 * abs(X).max()
    assert (X != 0).nnz == (Y != 0).nnz
    rp, _ = s1.pattern()
    rs = np.add.reduceat(A12.values(), A12.pattern()[0][:-1])
    ms = np.add.reduceat(M1.values(), rp[:-1])
    assert np.allclose(rs, ms, rtol=1e-12, atol=1e-20)
    assert abs(rs.sum() - 1.0) < 1e-12
