"""GPU parity: the HIP path through the C ABI against the oracle.

Bars (stated per assertion): CSR pattern and DOF data bit-exact; assembled
entries within 1e-13 of the row scale; right-hand sides within 1e-13 of the
vector scale; solutions within 1e-9 relative L2 of the oracle's AMGSolver
solution; L2/H1 errors agreeing to 3 digits (checked at 1e-6 relative).
"""
import numpy as np
import pytest

from tests.conftest import make_mesh, rel_err_rows

pytestmark = pytest.mark.gpu

CASES = [
    # kind, n, jitter, acc
    (1, 64, 0.0, 4),   # the 64x64 P1 Poisson configuration
    (2, 16, 0.2, 6),   # jittered P2 at the large configuration's accuracy
    (3, 6, 0.0, 2),    # tricube P1 tetrahedra
]


def _zeros(ctx, n):
    import afepack_amd as afe
    return afe.DeviceVector(ctx, n)


def _spaces(ctx, kind, n, jitter, seed=11):
    import afepack_amd as afe
    from oracle import oracle_py as orc
    c, ev, vm = make_mesh(kind, n, jitter, seed)
    return afe.Space(ctx, kind, c, ev, vm), orc.Space(kind, c, ev, vm), (c, ev, vm)


@pytest.mark.parametrize("kind,n,jitter,acc", CASES)
@pytest.mark.parametrize("op", [0, 1, 2])
def test_assembly_matches_oracle(ctx, kind, n, jitter, acc, op):
    import afepack_amd as afe
    from oracle import oracle_py as orc
    print(f"case kind={kind} n={n} op={op}", flush=True)
    sp, osp, _ = _spaces(ctx, kind, n, jitter)
    field = afe.field_random(8, sp.dim, 0.5, 5) if op == 2 else None
    ofield = orc.Field(field.k, field.phi, field.w, field.scale) if field else None
    A = afe.Matrix(sp)
    A.assemble(op, acc, field)
    OA = osp.build(op, acc, ofield)
    rp, col, oval = OA.csr()
    prp, pcol = sp.pattern()
    assert np.array_equal(prp, rp) and np.array_equal(pcol, col)
    val = A.values()
    if op == 2:
        # exp/sin of the device math library: within 1e-13 of the row scale
        assert rel_err_rows(val, oval, rp) <= 1e-13
    else:
        # the element kernels follow the reference's arithmetic: bit-exact
        assert np.array_equal(val, oval)


@pytest.mark.parametrize("kind,n,jitter,acc", CASES)
def test_rhs_and_dirichlet_match_oracle(ctx, kind, n, jitter, acc):
    import afepack_amd as afe
    from oracle import oracle_py as orc
    sp, osp, _ = _spaces(ctx, kind, n, jitter)
    A = afe.Matrix(sp)
    A.assemble(afe.OP_STIFF, acc)
    f = sp.rhs(acc, afe.F_POISSON)
    OA = osp.build(orc.OP_STIFF, acc)
    of = osp.rhs(acc, orc.F_POISSON)
    fh = f.numpy()
    assert np.abs(fh - of).max() <= 1e-13 * np.abs(of).max()
    A.apply_dirichlet(f)
    osp.apply_dirichlet(OA, of)
    rp, col, oval = OA.csr()
    assert rel_err_rows(A.values(), oval, rp) <= 1e-13
    assert np.abs(f.numpy() - of).max() <= 1e-13 * np.abs(of).max()
    # nodal right-hand side (interpolant of DOF values)
    rng = np.random.default_rng(3)
    nodal = rng.uniform(-1, 1, sp.n_dof)
    fn = sp.rhs(acc, afe.F_NODAL, nodal=nodal)
    ofn = osp.rhs(acc, orc.F_NODAL, nodal=nodal)
    assert np.abs(fn.numpy() - ofn).max() <= 1e-13 * np.abs(ofn).max()


def _system(ctx, kind, n, jitter, acc, lazy=False):
    import afepack_amd as afe
    from oracle import oracle_py as orc
    sp, osp, mesh = _spaces(ctx, kind, n, jitter)
    A = afe.Matrix(sp)
    A.assemble(afe.OP_STIFF, acc)
    f = sp.rhs(acc, afe.F_POISSON)
    A.apply_dirichlet(f)
    OA = osp.build(orc.OP_STIFF, acc)
    of = osp.rhs(acc, orc.F_POISSON)
    osp.apply_dirichlet(OA, of)
    return sp, osp, A, OA, f, of


def _oracle_solution(OA, of, lazy=False):
    """The oracle AMGSolver iterated (in chunks of 50 cycles, each chunk
    restarting from the previous x as the reference's solve does) until the
    L1 residual stops improving or is 1e-13 of the initial one."""
    from oracle import oracle_py as orc
    g = orc.Amg(OA, lazy=lazy)
    x = np.zeros(len(of))
    first = g.solve(x, of, smooth=3, step=49)
    r0, best = first[0], first[-1]
    for _ in range(60):
        if best <= 1e-13 * r0:
            break
        h = g.solve(x, of, smooth=3, step=49)
        if h[-1] >= 0.999 * best:
            break
        best = h[-1]
    return x, first, g


@pytest.mark.parametrize("kind,n,jitter,acc", CASES)
def test_hierarchy_on_device_matches_oracle(ctx, kind, n, jitter, acc):
    import afepack_amd as afe
    from oracle import oracle_py as orc
    sp, osp, A, OA, f, of = _system(ctx, kind, n, jitter, acc)
    g = afe.Amg(A)
    # the oracle coarsens the device-assembled matrix: every level must be
    # bit-identical (pattern, explicit zeros and values) to the uploaded one
    rp, col = sp.pattern()
    dev0 = orc.Mat.from_csr(sp.n_dof, sp.n_dof, rp, col, A.values())
    og = orc.Amg(dev0)
    info = g.info()
    assert info["levels"] == og.n_project + 1
    for l in range(info["levels"]):
        m, nn, lrp, lcol, lval = g.level_csr(l, 0)
        orp, ocol, oval = og.A(l).csr()
        assert np.array_equal(lrp, orp) and np.array_equal(lcol, ocol)
        assert np.array_equal(lval, oval)
    # and the oracle's own assembly coarsens to the same level sizes
    og2 = orc.Amg(OA)
    assert [og2.A(l).m for l in range(og2.n_project + 1)] == info["n"]


@pytest.mark.parametrize("kind,n,jitter,acc", CASES)
@pytest.mark.parametrize("mode", ["stationary", "pcg64", "pcg32"])
def test_solution_matches_oracle(ctx, kind, n, jitter, acc, mode):
    import afepack_amd as afe
    print(f"solve kind={kind} n={n} mode={mode}", flush=True)
    sp, osp, A, OA, f, of = _system(ctx, kind, n, jitter, acc)
    ox, hist, _ = _oracle_solution(OA, of)
    g = afe.Amg(A)
    x = _zeros(ctx, sp.n_dof)
    if mode == "stationary":
        # the reference iteration (V(3,3) cycles), run to a 1e-12 residual
        st = g.solve(x, f, step=5000, rtol=1e-12)
        assert st.rel_res_l2 <= 1e-12
    else:
        st = g.pcg(x, f, rtol=1e-12, max_it=300,
                   prec=afe.PREC_FP64 if mode == "pcg64" else afe.PREC_FP32)
        assert st.converged == 1 and st.rel_res_l2 <= 1e-12
    xs = x.numpy()
    assert np.linalg.norm(xs - ox) <= 1e-9 * np.linalg.norm(ox)
    # discretisation errors against u = prod sin(pi x_d) agree to 3 digits
    e2, e1 = sp.error_l2(x, acc), sp.error_h1(x, acc)
    oe2, oe1 = osp.l2_error(ox, acc), osp.h1_error(ox, acc)
    assert abs(e2 - oe2) <= 1e-6 * oe2 and abs(e1 - oe1) <= 1e-6 * oe1


def test_reference_solve_semantics(ctx):
    """afe_amg_solve runs exactly step+1 cycles (AMGSolver::solve) and reduces
    the L1 residual like the reference iteration does."""
    import afepack_amd as afe
    sp, osp, A, OA, f, of = _system(ctx, 1, 32, 0.0, 4)
    g = afe.Amg(A)
    x = _zeros(ctx, sp.n_dof)
    st = g.solve(x, f, step=7, mode=1)
    assert st.iterations == 8
    _, hist, _ = _oracle_solution(OA, of)
    # multicolour GS converges at a comparable rate: 8 cycles reach a
    # relative L1 residual within two orders of the lexicographic one
    assert st.rel_res_l1 < 100 * hist[8] / hist[0]


def _dense_levels(g):
    """Dense copies of the device hierarchy: level matrices, restrictions P
    and the stored patterns (explicit zeros included)."""
    L = g.info()["levels"]
    As, Ps, Gs = [], [], []
    for l in range(L):
        m, n, rp, col, val = g.level_csr(l, 0)
        D = np.zeros((m, n))
        S = np.zeros((m, n), dtype=bool)
        for i in range(m):
            D[i, col[rp[i]:rp[i + 1]]] += val[rp[i]:rp[i + 1]]
            S[i, col[rp[i]:rp[i + 1]]] = True
        As.append(D)
        Gs.append(S)
        if l < L - 1:
            m, n, rp, col, val = g.level_csr(l, 1)
            D = np.zeros((m, n))
            for i in range(m):
                D[i, col[rp[i]:rp[i + 1]]] += val[rp[i]:rp[i + 1]]
            Ps.append(D)
    return As, Ps, Gs


def _device_colours(g, Gs):
    """The device's colouring of each level, checked to be a colouring of
    the symmetrised stored pattern (no two coupled rows share a colour)."""
    out = []
    for l, S in enumerate(Gs):
        c = g.level_colours(l)
        G = (S | S.T) & ~np.eye(S.shape[0], dtype=bool)
        i, j = np.nonzero(G)
        assert not np.any(c[i] == c[j]), f"level {l}: coupled rows share a colour"
        assert c.min() == 0 and set(np.unique(c)) == set(range(c.max() + 1))
        out.append(c)
    return out


def _model_vcycle(As, Ps, colours, nu):
    """z = M b for the symmetric V(nu, nu)-cycle from zero: multicolour
    Gauss-Seidel in the given colour order (forward before, reverse after
    the coarse correction), restriction P, prolongation P^T, exact coarse
    solve."""
    L = len(As)
    orders = []
    for l in range(L - 1):
        c = colours[l]
        orders.append((np.concatenate([np.nonzero(c == k)[0] for k in range(c.max() + 1)]),
                       np.concatenate([np.nonzero(c == k)[0] for k in range(c.max(), -1, -1)])))

    def sweep(Am, x, b, order):
        for i in order:
            x[i] = (b[i] - Am[i] @ x + Am[i, i] * x[i]) / Am[i, i]

    def vc(l, b):
        if l == L - 1:
            return np.linalg.solve(As[l], b)
        x = np.zeros_like(b)
        for _ in range(nu):
            sweep(As[l], x, b, orders[l][0])
        x += Ps[l].T @ vc(l + 1, Ps[l] @ (b - As[l] @ x))
        for _ in range(nu):
            sweep(As[l], x, b, orders[l][1])
        return x

    return lambda b: vc(0, b)


@pytest.mark.parametrize("kind,n,jitter,acc,nmpo,nmpsd,lanes", [
    (1, 32, 0.0, 4, 50, 0.382, None),
    (1, 32, 0.0, 4, 400, 0.0, None),
    (3, 8, 0.05, 2, 50, 0.382, None),  # coarse levels with several lanes per row
    (3, 8, 0.05, 2, 50, 0.382, "8"),   # 8 and 16 lanes on every coarse level
    (3, 8, 0.05, 2, 50, 0.382, "16"),
])
@pytest.mark.parametrize("tail", [4096, 0])
def test_vcycle_matches_dense_model(ctx, kind, n, jitter, acc, nmpo, nmpsd, lanes, tail):
    """One symmetric FP64 V(1,1)-cycle from zero on the device against the
    dense numpy model of the same cycle, with the small levels in the
    single-workgroup tail (tail 4096) and as per-level launches (tail 0)."""
    import os
    import afepack_amd as afe
    print(f"case vcycle model kind={kind} n={n} nmpo={nmpo} lanes={lanes} tail={tail}", flush=True)
    sp, osp, A, OA, f, of = _system(ctx, kind, n, jitter, acc)
    if lanes:
        os.environ["AFE_ROW_LANES"] = lanes
    try:
        g = afe.Amg(A, afe.AmgParams(lazy=1, nmpo=nmpo, nmpsd=nmpsd, smooth=1))
    finally:
        os.environ.pop("AFE_ROW_LANES", None)
    g.set_tail_rows(tail)
    rl = [d["row_lanes"] for d in g.device_info()]
    if lanes:
        assert rl[0] == 1 and set(rl[1:]) == {int(lanes)}
    elif kind == 3:
        assert max(rl) > 1
    z = _zeros(ctx, sp.n_dof)
    g.vcycle(z, f, afe.PREC_FP64)
    As, Ps, Gs = _dense_levels(g)
    ref = _model_vcycle(As, Ps, _device_colours(g, Gs), 1)(f.numpy())
    assert np.linalg.norm(z.numpy() - ref) <= 1e-12 * np.linalg.norm(ref)


@pytest.mark.parametrize("kind,n,jitter,acc,smooth", [
    (1, 24, 0.1, 4, 0), (1, 24, 0.1, 4, 1), (1, 24, 0.1, 4, 2), (3, 8, 0.05, 2, 1)])
def test_pcg_iterates_match_model(ctx, kind, n, jitter, acc, smooth):
    """Three iterations of the FP64 flexible PCG on the device (several
    levels run as launches, the first level-0 colour from the zero guess, and
    no smoothing at all) against numpy flexible PCG with the dense V-cycle
    model: x_{k+1} = x_k + alpha p, alpha = (r, z) / (p, A p),
    beta = ((r, z_new) - (r, z_old)) / (r_old, z_old)."""
    import afepack_amd as afe
    print(f"case pcg model kind={kind} smooth={smooth}", flush=True)
    sp, osp, A, OA, f, of = _system(ctx, kind, n, jitter, acc)
    g = afe.Amg(A, afe.AmgParams(lazy=1, nmpo=50, nmpsd=0.382, smooth=max(smooth, 1), build_fp32=1))
    g.set_tail_rows(0)  # every level but the coarsest as separate launches
    g.set_smooth(smooth)
    As, Ps, Gs = _dense_levels(g)
    M = _model_vcycle(As, Ps, _device_colours(g, Gs), smooth)
    A0, b = As[0], f.numpy()
    x = np.zeros_like(b)
    r = b.copy()
    z = M(r)
    p = z.copy()
    rz = r @ z
    its = 3
    for _ in range(its):
        q = A0 @ p
        alpha = rz / (p @ q)
        x = x + alpha * p
        r = r - alpha * q
        zn = M(r)
        beta = (r @ zn - r @ z) / rz
        rz = r @ zn
        z = zn
        p = z + beta * p
    xd = _zeros(ctx, sp.n_dof)
    st = g.pcg(xd, f, rtol=1e-15, max_it=its, prec=afe.PREC_FP64, check_conv=False)
    assert st.iterations == its
    assert np.linalg.norm(xd.numpy() - x) <= 1e-11 * np.linalg.norm(x)


@pytest.mark.parametrize("kind,n,jitter,acc,modes,amp", [
    (3, 8, 0.05, 2, 16, 1.0),   # the 3D variable-coefficient configuration, small
    (2, 12, 0.0, 6, 8, 0.5),    # a batch member: P2, 8-mode field
])
def test_variable_coefficient_solution_matches_oracle(ctx, kind, n, jitter, acc, modes, amp):
    """-div(a grad u) = f, a = exp(g): assembly, rhs (f from the manufactured
    u), Dirichlet, mixed FP32-V-cycle / FP64 PCG to 1e-12 against the oracle's
    AMGSolver solution (1e-9), L2/H1 errors to 1e-6 relative."""
    import afepack_amd as afe
    from oracle import oracle_py as orc
    print(f"varcoef kind={kind} n={n}", flush=True)
    sp, osp, _ = _spaces(ctx, kind, n, jitter)
    field = afe.field_random(modes, sp.dim, amp, 23)
    ofield = orc.Field(field.k, field.phi, field.w, field.scale)
    A = afe.Matrix(sp)
    A.assemble(afe.OP_VARSTIFF, acc, field)
    f = sp.rhs(acc, afe.F_VARCOEF, field)
    A.apply_dirichlet(f)
    OA = osp.build(orc.OP_VARSTIFF, acc, ofield)
    of = osp.rhs(acc, orc.F_VARCOEF, ofield)
    osp.apply_dirichlet(OA, of)
    rp, col, oval = OA.csr()
    assert rel_err_rows(A.values(), oval, rp) <= 1e-13
    assert np.abs(f.numpy() - of).max() <= 1e-13 * np.abs(of).max()
    ox, _, _ = _oracle_solution(OA, of)
    g = afe.Amg(A, afe.AmgParams(lazy=1, smooth=1, build_fp32=1))
    x = _zeros(ctx, sp.n_dof)
    st = g.pcg(x, f, rtol=1e-12, max_it=400, prec=afe.PREC_FP32)
    assert st.converged == 1 and st.rel_res_l2 <= 1e-12
    xs = x.numpy()
    assert np.linalg.norm(xs - ox) <= 1e-9 * np.linalg.norm(ox)
    e2, e1 = sp.error_l2(x, acc), sp.error_h1(x, acc)
    oe2, oe1 = osp.l2_error(ox, acc), osp.h1_error(ox, acc)
    assert abs(e2 - oe2) <= 1e-6 * oe2 and abs(e1 - oe1) <= 1e-6 * oe1


# meshes just above 2^18 elements, where afe_assemble pipelines the element
# chunks and the gathers of the rows they complete (two streams)
PIPELINE_CASES = [
    (1, 364, 0.2, 4),  # 264,992 triangles
    (2, 364, 0.2, 6),
    (3, 36, 0.05, 2),  # 279,936 tetrahedra
]


@pytest.mark.parametrize("kind,n,jitter,acc", PIPELINE_CASES)
@pytest.mark.parametrize("shuffle", [False, True])
def test_pipelined_assembly_matches_oracle(ctx, kind, n, jitter, acc, shuffle):
    """Element chunks overlapping the gathers of completed rows give the
    oracle's values bit for bit (stiffness and mass; the variable
    coefficient within 1e-13 of the row scale), in the mesh's own element
    order and with the elements shuffled (a row then waits for late chunks)"""
    import afepack_amd as afe
    from oracle import oracle_py as orc
    print(f"case pipelined kind={kind} n={n} shuffle={shuffle}", flush=True)
    c, ev, vm = make_mesh(kind, n, jitter, 7)
    if shuffle:  # permute whole elements (rows of nv vertices)
        nv = 4 if kind == 3 else 3
        ev2 = np.asarray(ev).reshape(-1, nv)
        ev = np.ascontiguousarray(ev2[np.random.default_rng(3).permutation(len(ev2))]).reshape(-1)
    sp = afe.Space(ctx, kind, c, ev, vm)
    osp = orc.Space(kind, c, ev, vm)
    assert sp.n_elem >= 1 << 18
    field = afe.field_random(8, sp.dim, 0.5, 5)
    ofield = orc.Field(field.k, field.phi, field.w, field.scale)
    A = afe.Matrix(sp)
    for op in (0, 1, 2):
        A.assemble(op, acc, field if op == 2 else None)
        rp, col, oval = osp.build(op, acc, ofield if op == 2 else None).csr()
        val = A.values()
        if op == 2:
            assert rel_err_rows(val, oval, rp) <= 1e-13
        else:
            assert np.array_equal(val, oval)


@pytest.mark.parametrize("kind,n,jitter,acc", [PIPELINE_CASES[1], PIPELINE_CASES[2]])
def test_assembly_timing_entries_leave_the_assembly(ctx, kind, n, jitter, acc):
    """afe_assemble_timing (the assembly as it runs, pipelined on tetrahedra)
    and afe_assemble_kernel_timing (its kernels as whole-mesh launches)
    report positive times and the algorithmic bytes (coordinates, element
    vertex and DOF indices, CSR values), and leave the values of a plain
    assembly bit for bit"""
    import afepack_amd as afe
    print(f"case assembly timing kind={kind} n={n}", flush=True)
    c, ev, vm = make_mesh(kind, n, jitter, 3)
    sp = afe.Space(ctx, kind, c, ev, vm)
    A = afe.Matrix(sp)
    A.assemble(afe.OP_STIFF, acc)
    ref = A.values()
    ms = A.assemble_pipelined_ms(afe.OP_STIFF, acc, reps=2)
    assert ms > 0 and np.array_equal(A.values(), ref)
    te, tg, by = A.assemble_timing(afe.OP_STIFF, acc, reps=2)
    assert te > 0 and tg > 0 and np.array_equal(A.values(), ref)
    nv = 4 if kind == afe.P1_TET else 3
    assert by == len(c) * 8 + sp.n_elem * (nv + sp.n_loc) * 4 + sp.nnz * 8


@pytest.mark.parametrize("kind,n,jitter", [(2, 128, 0.2), (3, 36, 0.05)])
def test_caller_element_matrices_gathered_in_element_order(ctx, kind, n, jitter):
    """afe_assemble_elements (the caller's element matrices, element-major
    NL x NL, not symmetric) sums every CSR entry over its elements in
    ascending element order from zero: equal bit for bit to a sequential
    numpy accumulation (np.add.at in element order); tetrahedra go through
    the LDS-staged gather"""
    import afepack_amd as afe
    print(f"case caller element matrices kind={kind} n={n}", flush=True)
    c, ev, vm = make_mesh(kind, n, jitter, 9)
    sp = afe.Space(ctx, kind, c, ev, vm)
    nl = sp.n_loc
    ed = sp.dofs()[0].reshape(-1, nl).astype(np.int64)
    K = np.random.default_rng(4).standard_normal((sp.n_elem, nl, nl))
    A = afe.Matrix(sp)
    A.assemble_elements(K)
    val = A.values()
    rp, col = sp.pattern()
    rp = rp.astype(np.int64)
    nd = sp.n_dof
    rows = np.repeat(np.arange(nd, dtype=np.int64), np.diff(rp))
    keys = rows * nd + col.astype(np.int64)
    order = np.argsort(keys)
    ck = (ed[:, :, None] * nd + ed[:, None, :]).reshape(-1)  # element-major, (a, b) inside
    pos = order[np.searchsorted(keys[order], ck)]
    assert np.array_equal(keys[pos], ck)
    ref = np.zeros(len(col))
    np.add.at(ref, pos, K.reshape(-1))
    assert np.array_equal(val, ref)


def _assembly_cases_match_oracle():
    """(child process) every kind at the pipeline size, three operators, in
    the launch configuration the environment forces; prints one line per case"""
    import afepack_amd as afe
    from oracle import oracle_py as orc
    from tests.conftest import rel_err_rows
    c0 = afe.Context(0)
    for kind, n, jitter, acc in PIPELINE_CASES:
        c, ev, vm = make_mesh(kind, n, jitter, 7)
        sp = afe.Space(c0, kind, c, ev, vm)
        osp = orc.Space(kind, c, ev, vm)
        field = afe.field_random(8, sp.dim, 0.5, 5)
        ofield = orc.Field(field.k, field.phi, field.w, field.scale)
        A = afe.Matrix(sp)
        for op in (0, 1, 2):
            A.assemble(op, acc, field if op == 2 else None)
            rp, col, oval = osp.build(op, acc, ofield if op == 2 else None).csr()
            val = A.values()
            ok = rel_err_rows(val, oval, rp) <= 1e-13 if op == 2 else np.array_equal(val, oval)
            print(f"kind {kind} op {op} {'ok' if ok else 'MISMATCH'}", flush=True)
            assert ok


@pytest.mark.parametrize("chunks,xcd,layout,staged", [("8", "1", "element", "1"), ("1", "0", "entry", "0")])
def test_forced_assembly_launches_match_oracle(chunks, xcd, layout, staged):
    """The launch forms the defaults do not pick for a kind (pipelined
    triangles with XCD-ranged, LDS-staged gathers and element-major element
    matrices; tetrahedra in two whole-mesh launches with round-robin
    one-thread-per-row gathers and entry-major matrices) give the oracle's
    values too (a child process, since the forms are read from the
    environment once)"""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    env = dict(os.environ, AFE_ASM_CHUNKS=chunks, AFE_GATHER_XCD=xcd, AFE_KE_LAYOUT=layout,
               AFE_GATHER_STAGED=staged)
    print(f"case forced chunks={chunks} xcd={xcd} layout={layout} staged={staged}", flush=True)
    r = subprocess.run([sys.executable, "-c", "from tests.test_gpu_parity import _assembly_cases_match_oracle as f; f()"],
                       cwd=root, env=env, capture_output=True, text=True, timeout=100)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0 and r.stdout.count(" ok") == 9


def test_full_size_p2_assembly_matches_oracle(ctx):
    """The benchmark's own system, 1024x1024 squares x 2 P2 triangles
    (4,198,401 DOFs, 48,250,881 entries, jitter 0.2, seed 20261, acc 6):
    pattern and stiffness values bit-exact against the oracle's assembly of
    the same mesh"""
    import afepack_amd as afe
    from oracle import oracle_py as orc
    print("case full-size P2 assembly n=1024", flush=True)
    c, ev, vm = afe.mesh_square(1024, 0.2, 20261)
    sp = afe.Space(ctx, afe.P2_TRI, c, ev, vm)
    A = afe.Matrix(sp)
    A.assemble(afe.OP_STIFF, 6)
    val = A.values()
    prp, pcol = sp.pattern()
    del A
    osp = orc.Space(afe.P2_TRI, c, ev, vm)
    rp, col, oval = osp.build(orc.OP_STIFF, 6).csr()
    assert len(rp) == 4198402 and rp[-1] == 48250881
    assert np.array_equal(prp, rp) and np.array_equal(pcol, col)
    assert np.array_equal(val, oval)
