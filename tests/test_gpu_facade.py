"""GPU parity of the host-callback entry points and the C++ drop-in classes.

The quadrature-point entry points (afe_space_qpoints, afe_rhs_qp,
afe_apply_dirichlet_values, afe_error_qp) are compared with the oracle
directly; the C++ program afepack_amd/examples/facade_check drives the whole
path through StiffMatrix / MassMatrix / derived classes / L2Discretize /
BoundaryConditionAdmin / AMGSolver / AMGPreconditioner / Functional and its
outputs are compared with the oracle on the same mesh.

Bars: patterns, assembled matrices and the eliminated matrix bit-exact (the
host-evaluated derived class included); right-hand sides within 1e-13 of the
vector scale; PCG solutions within 1e-9 relative of a direct solve of the
oracle's system; error functionals within 1e-12 relative of the oracle's.
"""
import os
import subprocess

import numpy as np
import pytest

from tests.conftest import ROOT, make_mesh

pytestmark = pytest.mark.gpu

BIN = os.path.join(ROOT, "afepack_amd", "examples", "bin")


def _sinprod(P):
    return np.prod(np.sin(np.pi * P), axis=-1)


def _sinprod_grad(P):
    d = P.shape[-1]
    out = np.empty_like(P)
    for k in range(d):
        v = np.pi * np.cos(np.pi * P[..., k])
        for e in range(d):
            if e != k:
                v = v * np.sin(np.pi * P[..., e])
        out[..., k] = v
    return out


def _direct(OA, f):
    import scipy.sparse as sps
    import scipy.sparse.linalg as spl
    rp, col, val = OA.csr()
    n = len(rp) - 1
    A = sps.csr_matrix((val, col, rp), shape=(n, n))
    return spl.spsolve(A.tocsc(), f)


@pytest.mark.parametrize("kind,n,jitter", [(1, 32, 0.2), (2, 12, 0.2), (3, 5, 0.05)])
def test_qp_entry_points_match_oracle(ctx, kind, n, jitter):
    import afepack_amd as afe
    from oracle import oracle_py as orc
    print(f"case qp kind={kind} n={n}", flush=True)
    c, ev, vm = make_mesh(kind, n, jitter, 3)
    sp = afe.Space(ctx, kind, c, ev, vm)
    osp = orc.Space(kind, c, ev, vm)
    acc = 4
    pts, jxw = sp.qpoints(acc)
    _, qw = orc.quad(sp.dim, acc)
    assert pts.shape == (sp.n_elem, len(qw), sp.dim)
    # Jxw integrates 1 over the unit square / cube (the 11-point tetrahedron
    # rule has a negative centroid weight, so the sign is not checked)
    assert abs(jxw.sum() - 1.0) <= 1e-12
    # f at the points, as a host callback would give it: F_POISSON's form
    fq = sp.dim * np.pi ** 2 * _sinprod(pts)
    f = sp.rhs_qp(acc, fq).numpy()
    of = osp.rhs(acc, orc.F_POISSON)
    assert np.abs(f - of).max() <= 1e-13 * np.abs(of).max()
    # Dirichlet elimination with nonzero data on marks {1}
    A = afe.Matrix(sp)
    A.assemble(afe.OP_STIFF, acc)
    OA = osp.build(orc.OP_STIFF, acc)
    _, dm, dp = sp.dofs()
    P = dp.reshape(-1, sp.dim)
    g = 1.0 + 0.5 * P[:, 0] - 0.25 * P[:, 1] + np.sin(3 * P[:, 0]) * np.cos(2 * P[:, 1])
    df = afe.DeviceVector.from_numpy(ctx, of)
    A.apply_dirichlet_values([1], g, df)
    of2 = of.copy()
    osp.apply_dirichlet_values(OA, [1], g, of2)
    assert np.array_equal(A.values(), OA.csr()[2])
    fa = df.numpy()
    assert np.abs(fa - of2).max() <= 1e-13 * np.abs(of2).max()
    # errors of a solution against exact values / gradients at the points
    x = _direct(OA, of2)
    dx = afe.DeviceVector.from_numpy(ctx, x)
    # exact solution sinprod: its boundary data is zero, so compare the
    # functional itself on the same x
    e2 = sp.error_qp(dx, acc, _sinprod(pts))
    e1 = sp.error_qp(dx, acc, _sinprod_grad(pts), h1=True)
    o2 = osp.l2_error(x, acc)
    o1 = osp.h1_error(x, acc)
    assert abs(e2 - o2) <= 1e-12 * o2 and abs(e1 - o1) <= 1e-12 * o1


def _read(d, name, dtype=np.float64):
    return np.fromfile(os.path.join(d, name), dtype=dtype)


@pytest.mark.parametrize("kind,n", [(1, 24), (2, 10), (3, 5)])
def test_cpp_facade_matches_oracle(tmp_path, kind, n):
    import afepack_amd as afe
    from oracle import oracle_py as orc
    exe = os.path.join(BIN, "facade_check")
    if not os.path.exists(exe):
        pytest.fail("afepack_amd/examples/bin/facade_check is not built (run __graft_entry__.build())")
    print(f"case facade kind={kind} n={n}", flush=True)
    r = subprocess.run([exe, str(kind), str(n), str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    jitter = 0.05 if kind == 3 else 0.2
    c, ev, vm = make_mesh(kind, n, jitter, 7)
    osp = orc.Space(kind, c, ev, vm)
    acc = 4
    OK = osp.build(orc.OP_STIFF, acc)
    rp, col, kval = OK.csr()
    assert np.array_equal(_read(tmp_path, "rp.i64", np.int64), rp)
    assert np.array_equal(_read(tmp_path, "col.i64", np.int64), col)
    K = _read(tmp_path, "K.f64")
    assert np.array_equal(K, kval)
    # the derived class runs the same element matrix on the host: bitwise
    assert np.array_equal(_read(tmp_path, "KH.f64"), kval)
    assert np.array_equal(_read(tmp_path, "K2.f64"), 2.0 * kval)
    assert np.array_equal(_read(tmp_path, "M.f64"), osp.build(orc.OP_MASS, acc).csr()[2])
    of = osp.rhs(acc, orc.F_POISSON)
    f = _read(tmp_path, "f.f64")
    assert np.abs(f - of).max() <= 1e-13 * np.abs(of).max()
    g = _read(tmp_path, "g.f64")
    # apply(..., false): the boundary rows only, on a copy of the system
    OKn = osp.build(orc.OP_STIFF, acc)
    ofn = of.copy()
    osp.apply_dirichlet_values(OKn, [1], g, ofn, preserve_symmetry=False)
    assert np.array_equal(_read(tmp_path, "Kn.f64"), OKn.csr()[2])
    assert np.abs(_read(tmp_path, "fn.f64") - ofn).max() <= 1e-13 * np.abs(ofn).max()
    osp.apply_dirichlet_values(OK, [1], g, of)
    assert np.array_equal(_read(tmp_path, "Ka.f64"), OK.csr()[2])
    fa = _read(tmp_path, "fa.f64")
    assert np.abs(fa - of).max() <= 1e-13 * np.abs(of).max()
    assert not np.any(_read(tmp_path, "ua.f64"))  # apply leaves u untouched
    xd = _direct(OK, of)
    xp = _read(tmp_path, "xp.f64")
    assert np.linalg.norm(xp - xd) <= 1e-9 * np.linalg.norm(xd)
    # stationary AMGSolver::solve, 61 cycles: converging to the same solution
    x = _read(tmp_path, "x.f64")
    assert np.linalg.norm(x - xd) <= 1e-3 * np.linalg.norm(xd)
    # AMGPreconditioner::vmult from zero reduces the residual
    rpv, colv, vv = OK.csr()
    import scipy.sparse as sps
    A = sps.csr_matrix((vv, colv, rpv), shape=(len(rpv) - 1,) * 2)
    xv = _read(tmp_path, "xv.f64")
    assert np.linalg.norm(of - A @ xv) <= 0.5 * np.linalg.norm(of)
    l2, h1, its = open(os.path.join(tmp_path, "err.txt")).read().split()
    o2, o1 = osp.l2_error(xp, acc), osp.h1_error(xp, acc)
    assert abs(float(l2) - o2) <= 1e-12 * o2 and abs(float(h1) - o1) <= 1e-12 * o1

    # L2InnerProduct on one space: the mass matrix, bit for bit
    om = osp.build(orc.OP_MASS, acc).csr()[2]
    assert np.array_equal(_read(tmp_path, "L2.f64"), om)
    # L2InnerProduct of this space and a second one on the same mesh (the
    # other order in 2D, another P1 space in 3D): pattern, values and
    # nMaxCouplingDof of the oracle's two-space build; the host-evaluated
    # derived class doubling it gathers to exactly twice the values
    osp2 = orc.Space(kind if kind == 3 else 3 - kind, c, ev, vm)
    OX, omc = osp.build_pair(osp2, orc.OP_MASS, acc)
    xrp, xcol, xval = OX.csr()
    assert np.array_equal(_read(tmp_path, "L2x_rp.i64", np.int64), xrp)
    assert np.array_equal(_read(tmp_path, "L2x_col.i64", np.int64), xcol)
    assert np.array_equal(_read(tmp_path, "L2x.f64"), xval)
    assert np.array_equal(_read(tmp_path, "L2xd.f64"), 2.0 * xval)
    info = tuple(map(int, open(os.path.join(tmp_path, "infox.txt")).read().split()))
    assert info == (omc, osp.n_dof, osp2.n_dof, osp.n_dof, osp2.n_dof)
    # a user operator declared BilinearOperator<DIM, double> (the default
    # second value type), stiffness plus mass in one host loop
    from tests.conftest import rel_err_rows
    assert rel_err_rows(_read(tmp_path, "KM.f64"), osp.build(orc.OP_STIFF, acc).csr()[2] + om, rp) <= 1e-13
    # nMaxCouplingDof (BilinearOperator.templates.h:115-172): the largest sum
    # of element DOF counts over the elements holding a DOF, capped at nDof
    ed, _, _ = osp.dofs()
    nl = ed.size // (len(ev) // (4 if kind == 3 else 3))
    coupling = np.bincount(ed, minlength=len(rp) - 1) * nl
    mc, nd0, nd1 = (int(t) for t in open(os.path.join(tmp_path, "info.txt")).read().split())
    assert (mc, nd0, nd1) == (min(int(coupling.max()), len(rp) - 1), len(rp) - 1, len(rp) - 1)
    # AMGSolver with isSolveMostProjectExactly() = false (Gauss-Seidel on the
    # coarsest level): 401 cycles reach the direct solution
    xg = _read(tmp_path, "xg.f64")
    assert np.linalg.norm(xg - xd) <= 1e-9 * np.linalg.norm(xd)
    # moving mesh: the stiffness re-assembled on the moved vertices equals the
    # oracle's assembly there, and the lazy hierarchy refreshed on it equals
    # the oracle's lazyReinit of the same values (src/AMGSolver.cpp:505-545)
    cm = _read(tmp_path, "cm.f64")
    osp_m = orc.Space(kind, cm, ev, vm)
    km = osp_m.build(orc.OP_STIFF, acc).csr()[2]
    assert np.array_equal(_read(tmp_path, "Km.f64"), km)
    og = orc.Amg(orc.Mat.from_csr(len(rp) - 1, len(rp) - 1, rp, col, _read(tmp_path, "Ka.f64")), lazy=True)
    o_before = og.A(1).csr()[2].copy()
    assert og.lazy_reinit(og.mat.share_pattern(km)) == 0
    lz1 = _read(tmp_path, "lz1.f64")
    assert np.array_equal(lz1, og.A(1).csr()[2]) and not np.array_equal(lz1, o_before)
    # another SparsityPattern object: the reference's abort path was taken
    assert open(os.path.join(tmp_path, "foreign.txt")).read().split() == ["1"]
    # a host-assembled matrix refreshed twice in place (values 2 K, then 3 K)
    oh = orc.Amg(orc.Mat.from_csr(len(rp) - 1, len(rp) - 1, rp, col, 3.0 * km), lazy=True)
    assert np.array_equal(_read(tmp_path, "hl1.f64"), oh.A(1).csr()[2])


def test_moving_mesh_example_runs(tmp_path):
    """afepack_amd/examples/moving_mesh_cg: a user CG loop preconditioned by
    AMGPreconditioner on a moving P2 mesh with a variable coefficient; every
    position converges and the L2 error is at the P2 level for h = 1/32."""
    exe = os.path.join(BIN, "moving_mesh_cg")
    if not os.path.exists(exe):
        pytest.fail("afepack_amd/examples/bin/moving_mesh_cg is not built")
    print("case example moving_mesh_cg", flush=True)
    r = subprocess.run([exe, "32", "3"], capture_output=True, text=True, timeout=120, cwd=tmp_path)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    lines = [s for s in r.stdout.splitlines() if s.startswith("mesh ")]
    assert len(lines) == 3
    for s in lines:
        its = int(s.split("pcg ")[1].split()[0])
        assert 0 < its < 100
        l2 = float(s.split("L2 ")[1].split(",")[0])
        assert 0 < l2 < 1e-4


def _write_easymesh(base, c, ev, vm, flip):
    """EasyMesh .n/.s/.e files of a triangulation; elements listed in flip
    are written clockwise (the reader turns them back)"""
    C = c.reshape(-1, 2)
    E = ev.reshape(-1, 3).copy()
    E[flip] = E[flip][:, [0, 2, 1]]
    sides = {}
    for e in E:
        for k in range(3):
            a, b = sorted((int(e[(k + 1) % 3]), int(e[(k + 2) % 3])))
            sides.setdefault((a, b), 0)
            sides[(a, b)] += 1
    side_list = list(sides)
    sid = {s: i for i, s in enumerate(side_list)}
    with open(base + ".n", "w") as f:
        f.write(f"{len(C)} {len(E)} {len(side_list)} **(Nnd, Nee, Nsd)**\n")
        for i, p in enumerate(C):
            f.write(f"{i} {float(p[0])!r} {float(p[1])!r} {int(vm[i])}\n")
    with open(base + ".s", "w") as f:
        f.write(f"{len(side_list)}\n")
        for i, (a, b) in enumerate(side_list):
            f.write(f"{i} {a} {b} -1 -1 {1 if sides[(a, b)] == 1 else 0}\n")
    with open(base + ".e", "w") as f:
        f.write(f"{len(E)} {len(C)} {len(side_list)} **(Nee, Nnd, Nsd)**\n")
        for i, e in enumerate(E):
            s = [sid[tuple(sorted((int(e[(k + 1) % 3]), int(e[(k + 2) % 3]))))] for k in range(3)]
            f.write(f"{i} {e[0]} {e[1]} {e[2]} -1 -1 -1 {s[0]} {s[1]} {s[2]}\n")


@pytest.mark.parametrize("name,fmt", [("D", "easymesh"), ("Lshape", "easymesh"), ("Lshape", "mesh")])
@pytest.mark.parametrize("kind", [1, 2])
def test_cpp_facade_generated_meshes(tmp_path, kind, name, fmt):
    """Meshes made by the reference's own generator (tests/golden/easymesh),
    read by the drop-in's EasyMesh reader or, rewritten in the Mesh text
    format, by Mesh::readData: pattern and stiffness bit-exact and the
    right-hand side within 1e-13 of the oracle's on the same mesh, and the
    Dirichlet elimination on mark 1 -- with P2 edge DOFs marked by their
    sides (several boundary segments) -- bit-exact."""
    from oracle import oracle_py as orc
    from tests.test_mesh_ingest import GOLD, read_easymesh, write_mesh_format
    exe = os.path.join(BIN, "facade_check")
    if not os.path.exists(exe):
        pytest.fail("afepack_amd/examples/bin/facade_check is not built")
    print(f"case generated mesh {name} {fmt} kind={kind}", flush=True)
    coords, vmark, elem, sv, sm = read_easymesh(os.path.join(GOLD, name))
    src = os.path.join(GOLD, name)
    if fmt == "mesh":
        src = str(tmp_path / (name + ".mesh"))
        write_mesh_format(src, coords, vmark, elem, sv, sm, np.random.default_rng(11))
    out = tmp_path / "out"
    out.mkdir()
    r = subprocess.run([exe, str(kind), "0", str(out), src], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    osp = orc.Space(kind, coords.ravel(), elem.ravel(), vmark, sides=(sv, sm))
    OK = osp.build(orc.OP_STIFF, 4)
    rp, col, kval = OK.csr()
    assert np.array_equal(_read(out, "rp.i64", np.int64), rp)
    assert np.array_equal(_read(out, "col.i64", np.int64), col)
    assert np.array_equal(_read(out, "K.f64"), kval)
    of = osp.rhs(4, orc.F_POISSON)
    assert np.abs(_read(out, "f.f64") - of).max() <= 1e-13 * np.abs(of).max()
    g = _read(out, "g.f64")
    osp.apply_dirichlet_values(OK, [1], g, of)
    assert np.array_equal(_read(out, "Ka.f64"), OK.csr()[2])
    assert np.abs(_read(out, "fa.f64") - of).max() <= 1e-13 * np.abs(of).max()


@pytest.mark.parametrize("kind", [1, 2])
def test_cpp_facade_easymesh_input(tmp_path, kind):
    """EasyMesh ingestion (EasyMesh.templates.h:16-94): elements in a
    shuffled order, some written clockwise; the facade's pattern, matrix and
    right-hand side equal the oracle's on the re-oriented mesh bit for bit."""
    import afepack_amd as afe
    from oracle import oracle_py as orc
    exe = os.path.join(BIN, "facade_check")
    if not os.path.exists(exe):
        pytest.fail("afepack_amd/examples/bin/facade_check is not built")
    print(f"case easymesh kind={kind}", flush=True)
    c, ev, vm = afe.mesh_square(10, 0.2, 41)
    rng = np.random.default_rng(5)
    E = ev.reshape(-1, 3)[rng.permutation(len(ev) // 3)]
    flip = rng.uniform(size=len(E)) < 0.3
    base = str(tmp_path / "mesh")
    _write_easymesh(base, c, E.ravel(), vm, flip)
    out = tmp_path / "out"
    out.mkdir()
    r = subprocess.run([exe, str(kind), "0", str(out), base], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    osp = orc.Space(kind, c, np.ascontiguousarray(E.ravel()), vm)
    OK = osp.build(orc.OP_STIFF, 4)
    rp, col, kval = OK.csr()
    assert np.array_equal(_read(out, "rp.i64", np.int64), rp)
    assert np.array_equal(_read(out, "col.i64", np.int64), col)
    assert np.array_equal(_read(out, "K.f64"), kval)
    of = osp.rhs(4, orc.F_POISSON)
    f = _read(out, "f.f64")
    assert np.abs(f - of).max() <= 1e-13 * np.abs(of).max()
