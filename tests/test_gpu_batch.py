"""GPU tests of the batched solver (afe_batch_*).

Bars: every member's solution is bitwise identical alone, in a full batch
and when admitted while other members are mid-solve; batched solutions are
within 1e-9 relative L2 of a direct solve of the oracle's system (rtol 1e-12).
"""
import numpy as np
import pytest

from tests.conftest import make_mesh

pytestmark = pytest.mark.gpu

# kind, n, jitter, op, rhs
PROBLEMS = [
    (1, 24, 0.2, 0, "poisson"),
    (2, 12, 0.2, 0, "poisson"),
    (1, 40, 0.1, 2, "nodal"),
    (3, 5, 0.05, 2, "poisson"),
    (2, 20, 0.0, 2, "nodal"),
    (1, 8, 0.0, 0, "poisson"),
    (3, 6, 0.05, 2, "nodal", "16"),  # 16 lanes per row on its coarse levels
]


def _problem(ctx, kind, n, jitter, op, rhs, lanes=None, seed=0):
    import os
    import afepack_amd as afe
    c, ev, vm = make_mesh(kind, n, jitter, seed)
    sp = afe.Space(ctx, kind, c, ev, vm)
    field = afe.field_random(8, sp.dim, 0.5, seed) if op == 2 else None
    A = afe.Matrix(sp)
    A.assemble(op, 4, field)
    if rhs == "nodal":
        rng = np.random.default_rng(seed)
        b = sp.rhs(4, afe.F_NODAL, nodal=rng.uniform(-1, 1, sp.n_dof))
    else:
        b = sp.rhs(4, afe.F_POISSON)
    A.apply_dirichlet(b)
    if lanes:
        os.environ["AFE_ROW_LANES"] = lanes
    try:
        g = afe.Amg(A, afe.AmgParams(lazy=1, smooth=1, build_fp32=1))
    finally:
        os.environ.pop("AFE_ROW_LANES", None)
    if lanes:
        assert [d["row_lanes"] for d in g.device_info()][1:] == [int(lanes)] * (g.info()["levels"] - 1)
    return {"sp": sp, "A": A, "b": b, "g": g, "field": field}


@pytest.fixture(scope="module")
def problems(ctx):
    return [_problem(ctx, *p, seed=31 + i) for i, p in enumerate(PROBLEMS)]


def _solve_batch(ctx, probs, order, rtol, stagger=0):
    import afepack_amd as afe
    bt = afe.Batch(ctx, afe.PREC_FP32, smooth=1)
    xs, slots = {}, {}
    first = order[: len(order) - stagger] if stagger else order
    for i in first:
        xs[i] = afe.DeviceVector(ctx, probs[i]["sp"].n_dof)
        slots[i] = bt.add(probs[i]["g"], xs[i], probs[i]["b"], rtol=rtol, max_it=400)
    if stagger:
        for _ in range(3):
            bt.step()
        for i in order[len(order) - stagger:]:
            xs[i] = afe.DeviceVector(ctx, probs[i]["sp"].n_dof)
            slots[i] = bt.add(probs[i]["g"], xs[i], probs[i]["b"], rtol=rtol, max_it=400)
    bt.run()
    out = {}
    for i in order:
        done, st, code = bt.result(slots[i])
        assert done and code == 0 and st.converged == 1, f"member {i}: code {code}, res {st.rel_res_l2}"
        assert st.rel_res_l2 <= rtol
        out[i] = (xs[i].numpy(), st.iterations)
    bt.close()
    return out


def test_batch_alone_equals_mixed_bitwise(ctx, problems):
    print("case batch alone vs mixed", flush=True)
    idx = list(range(len(problems)))
    alone = {}
    for i in idx:
        alone.update(_solve_batch(ctx, problems, [i], 1e-10))
    mixed = _solve_batch(ctx, problems, idx, 1e-10)
    late = _solve_batch(ctx, problems, idx[::-1], 1e-10, stagger=3)
    for i in idx:
        assert np.array_equal(alone[i][0], mixed[i][0]), f"member {i} differs in the mix"
        assert np.array_equal(alone[i][0], late[i][0]), f"member {i} differs when admitted late"
        assert alone[i][1] == mixed[i][1] == late[i][1]


def test_batch_members_from_other_contexts(ctx, problems):
    """hierarchies built in two other contexts (own streams and locks, the
    server's set-up threads) join a batch of this context and give the same
    solutions bit for bit as the ones built here"""
    import afepack_amd as afe
    print("case batch members from other contexts", flush=True)
    idx = list(range(len(problems)))
    here = _solve_batch(ctx, problems, idx, 1e-10)
    others = [afe.Context(0), afe.Context(0)]
    theirs = [_problem(others[i % 2], *p, seed=31 + i) for i, p in enumerate(PROBLEMS)]
    away = _solve_batch(ctx, theirs, idx, 1e-10)
    for i in idx:
        assert np.array_equal(here[i][0], away[i][0]), f"member {i} differs when built in another context"
        assert here[i][1] == away[i][1]
    del theirs


@pytest.mark.parametrize("tail", [4096, 0])
def test_batch_solutions_match_direct(ctx, problems, tail):
    """tail 0: every level of every member but its coarsest runs in the
    shared per-level launches (several lanes per row on the 3D member's
    coarse levels), else the small levels run in the per-member tails."""
    import scipy.sparse as sps
    import scipy.sparse.linalg as spl
    print(f"case batch vs direct tail={tail}", flush=True)
    idx = list(range(len(problems)))
    try:
        for p in problems:
            p["g"].set_tail_rows(tail)
        res = _solve_batch(ctx, problems, idx, 1e-12)
    finally:
        for p in problems:
            p["g"].set_tail_rows(4096)
    for i in idx:
        p = problems[i]
        rp, col = p["sp"].pattern()
        A = sps.csr_matrix((p["A"].values(), col, rp), shape=(p["sp"].n_dof,) * 2)
        xd = spl.spsolve(A.tocsc(), p["b"].numpy())
        x = res[i][0]
        assert np.linalg.norm(x - xd) <= 1e-9 * np.linalg.norm(xd), f"member {i}"


def test_batch_matches_unbatched_iterations(ctx, problems):
    import afepack_amd as afe
    print("case batch vs afe_amg_pcg", flush=True)
    idx = list(range(len(problems)))
    res = _solve_batch(ctx, problems, idx, 1e-10)
    for i in idx:
        p = problems[i]
        x = afe.DeviceVector(ctx, p["sp"].n_dof)
        p["g"].set_smooth(1)
        st = p["g"].pcg(x, p["b"], rtol=1e-10, max_it=400, prec=afe.PREC_FP32)
        # same algorithm, different reduction blocking: same iteration count
        # up to one, solutions agree to the solve tolerance
        assert abs(st.iterations - res[i][1]) <= 1
        xr = x.numpy()
        assert np.linalg.norm(xr - res[i][0]) <= 1e-6 * np.linalg.norm(xr)


@pytest.fixture(scope="module")
def c4_members(ctx):
    """32 members drawn as the 256-problem batch is: n = round(exp(U[ln 32,
    ln 512])), even members P1 (accuracy 4) and odd P2 (accuracy 6), a =
    exp(8-mode field of amplitude 0.5), f uniform in [-1, 1] at the DOFs"""
    import afepack_amd as afe
    rng = np.random.default_rng(4321)
    sizes = np.rint(np.exp(rng.uniform(np.log(32), np.log(512), 32))).astype(int)
    out = []
    for i, n in enumerate(sizes):
        kind, acc = (afe.P1_TRI, 4) if i % 2 == 0 else (afe.P2_TRI, 6)
        c, ev, vm = afe.mesh_square(int(n), 0.0, 0)
        sp = afe.Space(ctx, kind, c, ev, vm)
        field = afe.field_random(8, 2, 0.5, 7000 + i)
        A = afe.Matrix(sp)
        A.assemble(afe.OP_VARSTIFF, acc, field)
        b = sp.rhs(acc, afe.F_NODAL, nodal=np.random.default_rng(8000 + i).uniform(-1, 1, sp.n_dof))
        A.apply_dirichlet(b)
        g = afe.Amg(A, afe.AmgParams(lazy=1, smooth=1, build_fp32=1))
        out.append({"sp": sp, "A": A, "b": b, "g": g})
    return out


def _run_members(ctx, probs, idx, tiles=(1, 1)):
    import afepack_amd as afe
    bt = afe.Batch(ctx, afe.PREC_FP32, smooth=1)
    bt.set_tiles(*tiles)
    xs, slots = {}, {}
    for i in idx:
        xs[i] = afe.DeviceVector(ctx, probs[i]["sp"].n_dof)
        slots[i] = bt.add(probs[i]["g"], xs[i], probs[i]["b"], rtol=1e-10, max_it=1000)
    bt.run()
    out = {}
    for i in idx:
        done, st, code = bt.result(slots[i])
        assert done and code == 0 and st.converged == 1 and st.rel_res_l2 <= 1e-10, (i, code)
        out[i] = xs[i].numpy()
    bt.close()
    return out


def test_c4_distribution_batch(ctx, c4_members):
    """32 members of the C4 distribution in one batch: launch buckets
    (slices per wavefront on level 0 and on the coarser levels) change no
    bit of any member; members solved alone equal the mix bit for bit; each
    member's true residual, computed here by scipy from its assembled matrix,
    is at most 1e-10 of its right-hand side."""
    import scipy.sparse as sps
    probs = c4_members
    idx = list(range(len(probs)))
    print(f"case c4 distribution: {sum(p['sp'].n_dof for p in probs)} DOFs", flush=True)
    base = _run_members(ctx, probs, idx)
    for tiles in ((4, 2), (2, 8)):
        other = _run_members(ctx, probs, idx, tiles)
        for i in idx:
            assert np.array_equal(other[i], base[i]), (tiles, i)
    for i in (0, 1, 17, 30):
        assert np.array_equal(_run_members(ctx, probs, [i])[i], base[i]), i
    for i in idx:
        p = probs[i]
        rp, col = p["sp"].pattern()
        A = sps.csr_matrix((p["A"].values(), col, rp), shape=(p["sp"].n_dof,) * 2)
        b = p["b"].numpy()
        assert np.linalg.norm(b - A @ base[i]) <= 1e-10 * np.linalg.norm(b), i
