"""GPU edge cases: the smallest meshes (one square / one cube), single-level
hierarchies, and the batch with members of one level next to deep ones."""
import numpy as np
import pytest

from tests.conftest import make_mesh

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind,n", [(1, 1), (2, 1), (3, 1), (1, 2), (2, 2), (3, 2)])
def test_tiny_meshes_match_oracle(ctx, kind, n):
    import afepack_amd as afe
    from oracle import oracle_py as orc
    print(f"case tiny kind={kind} n={n}", flush=True)
    c, ev, vm = make_mesh(kind, n, 0.0, 0)
    sp = afe.Space(ctx, kind, c, ev, vm)
    osp = orc.Space(kind, c, ev, vm)
    acc = 2 if kind == 3 else 4
    A = afe.Matrix(sp)
    A.assemble(afe.OP_STIFF, acc)
    OA = osp.build(orc.OP_STIFF, acc)
    rp, col, oval = OA.csr()
    prp, pcol = sp.pattern()
    assert np.array_equal(prp, rp) and np.array_equal(pcol, col)
    assert np.array_equal(A.values(), oval)
    f = sp.rhs(acc, afe.F_ONE)
    of = osp.rhs(acc, orc.F_ONE)
    A.apply_dirichlet(f)
    osp.apply_dirichlet(OA, of)
    xd = np.linalg.solve(OA.dense(), of)
    for lazy in (0, 1):
        g = afe.Amg(A, afe.AmgParams(lazy=lazy, smooth=1, build_fp32=1))
        x = afe.DeviceVector(ctx, sp.n_dof)
        st = g.pcg(x, f, rtol=1e-12, max_it=50, prec=afe.PREC_FP32)
        assert st.converged == 1
        assert np.linalg.norm(x.numpy() - xd) <= 1e-10 * max(1.0, np.linalg.norm(xd))
        x.zero()
        g.solve(x, f, step=20, rtol=1e-13)
        assert np.linalg.norm(x.numpy() - xd) <= 1e-10 * max(1.0, np.linalg.norm(xd))


def test_batch_with_single_level_members(ctx):
    import afepack_amd as afe
    print("case batch tiny+deep", flush=True)
    probs = []
    for kind, n in ((1, 2), (2, 1), (1, 48), (3, 2), (2, 24)):
        c, ev, vm = make_mesh(kind, n, 0.0, 0)
        sp = afe.Space(ctx, kind, c, ev, vm)
        A = afe.Matrix(sp)
        acc = 2 if kind == 3 else 4
        A.assemble(afe.OP_STIFF, acc)
        b = sp.rhs(acc, afe.F_POISSON)
        A.apply_dirichlet(b)
        g = afe.Amg(A, afe.AmgParams(lazy=1, smooth=1, build_fp32=1))
        probs.append((sp, A, b, g))
    assert any(p[3].info()["levels"] == 1 for p in probs)
    assert any(p[3].info()["levels"] > 2 for p in probs)
    bt = afe.Batch(ctx, afe.PREC_FP32, smooth=1)
    xs = [afe.DeviceVector(ctx, p[0].n_dof) for p in probs]
    slots = [bt.add(p[3], x, p[2], rtol=1e-10, max_it=200) for p, x in zip(probs, xs)]
    bt.run()
    for k, (p, x) in enumerate(zip(probs, xs)):
        done, st, code = bt.result(slots[k])
        assert done and code == 0 and st.rel_res_l2 <= 1e-10
    bt.close()


def test_live_timing_keeps_results(ctx):
    """The SpMV launched with dispatch-timed events gives the same iterate as
    the captured iteration, and one timing per CG iteration."""
    import afepack_amd as afe
    print("case live timing P2 n=64", flush=True)
    c, ev, vm = make_mesh(2, 64, 0.2, 3)
    sp = afe.Space(ctx, 2, c, ev, vm)
    A = afe.Matrix(sp)
    A.assemble(afe.OP_STIFF, 4)
    b = sp.rhs(4, afe.F_POISSON)
    A.apply_dirichlet(b)
    g = afe.Amg(A, afe.AmgParams(lazy=1, smooth=1, build_fp32=1))
    x0 = afe.DeviceVector(ctx, sp.n_dof)
    st0 = g.pcg(x0, b, rtol=1e-10, max_it=200, prec=afe.PREC_FP32)
    g.live_timing(True)
    x1 = afe.DeviceVector(ctx, sp.n_dof)
    st1 = g.pcg(x1, b, rtol=1e-10, max_it=200, prec=afe.PREC_FP32)
    ms, n = g.live_timing(False)
    assert st0.converged == 1 and st1.converged == 1
    assert st1.iterations == st0.iterations
    assert np.array_equal(x1.numpy(), x0.numpy())
    assert n == st1.iterations and ms > 0.0
    # stopped: another solve adds no launches
    x2 = afe.DeviceVector(ctx, sp.n_dof)
    g.pcg(x2, b, rtol=1e-10, max_it=200, prec=afe.PREC_FP32)
    ms2, n2 = g.live_timing(False)
    assert n2 == n and ms2 == ms


def test_part_timing_keeps_results(ctx):
    """Events between the two captured graphs of each CG iteration leave the
    iterate unchanged and time every iteration's two parts"""
    import afepack_amd as afe
    print("case part timing P2 n=64", flush=True)
    c, ev, vm = make_mesh(2, 64, 0.2, 3)
    sp = afe.Space(ctx, 2, c, ev, vm)
    A = afe.Matrix(sp)
    A.assemble(afe.OP_STIFF, 4)
    b = sp.rhs(4, afe.F_POISSON)
    A.apply_dirichlet(b)
    g = afe.Amg(A, afe.AmgParams(lazy=1, smooth=1, build_fp32=1))
    x0 = afe.DeviceVector(ctx, sp.n_dof)
    st0 = g.pcg(x0, b, rtol=1e-10, max_it=200, prec=afe.PREC_FP32)
    g.part_timing(True)
    x1 = afe.DeviceVector(ctx, sp.n_dof)
    st1 = g.pcg(x1, b, rtol=1e-10, max_it=200, prec=afe.PREC_FP32)
    p0, p1, n = g.part_timing(False)
    assert st1.iterations == st0.iterations and np.array_equal(x1.numpy(), x0.numpy())
    assert n == st1.iterations and p0 > 0.0 and p1 > 0.0
    assert (p0 + p1) * n <= 1e3 * st1.seconds * 1.05  # the parts lie inside the solve
    x2 = afe.DeviceVector(ctx, sp.n_dof)
    g.pcg(x2, b, rtol=1e-10, max_it=200, prec=afe.PREC_FP32)
    assert g.part_timing(False)[2] == n  # stopped


@pytest.mark.parametrize("kind,n,stream", [(2, 48, None), (3, 10, None), (2, 48, "0"), (3, 10, "0")])
def test_natural_columns_16_and_32_bit_agree(ctx, kind, n, stream):
    """The natural-order FP64 copy of level 0 stores 16-bit column offsets
    from a per-slice base when the columns of every slice span at most
    65535, and streamed levels (all of them with AFE_STREAM_BYTES=0) store
    colour-coded 16-bit columns for the Gauss-Seidel and residual kernels;
    with AFE_NO_C16 set both keep 32-bit columns.  Both forms decode to the
    same columns, so the PCG iterates are bitwise equal."""
    import os
    import afepack_amd as afe
    print(f"case c16 kind={kind} n={n} stream={stream}", flush=True)
    if stream is not None:
        os.environ["AFE_STREAM_BYTES"] = stream
    c, ev, vm = make_mesh(kind, n, 0.05 if kind == 3 else 0.2, 4)
    sp = afe.Space(ctx, kind, c, ev, vm)
    A = afe.Matrix(sp)
    acc = 2 if kind == 3 else 4
    A.assemble(afe.OP_STIFF, acc)
    b = sp.rhs(acc, afe.F_POISSON)
    A.apply_dirichlet(b)
    xs = []
    for no16 in (False, True):
        if no16:
            os.environ["AFE_NO_C16"] = "1"
        try:
            g = afe.Amg(A, afe.AmgParams(lazy=1, smooth=1, build_fp32=1))
        finally:
            os.environ.pop("AFE_NO_C16", None)
        x = afe.DeviceVector(ctx, sp.n_dof)
        st = g.pcg(x, b, rtol=1e-10, max_it=300, prec=afe.PREC_FP32)
        x64 = afe.DeviceVector(ctx, sp.n_dof)
        g.pcg(x64, b, rtol=1e-10, max_it=300, prec=afe.PREC_FP64)
        xst = afe.DeviceVector(ctx, sp.n_dof)
        g.solve(xst, b, step=4, mode=1)
        ms, by = g.kernel_timing(0, reps=2)
        ms1, by1 = g.kernel_timing(1, reps=1)
        xs.append((st.iterations, x.numpy(), by, x64.numpy(), xst.numpy(), by1))
        g.close()
    os.environ.pop("AFE_STREAM_BYTES", None)
    assert xs[0][0] == xs[1][0]
    assert np.array_equal(xs[0][3], xs[1][3]) and np.array_equal(xs[0][4], xs[1][4])
    if stream is not None:  # level 0 colour-coded: 2 bytes per entry less in the sweep
        assert xs[1][5] > xs[0][5]
    assert np.array_equal(xs[0][1], xs[1][1])
    nnz = sp.nnz
    assert xs[1][2] - xs[0][2] >= 2 * nnz - 4 * (sp.n_dof // 64 + 1) - 1  # 2 bytes per entry saved


@pytest.mark.parametrize("tail", [4096, 0])
def test_row_lanes_agree_with_one_lane_per_row(ctx, tail):
    """Small levels of long rows split each row over 2, 4, 8 or 16 lanes
    (per-lane partial sums added by a butterfly); with AFE_ROW_LANES=1 every
    level keeps one lane per row.  The two layouts hold the same matrices and
    differ only in the order of each row's sum: the V-cycles agree to FP64
    rounding and the PCG solves to the solve tolerance."""
    import os
    import afepack_amd as afe
    print(f"case row lanes tail={tail}", flush=True)
    c, ev, vm = make_mesh(3, 12, 0.05, 5)
    sp = afe.Space(ctx, 3, c, ev, vm)
    A = afe.Matrix(sp)
    A.assemble(afe.OP_VARSTIFF, 2, afe.field_random(8, 3, 0.5, 5))
    b = sp.rhs(2, afe.F_POISSON)
    A.apply_dirichlet(b)
    out = []
    for lanes in (None, "1"):
        if lanes:
            os.environ["AFE_ROW_LANES"] = lanes
        try:
            g = afe.Amg(A, afe.AmgParams(lazy=1, smooth=1, build_fp32=1))
        finally:
            os.environ.pop("AFE_ROW_LANES", None)
        g.set_tail_rows(tail)
        rl = [d["row_lanes"] for d in g.device_info()]
        z = afe.DeviceVector(ctx, sp.n_dof)
        g.vcycle(z, b, afe.PREC_FP64)
        x64 = afe.DeviceVector(ctx, sp.n_dof)
        s64 = g.pcg(x64, b, rtol=1e-11, max_it=300, prec=afe.PREC_FP64)
        x32 = afe.DeviceVector(ctx, sp.n_dof)
        s32 = g.pcg(x32, b, rtol=1e-11, max_it=300, prec=afe.PREC_FP32)
        out.append((rl, z.numpy(), s64, x64.numpy(), s32, x32.numpy()))
        g.close()
    (rl, z, s64, x64, s32, x32), (rl1, z1, t64, y64, t32, y32) = out
    assert rl[0] == 1 and max(rl) > 1 and set(rl1) == {1}
    assert np.linalg.norm(z - z1) <= 1e-13 * np.linalg.norm(z1)
    assert s64.converged == t64.converged == 1 and abs(s64.iterations - t64.iterations) <= 1
    assert s32.converged == t32.converged == 1 and abs(s32.iterations - t32.iterations) <= 1
    assert np.linalg.norm(x64 - y64) <= 1e-9 * np.linalg.norm(y64)
    assert np.linalg.norm(x32 - y32) <= 1e-9 * np.linalg.norm(y32)


@pytest.mark.parametrize("kind,n", [(1, 128), (3, 24)])
def test_tail_with_levels_beyond_lds(ctx, kind, n):
    """The single-workgroup tail keeps a level's vectors (and slice tables)
    in LDS when they fit and works on global memory otherwise: with every
    level in the tail (level 0 of 16641 / 15625 rows does not fit), the
    V-cycle agrees with per-level launches to rounding and the solves agree
    to the tolerance, in FP32 and FP64."""
    import afepack_amd as afe
    print(f"case big tail kind={kind} n={n}", flush=True)
    c, ev, vm = make_mesh(kind, n, 0.05 if kind == 3 else 0.2, 9)
    sp = afe.Space(ctx, kind, c, ev, vm)
    A = afe.Matrix(sp)
    acc = 2 if kind == 3 else 4
    A.assemble(afe.OP_STIFF, acc)
    b = sp.rhs(acc, afe.F_POISSON)
    A.apply_dirichlet(b)
    g = afe.Amg(A, afe.AmgParams(lazy=1, smooth=1, build_fp32=1))
    out = {}
    for tail in (1 << 22, 0):
        g.set_tail_rows(tail)
        assert g.device_info()[0]["tail"] == (0 if tail else g.info()["levels"] - 1)
        z = afe.DeviceVector(ctx, sp.n_dof)
        g.vcycle(z, b, afe.PREC_FP64)
        res = [z.numpy()]
        for prec in (afe.PREC_FP64, afe.PREC_FP32):
            x = afe.DeviceVector(ctx, sp.n_dof)
            st = g.pcg(x, b, rtol=1e-11, max_it=300, prec=prec)
            assert st.converged == 1
            res += [x.numpy(), st.iterations]
        out[tail] = res
    big, lau = out[1 << 22], out[0]
    assert np.linalg.norm(big[0] - lau[0]) <= 1e-12 * np.linalg.norm(lau[0])
    for k in (1, 3):
        assert abs(big[k + 1] - lau[k + 1]) <= 1
        assert np.linalg.norm(big[k] - lau[k]) <= 1e-9 * np.linalg.norm(lau[k])
    g.close()


def test_stationary_cycles_back_to_back_match_checked_cycles(ctx):
    """With rtol = 0 (the reference's AMGSolver::solve: step + 1 cycles, no
    early exit) the cycles run back to back and the residual is taken once;
    with a positive rtol that is never met the residual is checked after
    every cycle.  Both run the same cycles: the iterates are bitwise equal."""
    import afepack_amd as afe
    print("case stationary back to back", flush=True)
    c, ev, vm = make_mesh(2, 12, 0.2, 3)
    sp = afe.Space(ctx, 2, c, ev, vm)
    A = afe.Matrix(sp)
    A.assemble(afe.OP_STIFF, 6)
    b = sp.rhs(6, afe.F_POISSON)
    A.apply_dirichlet(b)
    g = afe.Amg(A)
    xs, sts = [], []
    for rtol in (0.0, 1e-300):
        x = afe.DeviceVector(ctx, sp.n_dof)
        sts.append(g.solve(x, b, step=5, mode=1, rtol=rtol))
        xs.append(x.numpy())
    assert sts[0].iterations == sts[1].iterations == 6
    assert np.array_equal(xs[0], xs[1])
    assert sts[0].res_l1 == sts[1].res_l1
