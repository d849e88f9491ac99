"""GPU tests of the solve server (afe_server_*).

Bars: a request's solution is bitwise the same solved alone and solved among
other requests (continuous batching); a moving-mesh re-solve of a cached
mesh (values-only reassembly + lazyReinit) is bitwise the same as a fresh
set-up of the moved mesh; every request reaches rtol; the cache stays within
its HBM budget by evicting idle meshes.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

P1, P2, TET = 1, 2, 3


def _ndof(kind, n):
    return (n + 1) ** 3 if kind == TET else ((2 * n + 1) ** 2 if kind == P2 else (n + 1) ** 2)


def test_server_alone_mix_moving_mesh():
    import afepack_amd as afe
    print("case server", flush=True)
    srv = afe.Server(0, hbm_budget=1 << 30, max_slots=16, setup_threads=4, smooth=1)
    try:
        base = dict(kind=P2, n=24, jitter=0.2, seed=5, acc=6, rtol=1e-10, keep_x=True)
        r0 = srv.submit(1, **base)
        x0, st0, lat0, c0 = srv.wait(r0, _ndof(P2, 24))
        assert c0 == 0 and st0.rel_res_l2 <= 1e-10
        field = afe.field_random(8, 2, 0.5, 9)
        ids = []
        # a mix: other meshes (some variable-coefficient), the same request
        # again, and moving-mesh updates of mesh 1 and of mesh 2
        ids.append(("other", srv.submit(2, kind=P1, n=40, jitter=0.1, seed=2, keep_x=True), P1, 40))
        ids.append(("other", srv.submit(3, kind=TET, n=8, jitter=0.05, seed=3, keep_x=True), TET, 8))
        ids.append(("var", srv.submit(4, kind=P2, n=16, op=afe.OP_VARSTIFF, fkind=afe.F_VARCOEF,
                                      field=field, acc=6, keep_x=True), P2, 16))
        ids.append(("same", srv.submit(1, **base), P2, 24))
        ids.append(("move1", srv.submit(1, move_amp=0.1, move_k=2, **base), P2, 24))
        ids.append(("move2", srv.submit(2, kind=P1, n=40, jitter=0.1, seed=2, move_amp=0.1, move_k=3,
                                        keep_x=True), P1, 40))
        for k in range(6):
            ids.append(("load", srv.submit(10 + k, kind=P1 if k % 2 else P2, n=20 + 6 * k, seed=k,
                                           keep_x=True), P1 if k % 2 else P2, 20 + 6 * k))
        res = {}
        for tag, rid, kind, n in ids:
            x, st, lat, code = srv.wait(rid, _ndof(kind, n))
            assert code == 0 and st.converged == 1 and st.rel_res_l2 <= 1e-10, (tag, code, st.rel_res_l2)
            res.setdefault(tag, []).append(x)
        assert np.array_equal(res["same"][0], x0)
        # fresh set-ups of the moved meshes under new keys
        f1 = srv.submit(101, move_amp=0.1, move_k=2, **base)
        f2 = srv.submit(102, kind=P1, n=40, jitter=0.1, seed=2, move_amp=0.1, move_k=3, keep_x=True)
        xf1 = srv.wait(f1, _ndof(P2, 24))[0]
        xf2 = srv.wait(f2, _ndof(P1, 40))[0]
        assert np.array_equal(res["move1"][0], xf1)
        assert np.array_equal(res["move2"][0], xf2)
        assert not np.array_equal(xf1, x0)
        info = srv.info()
        assert info["evictions"] == 0 and info["completed"] == len(ids) + 3
    finally:
        srv.close()


def test_server_budget_evicts_idle_meshes():
    import afepack_amd as afe
    print("case server budget", flush=True)
    srv = afe.Server(0, hbm_budget=32 << 20, max_slots=8, setup_threads=3, smooth=1)
    try:
        ids = [srv.submit(k, kind=P2 if k % 2 else P1, n=20 + 4 * k, seed=k) for k in range(10)]
        for rid in ids:
            _, st, lat, code = srv.wait(rid)
            assert code == 0 and st.rel_res_l2 <= 1e-10
        info = srv.info()
        print(info, flush=True)
        assert info["evictions"] > 0 and info["cached"] < 10
    finally:
        srv.close()


def test_server_stream_of_small_requests():
    """Set-ups running on several host threads while the engine steps the
    batch (and captures its graphs): every request of a stream of new and
    moving-mesh requests succeeds."""
    import afepack_amd as afe
    print("case server stream", flush=True)
    srv = afe.Server(0, hbm_budget=2 << 30, max_slots=12, setup_threads=6, smooth=1)
    try:
        ids = []
        for k in range(80):
            key = k % 25  # keys repeat: moving-mesh re-solves of cached meshes
            kind = (P1, P2, TET)[key % 3]
            n = (4 + key % 7) if kind == TET else (8 + 5 * (key % 9))
            ids.append((srv.submit(1000 + key, kind=kind, n=n, seed=key, move_amp=0.05 * (k // 25),
                                   move_k=1 + k % 3), kind, n))
        for rid, kind, n in ids:
            _, st, lat, code = srv.wait(rid)
            assert code == 0 and st.converged == 1 and st.rel_res_l2 <= 1e-10, (rid, code)
        assert srv.info()["completed"] == len(ids)
    finally:
        srv.close()


def _request_mesh(kind, n, jitter, seed, move_amp=0.0, move_k=1):
    import afepack_amd as afe
    c, ev, vm = afe.mesh_cube(n, jitter, seed) if kind == TET else afe.mesh_square(n, jitter, seed)
    if move_amp:
        c = afe.mesh_displace(3 if kind == TET else 2, c, vm, move_amp / n, move_k)
    return c, ev, vm


def _oracle_direct(kind, mesh, acc, op, fkind, field):
    """the oracle's assembly, right-hand side and zero Dirichlet elimination
    of the request, solved directly"""
    import scipy.sparse as sps
    import scipy.sparse.linalg as spl
    from oracle import oracle_py as orc
    c, ev, vm = mesh
    osp = orc.Space(kind, c, ev, vm)
    of_ = orc.Field(field.k, field.phi, field.w, field.scale) if field is not None else None
    OA = osp.build(op, acc, of_)
    f = osp.rhs(acc, fkind, of_)
    osp.apply_dirichlet(OA, f)
    rp, col, val = OA.csr()
    A = sps.csr_matrix((val, col, rp), shape=(len(rp) - 1,) * 2)
    return spl.spsolve(A.tocsc(), f)


def _c5_specs():
    """requests of every kind the server takes: structured P1 / P2 / 3D,
    variable coefficients, a mesh of the caller's (made by the reference's
    mesh generator), a new mesh of a cached mesh's connectivity (copied
    structure), moving-mesh re-solves of a structured and of a caller mesh"""
    import os
    import afepack_amd as afe
    from tests.test_mesh_ingest import GOLD, read_easymesh
    f2 = afe.field_random(8, 2, 0.5, 21)
    f3 = afe.field_random(16, 3, 1.0, 22)
    cD, vD, eD, _, _ = read_easymesh(os.path.join(GOLD, "D"))
    cD, eD = cD.ravel(), eD.ravel()
    cDm = afe.mesh_displace(2, cD, vD, 0.1 * 0.05, 2)
    base = [
        ("p1", dict(mesh_key=1, kind=P1, n=24, jitter=0.1, seed=1, acc=4), None),
        ("p2var", dict(mesh_key=2, kind=P2, n=12, jitter=0.2, seed=2, acc=6, op=afe.OP_VARSTIFF,
                       fkind=afe.F_VARCOEF, field=f2), None),
        ("tetvar", dict(mesh_key=3, kind=TET, n=6, jitter=0.05, seed=3, acc=2, op=afe.OP_VARSTIFF,
                        fkind=afe.F_VARCOEF, field=f3), None),
        ("own", dict(mesh_key=4, kind=P2, acc=6, mesh=(cD, eD, vD)), None),
    ]
    later = [
        ("clone", dict(mesh_key=5, kind=P1, n=24, jitter=0.15, seed=5, acc=4), None),
        ("move", dict(mesh_key=1, kind=P1, n=24, jitter=0.1, seed=1, acc=4, move_amp=0.1, move_k=2), "p1"),
        ("own_move", dict(mesh_key=4, kind=P2, acc=6, mesh=(cDm, eD, vD)), "own"),
    ]
    return base, later


def _spec_mesh(spec):
    if "mesh" in spec:
        return spec["mesh"]
    return _request_mesh(spec["kind"], spec["n"], spec["jitter"], spec["seed"], spec.get("move_amp", 0.0),
                         spec.get("move_k", 1))


def _ndof_of(spec):
    import afepack_amd as afe
    c, ev, vm = _spec_mesh(spec)
    return afe.Space(None, spec["kind"], c, ev, vm).n_dof


def _serve(srv, spec):
    s = dict(spec)
    key = s.pop("mesh_key")
    return srv.submit(key, rtol=1e-10, max_it=1000, keep_x=True, **s)


def test_served_requests_match_oracle_alone_and_in_the_mix():
    """C5 parity: every served solution within 1e-9 (relative L2) of a
    direct solve of the oracle's system for the same (moved) mesh; each
    request's solution in the mix bitwise equal to the same request served
    alone (a structure copy therefore equals a fresh set-up)."""
    import afepack_amd as afe
    print("case server parity", flush=True)
    base, later = _c5_specs()
    ndof = {tag: _ndof_of(spec) for tag, spec, _ in base + later}
    mix = {}
    srv = afe.Server(0, hbm_budget=4 << 30, max_slots=16, setup_threads=4, smooth=1)
    try:
        # the first mesh alone first: its cached entry lends its structure
        rid = _serve(srv, base[0][1])
        mix["p1"] = srv.wait(rid, ndof["p1"])
        ids = [(tag, _serve(srv, spec)) for tag, spec, _ in base[1:]]
        for tag, rid in ids:
            mix[tag] = srv.wait(rid, ndof[tag])
        ids = [(tag, _serve(srv, spec)) for tag, spec, _ in later]
        for tag, rid in ids:
            mix[tag] = srv.wait(rid, ndof[tag])
        tm = srv.timing()
        print(tm, flush=True)
        assert tm["n_cloned"] >= 1
    finally:
        srv.close()
    for tag, spec, _ in base + later:
        x, st, lat, code = mix[tag]
        assert code == 0 and st.converged == 1 and st.rel_res_l2 <= 1e-10, (tag, code, st.rel_res_l2)
        ref = _oracle_direct(spec["kind"], _spec_mesh(spec), spec["acc"], spec.get("op", afe.OP_STIFF),
                             spec.get("fkind", afe.F_POISSON), spec.get("field"))
        assert np.linalg.norm(x - ref) <= 1e-9 * np.linalg.norm(ref), tag
    # alone: each request in a server of its own (a re-solve after its base)
    specs = {tag: spec for tag, spec, _ in base + later}
    for tag, spec, before in base + later:
        srv = afe.Server(0, hbm_budget=4 << 30, max_slots=4, setup_threads=1, smooth=1)
        try:
            if before:
                srv.wait(_serve(srv, specs[before]), ndof[before])
            x, st, _, code = srv.wait(_serve(srv, spec), ndof[tag])
            assert code == 0
            if tag == "clone":
                assert srv.timing()["n_cloned"] == 0  # alone: a fresh set-up
            assert np.array_equal(x, mix[tag][0]), tag
        finally:
            srv.close()


def test_server_refuses_a_request_that_cannot_fit():
    """a request whose measured size exceeds what the budget can ever free
    fails with AFE_E_BUDGET instead of waiting forever"""
    import afepack_amd as afe
    print("case server refuse", flush=True)
    # the server's estimate for a 40 x 40 P1 mesh (96 nnz + 480 n_dof + 96
    # n_elem bytes): the budget exceeds it by 1 KB, less than the context's
    # own device buffers (quadrature tables alone take 4 KB)
    n = 40
    est = 96 * (7 * n * n + 6 * n + 1) + 480 * (n + 1) ** 2 + 96 * 2 * n * n
    srv = afe.Server(0, hbm_budget=est + 1024, max_slots=4, setup_threads=2, smooth=1)
    try:
        ok = srv.submit(1, kind=P1, n=4, rtol=1e-10)
        assert srv.wait(ok)[3] == 0
        # the estimate fits the budget, but not with the context's buffers:
        # refused (not waiting for memory that nothing will give back)
        big = srv.submit(2, kind=P1, n=n, rtol=1e-10)
        with pytest.raises(afe.AfeError) as err:
            srv.wait(big)
        assert err.value.code == -7
        # the server keeps serving
        again = srv.submit(3, kind=P1, n=6, rtol=1e-10)
        assert srv.wait(again)[3] == 0
    finally:
        srv.close()
