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
