"""Write the golden fixtures of tests/golden/ from the reference pieces that
build from their own sources (oracle/_ref: the template-element plugins and
the header-only SparsityPattern class; see oracle/Makefile).

Run in the development container (where the AFEPack tree is present):
    make -C oracle ref && python tests/golden/make_golden.py

Fixtures (data only):
  plugins.npz  -- basis values/gradients and coordinate maps evaluated by the
                  reference's C plugins at seeded points of seeded simplices
  patterns.npz -- CSR patterns built by the reference's SparsityPattern class
                  (reinit with the max coupling count, add per element pair,
                  compress) for the 8x8 P1, 4x4 P2 and 3^3 tetrahedron meshes,
                  and for two spaces on the 4x4 mesh (P1 x P2, P2 x P1)
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
REF = os.path.join(ROOT, "oracle", "_ref")

PD = C.POINTER(C.c_double)


def _plugin(name):
    return C.CDLL(os.path.join(REF, name))


def _vptr(verts):
    arr = (PD * len(verts))()
    for i, v in enumerate(verts):
        arr[i] = v.ctypes.data_as(PD)
    return arr


def plugin_fixture(rng):
    out = {}
    tri1, tri2, geo = _plugin("libtri1.so"), _plugin("libtri2.so"), _plugin("libtri_geo.so")
    tet1, tgeo = _plugin("libtet1.so"), _plugin("libtet_geo.so")
    tri_vol, tet_vol = _plugin("libtri_vol.so"), _plugin("libtet_vol.so")
    for f in (tri_vol.triangle_volume, tet_vol.tetrahedron_volume):
        f.restype = C.c_double
    geo.local_to_global_jacobian.restype = C.c_double
    tgeo.local_to_global_jacobian.restype = C.c_double
    ref_tri = [np.array(p, dtype=np.float64) for p in ((0., 0.), (1., 0.), (0., 1.))]
    ref_tet = [np.array(p, dtype=np.float64) for p in ((0., 0., 0.), (1., 0., 0.), (0., 1., 0.), (0., 0., 1.))]
    out["tri_ref_volume"] = np.array([tri_vol.triangle_volume(_vptr(ref_tri))])
    out["tet_ref_volume"] = np.array([tet_vol.tetrahedron_volume(_vptr(ref_tet))])
    # triangles: positively oriented, random
    T, Q = 16, 5
    tris = np.empty((T, 3, 2))
    for t in range(T):
        while True:
            v = rng.uniform(-1, 1, (3, 2))
            a = (v[1, 0] - v[0, 0]) * (v[2, 1] - v[0, 1]) - (v[1, 1] - v[0, 1]) * (v[2, 0] - v[0, 0])
            if a > 0.3:
                break
        tris[t] = v
    lp = rng.uniform(0, 0.5, (T, Q, 2))
    gp = np.empty((T, Q, 2))
    jac = np.empty(T)
    p1v, p1g = np.empty((T, Q, 3)), np.empty((T, Q, 3, 2))
    p2v, p2g = np.empty((T, Q, 6)), np.empty((T, Q, 6, 2))
    p1names = [("lambda_1", "gradient_lambda_1"), ("lambda_2", "gradient_lambda_2"),
               ("lambda_3", "gradient_lambda_3")]
    p2names = [("phi_1", "gradient_phi_1"), ("phi_2", "gradient_phi_2"), ("phi_3", "gradient_phi_3"),
               ("psi_1", "gradient_psi_1"), ("psi_2", "gradient_psi_2"), ("psi_3", "gradient_psi_3")]
    for t in range(T):
        gv = [np.ascontiguousarray(tris[t, i]) for i in range(3)]
        jac[t] = geo.local_to_global_jacobian(lp[t, 0].ctypes.data_as(PD), _vptr(ref_tri), _vptr(gv))
        for q in range(Q):
            l = np.ascontiguousarray(lp[t, q])
            g = np.zeros(2)
            geo.local_to_global(l.ctypes.data_as(PD), _vptr(ref_tri), _vptr(gv), g.ctypes.data_as(PD))
            gp[t, q] = g
            for k, (fv, fg) in enumerate(p1names):
                val, grad = np.zeros(1), np.zeros(2)
                getattr(tri1, fv)(g.ctypes.data_as(PD), _vptr(gv), val.ctypes.data_as(C.c_void_p))
                getattr(tri1, fg)(g.ctypes.data_as(PD), _vptr(gv), grad.ctypes.data_as(C.c_void_p))
                p1v[t, q, k], p1g[t, q, k] = val[0], grad
            for k, (fv, fg) in enumerate(p2names):
                val, grad = np.zeros(1), np.zeros(2)
                getattr(tri2, fv)(g.ctypes.data_as(PD), _vptr(gv), val.ctypes.data_as(C.c_void_p))
                getattr(tri2, fg)(g.ctypes.data_as(PD), _vptr(gv), grad.ctypes.data_as(C.c_void_p))
                p2v[t, q, k], p2g[t, q, k] = val[0], grad
    out.update(tri_vertices=tris, tri_local=lp, tri_global=gp, tri_jacobian=jac,
               p1_values=p1v, p1_gradients=p1g, p2_values=p2v, p2_gradients=p2g)
    # tetrahedra
    tets = np.empty((T, 4, 3))
    for t in range(T):
        while True:
            v = rng.uniform(-1, 1, (4, 3))
            d = np.linalg.det(np.array([v[1] - v[0], v[2] - v[0], v[3] - v[0]]))
            if d > 0.3:
                break
        tets[t] = v
    tlp = rng.uniform(0, 0.3, (T, Q, 3))
    tgp = np.empty((T, Q, 3))
    tjac = np.empty(T)
    tv, tg = np.empty((T, Q, 4)), np.empty((T, Q, 4, 3))
    names = [("lambda_%d" % (k + 1), "gradient_lambda_%d" % (k + 1)) for k in range(4)]
    for t in range(T):
        gv = [np.ascontiguousarray(tets[t, i]) for i in range(4)]
        tjac[t] = tgeo.local_to_global_jacobian(tlp[t, 0].ctypes.data_as(PD), _vptr(ref_tet), _vptr(gv))
        for q in range(Q):
            l = np.ascontiguousarray(tlp[t, q])
            g = np.zeros(3)
            tgeo.local_to_global(l.ctypes.data_as(PD), _vptr(ref_tet), _vptr(gv), g.ctypes.data_as(PD))
            tgp[t, q] = g
            for k, (fv, fg) in enumerate(names):
                val, grad = np.zeros(1), np.zeros(3)
                getattr(tet1, fv)(g.ctypes.data_as(PD), _vptr(gv), val.ctypes.data_as(C.c_void_p))
                getattr(tet1, fg)(g.ctypes.data_as(PD), _vptr(gv), grad.ctypes.data_as(C.c_void_p))
                tv[t, q, k], tg[t, q, k] = val[0], grad
    out.update(tet_vertices=tets, tet_local=tlp, tet_global=tgp, tet_jacobian=tjac,
               tet_values=tv, tet_gradients=tg)
    return out


def pattern_fixture():
    import afepack_amd as afe
    L = C.CDLL(os.path.join(REF, "libref_sparsity.so"))
    L.refsp_new_max.restype = C.c_void_p
    L.refsp_new_max.argtypes = [C.c_int64, C.c_int64, C.c_int64]
    L.refsp_add.argtypes = [C.c_void_p, C.c_int64, C.c_int64]
    L.refsp_compress.argtypes = [C.c_void_p]
    L.refsp_nnz.restype = C.c_int64
    L.refsp_nnz.argtypes = [C.c_void_p]
    L.refsp_get.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    L.refsp_free.argtypes = [C.c_void_p]
    out = {}
    for tag, kind, n in (("p1_8", 1, 8), ("p2_4", 2, 4), ("tet_3", 3, 3)):
        mesh = afe.mesh_cube(n) if kind == 3 else afe.mesh_square(n, 0.0, 0)
        from oracle import oracle_py as orc
        osp = orc.Space(kind, *mesh)
        ed, _, _ = osp.dofs()
        nl = osp.n_loc
        ed = ed.reshape(-1, nl)
        N = osp.n_dof
        cnt = np.zeros(N, dtype=np.int64)
        for e in ed:
            cnt[e] += nl
        maxc = min(int(cnt.max()), N)
        sp = L.refsp_new_max(N, N, maxc)
        for e in ed:
            for i in e:
                for j in e:
                    L.refsp_add(sp, int(i), int(j))
        L.refsp_compress(sp)
        nnz = L.refsp_nnz(sp)
        rp = np.empty(N + 1, dtype=np.int64)
        col = np.empty(nnz, dtype=np.int64)
        L.refsp_get(sp, rp.ctypes.data_as(C.c_void_p), col.ctypes.data_as(C.c_void_p))
        L.refsp_free(sp)
        out[tag + "_elem_dof"] = ed.astype(np.int32)
        out[tag + "_row_ptr"] = rp.astype(np.int32)
        out[tag + "_col"] = col.astype(np.int32)
    # two spaces on one mesh (BilinearOperator's same-mesh branch): the
    # reference class reinit to (N0, N1, max coupling), element e adds
    # (dof0_i, dof1_j) for all i, j, then compress
    for tag, k0, k1, n in (("p1p2_4", 1, 2, 4), ("p2p1_4", 2, 1, 4)):
        mesh = afe.mesh_square(n, 0.0, 0)
        from oracle import oracle_py as orc
        e0 = orc.Space(k0, *mesh).dofs()[0].reshape(2 * n * n, -1)
        s1 = orc.Space(k1, *mesh)
        e1 = s1.dofs()[0].reshape(2 * n * n, -1)
        N0, N1 = int(e0.max()) + 1, s1.n_dof
        cnt = np.zeros(N0, dtype=np.int64)
        for a in e0:
            cnt[a] += e1.shape[1]
        maxc = min(int(cnt.max()), N1)
        sp = L.refsp_new_max(N0, N1, maxc)
        for a, b in zip(e0, e1):
            for i in a:
                for j in b:
                    L.refsp_add(sp, int(i), int(j))
        L.refsp_compress(sp)
        nnz = L.refsp_nnz(sp)
        rp = np.empty(N0 + 1, dtype=np.int64)
        col = np.empty(nnz, dtype=np.int64)
        L.refsp_get(sp, rp.ctypes.data_as(C.c_void_p), col.ctypes.data_as(C.c_void_p))
        L.refsp_free(sp)
        out[tag + "_max_coupling"] = np.array([maxc], dtype=np.int32)
        out[tag + "_row_ptr"] = rp.astype(np.int32)
        out[tag + "_col"] = col.astype(np.int32)
    return out


def main():
    rng = np.random.default_rng(20261202)
    np.savez_compressed(os.path.join(HERE, "plugins.npz"), **plugin_fixture(rng))
    np.savez_compressed(os.path.join(HERE, "patterns.npz"), **pattern_fixture())
    print("wrote", os.listdir(HERE))


if __name__ == "__main__":
    main()
