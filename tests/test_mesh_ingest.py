"""Mesh ingestion (CPU): unstructured meshes made by the reference's own mesh
generator (tests/golden/easymesh/, written by make_easymesh_fixture.py) and
the reference's Mesh text format.

  EasyMesh files     read as EasyMesh.templates.h:16-94 reads them (clockwise
                     triangles turned, their side list swapped alike)
  DOF numbering      host-only afe_space_create_sides against a brute-force
                     first-touch numbering and the oracle (bit-exact)
  DOF boundary marks vertex DOFs take the node mark, P2 edge DOFs the mark of
                     their side in the .s file (FEMSpace.templates.h:689-701)
  Mesh text format   the C++ drop-in's Mesh::readData (Geometry.templates.h:
                     470-497) on a file written here in that format, with
                     geometries out of order: same vertices, elements and sides
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from tests.conftest import ROOT

GOLD = os.path.join(ROOT, "tests", "golden", "easymesh")


def read_easymesh(base):
    """nodes, marks, elements (counter-clockwise), sides (vertex pairs, marks)"""
    with open(base + ".n") as fh:
        nn = int(fh.readline().split()[0])
        rows = [fh.readline().split() for _ in range(nn)]
    coords = np.array([[float(r[1]), float(r[2])] for r in rows])
    vmark = np.array([int(r[3]) for r in rows], dtype=np.int32)
    with open(base + ".s") as fh:
        ns = int(fh.readline().split()[0])
        srows = [fh.readline().split() for _ in range(ns)]
    side_vert = np.zeros((ns, 2), dtype=np.int32)
    side_mark = np.zeros(ns, dtype=np.int32)
    for r in srows:
        i = int(r[0])
        side_vert[i] = (int(r[1]), int(r[2]))
        side_mark[i] = int(r[5])
    with open(base + ".e") as fh:
        ne = int(fh.readline().split()[0])
        erows = [fh.readline().split() for _ in range(ne)]
    elem = np.array([[int(r[1]), int(r[2]), int(r[3])] for r in erows], dtype=np.int32)
    p0, p1, p2 = coords[elem[:, 0]], coords[elem[:, 1]], coords[elem[:, 2]]
    cw = (p1[:, 0] - p0[:, 0]) * (p2[:, 1] - p0[:, 1]) - (p1[:, 1] - p0[:, 1]) * (p2[:, 0] - p0[:, 0]) < 0
    elem[cw] = elem[cw][:, [0, 2, 1]]
    return coords, vmark, elem, side_vert, side_mark


def brute_first_touch(kind, elem):
    ids, out = {}, []
    for e in elem:
        keys = [("v", int(p)) for p in e]
        if kind == 2:
            keys += [("e", tuple(sorted((int(e[(k + 1) % 3]), int(e[(k + 2) % 3]))))) for k in range(3)]
        row = []
        for key in keys:
            if key not in ids:
                ids[key] = len(ids)
            row.append(ids[key])
        out.append(row)
    return np.array(out, dtype=np.int32), len(ids)


@pytest.mark.parametrize("name", ["D", "Lshape"])
@pytest.mark.parametrize("kind", [1, 2])
def test_easymesh_fixture_numbering_and_marks(name, kind):
    import afepack_amd as afe
    from oracle import oracle_py as orc
    coords, vmark, elem, sv, sm = read_easymesh(os.path.join(GOLD, name))
    assert len(np.unique(vmark)) > 2 and sm.max() >= 4  # several boundary marks
    sp = afe.Space(None, kind, coords.ravel(), elem.ravel(), vmark, sides=(sv, sm))
    ed, dm, dp = sp.dofs()
    bed, nd = brute_first_touch(kind, elem)
    assert sp.n_dof == nd and np.array_equal(ed.reshape(len(elem), -1), bed)
    osp = orc.Space(kind, coords.ravel(), elem.ravel(), vmark, sides=(sv, sm))
    oed, odm, odp = osp.dofs()
    assert np.array_equal(oed, ed) and np.array_equal(odm, dm) and np.array_equal(odp, dp)
    rp, col = sp.pattern()
    orp, ocol, _ = osp.build(orc.OP_STIFF, 4).csr()
    assert np.array_equal(rp, orp) and np.array_equal(col, ocol)
    # marks: node marks on vertex DOFs, the .s file's side marks on edge DOFs
    E = ed.reshape(len(elem), -1)
    for j in range(3):
        assert np.array_equal(dm[E[:, j]], vmark[elem[:, j]])
    if kind == 2:
        smap = {tuple(sorted(map(int, s))): int(m) for s, m in zip(sv, sm)}
        want = np.array([[smap[tuple(sorted((int(e[(j + 1) % 3]), int(e[(j + 2) % 3]))))]
                          for j in range(3)] for e in elem])
        assert np.array_equal(dm[E[:, 3:]], want)
        # the rule without a side list differs on these meshes: boundary edge
        # DOFs then get mark 1 whatever their segment
        plain = afe.Space(None, kind, coords.ravel(), elem.ravel(), vmark)
        assert not np.array_equal(plain.dofs()[1], dm)


def test_side_list_missing_an_edge_is_an_error():
    import afepack_amd as afe
    coords, vmark, elem, sv, sm = read_easymesh(os.path.join(GOLD, "D"))
    with pytest.raises(afe.AfeError):
        afe.Space(None, 2, coords.ravel(), elem.ravel(), vmark, sides=(sv[1:], sm[1:]))


def write_mesh_format(path, coords, vmark, elem, sv, sm, rng):
    """the Mesh text format (Mesh::writeData layout) with the vertex
    geometries pointing at permuted points and every geometry list written
    in shuffled order"""
    nv = len(coords)
    perm = rng.permutation(nv)  # vertex geometry i sits at point perm[i]
    pts = np.empty_like(coords)
    pts[perm] = coords
    lines = [str(nv)] + [f"{float(p[0])!r} {float(p[1])!r}" for p in pts]

    def block(recs):
        order = rng.permutation(len(recs))
        out = [str(len(recs))]
        for k in order:
            idx, verts, bnd, bm = recs[k]
            out.append(f"{idx} {len(verts)} {' '.join(map(str, verts))} {len(bnd)} {' '.join(map(str, bnd))} {bm}")
        return out

    lines += block([(i, [int(perm[i])], [i], int(vmark[i])) for i in range(nv)])
    lines += block([(i, [int(a), int(b)], [int(a), int(b)], int(m)) for i, ((a, b), m) in enumerate(zip(sv, sm))])
    lines += block([(i, [int(x) for x in e], [0, 0, 0], 0) for i, e in enumerate(elem)])
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


DUMP_SRC = r"""
#include <cstdio>
#include <string>
#include <AFEPack/FEMSpace.h>
// reads a mesh (EasyMesh base, or a Mesh text file ending in .mesh) and
// prints its vertices, elements and sides as plain text
int main(int argc, char **argv) {
  Mesh<2, 2> m;
  EasyMesh em;
  const std::string f = argv[1];
  Mesh<2, 2> &r = f.size() > 5 && f.substr(f.size() - 5) == ".mesh" ? m : em;
  if (&r == &m) m.readData(f); else em.readData(f);
  std::printf("%d %d %d\n", r.n_point(), r.n_element(), (int)r.side_mark.size());
  for (int i = 0; i < r.n_point(); i++) std::printf("%.17g %.17g %d\n", r.coords[2 * i], r.coords[2 * i + 1], r.vmark[i]);
  for (int i = 0; i < r.n_element(); i++) std::printf("%d %d %d\n", r.elem[3 * i], r.elem[3 * i + 1], r.elem[3 * i + 2]);
  for (size_t i = 0; i < r.side_mark.size(); i++) std::printf("%d %d %d\n", r.side_vert[2 * i], r.side_vert[2 * i + 1], r.side_mark[i]);
  return 0;
}
"""


def _dump(exe, path):
    r = subprocess.run([exe, path], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    tok = r.stdout.split("\n")
    nv, ne, ns = map(int, tok[0].split())
    v = np.array([list(map(float, t.split())) for t in tok[1:1 + nv]])
    e = np.array([list(map(int, t.split())) for t in tok[1 + nv:1 + nv + ne]], dtype=np.int32)
    s = np.array([list(map(int, t.split())) for t in tok[1 + nv + ne:1 + nv + ne + ns]], dtype=np.int32)
    return v[:, :2], v[:, 2].astype(np.int32), e, s[:, :2], s[:, 2]


def test_cpp_mesh_readers(tmp_path):
    """the drop-in's EasyMesh and Mesh readers give the vertices, elements,
    marks and sides the reference's readers would"""
    if shutil.which("g++") is None:
        pytest.skip("no host C++ compiler")
    src = tmp_path / "dump.cpp"
    src.write_text(DUMP_SRC)
    exe = str(tmp_path / "dump")
    lib = os.path.join(ROOT, "afepack_amd")
    r = subprocess.run(["g++", "-std=c++17", "-O1", f"-I{ROOT}/afepack_amd/include", f"-I{ROOT}/include", str(src),
                        "-o", exe, f"-L{lib}", "-lafepack_amd", f"-Wl,-rpath,{lib}"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    coords, vmark, elem, sv, sm = read_easymesh(os.path.join(GOLD, "Lshape"))
    c1, m1, e1, s1, k1 = _dump(exe, os.path.join(GOLD, "Lshape"))
    assert np.array_equal(c1, coords) and np.array_equal(m1, vmark) and np.array_equal(e1, elem)
    assert np.array_equal(s1, sv) and np.array_equal(k1, sm)
    mpath = str(tmp_path / "L.mesh")
    write_mesh_format(mpath, coords, vmark, elem, sv, sm, np.random.default_rng(3))
    c2, m2, e2, s2, k2 = _dump(exe, mpath)
    assert np.array_equal(c2, coords) and np.array_equal(m2, vmark) and np.array_equal(e2, elem)
    assert np.array_equal(s2, sv) and np.array_equal(k2, sm)
