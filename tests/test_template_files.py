"""CPU: the drop-in's template-element readers (TemplateGeometry / TemplateDOF
/ BasisFunctionAdmin::readData, TemplateElement.templates.h) find their data
files as the reference does (AFEPACK_TEMPLATE_PATH, then ".",
src/Miscellaneous.cpp:145-185) and select the element from the files'
content, whatever the files are called:

  synthetic files in the reference's formats, renamed  -> P1 / P2 / P1 tet
  the reference's own template files under other names -> the same (read in
                                                          place, when present)
  content no supported element has, or DOF and basis files that disagree
                                                        -> a loud error
  files that are not found                              -> the file name
                                                          selects the element
"""
import os
import shutil
import subprocess

import pytest

from tests.conftest import ROOT

SRC = r"""
#include <cstdio>
#include <AFEPack/FEMSpace.h>
template <int D> int run(char **a) {
  TemplateGeometry<D> g; CoordTransform<D, D> c; TemplateDOF<D> d(g); BasisFunctionAdmin<double, D, D> b(d);
  g.readData(a[2]); c.readData(a[3]); d.readData(a[4]); b.readData(a[5]);
  TemplateElement<double, D, D> t;
  t.reinit(g, d, c, b);
  std::printf("kind %d\n", t.kind);
  return 0;
}
int main(int argc, char **argv) {
  try {
    return argv[1][0] == '3' ? run<3>(argv) : run<2>(argv);
  } catch (const std::exception &e) {
    std::printf("error: %s\n", e.what());
    return 3;
  }
}
"""


def _geo(tdim):
    verts = [[0.0] * tdim] + [[1.0 if j == i else 0.0 for j in range(tdim)] for i in range(tdim)]
    lines = ["# volume function", "./libmine.so", "  my_volume", "", str(tdim + 1)]
    lines += ["  " + " ".join(map(str, v)) for v in verts]
    return "\n".join(lines) + "\n"


def _dof(entries):
    return "\n".join([str(len(entries))] + [f"{d} {i} {c}" for d, i, c in entries]) + "\n"


def _bas(tdim, bases):
    out = [str(len(bases)), ""]
    for (d, i), order in bases:
        out += [f"{d} {i}", " ".join(["0.0"] * tdim), " ".join([str(order)] + ["0"] * tdim + ["0"]),
                "./libmine.bas.so", "  f g", ""]
    return "\n".join(out) + "\n"


def _p1(tdim):
    return _dof([(0, i, 1) for i in range(tdim + 1)]), _bas(tdim, [((0, i), 1) for i in range(tdim + 1)])


def _p2():
    ent = [(0, i, 1) for i in range(3)] + [(1, i, 1) for i in range(3)]
    return _dof(ent), _bas(2, [((d, i), 2) for d, i, _ in ent])


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("no host C++ compiler")
    d = tmp_path_factory.mktemp("tmplexe")
    (d / "t.cpp").write_text(SRC)
    lib = os.path.join(ROOT, "afepack_amd")
    r = subprocess.run(["g++", "-std=c++17", "-O1", f"-I{lib}/include", f"-I{ROOT}/include", str(d / "t.cpp"),
                        "-o", str(d / "t"), f"-L{lib}", "-lafepack_amd", f"-Wl,-rpath,{lib}"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    return str(d / "t")


def _run(exe, tdim, names, path, cwd):
    env = dict(os.environ)
    env.pop("AFEPACK_TEMPLATE_PATH", None)
    if path:
        env["AFEPACK_TEMPLATE_PATH"] = path
    r = subprocess.run([exe, str(tdim), *names], capture_output=True, text=True, timeout=60, env=env, cwd=cwd)
    return r.stdout.strip()


def test_renamed_synthetic_files_select_by_content(exe, tmp_path):
    lib = tmp_path / "lib"
    lib.mkdir()
    files = {"a.geo": _geo(2), "a.crd": "./libmine.so\n l2g g2l j1 j2\n"}
    files["p1.dof"], files["p1.bas"] = _p1(2)
    files["p2.dof"], files["p2.bas"] = _p2()
    files["t.geo"] = _geo(3)
    files["t.dof"], files["t.bas"] = _p1(3)
    for k, v in files.items():
        (lib / k).write_text(v)
    other = tmp_path / "elsewhere"
    other.mkdir()
    assert _run(exe, 2, ["a.geo", "a.crd", "p1.dof", "p1.bas"], f"/nonexistent:{lib}", other) == "kind 1"
    assert _run(exe, 2, ["a.geo", "a.crd", "p2.dof", "p2.bas"], str(lib), other) == "kind 2"
    assert _run(exe, 3, ["t.geo", "a.crd", "t.dof", "t.bas"], str(lib), other) == "kind 3"
    # found in the working directory when the path is not set
    assert _run(exe, 2, ["a.geo", "a.crd", "p2.dof", "p2.bas"], None, lib) == "kind 2"


def test_unsupported_or_inconsistent_content_is_an_error(exe, tmp_path):
    (tmp_path / "q.geo").write_text(_geo(2).replace("\n3\n", "\n4\n") + "  1.0 1.0\n")  # a quadrilateral
    (tmp_path / "a.geo").write_text(_geo(2))
    d1, b1 = _p1(2)
    d2, b2 = _p2()
    (tmp_path / "p1.dof").write_text(d1)
    (tmp_path / "p1.bas").write_text(b1)
    (tmp_path / "p2.dof").write_text(d2)
    ent = [(0, i, 1) for i in range(3)] + [(1, i, 1) for i in range(3)] + [(2, 0, 1)]
    (tmp_path / "p3.dof").write_text(_dof(ent))
    (tmp_path / "p3.bas").write_text(_bas(2, [((d, i), 3) for d, i, _ in ent]))
    out = _run(exe, 2, ["q.geo", "x", "p1.dof", "p1.bas"], str(tmp_path), tmp_path)
    assert out.startswith("error:") and "simplex" in out
    out = _run(exe, 2, ["a.geo", "x", "p2.dof", "p1.bas"], str(tmp_path), tmp_path)
    assert out.startswith("error:") and "disagree" in out
    out = _run(exe, 2, ["a.geo", "x", "p3.dof", "p3.bas"], str(tmp_path), tmp_path)
    assert out.startswith("error:") and "not supported" in out


def test_files_not_found_fall_back_to_the_name(exe, tmp_path):
    assert _run(exe, 2, ["triangle.tmp_geo", "triangle.crd_trs", "triangle.2.tmp_dof", "triangle.2.bas_fun"],
                None, tmp_path) == "kind 2"
    assert _run(exe, 3, ["tetrahedron.tmp_geo", "tetrahedron.crd_trs", "tetrahedron.1.tmp_dof",
                         "tetrahedron.1.bas_fun"], None, tmp_path) == "kind 3"
    assert _run(exe, 2, ["x.geo", "x.crd", "x.dof", "x.bas"], None, tmp_path).startswith("error:")


REF_TPL = "/root/reference/include/AFEPack/template"


@pytest.mark.skipif(not os.path.isdir(REF_TPL), reason="the AFEPack tree is not present")
@pytest.mark.parametrize("tdim,base,order,kind", [(2, "triangle", "1", 1), (2, "triangle", "2", 2),
                                                  (3, "tetrahedron", "1", 3)])
def test_reference_template_files_under_other_names(exe, tmp_path, tdim, base, order, kind):
    src = os.path.join(REF_TPL, base)
    shutil.copy(os.path.join(src, f"{base}.tmp_geo"), tmp_path / "g")
    shutil.copy(os.path.join(src, f"{base}.crd_trs"), tmp_path / "c")
    shutil.copy(os.path.join(src, f"{base}.{order}", f"{base}.{order}.tmp_dof"), tmp_path / "d")
    shutil.copy(os.path.join(src, f"{base}.{order}", f"{base}.{order}.bas_fun"), tmp_path / "b")
    assert _run(exe, tdim, ["g", "c", "d", "b"], str(tmp_path), tmp_path) == f"kind {kind}"
