"""Write the unstructured-mesh fixtures of tests/golden/easymesh/ with the
reference's own mesh generator, built from its source by oracle/Makefile
(oracle/_ref/easymesh, from external/easymesh/easymesh.c of the AFEPack tree).

Run in the development container (where the AFEPack tree is present):
    make -C oracle ref && python tests/golden/make_easymesh_fixture.py

Fixtures (data only): <name>.n / .s / .e, the generator's output for
  D        the unit-square domain input of the reference's Poisson example
           (examples/poisson_equation/D.d, read in place): 4 boundary
           segments marked 1..4, spacing 0.05
  Lshape   an L-shaped domain written here: 6 boundary segments marked 1..6,
           spacing 0.04 at the convex corners and 0.01 at the re-entrant one
"""
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
EXE = os.path.join(ROOT, "oracle", "_ref", "easymesh")
REF_D = "/root/reference/examples/poisson_equation/D.d"
OUT = os.path.join(HERE, "easymesh")

LSHAPE = """# L-shaped domain (0,1)^2 minus [0.5,1)x[0.5,1) #
6 # points #
0:  0.0  0.0  0.04  1
1:  1.0  0.0  0.04  2
2:  1.0  0.5  0.04  3
3:  0.5  0.5  0.01  4
4:  0.5  1.0  0.04  5
5:  0.0  1.0  0.04  6
6 # segments #
0:  0  1  1
1:  1  2  2
2:  2  3  3
3:  3  4  4
4:  4  5  5
5:  5  0  6
"""


def run(name, text):
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, name + ".d"), "w") as fh:
            fh.write(text)
        # the generator's main returns 1 after a successful run: success is
        # judged by its three output files
        subprocess.run([EXE, name], cwd=d, capture_output=True, timeout=300)
        for ext in (".n", ".s", ".e"):
            if not os.path.exists(os.path.join(d, name + ext)):
                sys.exit(f"easymesh wrote no {name}{ext}")
        for ext in (".n", ".s", ".e"):
            shutil.copy(os.path.join(d, name + ext), os.path.join(OUT, name + ext))


def main():
    if not os.path.exists(EXE):
        sys.exit("build oracle/_ref/easymesh first: make -C oracle ref")
    os.makedirs(OUT, exist_ok=True)
    with open(REF_D) as fh:
        run("D", fh.read())
    run("Lshape", LSHAPE)
    for f in sorted(os.listdir(OUT)):
        print(f, os.path.getsize(os.path.join(OUT, f)))


if __name__ == "__main__":
    main()
