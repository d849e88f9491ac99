"""GPU: the opt-in stream-ordered buffers (AFE_STREAM_ALLOC=1) give the same
bits as hipMalloc / hipFree buffers: assembly (triangles and pipelined
tetrahedra), right-hand side, Dirichlet elimination, lazy hierarchy,
mixed-precision PCG and a lazy refresh on moved coordinates, each form in a
child process (the switch is read once per process)."""
import hashlib
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _child():
    """prints one digest per case of the values and solutions"""
    import afepack_amd as afe
    from tests.conftest import make_mesh
    ctx = afe.Context(0)
    for kind, n, jit, acc in ((2, 48, 0.2, 6), (3, 36, 0.05, 2)):
        c, ev, vm = make_mesh(kind, n, jit, 5)
        sp = afe.Space(ctx, kind, c, ev, vm)
        A = afe.Matrix(sp)
        A.assemble(afe.OP_STIFF, acc)
        b = sp.rhs(acc, afe.F_POISSON)
        A.apply_dirichlet(b)
        g = afe.Amg(A, afe.AmgParams(lazy=1, smooth=1, build_fp32=1))
        x = afe.DeviceVector(ctx, sp.n_dof)
        st = g.pcg(x, b, rtol=1e-10, prec=afe.PREC_FP32)
        h = hashlib.sha256()
        h.update(A.values().tobytes())
        h.update(x.numpy().tobytes())
        # moved coordinates: values-only assembly and the lazy refresh
        sp.update_coords(np.ascontiguousarray(c * 1.01))
        A.assemble(afe.OP_STIFF, acc)
        b = sp.rhs(acc, afe.F_POISSON)
        A.apply_dirichlet(b)
        g.lazy_reinit(A)
        x.zero()
        g.pcg(x, b, rtol=1e-10, prec=afe.PREC_FP32)
        h.update(A.values().tobytes())
        h.update(x.numpy().tobytes())
        print(f"kind {kind} its {st.iterations} digest {h.hexdigest()}", flush=True)


def _run(flag):
    env = dict(os.environ, AFE_STREAM_ALLOC=flag)
    r = subprocess.run([sys.executable, "-c", "from tests.test_gpu_alloc import _child; _child()"],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=100)
    print(r.stdout[-1500:], r.stderr[-1500:], flush=True)
    assert r.returncode == 0
    return [ln for ln in r.stdout.splitlines() if ln.startswith("kind ")]


def test_stream_ordered_buffers_give_the_same_bits():
    print("case AFE_STREAM_ALLOC 0 vs 1", flush=True)
    a, b = _run("0"), _run("1")
    assert len(a) == 2 and a == b
