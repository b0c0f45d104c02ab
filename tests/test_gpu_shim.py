"""The drop-in C++ headers (include/gsap/*.h) used the way the reference's
DemoParticles uses Sap / RadixSort / ParticleRigidBody, from a compiled C++
program: results equal the oracle's reference-format output."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "shim", "_build", "shim_demo")


def test_reference_style_caller(tmp_path, dev):
    assert os.path.exists(EXE), "tests/shim/_build/shim_demo not built (run build())"
    rng = np.random.default_rng(12)
    n = 5000
    s = np.concatenate([rng.uniform(0, 25, (n, 3)), rng.uniform(0.2, 0.6, (n, 1))], 1).astype(np.float32)
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        f.write(np.int32(n).tobytes())
        f.write(s.tobytes())
    r = subprocess.run([EXE, str(inp), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = np.fromfile(out, np.uint8)
    np_, ne = np.frombuffer(raw[:8], np.int32)
    off = 8
    pairs = np.frombuffer(raw[off:off + 8 * np_], np.int32).reshape(-1, 2)
    off += 8 * np_
    ent = np.frombuffer(raw[off:off + 4 * ne], np.int32)
    off += 4 * ne
    keys = np.frombuffer(raw[off:off + 4 * n], np.float32)
    off += 4 * n
    vals = np.frombuffer(raw[off:off + 4 * n], np.uint32)
    sym, ent_ref = O.sym_csr(O.brute(s, kind=O.KIND_SPHERE))
    assert np.array_equal(pairs, sym) and np.array_equal(ent, ent_ref)
    o = np.argsort(s[:, 0], kind="stable")
    assert np.array_equal(keys, s[o, 0]) and np.array_equal(vals, o)
