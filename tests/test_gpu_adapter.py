"""The drop-in accelerator driven by HiDEMLib's own call expressions
(tests/adapter_callsites.cpp: generateNeighbours / generateLinks /
lAccel.sortSideData / generateIndexes / iterateCombinedPre+Post /
iterateLinks with the reference's argument lists) on a GPU: one link
stretched by 1 um gives equal and opposite axial forces of E W^2 / L * 1e-6."""
import os
import subprocess

import numpy as np
import pytest

from test_abi_cpu import build_callsites

pytestmark = pytest.mark.gpu


def test_adapter_callsites_run(tmp_path):
    exe = tmp_path / "callsites"
    r = build_callsites(exe)
    assert r.returncode == 0, r.stderr
    print("adapter call sites: run", flush=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    f0, f1 = (float(v) for v in out.stdout.split()[1:3])
    assert f0 == -f1
    np.testing.assert_allclose(abs(f0), 1e9 * 1.0 / 1.0 * 1e-6, rtol=1e-9)
