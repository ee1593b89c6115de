"""CPU checks of the product's contracts: without the built library every
entry point raises (there is no CPU fallback), and bench.py refuses more
than one GPU with a message instead of running."""
import json
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent


def test_missing_library_raises(pkg, monkeypatch, tmp_path):
    from importlib import import_module
    capi = import_module(pkg.__name__ + "._capi")
    monkeypatch.setattr(capi, "LIB_PATH", tmp_path / "libmeshing_hip.so")
    monkeypatch.setattr(capi, "_lib", None)
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        capi.lib()
    with pytest.raises(RuntimeError):
        pkg.isosurface(np.zeros((4, 4, 4), np.float32), pkg.MarchingCubes())


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]
