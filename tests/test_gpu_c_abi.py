"""The library through its C ABI alone: meshing.jl_amd/c_abi_check (built
from tests/c_abi_check.c by the package Makefile, no Python or torch in the
process) meshes a host Float32 field with each method and as a batch, and
its meshes must equal the CPU oracle's bit for bit."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from helpers import assert_same_mesh, noise

ROOT = Path(__file__).resolve().parent.parent
EXE = ROOT / "meshing.jl_amd" / "c_abi_check"


def read_meshes(path):
    data = path.read_bytes()
    out, off = [], 0
    while off < len(data):
        nv, nf = np.frombuffer(data, np.int64, 2, off)
        kind = int(np.frombuffer(data, np.int32, 1, off + 16)[0])
        off += 20
        dt = np.float32 if kind == 1 else np.float64
        V = np.frombuffer(data, dt, 3 * nv, off).reshape(nv, 3)
        off += V.nbytes
        F = np.frombuffer(data, np.int64, 3 * nf, off).reshape(nf, 3)
        off += F.nbytes
        out.append((V, F))
    return out


def run_check(tmp_path, f):
    f = np.asfortranarray(f.astype(np.float32))
    (tmp_path / "field.f32").write_bytes(f.tobytes(order="F"))
    return subprocess.run([str(EXE), *map(str, f.shape), str(tmp_path / "field.f32"),
                           str(tmp_path / "out.bin")], capture_output=True, text=True, timeout=120)


@pytest.mark.gpu
def test_c_caller_matches_oracle(tmp_path, oracle):
    f = noise((37, 41, 29), 9, np.float32)
    r = run_check(tmp_path, f)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "bad dtype: status 5" in r.stdout
    meshes = read_meshes(tmp_path / "out.bin")
    assert len(meshes) == 6
    for k, (name, kw) in enumerate((("mc", dict(iso=0)), ("mt", dict(iso=0, eps=1e-3)),
                                     ("sn", dict(iso=0)))):
        Vo, Fo = oracle.isosurface(f, name, **kw)
        assert len(Fo) > 0
        assert_same_mesh(*meshes[k], Vo, Fo)       # single call
        assert_same_mesh(*meshes[3 + k], Vo, Fo)   # the batch


def test_c_caller_links_and_reports_missing_device(tmp_path):
    # on a machine without a GPU the program still loads the library and
    # gets a status (not a crash) from msh_create
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present: the GPU test covers the program")
    if not EXE.exists():
        pytest.skip("c_abi_check is not built")
    r = run_check(tmp_path, noise((5, 5, 5), 1, np.float32))
    assert r.returncode == 1 and "msh_create: status" in r.stderr, r.stdout + r.stderr
