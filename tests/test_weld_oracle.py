"""CPU checks of the welding definition (oracle.weld): against a plain
dictionary walk in first-appearance order, and against the properties that
define it (welded vertices pairwise distinct and in first-appearance order,
every face keeps its three positions bit for bit)."""
import numpy as np
import pytest

from helpers import noise, sphere_field


def dict_weld(V, F):
    """First-appearance welding by walking the vertices with a dict."""
    seen, newid, keep = {}, [], []
    for i in range(V.shape[0]):
        k = V[i].tobytes()
        if k not in seen:
            seen[k] = len(keep)
            keep.append(i)
        newid.append(seen[k])
    newid = np.asarray(newid, np.int64)
    return V[keep], (newid[F - 1] + 1) if F.size else F.copy()


def check_properties(V, F, Vw, Fw):
    w = V.dtype.itemsize
    kb = Vw.view(np.dtype((np.void, 3 * w))).ravel()
    assert len(set(kb.tolist())) == Vw.shape[0], "welded vertices repeat"
    # first-appearance order: the first index of each welded vertex increases
    first = {}
    for i, k in enumerate(V.view(np.dtype((np.void, 3 * w))).ravel().tolist()):
        first.setdefault(k, i)
    idx = [first[k] for k in kb.tolist()]
    assert idx == sorted(idx)
    assert len(first) == Vw.shape[0]
    # faces keep their positions
    if F.size:
        assert np.array_equal(Vw[Fw - 1].view(np.uint8), V[F - 1].view(np.uint8))


def cases(oracle):
    out = []
    for dt in (np.float32, np.float64):
        out.append(oracle.isosurface(sphere_field(24, 0.6, dt), "mc"))
        out.append(oracle.isosurface(noise((20, 17, 23), 3, dt), "mc", iso=0.0))
        out.append(oracle.isosurface(noise((16, 16, 16), 5, dt), "sn"))
    rng = np.random.default_rng(7)
    pts = rng.standard_normal((50, 3))
    V = pts[rng.integers(0, 50, 400)]
    V[::7] = [-0.0, 0.0, 1.0]
    V[::11] = [0.0, 0.0, 1.0]
    V[5] = V[9] = [np.nan, 1.0, 2.0]
    F = rng.integers(1, 401, (300, 3))
    out.append((V, F))
    out.append((V.astype(np.float32), F))
    out.append((np.zeros((0, 3)), np.zeros((0, 3), np.int64)))
    out.append((np.ones((5, 3), np.float32), np.zeros((0, 3), np.int64)))
    return out


def test_weld_matches_dictionary_walk(oracle):
    for V, F in cases(oracle):
        Vw, Fw = oracle.weld(V, F)
        Vd, Fd = dict_weld(V, F)
        assert Vw.dtype == V.dtype
        assert np.array_equal(Vw.view(np.uint8), Vd.view(np.uint8))
        assert np.array_equal(Fw, Fd)
        check_properties(V, F, Vw, Fw)


def test_weld_merges_marching_cubes_copies(oracle):
    # the cubes around a crossing edge each emit a copy of its vertex; welding
    # leaves at least one vertex per crossing edge and removes most copies
    V, F = oracle.isosurface(sphere_field(32, 0.5, np.float64), "mc")
    Vw, Fw = oracle.weld(V, F)
    assert Vw.shape[0] < V.shape[0] // 3
    assert Fw.shape == F.shape


def test_weld_bitwise_equality(oracle):
    V = np.array([[0.0, 0, 0], [-0.0, 0, 0], [0.0, 0, 0]])
    Vw, Fw = oracle.weld(V, np.array([[1, 2, 3]]))
    assert Vw.shape[0] == 2 and Fw.tolist() == [[1, 2, 1]]


def test_weld_rejects_bad_indices(oracle):
    with pytest.raises(ValueError):
        oracle.weld(np.zeros((3, 3)), np.array([[1, 2, 4]]))
    with pytest.raises(ValueError):
        oracle.weld(np.zeros((3, 3)), np.array([[0, 1, 2]]))
