"""Writes the golden fixtures of this directory from the CPU oracle.

Each fixture holds the input field (x-fastest, indexed [x, y, z]), the
method parameters and axis ends, and the expected (vertices, faces) with
1-based faces, as `isosurface(A, method, X, Y, Z)` returns them
(src/isosurface.jl:30-32, src/marching_cubes.jl:27-64,
src/marching_tetrahedra.jl:129-163).  They are written only after the
oracle passes its known-answer and invariant tests, and they freeze its
output so that later edits of the oracle or the library cannot drift
silently.

    python tests/golden/make_golden.py
"""
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))


def grid(n, lo, hi):
    r = np.linspace(lo, hi, n)
    return np.meshgrid(r, r, r, indexing="ij")


def sphere(n, radius, dtype):
    x, y, z = grid(n, -1.0, 1.0)
    return (np.sqrt(x * x + y * y + z * z) - radius).astype(dtype)


def gyroid(n, dtype):
    x, y, z = grid(n, -2 * np.pi, 2 * np.pi)
    return (np.cos(x) * np.sin(y) + np.cos(y) * np.sin(z) + np.cos(z) * np.sin(x)).astype(dtype)


def signs(shape, seed, dtype):
    """Random +-1 (and a few zeros) samples: most of the 256 cube cases."""
    rng = np.random.default_rng(seed)
    return rng.choice(np.array([-1.0, -0.5, 0.0, 0.5, 1.0]), size=shape).astype(dtype)


def torus(n, dtype):
    x, y, z = grid(n, -1.0, 1.0)
    q = np.sqrt(x * x + y * y) - 0.6
    return (np.sqrt(q * q + z * z) - 0.2).astype(dtype)


# name -> (field builder, method, iso, eps, X, Y, Z)
CASES = {
    "mc_sphere32_f32": (lambda: sphere(32, 0.5, np.float32), "mc", 0, None,
                        (-1, 1), (-1, 1), (-1, 1)),
    "mc_torus28_f64_iso": (lambda: torus(28, np.float64), "mc", 0.013, None,
                           (-1.0, 1.0), (-1.0, 1.0), (-1.0, 1.0)),
    "mc_signs9x8x7_f32": (lambda: signs((9, 8, 7), 7, np.float32), "mc", 0.0, None,
                          (0, 10), (np.float32(-0.3), np.float32(0.7)), (-2.5, 3.0)),
    "mt_gyroid16_f32": (lambda: gyroid(16, np.float32), "mt", 0, 1e-3,
                        (-2 * np.pi, 2 * np.pi), (-2 * np.pi, 2 * np.pi), (-2 * np.pi, 2 * np.pi)),
    "mt_signs9x8x7_f64": (lambda: signs((9, 8, 7), 11, np.float64), "mt", 0.0, 1e-3,
                          (-1, 1), (-1, 1), (-1, 1)),
    "sn_gyroid16_f32": (lambda: gyroid(16, np.float32), "sn", 0, None,
                        (-2 * np.pi, 2 * np.pi), (-2 * np.pi, 2 * np.pi), (-2 * np.pi, 2 * np.pi)),
    "sn_torus28_f64": (lambda: torus(28, np.float64), "sn", 0.0, None,
                       (-1, 1), (-1, 1), (-1, 1)),
}


def axis_repr(a):
    """Axis end pair -> [(type, value)] keeping Int / Float32 / Float64."""
    out = []
    for v in a:
        if isinstance(v, (int, np.integer)):
            out.append(["int64", int(v)])
        elif isinstance(v, np.float32):
            out.append(["float32", float(v)])
        else:
            out.append(["float64", float(v)])
    return out


def main():
    import __graft_entry__ as entry
    O = entry.load_oracle()
    for name, (build, method, iso, eps, X, Y, Z) in CASES.items():
        f = build()
        kw = dict(iso=iso, X=X, Y=Y, Z=Z)
        if eps is not None:
            kw["eps"] = eps
        V, F = O.isosurface(f, method, **kw)
        meta = dict(method=method, iso=axis_repr([iso])[0], eps=eps,
                    X=axis_repr(X), Y=axis_repr(Y), Z=axis_repr(Z))
        np.savez_compressed(HERE / f"{name}.npz", field=f, V=V, F=F,
                            meta=np.array(json.dumps(meta)))
        print(f"{name}: field {f.shape} {f.dtype}, {len(V)} vertices ({V.dtype}), {len(F)} faces")


if __name__ == "__main__":
    main()
