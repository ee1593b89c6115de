"""numpy statements of the synthetic fields (the definitions the device
generator follows), used to check msh_generate_field.

All fields are indexed [x, y, z] on an nx x ny x nz grid; the sample at
index i of an axis of n samples sits at lin(i, n, a, b) =
(1 - t) a + t b with t = i / max(n - 1, 1).

- sphere  sqrt(x^2 + y^2 + z^2) - p0 on [-1, 1]^3
- torus   sqrt((sqrt(x^2 + y^2) - p0)^2 + z^2) - p1 on [-1, 1]^3
- gyroid  cos x sin y + cos y sin z + cos z sin x on [-p0, p0]^3
- value noise: a 32^3 lattice of Uniform(-1, 1) values spanning the unit
  cube, trilinearly interpolated, plus 0.25 x a 128^3 lattice (seed xor
  0xA5A5A5A5A5A5A5A5); lattice value l = 2 u01(splitmix64(seed xor
  l_index * 0xD1B54A32D192ED03)) - 1 with l_index = a + L (b + L c)
- blobs: 0.5 - sum of 16 Gaussians exp(-|p - c_b|^2 / (2 s_b^2)) with
  centres in [-0.6, 0.6]^3 and s_b in [0.05, 0.2] drawn from
  splitmix64(4 h + 0..3), h = 16 seed + b.

Every operation is written in the order the generator performs it, in
Float64, so the integer hashing and the rational arithmetic agree bit for
bit; cos, sin and exp may differ from the device's by an ulp.
"""
import numpy as np

M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x):
    """SplitMix64's output for state x (Steele, Lea and Flood 2014): one
    golden-ratio increment, then the variant-13 finaliser."""
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def u01(h):
    """The top 53 bits of h as a double in [0, 1)."""
    return (np.asarray(h, dtype=np.uint64) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


def lin(n, a, b):
    d = max(n - 1, 1)
    t = np.arange(n, dtype=np.float64) / float(d)
    return (1.0 - t) * a + t * b


def _axes(shape, a, b):
    x, y, z = (lin(n, a, b) for n in shape)
    return np.meshgrid(x, y, z, indexing="ij")


def value_noise(shape, L, seed, index=None):
    """One octave on an L^3 lattice; `index` optionally picks the samples
    of each axis (three index arrays), so a slice of a large grid can be
    evaluated alone."""
    seed = np.uint64(seed)
    if index is None:
        index = [np.arange(n) for n in shape]
    lo, fr = [], []
    for n, idx in zip(shape, index):
        d = max(n - 1, 1)
        u = (np.asarray(idx, dtype=np.int64) * (L - 1)).astype(np.float64) / float(d)
        a = np.clip(np.floor(u).astype(np.int64), 0, L - 2)
        lo.append(a)
        fr.append(u - a.astype(np.float64))
    acc = np.zeros(tuple(len(a) for a in lo), dtype=np.float64)
    for c in range(8):
        dx, dy, dz = c & 1, (c >> 1) & 1, (c >> 2) & 1
        ix = (lo[0] + dx).astype(np.uint64)[:, None, None]
        iy = (lo[1] + dy).astype(np.uint64)[None, :, None]
        iz = (lo[2] + dz).astype(np.uint64)[None, None, :]
        Lu = np.uint64(L)
        with np.errstate(over="ignore"):
            li = ix + Lu * (iy + Lu * iz)
            val = 2.0 * u01(splitmix64(seed ^ (li * np.uint64(0xD1B54A32D192ED03)))) - 1.0
        wx = (fr[0] if dx else 1.0 - fr[0])[:, None, None]
        wy = (fr[1] if dy else 1.0 - fr[1])[None, :, None]
        wz = (fr[2] if dz else 1.0 - fr[2])[None, None, :]
        w = (wx * wy) * wz
        acc = acc + w * val
    return acc


def value_noise_at(shape, seed, index=None):
    """The value-noise field (both octaves), optionally at chosen samples."""
    return value_noise(shape, 32, seed, index) + 0.25 * value_noise(
        shape, 128, np.uint64(seed) ^ np.uint64(0xA5A5A5A5A5A5A5A5), index)


def field(kind, shape, seed=0, p0=None, p1=None):
    """The Float64 values of a synthetic field (round to Float32 for a
    Float32 field)."""
    if kind == "sphere":
        p0 = 0.5 if p0 is None else p0
        x, y, z = _axes(shape, -1.0, 1.0)
        return np.sqrt(x * x + y * y + z * z) - p0
    if kind == "torus":
        p0 = 0.6 if p0 is None else p0
        p1 = 0.2 if p1 is None else p1
        x, y, z = _axes(shape, -1.0, 1.0)
        q = np.sqrt(x * x + y * y) - p0
        return np.sqrt(q * q + z * z) - p1
    if kind == "gyroid":
        p0 = 2 * np.pi if p0 is None else p0
        x, y, z = _axes(shape, -p0, p0)
        return np.cos(x) * np.sin(y) + np.cos(y) * np.sin(z) + np.cos(z) * np.sin(x)
    if kind == "value_noise":
        return value_noise_at(shape, seed)
    if kind == "blobs":
        x, y, z = _axes(shape, -1.0, 1.0)
        acc = np.zeros(shape, dtype=np.float64)
        with np.errstate(over="ignore"):
            for b in range(16):
                h = np.uint64(seed) * np.uint64(16) + np.uint64(b)
                r = [u01(splitmix64(np.uint64(4) * h + np.uint64(q))) for q in range(4)]
                cx, cy, cz = (1.2 * float(r[q]) - 0.6 for q in range(3))
                sg = 0.05 + 0.15 * float(r[3])
                d2 = (x - cx) * (x - cx) + (y - cy) * (y - cy) + (z - cz) * (z - cz)
                acc = acc + np.exp(-d2 / (2.0 * sg * sg))
        return 0.5 - acc
    raise ValueError(kind)
