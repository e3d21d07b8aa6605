"""Synthetic inputs of the benchmark configurations (build-defined generators,
no files): points, grids and subdomain layouts.

cfg1  20,000 points uniform in [0,10]^3 (SplitMix64, seed 1), Shinozuka,
      corrL 1, Nmc 4, seed 1, xRange 10.
cfg2  128^3 grid, spacing 0.1, x fastest, corrL 0.5, isotropic, lognormal
      (avg 1, var 0.25), Nmc 64, seed 7.
cfg3  2,000,000 points of an 8-component Gaussian mixture (centres U[0,100]^3,
      std 5), corrL (2,2,0.5), randomization with N_k 4096 drawn from the
      spectral density, Nmc 256, seed 11.
cfg4  512x512x256 grid, spacing 0.25, 4x4x2 uneven subdomains with 7-point
      overlaps, Shinozuka per subdomain (seed 100 + field index, last axis
      fastest), Nmc 16, variance-preserving blend.
"""
from __future__ import annotations

import numpy as np

_M64 = (1 << 64) - 1


def splitmix64(seed: int, count: int) -> np.ndarray:
    """count 64-bit outputs of SplitMix64 (Steele, Lea, Flood 2014)."""
    out = np.empty(count, dtype=np.uint64)
    s = np.uint64(seed & _M64)
    golden = np.uint64(0x9E3779B97F4A7C15)
    m1 = np.uint64(0xBF58476D1CE4E5B9)
    m2 = np.uint64(0x94D049BB133111EB)
    idx = np.arange(1, count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = s + golden * idx
        z = (z ^ (z >> np.uint64(30))) * m1
        z = (z ^ (z >> np.uint64(27))) * m2
        z = z ^ (z >> np.uint64(31))
    out[:] = z
    return out


def uniform01(seed: int, count: int) -> np.ndarray:
    """U[0,1) with 53-bit mantissas from SplitMix64."""
    return (splitmix64(seed, count) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def cfg1_points(n: int = 20000, seed: int = 1) -> np.ndarray:
    """(n, 3) uniform in [0,10)^3, coordinate-fastest draw order."""
    return uniform01(seed, 3 * n).reshape(n, 3) * 10.0


def grid_points(nstep, step, origin=(0.0, 0.0, 0.0)) -> np.ndarray:
    """Structured grid with x fastest (setGrid order, math_RF.f90:70-88)."""
    d = len(nstep)
    axes = [origin[i] + step[i] * np.arange(nstep[i], dtype=np.float64) for i in range(d)]
    mesh = np.meshgrid(*axes[::-1], indexing="ij")        # last axis slowest
    return np.stack([m.reshape(-1) for m in mesh[::-1]], axis=1)


def gaussian_mixture_points(n: int = 2_000_000, seed: int = 11, components: int = 8,
                            box: float = 100.0, std: float = 5.0) -> np.ndarray:
    """(n, 3) points: component c uniform, centre U[0,box]^3, isotropic std."""
    u = uniform01(seed, 3 * components + 4 * n)
    centres = u[: 3 * components].reshape(components, 3) * box
    rest = u[3 * components:]
    comp = np.minimum((rest[:n] * components).astype(np.int64), components - 1)
    u1 = rest[n: 2 * n]
    u2 = rest[2 * n: 3 * n]
    u3 = rest[3 * n: 4 * n]
    # Box-Muller, three normals from three uniforms pairs (u1,u2), (u3,u2+.5)
    r1 = np.sqrt(-2.0 * np.log1p(-u1))
    r3 = np.sqrt(-2.0 * np.log1p(-u3))
    z = np.stack([r1 * np.cos(2 * np.pi * u2), r1 * np.sin(2 * np.pi * u2),
                  r3 * np.cos(2 * np.pi * u3)], axis=1)
    return centres[comp] + std * z


def cfg4_layout(groupN=(512, 512, 256), parts=(4, 4, 2), nOvl: int = 7):
    """Uneven tensor-product subdomains of a group grid: cores of
    round(m*N/p) .. round((m+1)*N/p)-1 per axis, extended into each
    neighbour by ceil/floor of nOvl/2 so adjacent subdomains share exactly
    nOvl points; boundary subdomains are not extended outwards and are
    therefore smaller.  Subdomains are numbered with the last axis fastest
    (prepare_Localization's setGrid(..., inverse=.true.),
    localization_RF.f90:152).  Returns (origins, sizes, faceMasks)."""
    d = len(groupN)
    lo_ext, hi_ext = nOvl // 2, nOvl - nOvl // 2
    per_axis = []
    for a in range(d):
        cuts = [round(m * groupN[a] / parts[a]) for m in range(parts[a] + 1)]
        iv = []
        for m in range(parts[a]):
            o = cuts[m] - (lo_ext if m > 0 else 0)
            e = cuts[m + 1] + (hi_ext if m < parts[a] - 1 else 0)
            iv.append((o, e - o, m > 0, m < parts[a] - 1))
        per_axis.append(iv)
    origins, sizes, masks = [], [], []
    import itertools
    for idx in itertools.product(*[range(p) for p in parts]):   # last axis fastest
        o, s, m = [], [], 0
        for a in range(d):
            iv = per_axis[a][idx[a]]
            o.append(iv[0]); s.append(iv[1])
            if iv[2]:
                m |= 1 << (2 * a)
            if iv[3]:
                m |= 1 << (2 * a + 1)
        origins.append(o); sizes.append(s); masks.append(m)
    return np.array(origins, np.int32), np.array(sizes, np.int32), np.array(masks, np.int32)
