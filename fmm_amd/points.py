"""Seeded point sets and charges of the benchmarks (numpy).

The generator is fmm-validate's SplitMix64 (fmm-validate/src/rng.rs:29-45): draw k of a
stream is mix(state + (k + 1) * 0x9e3779b97f4a7c15), so a run of draws is computed at
once.  Distributions follow fmm-validate/src/points.rs (cube :27-31, Box-Muller normal
:142-147, Gaussian clusters truncated at 4 sigma :114-135) and the benchmark sample is
fmm-bench's partial Fisher-Yates (fmm-bench/src/measure.rs:52-57).
"""
from __future__ import annotations

import math

import numpy as np

BENCH_SEED = 0x00BE9C06
_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


class SplitMix64:
    def __init__(self, seed: int):
        self.state = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)

    def next_u64(self, n: int) -> np.ndarray:
        with np.errstate(over="ignore"):
            k = np.arange(1, n + 1, dtype=np.uint64)
            z = self.state + k * _GAMMA
            self.state = self.state + np.uint64(n) * _GAMMA
            z = (z ^ (z >> np.uint64(30))) * _M1
            z = (z ^ (z >> np.uint64(27))) * _M2
            return z ^ (z >> np.uint64(31))

    def uniform(self, n: int) -> np.ndarray:
        """Top 53 bits scaled to [0, 1)."""
        return (self.next_u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)

    def normals(self, n: int) -> np.ndarray:
        """Box-Muller, cosine branch: draws (u1, u2) per value."""
        u = self.uniform(2 * n)
        return np.sqrt(-2.0 * np.log(1.0 - u[0::2])) * np.cos(2.0 * 3.141592653589793 * u[1::2])

    def charges(self, n: int) -> np.ndarray:
        """Uniform in [-1, 1) (fmm-validate/src/points.rs:137-139)."""
        return -1.0 + 2.0 * self.uniform(n)

    def cube(self, n: int, centre=(0.5, 0.5, 0.5), half=0.5) -> np.ndarray:
        u = self.uniform(3 * n).reshape(n, 3)
        return np.asarray(centre, np.float64) + half * (-1.0 + 2.0 * u)

    def sample(self, n: int, sampled: int) -> np.ndarray:
        """The first `sampled` entries of a partial Fisher-Yates shuffle of 0..n."""
        m = min(sampled, n)
        idx = np.arange(n, dtype=np.int64)
        r = self.next_u64(m) >> np.uint64(32)
        for i in range(m):
            j = i + int((int(r[i]) * (n - i)) >> 32)
            idx[i], idx[j] = idx[j], idx[i]
        return idx[:m].copy()

    def gaussian_blob(self, n: int, centre, width: float) -> np.ndarray:
        """n points of one Gaussian cluster truncated at 4 sigma (rejected triples drawn
        again), in the order the sequential generator gives them."""
        out = np.empty((0, 3))
        while len(out) < n:
            want = n - len(out)
            batch = want + max(16, want // 50)
            before = self.state
            z = self.normals(3 * batch).reshape(batch, 3)
            ok = np.einsum("ij,ij->i", z, z) <= 16.0
            take = np.nonzero(ok)[0][:want]
            if len(take) < want:
                out = np.vstack([out, z[take]])
                continue
            # rewind to just after the last triple used
            used = int(take[-1]) + 1
            with np.errstate(over="ignore"):
                self.state = before + np.uint64(6 * used) * _GAMMA
            out = np.vstack([out, z[take]])
        return np.asarray(centre, np.float64) + width * out


def _next_accepted(ok: np.ndarray) -> np.ndarray:
    """nxt[j] = the first j' >= j with j' = j (mod 3) and ok[j'] (len(ok) where none)."""
    n = len(ok)
    nxt = np.full(n + 3, n, dtype=np.int64)
    for c in range(3):
        idx = np.arange(c, n, 3)
        cand = np.where(ok[idx], idx, n)
        nxt[idx] = np.minimum.accumulate(cand[::-1])[::-1]
    return nxt


def sphere(rng: SplitMix64, n: int, centre=(0.5, 0.5, 0.5), radius=0.5) -> np.ndarray:
    """points::sphere (fmm-validate/src/points.rs:48-57): in_unit_ball with norm >= 1e-3,
    normalised, one triple of draws per try."""
    out = np.empty((0, 3))
    while len(out) < n:
        want = n - len(out)
        before = rng.state
        u = rng.uniform(3 * (want + want // 2 + 64))
        t = u.reshape(-1, 3)
        v = -1.0 + 2.0 * t
        s = (v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1]) + v[:, 2] * v[:, 2]
        ok = np.nonzero((s <= 1.0) & (s >= 1e-3 * 1e-3))[0][:want]
        with np.errstate(over="ignore"):
            rng.state = before + np.uint64(3 * (int(ok[-1]) + 1 if len(ok) == want else len(t))) * _GAMMA
        vv = v[ok]
        out = np.vstack([out, vv / np.sqrt(s[ok])[:, None]])
    return np.asarray(centre, np.float64) + radius * out


def plummer(rng: SplitMix64, n: int, centre=(0.5, 0.5, 0.5), scale=0.1) -> np.ndarray:
    """points::plummer (fmm-validate/src/points.rs:60-100): per point one draw u for the
    radius, r = a / sqrt((m u)^(-2/3) - 1) with the mass m inside 10a, then a direction by
    in_unit_ball (norm >= 1e-3)."""
    x = 10.0
    mass = x ** 3 / (1.0 + x * x) ** 1.5
    pts = np.empty((n, 3))
    done = 0
    while done < n:
        want = n - done
        before = rng.state
        count = 8 * want + 256
        u = rng.uniform(count)
        v = -1.0 + 2.0 * u
        s = np.full(count, 2.0)
        s[: count - 2] = (v[:-2] * v[:-2] + v[1:-1] * v[1:-1]) + v[2:] * v[2:]
        ok = (s <= 1.0) & (s >= 1e-3 * 1e-3)
        nxt = _next_accepted(ok)
        starts = []
        pos = 0
        while len(starts) < want:
            if pos + 1 >= count:
                break
            j = int(nxt[pos + 1])
            if j + 3 > count:
                break
            starts.append((pos, j))
            pos = j + 3
        st = np.array(starts, dtype=np.int64).reshape(-1, 2)
        k = len(st)
        um = mass * u[st[:, 0]]
        # libm pow per value (numpy's vector pow may round differently)
        pw = np.array([math.pow(x, -2.0 / 3.0) if x > 0.0 else 2.0 for x in um.tolist()])
        r = np.where(um > 0.0, scale / np.sqrt(pw - 1.0), 0.0)
        d = np.stack([v[st[:, 1]], v[st[:, 1] + 1], v[st[:, 1] + 2]], axis=1)
        d /= np.sqrt(s[st[:, 1]])[:, None]
        pts[done:done + k] = np.asarray(centre, np.float64) + r[:, None] * d
        done += k
        with np.errstate(over="ignore"):
            rng.state = before + np.uint64(pos) * _GAMMA
    return pts


def bench_cube(n: int, sampled: int = 4096, charges: str = "normal", seed: int = BENCH_SEED):
    """The benchmark problem of fmm-bench/src/measure.rs:35-67: n points uniform in the
    unit cube, then charges from the same stream (N(0,1) by Box-Muller, or uniform in
    [-1, 1)), then the sampled targets."""
    rng = SplitMix64(seed)
    pts = rng.cube(n)
    q = rng.normals(n) if charges == "normal" else rng.charges(n)
    sample = rng.sample(n, sampled)
    return pts, q, sample


def batched_clouds(nclouds: int = 256, nmin: int = 2000, nmax: int = 256000, seed: int = BENCH_SEED + 5):
    """Independent clouds of log-uniform sizes in [nmin, nmax], each one Gaussian blob
    (width 0.1, centre uniform in [0.3, 0.7]^3), with N(0,1) charges."""
    rng = SplitMix64(seed)
    u = rng.uniform(nclouds)
    sizes = np.floor(np.exp(np.log(nmin) + u * (np.log(nmax) - np.log(nmin)))).astype(np.int64)
    clouds, charges = [], []
    for n in sizes:
        centre = 0.3 + 0.4 * rng.uniform(3)
        clouds.append(rng.gaussian_blob(int(n), centre, 0.1))
        charges.append(rng.normals(int(n)))
    return clouds, charges


def nonuniform(n_sphere: int = 2_000_000, n_plummer: int = 1_000_000, seed: int = BENCH_SEED):
    """The non-uniform configuration: n_sphere points on the surface of the sphere inscribed
    in the unit cube, then n_plummer Plummer points (a = 0.1, truncated at 10a) about its
    centre, then charges uniform in [0, 1), all from one stream."""
    rng = SplitMix64(seed)
    a = sphere(rng, n_sphere, (0.5, 0.5, 0.5), 0.5)
    b = plummer(rng, n_plummer, (0.5, 0.5, 0.5), 0.1)
    q = rng.uniform(n_sphere + n_plummer)
    return np.vstack([a, b]), q


def helmholtz_cube(n: int, sampled: int = 1000, seed: int = BENCH_SEED):
    """The Helmholtz benchmark problem: n points uniform in the unit cube (the generator of
    bench_cube), then complex charges with real and imaginary parts N(0,1) from the same
    stream (all real parts, then all imaginary parts), then the sampled targets."""
    rng = SplitMix64(seed)
    pts = rng.cube(n)
    re = rng.normals(n)
    im = rng.normals(n)
    sample = rng.sample(n, sampled)
    return pts, re + 1j * im, sample
