"""Seeded initial fields for the benchmark and test workloads.

The reference reads w_A, w_B from phi_<n>.dat (init_cuda.cu:435-456) and has
no generator of its own; these generators make those inputs reproducible.
Random numbers: SplitMix64 (Steele, Lea, Flood 2014), one 64-bit output per
index i of a stream with seed s:
    z = s + (i+1)*0x9E3779B97F4A7C15 (mod 2^64)
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9
    z = (z ^ (z >> 27)) * 0x94D049BB133111EB
    z =  z ^ (z >> 31)
    u = (z >> 11) * 2^-53                  in [0, 1)
Noise eta = a*(2u - 1); indices 0..N-1 feed w_A, N..2N-1 feed w_B.
Fields: w_A = +A*S(r) + eta_A, w_B = -A*S(r) + eta_B, with x = i*lx/Nx etc.
"""
import numpy as np

_G = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix_uniform(seed, n, start=0):
    """u_i in [0,1) for i = start .. start+n-1 of stream `seed`."""
    i = np.arange(start + 1, start + n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + i * _G
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def coords(nx, ny, nz, lx, ly, lz):
    """x, y, z of every grid point in [z][y][x] order (flattened)."""
    x = np.arange(nx) * (lx / nx)
    y = np.arange(ny) * (ly / ny)
    z = np.arange(nz) * (lz / nz)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    return X.ravel(), Y.ravel(), Z.ravel()


def seed_shape(kind, nx, ny, nz, lx, ly, lz):
    """S(r) of the named morphology seed."""
    X, Y, Z = coords(nx, ny, nz, lx, ly, lz)
    kx, ky, kz = 2 * np.pi / lx, 2 * np.pi / ly, 2 * np.pi / lz
    if kind == "none":
        return np.zeros(nx * ny * nz)
    if kind == "lamellar":       # cos(2 pi x / L)
        return np.cos(kx * X)
    if kind == "hexagonal":      # commensurate star (1,0,0), (0,1,0), (1,1,0)
        return np.cos(kx * X) + np.cos(ky * Y) + np.cos(kx * X + ky * Y)
    if kind == "gyroid":         # single-wavevector gyroid level-set function
        return (np.sin(kx * X) * np.cos(ky * Y) + np.sin(ky * Y) * np.cos(kz * Z)
                + np.sin(kz * Z) * np.cos(kx * X))
    raise ValueError(f"unknown seed kind {kind!r}")


def seeded_fields(nx, ny, nz, lx, ly, lz, kind, amplitude, noise, seed):
    """(w_A, w_B) float64 arrays of N = nx*ny*nz points."""
    N = nx * ny * nz
    S = seed_shape(kind, nx, ny, nz, lx, ly, lz)
    u = splitmix_uniform(seed, 2 * N)
    eta = noise * (2.0 * u - 1.0)
    wA = amplitude * S + eta[:N]
    wB = -amplitude * S + eta[N:]
    return np.ascontiguousarray(wA), np.ascontiguousarray(wB)


# The five benchmark workloads (grid, box, chiN, f, ds, seed) -- one place for all callers.
WORKLOADS = {
    "c1_lamellar_32": dict(n=32, L=4.0, chiN=15.0, fA=0.5, ds=0.01, kind="lamellar", amp=1.0,
                           noise=0.1, seed=1, iters=300),
    "c2_cylinder_64": dict(n=64, L=4.2, chiN=20.0, fA=0.3, ds=0.005, kind="hexagonal", amp=1.0,
                           noise=0.05, seed=2, iters=500, anderson_m=10),
    "c3_gyroid_128": dict(n=128, L=3.8, chiN=20.0, fA=0.35, ds=0.005, kind="gyroid", amp=1.0,
                          noise=0.02, seed=3, anderson_m=20),
    "c5_adjoint_64": dict(n=64, L=4.0, chiN=18.0, fA=0.5, ds=0.01, kind="none", amp=0.0,
                          noise=0.1, seed=5, iters=50),
}


def workload_fields(name, n=None):
    w = WORKLOADS[name]
    n = n or w["n"]
    return seeded_fields(n, n, n, w["L"], w["L"], w["L"], w["kind"], w["amp"], w["noise"], w["seed"])


def sweep_instances(count=256, seed=4, sizes=(32, 48, 64), chi=(12.0, 30.0), f=(0.25, 0.5), L=(3.5, 5.0)):
    """Seeded instance list (n, L, chiN, fA) of the phase-diagram sweep:
    grid drawn from `sizes`, chiN, fA and L uniform in their ranges."""
    u = splitmix_uniform(seed, 4 * count).reshape(count, 4)
    out = []
    for i in range(count):
        n = sizes[min(int(u[i, 0] * len(sizes)), len(sizes) - 1)]
        out.append((n, L[0] + (L[1] - L[0]) * u[i, 3], chi[0] + (chi[1] - chi[0]) * u[i, 1],
                    f[0] + (f[1] - f[0]) * u[i, 2]))
    return out


def sweep_fields(instances, noise=0.1, seed=400):
    """seeded +-noise fields of every sweep instance (instance i uses stream seed + i)"""
    return [seeded_fields(n, n, n, L, L, L, "none", 0.0, noise, seed + i) for i, (n, L, c, f) in enumerate(instances)]
