"""Generate the Poisson golden fixtures (independent of the oracle and the product).

For a random star velocity u* on a small MAC grid, the pressure increment dp of
the SMAC step solves A dp = b with b = -(1/dt) D u* (G_PressureSolverBase.cu:49-118)
and A = -D beta G (beta = 1, zero flux through walls; G_PCGSolver.cu:126-212).
Here A is assembled from Kronecker products of 1-D second differences and solved
with scipy.sparse.linalg.spsolve after pinning one cell; dp is stored mean-free.

Run: python tests/golden/make_golden.py   (writes the .npz files next to it)
"""
import os

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

HERE = os.path.dirname(os.path.abspath(__file__))


def lap1d(n, h, periodic):
    main = np.full(n, 2.0)
    if not periodic:
        main[0] = main[-1] = 1.0
    L = sp.diags([main, -np.ones(n - 1), -np.ones(n - 1)], [0, -1, 1], format="lil")
    if periodic:
        L[0, n - 1] = -1.0
        L[n - 1, 0] = -1.0
    return (L / (h * h)).tocsr()


def make(name, n, periodic, size, dt, seed):
    rng = np.random.default_rng(seed)
    nx, ny, nz = n
    h = [size[d] / n[d] for d in range(3)]
    I = [sp.identity(n[d], format="csr") for d in range(3)]
    A = (sp.kron(I[2], sp.kron(I[1], lap1d(nx, h[0], periodic[0])))
         + sp.kron(I[2], sp.kron(lap1d(ny, h[1], periodic[1]), I[0]))
         + sp.kron(lap1d(nz, h[2], periodic[2]), sp.kron(I[1], I[0]))).tocsr()
    # star velocities in the reference layout (z, y, x); wall faces zero
    shapes = {0: (nz + 2, ny + 2, nx + 3), 1: (nz + 2, ny + 3, nx + 2), 2: (nz + 3, ny + 2, nx + 2)}
    us = []
    for a in range(3):
        f = np.zeros(shapes[a])
        sl = [slice(1, nz + 1), slice(1, ny + 1), slice(1, nx + 1)]
        # faces 1..n along a (periodic) or 2..n (walls: 1 and n+1 are walls)
        lo = 1 if periodic[a] else 2
        sl[2 - a] = slice(lo, n[a] + 1)
        f[tuple(sl)] = rng.standard_normal(f[tuple(sl)].shape)
        if periodic[a]:  # duplicate face n+1 = face 1 for the divergence below
            src = [slice(1, nz + 1), slice(1, ny + 1), slice(1, nx + 1)]
            dst = list(src)
            src[2 - a] = slice(1, 2)
            dst[2 - a] = slice(n[a] + 1, n[a] + 2)
            f[tuple(dst)] = f[tuple(src)]
        us.append(f)
    u, v, w = us
    div = ((u[1:nz + 1, 1:ny + 1, 2:nx + 2] - u[1:nz + 1, 1:ny + 1, 1:nx + 1]) / h[0]
           + (v[1:nz + 1, 2:ny + 2, 1:nx + 1] - v[1:nz + 1, 1:ny + 1, 1:nx + 1]) / h[1]
           + (w[2:nz + 2, 1:ny + 1, 1:nx + 1] - w[1:nz + 1, 1:ny + 1, 1:nx + 1]) / h[2])
    b = (-(1.0 / dt) * div).ravel()
    b = b - b.mean()
    Ar = A[1:, 1:]
    x = np.zeros(b.size)
    x[1:] = spla.spsolve(Ar.tocsc(), b[1:])
    x -= x.mean()
    res = A @ x - b
    assert np.max(np.abs(res)) < 1e-9 * np.max(np.abs(b)), np.max(np.abs(res))
    np.savez_compressed(os.path.join(HERE, name), n=np.array(n), periodic=np.array(periodic), size=np.array(size),
                        dt=np.array(dt), ustar=u, vstar=v, wstar=w, dp=x.reshape(nz, ny, nx))
    print(name, "ok")


if __name__ == "__main__":
    make("poisson_periodic_8.npz", (8, 8, 8), (1, 1, 1), (2.0, 2.0, 2.0), 1e-2, 101)
    make("poisson_walls_8x6x4.npz", (8, 6, 4), (0, 0, 0), (1.0, 0.75, 0.5), 5e-3, 102)
