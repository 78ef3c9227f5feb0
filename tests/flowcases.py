"""Shared synthetic cases for the parity tests: meshes and smooth non-uniform flow states."""
import numpy as np

P_INF = 1.0
T_INF = 1.0 / 1.4          # rho = 1.4, speed of sound 1 (the reference's example cases)


def c1_elements(gr, n=20, seed=2024):
    """Unit box, n^3 hexes split into 6 tets, interior nodes jittered by U[-0.1h, 0.1h]."""
    return gr.Elements.box(n, n, n, tets=True, jitter=0.1, seed=seed)


def ogrid_elements(gr, nt=256, nr=128, nz=4):
    """Cylinder O-grid, r 0.5 -> 10 (20 radii), radial ratio 1.05, dz = 2*pi*0.5/256."""
    return gr.Elements.ogrid(nt, nr, nz, 0.5, 10.0, 1.05, 2 * np.pi * 0.5 / 256)


def perturbed_state(centers, mach=(0.5, 0.0, 0.0), amp=0.1):
    """Smooth state around the freestream: rho, rhou, rhoe at the given points (n,3)."""
    x, y, z = centers[:, 0], centers[:, 1], centers[:, 2]
    s = np.sin(2 * np.pi * x) * np.sin(2 * np.pi * y) * np.sin(2 * np.pi * z)
    rho = 1.4 * (1 + amp * s)
    u = np.stack([mach[0] * (1 + amp * np.cos(2 * np.pi * y)),
                  mach[1] + amp * 0.3 * np.sin(2 * np.pi * z),
                  mach[2] + amp * 0.2 * np.cos(2 * np.pi * x)], axis=1)
    p = P_INF * (1 + 0.5 * amp * np.cos(2 * np.pi * (x + y)))
    rhou = rho[:, None] * u
    rhoe = p / 0.4 + 0.5 * rho * np.sum(u * u, axis=1)
    return rho, rhou, rhoe


def farfield_bcs(names, mach):
    return [(n, "farfield", mach) for n in names]
