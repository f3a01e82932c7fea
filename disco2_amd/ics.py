"""Seeded initial states for the benchmark and test problems.

The reference selects its initial data through a function-pointer plugin
(src/Cell/cell_init_function_ptr.c:6-46).  The benchmark problems use initial
states of their own, defined here on the cell centres of a given mesh:

* ``vortex``      -- isentropic (Yee) vortex on a uniform background;
* ``binary_disk`` -- Keplerian disk around the binary's total mass with a
                     sound speed set by a Mach number;
* ``disk3d``      -- vertically isothermal Keplerian disk with a Gaussian
                     density profile in z.

Noise is drawn from a counter-based generator (splitmix64 of seed, disk id
and zone index), so the state of one zone does not depend on the mesh size
or on how many zones are generated together.

Zones are ordered layer k, annulus i, cell j (ghosts included); a primitive
array is q-major: ``prim[q, zone]`` with q = (rho, P, v_r, omega, v_z, ...).
"""
from __future__ import annotations

import numpy as np

RHO, PPP, URR, UPP, UZZ = range(5)
_MASK64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x: np.ndarray) -> np.ndarray:
    """splitmix64 finaliser (Steele, Lea & Flood 2014) on a uint64 array."""
    with np.errstate(over="ignore"):
        z = (x + np.uint64(0x9E3779B97F4A7C15)) & _MASK64
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _MASK64
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _MASK64
        return z ^ (z >> np.uint64(31))


def uniform01(seed: int, disk: int, n: int) -> np.ndarray:
    """n uniform deviates in [0,1) for zones 0..n-1 of one disk."""
    key = splitmix64(np.array([(seed << 20) ^ disk], dtype=np.uint64))[0]
    idx = np.arange(n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        h = splitmix64(idx ^ key)
    return (h >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def zone_coords(geom: dict):
    """Cell-centre (r, phi, z) of every zone of a mesh description.

    ``geom`` holds N_p[NR], r_faces[NR+1], z_faces[NZ+1], tiph[Nz], dphi[Nz];
    annulus i spans r_faces[i]..r_faces[i+1] (the reference's sim_FacePos
    offset, src/Sim/sim_access.c:26-36)."""
    N_p = np.asarray(geom["N_p"], dtype=np.int64)
    rf = np.asarray(geom["r_faces"], dtype=np.float64)
    zf = np.asarray(geom["z_faces"], dtype=np.float64)
    NR, NZ = len(N_p), len(zf) - 1
    rc_ann = 0.5 * (rf[1:] + rf[:-1])
    zc_lay = 0.5 * (zf[1:] + zf[:-1])
    r_layer = np.repeat(rc_ann, N_p)
    r = np.tile(r_layer, NZ)
    z = np.repeat(zc_lay, int(N_p.sum()))
    phi = np.asarray(geom["tiph"]) - 0.5 * np.asarray(geom["dphi"])
    return r, phi, z


def vortex(geom: dict, gamma: float, seed: int = 1, disk: int = 0, beta: float = 5.0,
           x0: float = 2.5, y0: float = 0.0, noise: float = 1e-3, nq: int = 5) -> np.ndarray:
    r, phi, _ = zone_coords(geom)
    x, y = r * np.cos(phi), r * np.sin(phi)
    dx, dy = x - x0, y - y0
    rr2 = dx * dx + dy * dy
    f = np.exp(0.5 * (1.0 - rr2))
    du = -beta / (2 * np.pi) * f * dy
    dv = beta / (2 * np.pi) * f * dx
    T = 1.0 - (gamma - 1.0) * beta * beta / (8.0 * gamma * np.pi ** 2) * np.exp(1.0 - rr2)
    rho = T ** (1.0 / (gamma - 1.0))
    P = rho ** gamma
    rho = rho * (1.0 + noise * (2.0 * uniform01(seed, disk, len(r)) - 1.0))
    vr = du * np.cos(phi) + dv * np.sin(phi)
    vp = -du * np.sin(phi) + dv * np.cos(phi)
    prim = np.zeros((nq, len(r)))
    prim[RHO], prim[PPP], prim[URR], prim[UPP] = rho, P, vr, vp / r
    for q in range(5, nq):  # passive scalars: mark the vortex core at widths 1, 1/2, 1/3
        prim[q] = np.exp(-(q - 4) * rr2)
    return prim


def binary_disk(geom: dict, gamma: float, mach: float = 20.0, sigma0: float = 1.0,
                seed: int = 1, disk: int = 0, noise: float = 0.01, nq: int = 5) -> np.ndarray:
    r, _, _ = zone_coords(geom)
    sig = sigma0 * (1.0 + noise * (2.0 * uniform01(seed, disk, len(r)) - 1.0))
    vk = r ** -0.5
    cs = vk / mach
    prim = np.zeros((nq, len(r)))
    prim[RHO] = sig
    prim[PPP] = sig * cs * cs / gamma
    prim[UPP] = r ** -1.5
    return prim


def disk3d(geom: dict, gamma: float, H: float = 0.1, seed: int = 1, disk: int = 0,
           noise: float = 0.01, rho_floor: float = 1e-5, nq: int = 5) -> np.ndarray:
    r, _, z = zone_coords(geom)
    om = r ** -1.5
    rho = np.exp(-z * z / (2.0 * H * H)) * (1.0 + noise * (2.0 * uniform01(seed, disk, len(r)) - 1.0))
    rho = np.maximum(rho, rho_floor)
    ciso = H * om
    prim = np.zeros((nq, len(r)))
    prim[RHO] = rho
    prim[PPP] = rho * ciso * ciso
    prim[UPP] = om
    return prim


def make_ic(kind: str, geom: dict, params: dict, **kw) -> np.ndarray:
    nq = params["NUM_C"] + params["NUM_N"]
    gamma = params["Adiabatic_Index"]
    if kind == "vortex":
        return vortex(geom, gamma, nq=nq, **kw)
    if kind == "binary_disk":
        return binary_disk(geom, gamma, nq=nq, **kw)
    if kind == "disk3d":
        return disk3d(geom, gamma, rho_floor=params["Rho_Floor"], nq=nq, **kw)
    raise ValueError(f"unknown initial state {kind!r}")
