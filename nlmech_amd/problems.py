"""Synthetic specimens for the explicit-dynamics workloads.

A specimen is what NLMech reads from its YAML deck and mesh file: node
coordinates and volumes, the Model / Material / Displacement_BC / Force_BC /
Fracture decks.  Grids follow the reference mesh tool's conventions
(tools/mesh/src/2d/fe2D.cpp:115-198): node (i, j) at x0 + (i h, j h), id
j*nx + i, volume h^2 halved on each boundary edge (quartered at corners).  A
seeded SplitMix64 counter hash perturbs every coordinate by up to jitter*h.
"""
from __future__ import annotations

import dataclasses
from dataclasses import dataclass, field

import numpy as np

# Workloads: plate side in nodes, mesh size, horizon/h ratio, step, steps.
PLATE_2D = {
    # 100x100 nodes, h = 1e-3 m, horizon 3h, 1000 steps of 1e-7 s
    "plate2d_1e4": dict(nx=100, ny=100, h=1e-3, ratio=3, dt=1e-7, nsteps=1000),
    # 1000x1000 nodes, h = 1e-4 m, horizon 4h, 2000 steps of 1e-8 s
    "plate2d_1e6": dict(nx=1000, ny=1000, h=1e-4, ratio=4, dt=1e-8,
                        nsteps=2000),
}


@dataclass
class BC:
    """One Displacement_BC / Force_BC set (inp::BCData)."""
    region: str
    dirs: list
    time_fn: str
    time_params: list
    space_fn: str = "constant"
    space_params: list = field(default_factory=lambda: [1.0])
    x1: float = 0.0
    x2: float = 0.0
    y1: float = 0.0
    y2: float = 0.0
    z1: float = 0.0
    z2: float = 0.0
    r1: float = 0.0
    r2: float = 0.0
    theta: float = 0.0


@dataclass
class CrackSpec:
    """One Fracture: Cracks: Set (inp::EdgeCrack)."""
    o: int
    pb: tuple
    pt: tuple
    theta: float = 0.0
    activation_time: float = -1.0


@dataclass
class Problem:
    name: str
    dim: int
    x: np.ndarray          # (n, 3) float64 reference coordinates
    vol: np.ndarray        # (n,) float64
    horizon: float
    h: float
    dt: float
    nsteps: int
    material: dict         # MaterialDeck fields
    ubc: list = field(default_factory=list)
    fbc: list = field(default_factory=list)
    cracks: list = field(default_factory=list)
    interior: np.ndarray | None = None
    shape: tuple = ()
    v0: np.ndarray | None = None   # (n, 3) initial velocity, None: zero
    # node-list displacement BCs (nodes, dof masks, rates[k, 3])
    node_bc: tuple | None = None
    specimens: list = field(default_factory=list)  # ensemble members

    @property
    def n(self) -> int:
        return int(self.x.shape[0])


def splitmix64(counter: np.ndarray, seed: int) -> np.ndarray:
    """SplitMix64 output for each counter value (Steele, Lea, Flood 2014)."""
    with np.errstate(over="ignore"):
        z = (np.uint64(seed) + (counter.astype(np.uint64) + np.uint64(1))
             * np.uint64(0x9E3779B97F4A7C15))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(counter: np.ndarray, seed: int) -> np.ndarray:
    return (splitmix64(counter, seed) >> np.uint64(11)).astype(np.float64) * (
        1.0 / 9007199254740992.0)


def rnp_material(dim=2, E=3.24e9, Gc=500.0, density=1200.0, check_sc=1.0,
                 inf_type=0, inf_params=(1.0,), plane_strain=False,
                 irreversible=True, no_penetration=False) -> dict:
    return dict(dim=dim, E=E, K=-1.0, Gc=Gc, KIc=-1.0, density=density,
                plane_strain=int(plane_strain), compute_from_classical=1,
                bond_params=[], inf_type=inf_type,
                inf_params=list(inf_params), check_sc_factor=check_sc,
                irreversible=int(irreversible),
                no_penetration=int(no_penetration))


def grid2d(nx: int, ny: int, h: float, jitter: float = 1e-6, seed: int = 7,
           x0: float = 0.0, y0: float = 0.0):
    ii, jj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")
    ii = ii.ravel()
    jj = jj.ravel()
    n = nx * ny
    x = np.zeros((n, 3))
    x[:, 0] = x0 + ii.astype(np.float64) * h
    x[:, 1] = y0 + jj.astype(np.float64) * h
    if jitter:
        cnt = np.arange(2 * n, dtype=np.uint64)
        r = uniform01(cnt, seed).reshape(n, 2)
        x[:, :2] += (2.0 * r - 1.0) * (jitter * h)
    vol = np.full(n, h * h)
    vol[(ii == 0) | (ii == nx - 1)] *= 0.5
    vol[(jj == 0) | (jj == ny - 1)] *= 0.5
    return x, vol


def plate2d(nx: int, ny: int, h: float, ratio: int, dt: float, nsteps: int,
            notch_len: float = 0.02, pull: float = 0.5, layers: int = 3,
            jitter: float = 1e-6, seed: int = 7, name: str | None = None,
            **mat) -> Problem:
    """Square plate with a left-edge notch at mid-height and top/bottom layers
    pulled apart at +/- pull m/s (Displacement_BC, linear in time, dof 2).

    The notch is an Orientation 1 crack from (-h/2, ym) to (notch_len - h/2,
    ym), ym = (ny-1) h / 2 the mid-height between two node rows; the bonds
    crossing it are broken at t = 0.
    """
    x, vol = grid2d(nx, ny, h, jitter, seed)
    horizon = ratio * h
    xmax = (nx - 1) * h
    ytop = (ny - 1) * h
    ubc = []
    if layers > 0 and pull != 0.0:
        ubc.append(BC("rectangle", [2], "linear", [pull],
                      x1=-h, x2=xmax + h, y1=ytop - (layers - 1) * h - 0.5 * h,
                      y2=ytop + 0.5 * h))
        ubc.append(BC("rectangle", [2], "linear", [-pull],
                      x1=-h, x2=xmax + h, y1=-0.5 * h,
                      y2=(layers - 1) * h + 0.5 * h))
    cracks = []
    if notch_len > 0:
        ym = 0.5 * ytop
        cracks.append(CrackSpec(1, (-0.5 * h, ym), (notch_len - 0.5 * h, ym)))
    m = rnp_material(dim=2, **mat)
    return Problem(name=name or f"plate2d_{nx}x{ny}", dim=2, x=x, vol=vol,
                   horizon=horizon, h=h, dt=dt, nsteps=nsteps, material=m,
                   ubc=ubc, cracks=cracks, shape=(nx, ny))


def es_material(dim=3, K=2.16e9, G=1.30e9, E=-1.0, density=1200.0,
                plane_strain=False, sc=1e-3) -> dict:
    """ElasticState deck (Material: Type: ElasticState, K, G; E in 1D) plus
    the critical stretch sc of the explicit model's bond breaking (sc <= 0:
    no breaking)."""
    return dict(model="elastic_state", dim=dim, K=K, G=G, E=E,
                density=density, plane_strain=int(plane_strain), sc=sc)


def gaussian(counter: np.ndarray, seed: int) -> np.ndarray:
    """Standard normal variates from two SplitMix64 streams (Box-Muller)."""
    u1 = uniform01(counter, seed ^ 0x5DEECE66D)
    u2 = uniform01(counter, seed ^ 0x2545F4914F6CDD1D)
    return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)


def grid3d(nx: int, ny: int, nz: int, h: float, jitter: float = 1e-6,
           seed: int = 7):
    """Node (i, j, k) at (i h, j h, k h), id (k ny + j) nx + i, volume h^3
    halved on each boundary face it lies on (the 3D analogue of the mesh
    tool's 2D rule, fe2D.cpp:148-158)."""
    kk, jj, ii = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx),
                             indexing="ij")
    ii, jj, kk = ii.ravel(), jj.ravel(), kk.ravel()
    n = nx * ny * nz
    x = np.empty((n, 3))
    x[:, 0] = ii.astype(np.float64) * h
    x[:, 1] = jj.astype(np.float64) * h
    x[:, 2] = kk.astype(np.float64) * h
    if jitter:
        r = uniform01(np.arange(3 * n, dtype=np.uint64), seed).reshape(n, 3)
        x += (2.0 * r - 1.0) * (jitter * h)
    vol = np.full(n, h * h * h)
    vol[(ii == 0) | (ii == nx - 1)] *= 0.5
    vol[(jj == 0) | (jj == ny - 1)] *= 0.5
    vol[(kk == 0) | (kk == nz - 1)] *= 0.5
    return x, vol


def cube3d(nx: int, h: float, ratio: int, dt: float, nsteps: int,
           pull: float = 1.0, layers: int = 3, sigma_v: float = 1e-3,
           jitter: float = 1e-6, seed: int = 7, name: str | None = None,
           ny: int | None = None, nz: int | None = None, **mat) -> Problem:
    """Block of linear state-based material (ElasticState) with its z faces
    pulled apart at +/- pull m/s (layers node planes each, dof 3, linear in
    time) and a seeded Gaussian initial velocity of standard deviation
    sigma_v in every component."""
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    x, vol = grid3d(nx, ny, nz, h, jitter, seed)
    n = x.shape[0]
    xmax, ymax, zmax = (nx - 1) * h, (ny - 1) * h, (nz - 1) * h
    ubc = []
    if layers > 0 and pull != 0.0:
        box = dict(x1=-h, x2=xmax + h, y1=-h, y2=ymax + h)
        ubc.append(BC("cuboid", [3], "linear", [pull],
                      z1=zmax - (layers - 1) * h - 0.5 * h, z2=zmax + 0.5 * h,
                      **box))
        ubc.append(BC("cuboid", [3], "linear", [-pull], z1=-0.5 * h,
                      z2=(layers - 1) * h + 0.5 * h, **box))
    v0 = None
    if sigma_v:
        v0 = sigma_v * gaussian(np.arange(3 * n, dtype=np.uint64),
                                seed + 1).reshape(n, 3)
    m = es_material(dim=3, **mat)
    return Problem(name=name or f"cube3d_{nx}x{ny}x{nz}", dim=3, x=x, vol=vol,
                   horizon=ratio * h, h=h, dt=dt, nsteps=nsteps, material=m,
                   ubc=ubc, shape=(nx, ny, nz), v0=v0)


def qs_plate(nx: int, h: float = 1e-3, ratio: int = 3, pull: float = 1e-5,
             layers: int = 3, nsteps: int = 20, jitter: float = 1e-6,
             seed: int = 7, notch_len: float = 0.0, notch_angle: float = 0.0,
             name: str | None = None, **mat) -> Problem:
    """Quasi-static specimen: square plate of nx x nx nodes, the bottom
    `layers` node rows clamped (both dofs), the top rows held at u_x = 0 and
    displaced by u_y = pull * t (one load step = pull with dt = 1); optional
    pre-notch from the left edge at mid-height (angle in radians)."""
    x, vol = grid2d(nx, nx, h, jitter, seed)
    xmax = ytop = (nx - 1) * h
    ylo = ytop - (layers - 1) * h - 0.5 * h
    ubc = [BC("rectangle", [1, 2], "constant", [0.0], x1=-h, x2=xmax + h,
              y1=-0.5 * h, y2=(layers - 1) * h + 0.5 * h),
           BC("rectangle", [1], "constant", [0.0], x1=-h, x2=xmax + h,
              y1=ylo, y2=ytop + 0.5 * h),
           BC("rectangle", [2], "linear", [pull], x1=-h, x2=xmax + h,
              y1=ylo, y2=ytop + 0.5 * h)]
    cracks = []
    if notch_len > 0:
        ym = 0.5 * ytop
        if notch_angle == 0.0:
            cracks.append(CrackSpec(1, (-0.5 * h, ym),
                                    (notch_len - 0.5 * h, ym)))
        else:
            ca, sa = np.cos(notch_angle), np.sin(notch_angle)
            pb = (-0.5 * h, ym)
            pt = (pb[0] + notch_len * ca, pb[1] + notch_len * sa)
            cracks.append(CrackSpec(0, pb, pt, theta=notch_angle))
    m = rnp_material(dim=2, **mat)
    return Problem(name=name or f"qs_plate{nx}", dim=2, x=x, vol=vol,
                   horizon=ratio * h, h=h, dt=1.0, nsteps=nsteps, material=m,
                   ubc=ubc, cracks=cracks, shape=(nx, nx))


def ensemble(count: int = 256, lo: int = 50, hi: int = 400, h: float = 1e-3,
             ratio: int = 3, dt: float = 1e-7, nsteps: int = 1000,
             pull: float = 0.5, layers: int = 3, seed: int = 3,
             as_sets: bool = False, name: str | None = None) -> Problem:
    """Independent square RNP plates of uneven sides (U{lo..hi} nodes), each
    with an angled pre-notch from its left edge at mid-height (length
    U[0.1, 0.3] L, angle U[-30, 30] degrees) and its top / bottom `layers`
    rows pulled at +/- pull m/s (dof 2), packed into one specimen: the plates
    sit on a square grid of cells (hi + 8) h wide, far apart compared to the
    horizon, so no bond joins two plates.  The pulls are node-list BCs
    (as_sets: one rectangle BC set per pulled layer, at most 7 plates)."""
    rng = np.random.default_rng(seed)
    sides = rng.integers(lo, hi + 1, size=count)
    lens = rng.uniform(0.1, 0.3, size=count)
    angles = np.radians(rng.uniform(-30.0, 30.0, size=count))
    per_row = int(np.ceil(np.sqrt(count)))
    cell = (hi + 8) * h
    xs, vols, cracks, ubc, specs = [], [], [], [], []
    bc_nodes, bc_rates = [], []
    base = 0
    for k in range(count):
        L = int(sides[k])
        ox, oy = (k % per_row) * cell, (k // per_row) * cell
        x, vol = grid2d(L, L, h, 1e-6, seed * 100003 + k, x0=ox, y0=oy)
        ytop = oy + (L - 1) * h
        ym = oy + 0.5 * (L - 1) * h
        pb = (ox - 0.5 * h, ym)
        ln = lens[k] * (L - 1) * h
        pt = (pb[0] + ln * np.cos(angles[k]), pb[1] + ln * np.sin(angles[k]))
        cracks.append(CrackSpec(0, pb, pt, theta=float(angles[k])))
        jj = np.arange(L * L) // L
        top = np.nonzero(jj >= L - layers)[0]
        bot = np.nonzero(jj < layers)[0]
        if as_sets:
            xmax = ox + (L - 1) * h
            ubc.append(BC("rectangle", [2], "linear", [pull], x1=ox - h,
                          x2=xmax + h, y1=ytop - (layers - 1) * h - 0.5 * h,
                          y2=ytop + 0.5 * h))
            ubc.append(BC("rectangle", [2], "linear", [-pull], x1=ox - h,
                          x2=xmax + h, y1=oy - 0.5 * h,
                          y2=oy + (layers - 1) * h + 0.5 * h))
        bc_nodes += [base + top, base + bot]
        bc_rates += [np.full(top.size, pull), np.full(bot.size, -pull)]
        specs.append(dict(offset=base, side=L, origin=(ox, oy),
                          notch=(float(ln), float(angles[k]))))
        xs.append(x)
        vols.append(vol)
        base += L * L
    x = np.concatenate(xs)
    vol = np.concatenate(vols)
    node_bc = None
    if not as_sets:
        nodes = np.concatenate(bc_nodes).astype(np.int64)
        rates = np.zeros((nodes.size, 3))
        rates[:, 1] = np.concatenate(bc_rates)
        node_bc = (nodes, np.full(nodes.size, 2, np.uint8), rates)
    return Problem(name=name or f"ensemble{count}", dim=2, x=x, vol=vol,
                   horizon=ratio * h, h=h, dt=dt, nsteps=nsteps,
                   material=rnp_material(dim=2), ubc=ubc, cracks=cracks,
                   node_bc=node_bc, specimens=specs)


def ensemble_member(p: Problem, k: int) -> Problem:
    """Plate k of an ensemble as a specimen of its own (same nodes, notch and
    pulls)."""
    sp = p.specimens[k]
    a, L = sp["offset"], sp["side"]
    sl = slice(a, a + L * L)
    nb = None
    if p.node_bc is not None:
        nodes, dofs, rates = p.node_bc
        m = (nodes >= a) & (nodes < a + L * L)
        nb = (nodes[m] - a, dofs[m], rates[m])
    return dataclasses.replace(p, name=f"{p.name}_m{k}", x=p.x[sl].copy(),
                               vol=p.vol[sl].copy(), cracks=[p.cracks[k]],
                               ubc=[], node_bc=nb, specimens=[sp])


# Quasi-static workload: 300 x 300 nodes, h = 1e-3 m, horizon 3h, 20 load
# steps of 1e-5 m on the top rows, bottom rows clamped.
QS_2D = {"qs_plate300": dict(nx=300, h=1e-3, ratio=3, pull=1e-5, nsteps=20)}


# 3D ElasticState block: side in nodes, mesh size, horizon/h, step, steps.
CUBE_3D = {
    # 160^3 nodes, h = 6.25e-4 m, horizon 3h, 500 steps
    "cube3d_4e6": dict(nx=160, h=6.25e-4, ratio=3, dt=5e-8, nsteps=500),
}


def workload(key: str, **over) -> Problem:
    if key == "ensemble256":
        return ensemble(name=key, **over)
    if key in QS_2D:
        spec = dict(QS_2D[key])
        spec.update(over)
        return qs_plate(name=key, **spec)
    if key in CUBE_3D:
        spec = dict(CUBE_3D[key])
        spec.update(over)
        return cube3d(name=key, **spec)
    spec = dict(PLATE_2D[key])
    spec.update(over)
    return plate2d(name=key, **spec)


def with_(p: Problem, **kw) -> Problem:
    return dataclasses.replace(p, **kw)
