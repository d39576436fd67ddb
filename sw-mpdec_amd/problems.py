"""Driver-side set-up of sw-mpdec problems: mesh, quadrature, seeded initial
states and bathymetry-gradient tables.

This is the caller's side of the drop-in boundary (the reference keeps it in
its Fortran driver: src/mesh.f90 BuildMesh, src/finitevolume.f90
FillInitialConditions / InitializeWBVariables, src/equation.f90 ExactFunction
and Bathymetry_X/Y).  The library only ever receives the arrays built here.

Layouts (all float64):
  U        (ny, nx, 3) C order == Fortran U(1:3, 1:nx, 1:ny)
  dbdx     (16, ny, nx): plane yg*4+xg holds d b/dx at MeshGP(:, i, j, xg, yg)
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

GRAVITY = 9.81


def gauss4():
    """Nodes and weights of the 4-point Gauss rule on [-1/2, 1/2], evaluated
    with the reference's operation order (src/mesh.f90:127-129)."""
    s65 = math.sqrt(6. / 5.)
    w = [(18. - math.sqrt(30.)) / 72., (18. + math.sqrt(30.)) / 72.,
         (18. + math.sqrt(30.)) / 72., (18. - math.sqrt(30.)) / 72.]
    nodes = [-0.5 * math.sqrt(3. / 7. + 2. / 7. * s65), -0.5 * math.sqrt(3. / 7. - 2. / 7. * s65),
             0.5 * math.sqrt(3. / 7. - 2. / 7. * s65), 0.5 * math.sqrt(3. / 7. + 2. / 7. * s65)]
    return np.array(nodes), np.array(w)


@dataclass
class Mesh:
    """Uniform periodic mesh on [x0, x1]^2, as BuildMesh (src/mesh.f90:28-146)."""
    n: int
    x0: float = 0.0
    x1: float = 1.0
    dx: float = field(init=False)
    bary: np.ndarray = field(init=False)   # (n,) cell centres along one axis
    gp: np.ndarray = field(init=False)     # (n, 4) Gauss points along one axis
    wbnd: np.ndarray = field(init=False)   # WeightsGPBnd(1:4)

    def __post_init__(self):
        sx = abs(self.x1 - self.x0)
        self.dx = abs(sx) / float(self.n)
        nodes, w = gauss4()
        self.wbnd = w
        idx = np.arange(self.n, dtype=np.float64)
        # MeshNodes = X0 + REAL(ii)*DX ; MeshBary = MeshNodes(ii-1) + 0.5*DX
        mesh_nodes = self.x0 + idx * self.dx
        self.bary = mesh_nodes + 0.5 * self.dx
        # MeshGP = MeshBary + quadNodes1D(iGP)*DX
        self.gp = self.bary[:, None] + nodes[None, :] * self.dx


def cell_average(mesh: Mesh, f, mesh_y: Mesh | None = None) -> np.ndarray:
    """16-point Gauss cell averages of f(x, y) -> (ny, nx, 3), accumulated in
    the order of FillInitialConditions (iGP outer, jGP inner).  mesh is the x
    axis, mesh_y the y axis (default: the same)."""
    my = mesh_y or mesh
    out = None
    for ig in range(4):
        for jg in range(4):
            # x varies along axis 1 (i), y along axis 0 (j); f broadcasts
            X = mesh.gp[None, :, ig]
            Y = my.gp[:, jg][:, None]
            val = np.broadcast_to(f(X, Y), (my.n, mesh.n, 3))
            w = mesh.wbnd[ig] * mesh.wbnd[jg]
            out = w * val if out is None else out + w * val
    return out


def _periodic_gauss_1d(x, c, sig, L=1.0, images=(-1, 0, 1)):
    """sum_a exp(-(x - c - a L)^2 / sig^2) and its x-derivative."""
    g = np.zeros_like(x)
    dg = np.zeros_like(x)
    for a in images:
        d = x - c - a * L
        e = np.exp(-(d * d) / (sig * sig))
        g += e
        dg += e * (-2.0 * d / (sig * sig))
    return g, dg


@dataclass
class GaussBumps:
    """b(x, y) = sum_k A_k exp(-|x - c_k|^2 / s_k^2) summed over the 3x3
    periodic images (separable, so cheap on any grid)."""
    A: np.ndarray
    cx: np.ndarray
    cy: np.ndarray
    s: np.ndarray

    @staticmethod
    def seeded(rng: np.random.Generator, k: int = 8, max_b: float = 1.0):
        while True:
            A = rng.uniform(0.1, 0.4, k)
            s = rng.uniform(0.05, 0.2, k)
            cx = rng.uniform(0.0, 1.0, k)
            cy = rng.uniform(0.0, 1.0, k)
            gb = GaussBumps(A, cx, cy, s)
            # reject bathymetry reaching the free surface (would dry cells)
            xs = np.linspace(0.0, 1.0, 257)
            if np.max(gb.value(xs[None, :], xs[:, None])) < max_b:
                return gb

    def _terms(self, x, y):
        for k in range(len(self.A)):
            gx, dgx = _periodic_gauss_1d(x, self.cx[k], self.s[k])
            gy, dgy = _periodic_gauss_1d(y, self.cy[k], self.s[k])
            yield self.A[k], gx, dgx, gy, dgy

    def value(self, x, y):
        return sum(A * gx * gy for A, gx, _, gy, _ in self._terms(x, y))

    def grad(self, x, y):
        bx = sum(A * dgx * gy for A, _, dgx, gy, _ in self._terms(x, y))
        by = sum(A * gx * dgy for A, gx, _, gy, dgy in self._terms(x, y))
        return bx, by


def bathy_tables(mesh: Mesh, grad, mesh_y: Mesh | None = None) -> tuple[np.ndarray, np.ndarray]:
    """d b/dx, d b/dy at the 4x4 Gauss points of every cell, planes yg*4+xg,
    (16, ny, nx).  grad(x, y) -> (bx, by) on broadcast coordinate arrays."""
    my = mesh_y or mesh
    bx = np.empty((16, my.n, mesh.n))
    by = np.empty((16, my.n, mesh.n))
    for yg in range(4):
        for xg in range(4):
            X = mesh.gp[None, :, xg]
            Y = my.gp[:, yg][:, None]
            gx, gy = grad(X, Y)
            bx[yg * 4 + xg] = np.broadcast_to(gx, (my.n, mesh.n))
            by[yg * 4 + xg] = np.broadcast_to(gy, (my.n, mesh.n))
    return bx, by


def bathy_tables_ref_flag1(mesh: Mesh, mesh_y: Mesh | None = None) -> tuple[np.ndarray, np.ndarray]:
    """Tables of the reference's own BathymetryFlag=1 (src/equation.f90:461-494),
    0.1*2*PI*COS(2PI x)*COS(2PI y) and -0.1*2*PI*SIN(2PI x)*SIN(2PI y), with
    libm cos/sin and the reference's operation order, point by point."""
    my = mesh_y or mesh
    pi = math.acos(-1.0)
    bx = np.empty((16, my.n, mesh.n))
    by = np.empty((16, my.n, mesh.n))
    for yg in range(4):
        for xg in range(4):
            for j in range(my.n):
                y = float(my.gp[j, yg])
                cy, sy = math.cos(2. * pi * y), math.sin(2. * pi * y)
                for i in range(mesh.n):
                    x = float(mesh.gp[i, xg])
                    bx[yg * 4 + xg, j, i] = 0.1 * 2. * pi * math.cos(2. * pi * x) * cy
                    by[yg * 4 + xg, j, i] = -(0.1 * 2. * pi * math.sin(2. * pi * x) * sy)
    return bx, by


def cons(h, u, v):
    h, u, v = np.broadcast_arrays(h, u, v)
    U = np.empty(h.shape + (3,))
    U[..., 0] = h
    U[..., 1] = h * u
    U[..., 2] = h * v
    return U


# ---- problem families -----------------------------------------------------

@dataclass
class Problem:
    """One solve: initial state, optional bathymetry and well-balanced state,
    on nx = n by ny = rows cells of the periodic domain [x0, x1] x [y0, y1]
    (y0, y1 default to x0, x1)."""
    n: int
    U0: np.ndarray                  # (rows, n, 3)
    wellbalanced: bool = False
    Uwb: np.ndarray | None = None   # (rows, n, 3)
    dbdx: np.ndarray | None = None  # (16, rows, n)
    dbdy: np.ndarray | None = None
    gravity: float = GRAVITY
    cfl: float = 0.5
    steps: int = 1
    kind: str = ""
    x0: float = 0.0
    x1: float = 1.0
    ny: int = 0                     # 0: square (ny = n)
    y0: float | None = None         # y extent (None: that of x)
    y1: float | None = None

    @property
    def rows(self) -> int:
        return self.ny or self.n

    @property
    def mesh(self) -> Mesh:
        """x axis"""
        return Mesh(self.n, self.x0, self.x1)

    @property
    def mesh_y(self) -> Mesh:
        y0 = self.x0 if self.y0 is None else self.y0
        y1 = self.x1 if self.y1 is None else self.y1
        return Mesh(self.rows, y0, y1)


# The reference's own setups per InitialCondition (src/parameters.f90:62-125):
# MESH_X0, MESH_X1, Gravity and BathymetryFlag.
REFERENCE_SETUPS = {
    1: ((0.0, 0.0), (3.0, 3.0), 9.81, 0),     # unsteady smooth vortex
    2: ((0.0, 0.0), (1.0, 1.0), 9.81, 1),     # lake at rest
    3: ((-5.0, -2.0), (5.0, 2.0), 9.8, 2),    # perturbation of a wet-dry lake at rest
    4: ((0.0, 0.0), (40.0, 40.0), 9.8, 0),    # circular dam break 1
    5: ((0.0, 0.0), (50.0, 50.0), 9.8, 0),    # circular dam break 2
    6: ((-5.0, -2.0), (5.0, 2.0), 9.8, 2),    # wave over a dry island
}


@dataclass
class ReferenceSetup:
    """One of the reference's initial conditions on nx x ny cells of its own
    domain, with the data to be made on the device (swm_fill_initial_data)
    rather than by the host.  wellbalanced: the WELLBALANCED variant
    (initial conditions 2 and 3)."""
    ic: int
    nx: int
    ny: int = 0
    wellbalanced: bool = False
    cfl: float = 0.5
    steps: int = 1

    def __post_init__(self):
        self.ny = self.ny or self.nx
        self.x0, self.x1, self.gravity, self.bathy_flag = REFERENCE_SETUPS[self.ic]
        self.mesh = Mesh(self.nx, self.x0[0], self.x1[0])
        self.mesh_y = Mesh(self.ny, self.x0[1], self.x1[1])

    @property
    def n(self) -> int:
        return self.nx

    @property
    def rows(self) -> int:
        return self.ny


def vortex(n: int, seed: int, steps: int = 50, cfl: float = 0.5, ny: int = 0) -> Problem:
    """Smooth vortex on the periodic unit square, flat bottom:
    h = 1 - eps^2/(2g) exp(1 - r^2), (u, v) = eps exp((1 - r^2)/2) (-y', x'),
    x' = (x - 1/2)/0.1, eps ~ U[0.5, 1.5] from the seed; n x ny cells."""
    rng = np.random.default_rng(seed)
    eps = rng.uniform(0.5, 1.5)
    mesh = Mesh(n)
    mesh_y = Mesh(ny or n)
    g = GRAVITY

    def h_u_v(X, Y):
        xp = (X - 0.5) / 0.1
        yp = (Y - 0.5) / 0.1
        r2 = xp * xp + yp * yp
        h = 1.0 - eps * eps / (2.0 * g) * np.exp(1.0 - r2)
        s = eps * np.exp(0.5 * (1.0 - r2))
        return h, -s * yp, s * xp

    U0 = cell_average(mesh, lambda X, Y: cons(*h_u_v(X, Y)), mesh_y)
    return Problem(n, U0, cfl=cfl, steps=steps, kind="vortex", ny=ny)


def lake_bumps(n: int, seed: int, steps: int = 100, cfl: float = 0.5,
               amp: float = 1e-3) -> Problem:
    """Lake at rest over 8 seeded Gaussian bumps plus a seeded Gaussian
    surface perturbation of amplitude amp, well-balanced variant."""
    rng = np.random.default_rng(seed)
    gb = GaussBumps.seeded(rng)
    pcx, pcy = rng.uniform(0.0, 1.0, 2)
    ps = rng.uniform(0.05, 0.2)
    mesh = Mesh(n)

    def wb_state(X, Y):
        b = gb.value(X, Y)
        z = np.zeros_like(X)
        return cons(1.0 - b, z, z)

    def ic_state(X, Y):
        b = gb.value(X, Y)
        gx, _ = _periodic_gauss_1d(X, pcx, ps)
        gy, _ = _periodic_gauss_1d(Y, pcy, ps)
        z = np.zeros_like(X)
        return cons(1.0 - b + amp * gx * gy, z, z)

    Uwb = cell_average(mesh, wb_state)
    U0 = cell_average(mesh, ic_state)
    bx, by = bathy_tables(mesh, gb.grad)
    return Problem(n, U0, wellbalanced=True, Uwb=Uwb, dbdx=bx, dbdy=by, cfl=cfl,
                   steps=steps, kind="lake_bumps")


def dam_break(n: int, steps: int = 200, cfl: float = 0.5, radius: float = 0.25,
              dry: float = 0.0) -> Problem:
    """Circular dam break on the periodic unit square, flat bottom:
    h = 1 inside radius 0.25 of the centre, h = dry outside, u = v = 0."""
    mesh = Mesh(n)

    def st(X, Y):
        r2 = (X - 0.5) ** 2 + (Y - 0.5) ** 2
        h = np.where(r2 < radius * radius, 1.0, dry)
        z = np.zeros_like(h)
        return cons(h, z, z)

    U0 = cell_average(mesh, st)
    return Problem(n, U0, cfl=cfl, steps=steps, kind="dam_break")


def lake_flag1(n: int, wellbalanced: bool = True, pert: float = 1e-3, ny: int = 0) -> Problem:
    """The reference's own lake at rest over BathymetryFlag=1 bathymetry,
    b = 0.1 sin(2 pi x) cos(2 pi y) (src/equation.f90:135-140, :429-430), plus
    a Gaussian surface perturbation of amplitude pert; n x ny cells."""
    m = Mesh(n)
    my = Mesh(ny or n)
    pi = math.acos(-1.0)

    def st(X, Y, base):
        b = 0.1 * np.sin(2 * pi * X) * np.cos(2 * pi * Y)
        h = 1.0 - b
        if not base:
            h = h + pert * np.exp(-((X - 0.3) ** 2 + (Y - 0.6) ** 2) / 0.01)
        z = np.zeros_like(h)
        return cons(h, z, z)

    bx, by = bathy_tables_ref_flag1(m, my)
    return Problem(n, cell_average(m, lambda X, Y: st(X, Y, False), my), wellbalanced=wellbalanced,
                   Uwb=cell_average(m, lambda X, Y: st(X, Y, True), my) if wellbalanced else None,
                   dbdx=bx, dbdy=by, kind="lake_flag1", ny=ny)


def mixed_batch(count: int, seed: int, nmin: int = 32, nmax: int = 512,
                smin: int = 20, smax: int = 200) -> list[Problem]:
    """Seeded mix of uneven solves: N = round(exp(U[ln nmin, ln nmax])),
    steps ~ U{smin..smax}, 50/50 vortex and lake-at-rest-over-bumps."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
        steps = int(rng.integers(smin, smax + 1))
        s = int(rng.integers(0, 2**31 - 1))
        if rng.uniform() < 0.5:
            out.append(vortex(n, s, steps=steps))
        else:
            out.append(lake_bumps(n, s, steps=steps))
    return out
