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


# ---- the exponential of the benchmark families -----------------------------
#
# The families below (vortex, lake over Gaussian bumps) are defined with this
# exponential rather than the platform's: libm and NumPy's vectorised exp
# differ between hosts and CPU instruction sets in the last bit, and the
# WENO5 weights turn last-bit differences of the data into 1e-9 relative
# differences within a few steps.  det_exp uses only IEEE +, -, *, exact
# scaling by powers of two and rounding by the 1.5*2^52 shift, in one fixed
# order, so NumPy here, the device kernel (k_family_data) and any C build
# without contraction give the same bits.  Method: Cody-Waite reduction
# x = k ln2 + r, |r| <= ln2/2 (ln2 split into a 32-bit head, so k*LN2_HI is
# exact), the rounding residual of r carried separately, and the Taylor
# polynomial of degree 13 of exp(r) - 1 - r in Horner form; measured error
# below 0.85 ulp against a 60-digit exp.
EXP_INV_LN2 = float.fromhex("0x1.71547652b82fep+0")
EXP_LN2_HI = float.fromhex("0x1.62e42fee00000p-1")
EXP_LN2_LO = float.fromhex("0x1.a39ef35793c76p-33")
EXP_SHIFT = float.fromhex("0x1.8p52")
EXP_COEF = tuple(float.fromhex(h) for h in (        # 1/k!, k = 2..13
    "0x1.0000000000000p-1", "0x1.5555555555555p-3", "0x1.5555555555555p-5",
    "0x1.1111111111111p-7", "0x1.6c16c16c16c17p-10", "0x1.a01a01a01a01ap-13",
    "0x1.a01a01a01a01ap-16", "0x1.71de3a556c734p-19", "0x1.27e4fb7789f5cp-22",
    "0x1.ae64567f544e4p-26", "0x1.1eed8eff8d898p-29", "0x1.6124613a86d09p-33"))
EXP_MAX = 709.8      # above: +inf
EXP_MIN = -746.0     # below: +0


def det_exp(x):
    """exp(x) by the fixed method above (float64 scalar or array)."""
    x = np.asarray(x, dtype=np.float64)
    xc = np.clip(np.where(np.isnan(x), 0.0, x), EXP_MIN, EXP_MAX)
    t = xc * EXP_INV_LN2 + EXP_SHIFT
    k = t - EXP_SHIFT
    hi = xc - k * EXP_LN2_HI
    lo = k * EXP_LN2_LO
    r = hi - lo
    dr = (hi - r) - lo
    q = EXP_COEF[-1]
    for c in EXP_COEF[-2::-1]:
        q = c + r * q
    p = 1.0 + (r + (dr + (r * r) * q))
    ki = k.astype(np.int64)
    # p * 2^k rounded once: exact scalings by powers of two, the last one
    # rounding into the subnormal range or overflowing
    big, small = ki > 1000, ki < -1000
    k1 = np.where(big, ki - 100, np.where(small, ki + 100, ki))
    y = p * np.ldexp(1.0, k1)
    with np.errstate(over="ignore"):
        y = np.where(big, y * np.ldexp(1.0, 100), np.where(small, y * np.ldexp(1.0, -100), y))
    y = np.where(x > EXP_MAX, np.inf, np.where(x < EXP_MIN, 0.0, y))
    y = np.where(np.isnan(x), x, y)
    return y if y.ndim else float(y)


def _periodic_gauss_1d(x, c, sig, L=1.0, images=(-1, 0, 1)):
    """sum_a exp(-(x - c - a L)^2 / sig^2) and its x-derivative, summed over
    the images a = -1, 0, 1 in that order."""
    g = np.zeros_like(x)
    dg = np.zeros_like(x)
    for a in images:
        d = x - c - a * L
        e = det_exp(-(d * d) / (sig * sig))
        g = g + e
        dg = dg + e * (-2.0 * d / (sig * sig))
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
    def seeded(rng: np.random.Generator, k: int = 8, max_b: float = 0.95):
        """Draws until the bathymetry stays below max_b, below the free
        surface h + b = 1 with a margin.  The maximum is taken on a 1025^2
        grid (spacing 1/1024); a bump's curvature is at most 2 A / s^2 <= 320,
        so the true maximum exceeds the sampled one by less than
        320 (1/1024)^2 / 8 < 4e-5, far inside the margin: every accepted draw
        keeps h >= 1 - max_b > 0 at every point."""
        xs = np.linspace(0.0, 1.0, 1025)
        while True:
            A = rng.uniform(0.1, 0.4, k)
            s = rng.uniform(0.05, 0.2, k)
            cx = rng.uniform(0.0, 1.0, k)
            cy = rng.uniform(0.0, 1.0, k)
            gb = GaussBumps(A, cx, cy, s)
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

# family kinds (include/swm.h SWM_FAMILY_*)
FAMILY_VORTEX = 1
FAMILY_LAKE_BUMPS = 2
FAMILY_DAM_BREAK = 3

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
    # the family's parameters, from which the device makes the same arrays
    # bitwise (swm_fill_family); None for data defined only by the arrays
    family: dict | None = None

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


@dataclass
class FamilySpec:
    """A problem of one of the benchmark families given by its parameters
    only.  The device makes its arrays from them (swm_fill_family,
    swm_job_desc.family), so a request carries no host data; build() makes
    the same arrays, bitwise, on the host.  kind: FAMILY_*; params: the
    family's parameters (the fields of swm_family); n x ny cells of the
    periodic unit square."""
    kind: int
    n: int
    params: dict
    ny: int = 0
    steps: int = 1
    cfl: float = 0.5
    gravity: float = GRAVITY

    @property
    def rows(self) -> int:
        return self.ny or self.n

    @property
    def wellbalanced(self) -> bool:
        return self.kind == FAMILY_LAKE_BUMPS

    @property
    def mesh(self) -> Mesh:
        return Mesh(self.n)

    @property
    def mesh_y(self) -> Mesh:
        return Mesh(self.rows)

    @property
    def name(self) -> str:
        return {FAMILY_VORTEX: "vortex", FAMILY_LAKE_BUMPS: "lake_bumps",
                FAMILY_DAM_BREAK: "dam_break"}[self.kind]

    def build(self) -> Problem:
        """The host arrays of this problem (cell averages at the 4x4 Gauss
        points; for the lake also the well-balanced state and the bathymetry
        tables)."""
        mesh, mesh_y, q, g = self.mesh, self.mesh_y, self.params, self.gravity
        common = dict(cfl=self.cfl, steps=self.steps, kind=self.name, ny=self.ny,
                      gravity=g, family={"kind": self.kind, **q})
        if self.kind == FAMILY_VORTEX:
            eps = q["eps"]

            def h_u_v(X, Y):
                xp = (X - 0.5) / 0.1
                yp = (Y - 0.5) / 0.1
                r2 = xp * xp + yp * yp
                h = 1.0 - eps * eps / (2.0 * g) * det_exp(1.0 - r2)
                s = eps * det_exp(0.5 * (1.0 - r2))
                return h, -s * yp, s * xp

            U0 = cell_average(mesh, lambda X, Y: cons(*h_u_v(X, Y)), mesh_y)
            return Problem(self.n, U0, **common)
        if self.kind == FAMILY_LAKE_BUMPS:
            gb = GaussBumps(np.array(q["A"]), np.array(q["cx"]), np.array(q["cy"]),
                            np.array(q["s"]))
            amp, pcx, pcy, ps = q["amp"], q["px"], q["py"], q["ps"]

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

            Uwb = cell_average(mesh, wb_state, mesh_y)
            U0 = cell_average(mesh, ic_state, mesh_y)
            assert U0[..., 0].min() > 0.0 and Uwb[..., 0].min() > 0.0
            bx, by = bathy_tables(mesh, gb.grad, mesh_y)
            return Problem(self.n, U0, wellbalanced=True, Uwb=Uwb, dbdx=bx, dbdy=by, **common)
        radius, h_in, h_out = q["radius"], q["h_in"], q["h_out"]

        def st(X, Y):
            r2 = (X - 0.5) ** 2 + (Y - 0.5) ** 2
            h = np.where(r2 < radius * radius, h_in, h_out)
            z = np.zeros_like(h)
            return cons(h, z, z)

        return Problem(self.n, cell_average(mesh, st, mesh_y), **common)


def vortex_spec(n: int, seed: int, steps: int = 50, cfl: float = 0.5, ny: int = 0) -> FamilySpec:
    """Smooth vortex on the periodic unit square, flat bottom:
    h = 1 - eps^2/(2g) exp(1 - r^2), (u, v) = eps exp((1 - r^2)/2) (-y', x'),
    x' = (x - 1/2)/0.1, eps ~ U[0.5, 1.5] from the seed; n x ny cells."""
    rng = np.random.default_rng(seed)
    eps = float(rng.uniform(0.5, 1.5))
    return FamilySpec(FAMILY_VORTEX, n, {"eps": eps}, ny=ny, steps=steps, cfl=cfl)


def lake_bumps_spec(n: int, seed: int, steps: int = 100, cfl: float = 0.5,
                    amp: float = 1e-3) -> FamilySpec:
    """Lake at rest over 8 seeded Gaussian bumps plus a seeded Gaussian
    surface perturbation of amplitude amp, well-balanced variant."""
    rng = np.random.default_rng(seed)
    gb = GaussBumps.seeded(rng)
    pcx, pcy = rng.uniform(0.0, 1.0, 2)
    ps = rng.uniform(0.05, 0.2)
    q = {"A": gb.A.tolist(), "cx": gb.cx.tolist(), "cy": gb.cy.tolist(), "s": gb.s.tolist(),
         "amp": float(amp), "px": float(pcx), "py": float(pcy), "ps": float(ps)}
    return FamilySpec(FAMILY_LAKE_BUMPS, n, q, steps=steps, cfl=cfl)


def dam_break_spec(n: int, steps: int = 200, cfl: float = 0.5, radius: float = 0.25,
                   dry: float = 0.0) -> FamilySpec:
    """Circular dam break on the periodic unit square, flat bottom:
    h = 1 inside radius 0.25 of the centre, h = dry outside, u = v = 0."""
    return FamilySpec(FAMILY_DAM_BREAK, n, {"radius": float(radius), "h_in": 1.0,
                                            "h_out": float(dry)}, steps=steps, cfl=cfl)


def vortex(n: int, seed: int, steps: int = 50, cfl: float = 0.5, ny: int = 0) -> Problem:
    """vortex_spec(...) with its host arrays."""
    return vortex_spec(n, seed, steps, cfl, ny).build()


def lake_bumps(n: int, seed: int, steps: int = 100, cfl: float = 0.5,
               amp: float = 1e-3) -> Problem:
    """lake_bumps_spec(...) with its host arrays."""
    return lake_bumps_spec(n, seed, steps, cfl, amp).build()


def dam_break(n: int, steps: int = 200, cfl: float = 0.5, radius: float = 0.25,
              dry: float = 0.0) -> Problem:
    """dam_break_spec(...) with its host arrays."""
    return dam_break_spec(n, steps, cfl, radius, dry).build()


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
                smin: int = 20, smax: int = 200) -> list[FamilySpec]:
    """Seeded mix of uneven solves: N = round(exp(U[ln nmin, ln nmax])),
    steps ~ U{smin..smax}, 50/50 vortex and lake-at-rest-over-bumps, as
    family parameters (the device makes the arrays; .build() makes them on
    the host)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
        steps = int(rng.integers(smin, smax + 1))
        s = int(rng.integers(0, 2**31 - 1))
        if rng.uniform() < 0.5:
            out.append(vortex_spec(n, s, steps=steps))
        else:
            out.append(lake_bumps_spec(n, s, steps=steps))
    return out
