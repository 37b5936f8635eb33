"""Synthetic Taylor-Hood meshes, boundary lines and flow parameters for the UVS2D solve.

A *deck* is the input of one steady Newton solve of the depth-averaged shallow-water
system: a P2/P1 (Taylor-Hood) triangle or serendipity-quad mesh, the control lines that
carry the inlet discharge (QINLET) and the outlet water level (OUTLET), per-element
roughness and eddy viscosity, and the initial state.  The same deck feeds the CPU oracle,
the GPU library and, in the oracle container, the reference program itself
(``write_reference_deck`` writes the reference's keyword files: region file
``Grid.cpp:137``, control file ``Grid.cpp:378``, initial file ``Grid.cpp:510``, project
file ``Project.cpp:1296``, time-step file ``Timeint.cpp:1466``, material file
``Type.cpp:130``).

Node and element conventions follow the reference (``Shape.cpp``/``Triangle.cpp``):
triangles list corners counter-clockwise then the midsides 01, 12, 20; quads list
corners then midsides 01, 12, 23, 30.  All indices here are 0-based; the reference's
files are 1-based.
"""
from __future__ import annotations

import dataclasses
import math
import os

import numpy as np

INLET_LINE = 1
OUTLET_LINE = 2


@dataclasses.dataclass
class Deck:
    x: np.ndarray            # [np] float64
    y: np.ndarray
    z: np.ndarray            # bed elevation (zor)
    elems: np.ndarray        # [ne, 8] int32, -1 padded
    nnd: np.ndarray          # [ne] int32 (6 triangle, 8 quad)
    manning: np.ndarray      # [ne] float64, Manning n of each element
    vt: np.ndarray           # [ne] float64, constant eddy viscosity of each element
    lines: dict              # line no -> int32 [nseg, 3] (corner, corner, midside)
    U0: np.ndarray           # initial state [np]
    V0: np.ndarray
    S0: np.ndarray
    Q_inlet: float = 100.0   # QINLET discharge (m3/s)
    S_inlet: float = 2.1     # QINLET reference water level used for the q distribution
    S_outlet: float = 2.0    # OUTLET water level
    name: str = "deck"
    veg_dp: np.ndarray | None = None   # [ne] stem diameter of emergent vegetation (m), 0: none
    veg_sp: np.ndarray | None = None   # [ne] stem spacing (m)
    turb: int = 1                      # turbulence model code ($TM_TURBULENCE): 1 constant vt,
                                       # 2 algebraic; +100 vt of the element is a lower bound
    turb_st: np.ndarray | None = None  # [ne] st of the algebraic model (None: 1.0)
    turb_kdest: np.ndarray | None = None   # [ne] KDest of the algebraic model (None: 0.6)
    turb_estx: np.ndarray | None = None    # [ne] estx / esty of the anisotropic (Elder) model,
    turb_esty: np.ndarray | None = None    # tens digit 1 of turb (None: 0)
    hydrograph: np.ndarray | None = None   # [k, 3] inlet hydrograph rows (t s, Q, S) (QTINLET)

    @property
    def n_nodes(self) -> int:
        return int(self.x.shape[0])

    @property
    def n_elems(self) -> int:
        return int(self.elems.shape[0])


# ---------------------------------------------------------------------------------------
# bathymetry noise

def gaussian_random_field(xs: np.ndarray, ys: np.ndarray, sigma: float, corr: float,
                          seed: int) -> np.ndarray:
    """Seeded stationary Gaussian field with covariance sigma^2 exp(-r^2 / (2 corr^2)).

    Spectral synthesis on a regular grid (xs along axis 1, ys along axis 0) that is
    zero-padded by 4 correlation lengths on every side so that the periodic wrap does not
    correlate opposite boundaries.  Returns an array of shape (len(ys), len(xs)).
    """
    if sigma == 0.0:
        return np.zeros((len(ys), len(xs)))
    dx = float(xs[1] - xs[0]) if len(xs) > 1 else 1.0
    dy = float(ys[1] - ys[0]) if len(ys) > 1 else 1.0
    pad_x = int(math.ceil(4.0 * corr / dx))
    pad_y = int(math.ceil(4.0 * corr / dy))
    nx = len(xs) + 2 * pad_x
    ny = len(ys) + 2 * pad_y
    rng = np.random.default_rng(seed)
    white = rng.standard_normal((ny, nx))
    kx = 2.0 * np.pi * np.fft.fftfreq(nx, d=dx)
    ky = 2.0 * np.pi * np.fft.fftfreq(ny, d=dy)
    k2 = ky[:, None] ** 2 + kx[None, :] ** 2
    # amplitude = sqrt(spectral density) of the Gaussian covariance, on the discrete grid
    amp = np.sqrt(2.0 * np.pi * corr * corr * np.exp(-0.5 * corr * corr * k2) / (dx * dy))
    field = np.real(np.fft.ifft2(np.fft.fft2(white) * amp))
    field = field[pad_y:pad_y + len(ys), pad_x:pad_x + len(xs)]
    std = field.std()
    if std > 0:
        field *= sigma / std
    return field


# ---------------------------------------------------------------------------------------
# structured channel (configs: straight channel 1000 m x 100 m, 200 x 20 cells)

def channel(nx: int = 200, ny: int = 20, length: float = 1000.0, width: float = 100.0,
            slope: float = 1.0e-4, z_out: float = 0.0, noise_sigma: float = 0.05,
            noise_corr: float = 20.0, seed: int = 0, manning: float = 0.03,
            vt: float = 0.1, Q: float = 100.0, S_out: float = 2.0, S_in: float = 2.1,
            U_init: float = 0.5, quads: bool = False, name: str | None = None,
            veg_frac: float = 0.0, veg_band: float = 0.3, veg_dp=(0.01, 0.05),
            veg_density=(0.5, 5.0), veg_seed: int = 0, turb: int = 1, turb_st=None,
            turb_kdest=None, turb_seed: int = 0, turb_estx=None, turb_esty=None) -> Deck:
    """Straight rectangular channel on a (2nx+1) x (2ny+1) lattice.

    Each cell becomes two six-node triangles split along its (i,j)->(i+2,j+2) diagonal,
    or, with ``quads``, one eight-node serendipity quad (cell centres are then unused and
    dropped).  Elements are listed column by column along x with the front across the
    width (the order that a front-width minimisation gives for a channel).  Bed:
    ``z = z_out + slope*(length - x) + GRF``; midside nodes carry the mean of their
    corners so the bed is piecewise linear.  Inlet line at x=0, outlet line at x=length.
    With ``veg_frac`` > 0, a seeded fraction of the elements in the outer ``veg_band`` of the
    width on either bank (the floodplain strips) carries emergent vegetation: stem diameter
    U[veg_dp] m and density U[veg_density] per m^2 (spacing = density^-1/2).
    ``turb`` selects the turbulence model code; ``turb_st`` / ``turb_kdest`` (and
    ``turb_estx`` / ``turb_esty`` of the anisotropic model) are a value or a (low, high)
    range sampled per element with ``turb_seed`` (None: model defaults).
    """
    NX, NY = 2 * nx + 1, 2 * ny + 1
    xs = np.linspace(0.0, length, NX)
    ys = np.linspace(0.0, width, NY)
    X, Y = np.meshgrid(xs, ys)                        # [NY, NX]
    Zc = z_out + slope * (length - X) + gaussian_random_field(xs, ys, noise_sigma,
                                                              noise_corr, seed)
    # piecewise-linear bed: odd lattice lines take the mean of their neighbours
    Z = Zc.copy()
    Z[:, 1::2] = 0.5 * (Zc[:, 0:-1:2] + Zc[:, 2::2])
    Z[1::2, :] = 0.5 * (Z[0:-1:2, :] + Z[2::2, :])
    if not quads:
        # the diagonal midside (cell centre) is the mean of the diagonal corners
        Z[1::2, 1::2] = 0.5 * (Zc[0:-1:2, 0:-1:2] + Zc[2::2, 2::2])

    def nid(i, j):
        return j * NX + i

    elems = []
    for i in range(nx):
        for j in range(ny):
            i0, j0 = 2 * i, 2 * j
            a, b = nid(i0, j0), nid(i0 + 2, j0)
            c, d = nid(i0 + 2, j0 + 2), nid(i0, j0 + 2)
            if quads:
                elems.append([a, b, c, d, nid(i0 + 1, j0), nid(i0 + 2, j0 + 1),
                              nid(i0 + 1, j0 + 2), nid(i0, j0 + 1)])
            else:
                m = nid(i0 + 1, j0 + 1)
                elems.append([a, b, c, nid(i0 + 1, j0), nid(i0 + 2, j0 + 1), m, -1, -1])
                elems.append([a, c, d, m, nid(i0 + 1, j0 + 2), nid(i0, j0 + 1), -1, -1])
    elems = np.asarray(elems, dtype=np.int64)
    x, y, z = X.ravel(), Y.ravel(), Z.ravel()
    inlet = [[nid(0, 2 * j), nid(0, 2 * j + 2), nid(0, 2 * j + 1)] for j in range(ny)]
    outlet = [[nid(2 * nx, 2 * j), nid(2 * nx, 2 * j + 2), nid(2 * nx, 2 * j + 1)]
              for j in range(ny)]
    lines = {INLET_LINE: np.asarray(inlet, np.int64), OUTLET_LINE: np.asarray(outlet, np.int64)}
    if quads:
        used = np.zeros(x.shape[0], bool)
        used[elems.ravel()] = True
        remap = -np.ones(x.shape[0], np.int64)
        remap[used] = np.arange(int(used.sum()))
        elems = remap[elems]
        lines = {k: remap[v] for k, v in lines.items()}
        x, y, z = x[used], y[used], z[used]
    nnd = np.full(elems.shape[0], 8 if quads else 6, np.int32)
    ne = elems.shape[0]
    npn = x.shape[0]
    veg_d = veg_s = None
    if veg_frac > 0.0:
        ncn = 4 if quads else 3
        yc = y[elems[:, :ncn]].mean(1)
        bank = np.abs(yc - 0.5 * width) > (1.0 - veg_band) * 0.5 * width
        rng = np.random.default_rng(veg_seed + 7)
        pick = bank & (rng.random(ne) < veg_frac)
        veg_d = np.where(pick, rng.uniform(veg_dp[0], veg_dp[1], ne), 0.0)
        dens = rng.uniform(veg_density[0], veg_density[1], ne)
        veg_s = np.where(pick, 1.0 / np.sqrt(dens), 0.0)
    rng_t = np.random.default_rng(turb_seed + 11)

    def per_elem(v):
        if v is None:
            return None
        if np.ndim(v) == 0:
            return np.full(ne, float(v))
        return rng_t.uniform(v[0], v[1], ne)

    t_st, t_kd = per_elem(turb_st), per_elem(turb_kdest)
    t_ex, t_ey = per_elem(turb_estx), per_elem(turb_esty)
    return Deck(x=x.astype(np.float64), y=y.astype(np.float64), z=z.astype(np.float64),
                elems=elems.astype(np.int32), nnd=nnd,
                manning=np.full(ne, manning), vt=np.full(ne, vt),
                lines={k: v.astype(np.int32) for k, v in lines.items()},
                U0=np.full(npn, U_init), V0=np.zeros(npn), S0=np.full(npn, S_out),
                Q_inlet=Q, S_inlet=S_in, S_outlet=S_out,
                name=name or f"channel_{nx}x{ny}{'_quad' if quads else ''}",
                veg_dp=veg_d, veg_sp=veg_s, turb=turb, turb_st=t_st, turb_kdest=t_kd,
                turb_estx=t_ex, turb_esty=t_ey)


# ---------------------------------------------------------------------------------------
# meandering channel with floodplains (GPU-scale workload)

def meander(n_along: int = 3200, n_across: int = 320, reach: float = 10000.0,
            amplitude: float = 300.0, wavelength: float = 2000.0, width: float = 900.0,
            main_width: float = 120.0, main_depth: float = 3.0, bank_slope: float = 0.2,
            slope: float = 2.0e-4, noise_sigma: float = 0.5, noise_corr: float = 50.0,
            jitter: float = 0.2, seed: int = 0, n_range=(0.025, 0.045), vt: float = 0.1,
            Q: float = 600.0, S_out: float = 4.2, U_init: float = 0.4,
            plain_rise: float = 0.0, name: str | None = None) -> Deck:
    """Meandering channel with floodplains on a curvilinear, jittered Taylor-Hood mesh.

    Centreline ``y_c = amplitude*sin(2*pi*s/wavelength)`` over ``reach`` metres.  The
    cross-section is a trapezoidal main channel (``main_width`` wide at the bed,
    ``main_depth`` deep, banks of slope ``bank_slope``) cut into floodplains that span
    ``width`` in total, rising by ``plain_rise`` m per m away from the main channel (with
    a rise, the outer floodplains lie above the initial water level and dry, and the initial
    velocity ``U_init`` is scaled by the depth over ``main_depth``, at most 1, so that the
    shallow fringe does not start at full speed).  The lattice
    follows the centreline and interior corner nodes
    are jittered by ``jitter`` of the local spacing, so element
    shapes and sizes vary (unstructured connectivity is not needed for the solver's cost
    model: row widths and the frontal ordering depend on the element adjacency, which is
    that of a triangulated lattice).  The diagonal of every cell alternates, giving
    vertices of degree 4, 6 and 8.  Bathymetry adds a seeded Gaussian random field;
    Manning n is drawn per element from ``U[n_range]``.  ``n_along*n_across*2`` triangles.
    Lateral offsets are vertical (a shear of the straight lattice), so the section
    measured normal to the centreline narrows by cos(slope angle) at the crossings.
    """
    nx, ny = n_along, n_across
    NX, NY = 2 * nx + 1, 2 * ny + 1
    rng = np.random.default_rng(seed + 1)
    s = np.linspace(0.0, reach, NX)
    t = np.linspace(-0.5 * width, 0.5 * width, NY)
    yc = amplitude * np.sin(2.0 * np.pi * s / wavelength)
    dyc = amplitude * 2.0 * np.pi / wavelength * np.cos(2.0 * np.pi * s / wavelength)
    nrm = np.sqrt(1.0 + dyc * dyc)
    tx, ty = 1.0 / nrm, dyc / nrm                  # unit tangent
    ox, oy = -ty, tx                               # unit normal (left)
    S_, T_ = np.meshgrid(np.arange(NX), np.arange(NY))   # lattice indices [NY, NX]
    # jitter interior corner nodes in lattice units (corners are even lattice indices)
    js = np.zeros((NY, NX))
    jt = np.zeros((NY, NX))
    corner = (S_ % 2 == 0) & (T_ % 2 == 0)
    interior = (S_ > 0) & (S_ < NX - 1) & (T_ > 0) & (T_ < NY - 1)
    sel = corner & interior
    js[sel] = rng.uniform(-jitter, jitter, int(sel.sum())) * 2.0
    jt[sel] = rng.uniform(-jitter, jitter, int(sel.sum())) * 2.0
    ds = s[1] - s[0]
    dt_ = t[1] - t[0]
    sc = S_ * ds + js * ds
    tc = (T_ - (NY - 1) / 2.0) * dt_ + jt * dt_
    # sheared mapping: lateral offsets are taken in y, so cells can never fold over at
    # the inner bends however tight the meander (a normal offset would invert cells
    # wherever |t| exceeds the bend radius)
    Xc = sc.copy()
    Yc = np.interp(sc, s, yc) + tc
    del tx, ty, ox, oy, dyc, nrm
    # bed: valley slope + trapezoidal section + GRF, evaluated at corner positions
    half_bed = 0.5 * main_width
    dist = np.abs(tc)
    bank = np.clip((dist - half_bed) * bank_slope, 0.0, main_depth)
    plain = plain_rise * np.clip(dist - half_bed - main_depth / bank_slope, 0.0, None)
    zsec = -main_depth + bank + plain
    grf = gaussian_random_field(np.arange(NX) * ds, np.arange(NY) * dt_, noise_sigma,
                                noise_corr, seed)
    Zc = slope * (reach - sc) + zsec + grf
    X = Xc.copy()
    Y = Yc.copy()
    Z = Zc.copy()

    # midside nodes are the straight-edge midpoints of their corners
    def mid(A, i0, j0, i1, j1):
        return 0.5 * (A[j0, i0] + A[j1, i1])

    # horizontal edge midsides (odd i, even j)
    for A in (X, Y, Z):
        A[0::2, 1::2] = 0.5 * (A[0::2, 0:-1:2] + A[0::2, 2::2])
        A[1::2, 0::2] = 0.5 * (A[0:-1:2, 0::2] + A[2::2, 0::2])
    # cell centres are diagonal midsides; the diagonal alternates per cell
    ci = np.arange(nx)
    cj = np.arange(ny)
    CI, CJ = np.meshgrid(ci, cj)                  # [ny, nx]
    flip = ((CI + CJ) % 2) == 1                   # diagonal (i,j+2)->(i+2,j) when flipped
    for A in (X, Y, Z):
        a = A[0:-1:2, 0:-1:2]
        b = A[0:-1:2, 2::2]
        c = A[2::2, 2::2]
        d = A[2::2, 0:-1:2]
        A[1::2, 1::2] = np.where(flip, 0.5 * (b + d), 0.5 * (a + c))
    del mid

    base = np.arange(NX * NY).reshape(NY, NX)
    i0 = 2 * CI.T.ravel()                          # column-major: along x outer
    j0 = 2 * CJ.T.ravel()
    fl = flip.T.ravel()
    a = base[j0, i0]
    b = base[j0, i0 + 2]
    c = base[j0 + 2, i0 + 2]
    d = base[j0 + 2, i0]
    mab = base[j0, i0 + 1]
    mbc = base[j0 + 1, i0 + 2]
    mcd = base[j0 + 2, i0 + 1]
    mda = base[j0 + 1, i0]
    m = base[j0 + 1, i0 + 1]
    neg = -np.ones_like(a)
    # unflipped: (a,b,c) + (a,c,d); flipped: (a,b,d) + (b,c,d)
    t1 = np.where(fl[:, None], np.stack([a, b, d, mab, m, mda, neg, neg], 1),
                  np.stack([a, b, c, mab, mbc, m, neg, neg], 1))
    t2 = np.where(fl[:, None], np.stack([b, c, d, mbc, mcd, m, neg, neg], 1),
                  np.stack([a, c, d, m, mcd, mda, neg, neg], 1))
    elems = np.empty((2 * t1.shape[0], 8), np.int64)
    elems[0::2] = t1
    elems[1::2] = t2
    ne = elems.shape[0]
    inlet = np.stack([base[0:-1:2, 0], base[2::2, 0], base[1::2, 0]], 1)
    outlet = np.stack([base[0:-1:2, -1], base[2::2, -1], base[1::2, -1]], 1)
    npn = NX * NY
    lo, hi = n_range
    man = np.random.default_rng(seed + 2).uniform(lo, hi, ne)
    U0 = np.full(npn, U_init)
    if plain_rise > 0.0:
        U0 *= np.clip((S_out - Z.ravel()) / main_depth, 0.0, 1.0)
    return Deck(x=X.ravel().astype(np.float64), y=Y.ravel().astype(np.float64),
                z=Z.ravel().astype(np.float64), elems=elems.astype(np.int32),
                nnd=np.full(ne, 6, np.int32), manning=man, vt=np.full(ne, vt),
                lines={INLET_LINE: inlet.astype(np.int32), OUTLET_LINE: outlet.astype(np.int32)},
                U0=U0, V0=np.zeros(npn), S0=np.full(npn, S_out),
                Q_inlet=Q, S_inlet=S_out + slope * reach, S_outlet=S_out,
                name=name or f"meander_{nx}x{ny}")


# ---------------------------------------------------------------------------------------
# reference-format writer (used only by the oracle harness in the oracle container)

def write_reference_deck(deck: Deck, outdir: str, solver: str = "1 6 1 0 60 1000 1.0e-10",
                         fgmres_solver: str = "2 7 1 0 60 30 1000 1.0e-10",
                         tm_solver: int = 1, maxiter: int = 12,
                         convergence: str = "1.0e-7 1.0e-7 1.0e-4 1.0e-4",
                         relax: str = "2 0.001 0.1 10.0 10.0 0.1",
                         dryrew: str = "2 1000 0.001 0.005 100 5", tm_steps: int = 1,
                         dt: float = 1000000.0, theta: float = 1.0,
                         stationary: int = 1, relax_time: float = 0.0) -> None:
    """Write the reference's keyword input files for ``deck`` into ``outdir``.

    One material per distinct (Manning n, vt, stem diameter, stem spacing, st, KDest, estx,
    esty).
    Values are written with %.17g so that the reference reads back the exact doubles used
    by the oracle.  ``tm_steps`` time steps of ``dt`` seconds with weight ``theta``;
    ``stationary`` 0 selects the unsteady flow cycle; ``relax_time`` the relaxed time step
    of relaxation methods 3 and 4 (relaxTimeFlow; 0: the time step).
    """
    os.makedirs(outdir, exist_ok=True)
    ne = deck.n_elems
    vd = deck.veg_dp if deck.veg_dp is not None else np.zeros(ne)
    vs = deck.veg_sp if deck.veg_sp is not None else np.zeros(ne)
    st = deck.turb_st if deck.turb_st is not None else np.full(ne, 1.0)
    kd = deck.turb_kdest if deck.turb_kdest is not None else np.full(ne, 0.6)
    ex = deck.turb_estx if deck.turb_estx is not None else np.zeros(ne)
    ey = deck.turb_esty if deck.turb_esty is not None else np.zeros(ne)
    keys = list(zip(deck.manning.tolist(), deck.vt.tolist(), vd.tolist(), vs.tolist(),
                    st.tolist(), kd.tolist(), ex.tolist(), ey.tolist()))
    mat_of = {p: i + 1 for i, p in enumerate(sorted(set(keys)))}
    with open(os.path.join(outdir, "region.dat"), "w") as f:
        f.write(f"{deck.n_nodes} {deck.n_elems} 0\n")
        for i in range(deck.n_nodes):
            f.write(f"{i + 1} {deck.x[i]:.17g} {deck.y[i]:.17g} {deck.z[i]:.17g}\n")
        for e in range(deck.n_elems):
            k = int(deck.nnd[e])
            mat = mat_of[keys[e]]
            shape = "tri" if k == 6 else "quad"
            ids = " ".join(str(int(v) + 1) for v in deck.elems[e, :k])
            f.write(f"{e + 1} {mat} {shape} {ids}\n")
    segs = [(no, s) for no in sorted(deck.lines) for s in deck.lines[no]]
    with open(os.path.join(outdir, "control.dat"), "w") as f:
        f.write(f"0 {len(segs)}\n")
        for k, (no, s) in enumerate(segs):
            f.write(f"{k + 1} {no} line {int(s[0]) + 1} {int(s[1]) + 1} {int(s[2]) + 1}\n")
    with open(os.path.join(outdir, "init.dat"), "w") as f:
        f.write("# 0.0 x 40000 U V S\n")
        f.write(f"{deck.n_nodes} 0\n")
        for i in range(deck.n_nodes):
            f.write(f"{i + 1} {deck.U0[i]:.17g} {deck.V0[i]:.17g} {deck.S0[i]:.17g}\n")
    with open(os.path.join(outdir, "mat.dat"), "w") as f:
        f.write("$RISMO2D 40000\n")
        for (n, vt, dp, sp, a, b, cx, cy), no in mat_of.items():
            f.write(f"$ROUGHNESS {no} 5 {n:.17g} 0 0.0 {dp:.17g} {sp:.17g}\n")
        for (n, vt, dp, sp, a, b, cx, cy), no in mat_of.items():
            # no vt st estx esty KDest lm betaSf (TYPE::Import_40000)
            f.write(f"$TURBULENCE {no} {vt:.17g} {a:.17g} {cx:.17g} {cy:.17g} {b:.17g} 0.0 0.0\n")
    if deck.hydrograph is not None:
        inlet_lines = "".join(f"$TM_LINE QTINLET {INLET_LINE} {q:.17g} {sv:.17g} {t:.17g}\n"
                              for t, q, sv in np.asarray(deck.hydrograph, np.float64))
    else:
        inlet_lines = f"$TM_LINE QINLET {INLET_LINE} {deck.Q_inlet:.17g} {deck.S_inlet:.17g}\n"
    with open(os.path.join(outdir, "tm.dat"), "w") as f:
        f.write(f"$RISMO2D 40100\n$TM_STEPS 1 {tm_steps} 0\n"
                f"$TM_INTERVAL {dt:.17g} {relax_time:.17g} 0.0\n"
                f"$TM_WEIGHT {theta:.17g} 1.0 1.0\n$TM_STEP_NO 1\n$TM_CYCLE 1\n"
                f"$TM_STATIONARY {stationary}\n"
                f"$TM_TURBULENCE {deck.turb}\n$TM_MAXITER {maxiter}\n$TM_SOLVER {tm_solver}\n"
                + inlet_lines +
                f"$TM_LINE OUTLET {OUTLET_LINE} {deck.S_outlet:.17g} 1000.0 0.0 0 0.0\n")
    with open(os.path.join(outdir, "input.dat"), "w") as f:
        f.write("$RISMO2D 40700\n$REPORTLEVEL 3\n$REPORTFILE report.txt\n"
                "$TIMESTEPFILE tm.dat\n$MATERIALFILE mat.dat\n$REGIONFILE region.dat\n"
                "$CONTROLFILE control.dat\n$ASC_INITFILE init.dat\n"
                f"$SOLVER {solver}\n$SOLVER {fgmres_solver}\n"
                f"$CONVERGENCE {convergence}\n$RELAX {relax}\n$DRYREW {dryrew}\n"
                "$GPDEGREE 5\n")


# ---------------------------------------------------------------------------------------
# reading the reference's keyword input files

class DeckFormatError(ValueError):
    """An input file that the reference's readers accept but this solve path does not
    support, or one they would reject."""


def _data_lines(path: str, skip_first: bool = False):
    """Lines of a reference input file the way ASCIIFILE::nextLine serves them
    (Asciifile.cpp:84-102): lines whose first character is '*' or '#' are comments.  With
    skip_first the first line (read before by ASCIIFILE::next, whatever it holds) is not
    served."""
    with open(path) as f:
        if skip_first:
            f.readline()
        for line in f:
            if line[:1] in ("*", "#"):
                continue
            yield line


# corner nodes of every midside node (Shape getCornerNodes: tri 01 12 20, quad 01 12 23 30)
_MIDSIDE_CORNERS = {6: ((3, 0, 1), (4, 1, 2), (5, 2, 0)),
                    8: ((4, 0, 1), (5, 1, 2), (6, 2, 3), (7, 3, 0))}


def _read_region(path: str):
    """GRID::InputRegion (Grid.cpp:137-376): 'np ne ndata', np node lines 'name x y z',
    ne element lines 'name mat shape nodes...' of which only tri (6 nodes) and quad (8
    nodes) are elements; elements are indexed by name; midside bed elevations are set to
    the mean of their corners."""
    it = _data_lines(path)
    head = next(it).split()
    npn, ne = int(head[0]), int(head[1])
    x = np.zeros(npn); y = np.zeros(npn); z = np.zeros(npn)
    seen = np.zeros(npn, bool)
    for _ in range(npn):
        t = next(it).split()
        name = int(t[0])
        if not 1 <= name <= npn:
            raise DeckFormatError(f"{path}: node numbers must be between 1 and {npn} ({name})")
        x[name - 1], y[name - 1], z[name - 1] = float(t[1]), float(t[2]), float(t[3])
        seen[name - 1] = True
    if not seen.all():
        raise DeckFormatError(f"{path}: {int((~seen).sum())} nodes missing")
    elems = {}
    for _ in range(ne):
        t = next(it).split()
        name, mat, shape = int(t[0]), int(t[1]), t[2]
        if shape not in ("tri", "quad"):
            continue                      # the region reader keeps only tri and quad
        if not 1 <= name <= ne:
            raise DeckFormatError(f"{path}: element numbers must be between 1 and {ne}")
        k = 6 if shape == "tri" else 8
        nodes = [int(v) for v in t[3:3 + k]]
        if len(nodes) != k or min(nodes) < 1 or max(nodes) > npn:
            raise DeckFormatError(f"{path}: element {name}: {k} node numbers in 1..{npn} expected")
        elems[name] = (mat, k, nodes)
    names = sorted(elems)
    el = np.full((len(names), 8), -1, np.int32)
    nnd = np.zeros(len(names), np.int32)
    mat = np.zeros(len(names), np.int64)
    for e, name in enumerate(names):
        m, k, nodes = elems[name]
        el[e, :k] = np.asarray(nodes) - 1
        nnd[e] = k
        mat[e] = m
    zor = z.copy()
    for e in range(len(names)):
        for i, a, b in _MIDSIDE_CORNERS[int(nnd[e])]:
            z[el[e, i]] = 0.5 * (zor[el[e, a]] + zor[el[e, b]])
    return x, y, z, el, nnd, mat


def _read_control(path: str):
    """GRID::InputControl (Grid.cpp:378-508): 'np ne', np node lines (skipped), ne element
    lines 'name line_no shape n0 n1 n2' of which only shape 'line' counts; segments are
    kept per line number in file order."""
    it = _data_lines(path)
    head = next(it).split()
    npc, nec = int(head[0]), int(head[1])
    for _ in range(npc):
        next(it)
    lines: dict = {}
    for _ in range(nec):
        t = next(it).split()
        if t[2] != "line":
            continue
        lines.setdefault(int(t[1]), []).append([int(t[3]) - 1, int(t[4]) - 1, int(t[5]) - 1])
    return {k: np.asarray(v, np.int32) for k, v in lines.items()}


def _read_initial(path: str, z: np.ndarray):
    """GRID::InputInitial, ASCII (Grid.cpp:510-760): a header line '#time x release vars...'
    (release >= 40000: the variables of the columns by name), 'np ne', then per node 'name
    values...'.  Unset values: U = V = 0, S = bed elevation."""
    with open(path) as f:
        header = f.readline()
    toks = header[1:].replace(",", " ").split()
    release = int(toks[2]) if len(toks) > 2 else 0
    names = toks[3:] if release >= 40000 else []
    it = _data_lines(path, skip_first=True)
    t = next(it).split()
    npi = int(t[0])
    if npi != len(z):
        raise DeckFormatError(f"{path}: wrong number of nodes in file ({npi}, region {len(z)})")
    U = np.zeros(len(z)); V = np.zeros(len(z)); S = z.copy()
    for _ in range(npi):
        t = next(it).split()
        name = int(t[0])
        vals = {"U": 0.0, "V": 0.0, "S": z[name - 1]}
        if release >= 40000:
            for var, tok in zip(names, t[1:]):
                vals[var] = float(tok)
        elif release >= 280:
            vals["U"], vals["V"], vals["S"] = (float(v) for v in t[1:4])
        else:
            vals["U"], vals["V"], vals["S"] = float(t[1]), float(t[2]), float(t[4])
        U[name - 1], V[name - 1], S[name - 1] = vals["U"], vals["V"], vals["S"]
    return U, V, S


def _keyword_lines(path: str):
    """Keyword lines of a $RISMO2D file: (release, [(keyword, tokens)]), leading blanks
    removed, comments skipped."""
    with open(path) as f:
        first = f.readline().split()
    release = int(first[1]) if len(first) > 1 and first[0] == "$RISMO2D" else 0
    out = []
    for n, line in enumerate(_data_lines(path)):
        s = line.lstrip(" ")
        if n == 0 and s.startswith("$RISMO2D"):
            continue
        if s.startswith("$"):
            t = s[1:].split()
            if t:
                out.append((t[0], t[1:]))
    return release, out


_MANNING = 5                                   # TYPE rtype kMANN (Type.h:120-123)


def _read_materials(path: str):
    """TYPE::Import_40000 (Type.cpp:434-700): per material no, $ROUGHNESS (law, coefficient,
    vegetation stem diameter / spacing in the release's layout), $VEGETATION and
    $TURBULENCE (vt st estx esty KDest lm betaSf).  A material starts from the TYPE
    defaults (vt 1e-4, st 1.0, KDest 0.6; Type.cpp:55-74); lines with no <= 0 set defaults
    after the materials exist and so change none of them."""
    release, kws = _keyword_lines(path)
    if release < 40000:
        raise DeckFormatError(f"{path}: material file release {release} (40000 or later supported)")
    mats: dict = {}

    def m(no):
        return mats.setdefault(no, dict(rtype=3, rcoef=0.0, dp=0.0, sp=0.0, vt=1.0e-4, st=1.0,
                                        kdest=0.6, estx=0.0, esty=0.0))
    laws = {"COWI": 2, "NIKU": 3, "CHEZ": 4, "MANN": 5}
    for key, t in kws:
        if key == "ROUGHNESS" and int(t[0]) > 0:
            r = m(int(t[0]))
            r["rtype"], r["rcoef"] = int(t[1]), float(t[2])
            if release >= 40075:
                r["dp"], r["sp"], dp1 = float(t[10]), float(t[11]), float(t[13])
            elif release >= 40070:
                r["dp"], r["sp"], dp1 = float(t[7]), float(t[8]), float(t[10])
            elif release >= 40060:
                r["dp"], r["sp"], dp1 = float(t[5]), float(t[6]), float(t[8])
            else:
                r["dp"], r["sp"], dp1 = float(t[5]), float(t[6]), 0.0
            if dp1 > 0.0:
                raise DeckFormatError(f"{path}: material {t[0]}: a second vegetation layer "
                                      "is not supported")
        elif key == "BOTTOM" and int(t[0]) > 0:
            r = m(int(t[0]))
            r["rtype"], r["rcoef"] = laws.get(t[1], 0), float(t[2])
        elif key == "VEGETATION" and int(t[0]) > 0:
            r = m(int(t[0]))
            r["dp"], r["sp"] = float(t[1]), float(t[2])
            if len(t) > 4 and float(t[4]) > 0.0:
                raise DeckFormatError(f"{path}: material {t[0]}: a second vegetation layer "
                                      "is not supported")
        elif key == "TURBULENCE" and int(t[0]) > 0:
            r = m(int(t[0]))
            r["vt"], r["st"], r["kdest"] = float(t[1]), float(t[2]), float(t[5])
            r["estx"], r["esty"] = float(t[3]), float(t[4])
    return mats


_UVS_CYCLE = 1                 # kUVSCyc (Project.h:102): the UVS Taylor-Hood flow cycle
# per-cycle lists of a boundary set and the reference's defaults (BCONSET::BCONSET)
_TM_CYCLE_LISTS = {"TM_STATIONARY": "stationary", "TM_TURBULENCE": "turb",
                   "TM_DISPERSION": "disp", "TM_MAXITER": "maxiter", "TM_SOLVER": "solver"}


def _read_timesteps(path: str):
    """The time-step file (TIMEINT::Input_40100, Timeint.cpp:1466-2340): $TM_STEPS,
    $TM_INTERVAL, $TM_WEIGHT, one boundary set ($TM_STEP_NO 1) with its cycle list
    ($TM_CYCLE) and the per-cycle lists $TM_STATIONARY, $TM_TURBULENCE, $TM_DISPERSION,
    $TM_MAXITER and $TM_SOLVER, and the line conditions $TM_LINE QINLET no Q S (one value,
    later lines overwrite), QTINLET no Q S t (hydrograph rows) and OUTLET no S x y ...
    (Timeint.cpp:2230-2330).

    A per-cycle list is read as the reference reads it (Timeint.cpp:2021-2135): its
    tokens are the values of cycles 0, 1, ...; the cycles of the set's $TM_CYCLE list read
    so far beyond the last token take the last token's value.  The flow cycle's entry is
    kept.  $TM_OUTPUT and $TM_RESET_STATIST (output and statistics only) are read past.
    Anything else (more than one boundary set, another cycle than the flow cycle 1,
    dispersion, $TM_SETTIME, boundary conditions by $TM_BOUND_* / $TM_NODE / periodic
    lines) raises DeckFormatError, as does a set without $TM_CYCLE or $TM_MAXITER (the
    reference then runs no Newton step: BCONSET::BCONSET leaves cycit = 0)."""
    release, kws = _keyword_lines(path)
    run = dict(release=release, steps=1, dt=None, theta=1.0, stationary=1, turb=1, maxiter=None)
    cond: dict = {}
    sets = 0
    cycles: list = []
    lists: dict = {}
    for key, t in kws:
        if key == "TM_STEPS":
            run["steps"] = int(t[1]) - int(t[0]) + 1
        elif key == "TM_INTERVAL":
            # deltaTm relaxTmFlow ...; relaxTmFlow <= 0 or above deltaTm: deltaTm
            # (Timeint.cpp:697-719)
            run["dt"] = float(t[0])
            rt = float(t[1]) if len(t) > 1 else 0.0
            run["relax_time"] = rt if 0.0 < rt <= run["dt"] else run["dt"]
        elif key == "TM_WEIGHT":
            run["theta"] = float(t[0])
        elif key in ("TM_OUTPUT", "TM_RESET_STATIST"):
            continue
        elif key == "TM_STEP_NO":
            sets += 1
            if sets > 1:
                raise DeckFormatError(f"{path}: more than one boundary set ($TM_STEP_NO) is "
                                      "not supported")
            cycles = []                          # ncyc = 0 (Timeint.cpp:1998)
        elif key == "TM_CYCLE":
            cycles.extend(int(v) for v in t)
        elif key in _TM_CYCLE_LISTS:
            vals = [int(v) for v in t]
            if not vals:
                raise DeckFormatError(f"{path}: ${key} without values")
            # tokens for cycles 0..n-1, the last one repeated up to the cycles read so far
            vals += [vals[-1]] * max(0, len(cycles) - len(vals))
            lists[_TM_CYCLE_LISTS[key]] = vals
        elif key == "TM_LINE" and int(t[1]) > 0:
            kind, no = t[0], int(t[1])
            c = cond.setdefault((kind, no), [])
            vals = [float(v) for v in t[2:]]
            if kind == "QINLET":
                c[:] = [vals[:2]]
            elif kind in ("QTINLET", "OUTLET"):
                c.append(vals)
            else:
                raise DeckFormatError(f"{path}: line condition {kind} is not supported "
                                      "(QINLET, QTINLET, OUTLET)")
        elif key.startswith("TM_") and key != "TM_LINE":
            raise DeckFormatError(f"{path}: ${key} is not supported")
    if not cycles:
        raise DeckFormatError(f"{path}: no $TM_CYCLE: the reference runs no cycle")
    if cycles != [_UVS_CYCLE]:
        raise DeckFormatError(f"{path}: $TM_CYCLE {' '.join(map(str, cycles))}: only the flow "
                              f"cycle {_UVS_CYCLE} alone is supported")
    if "maxiter" not in lists:
        raise DeckFormatError(f"{path}: no $TM_MAXITER: the reference runs no Newton step")
    if lists.get("disp", [0])[0] != 0:
        raise DeckFormatError(f"{path}: $TM_DISPERSION {lists['disp'][0]} is not supported")
    for k in ("stationary", "turb", "maxiter"):
        if k in lists:
            run[k] = lists[k][0]
    return run, cond


def write_reference_restart(path: str, deck: Deck, U, V, S, time: str = "0.0",
                            dUdt=None, dVdt=None, dSdt=None, release: int = 40700,
                            exact: bool = False) -> None:
    """Write a state as the reference's ASCII restart file (GRID::OutputData, Grid.cpp:1183-
    1244): header '# time Release n U,V,S,dUdt,dVdt,dSdt,K,D,C,qb,Zb', 'np ne', per node
    its name and the eleven values, per element its name and U V P dz (element means of
    the corner values; dz 0).  The reference prints %14.6le; ``exact`` prints %.17g so that
    the state reads back bit for bit.  The file is also an initial file ($ASC_INITFILE) for
    the reference and for read_reference_deck."""
    n = deck.n_nodes
    z0 = np.zeros(n)
    cols = [np.asarray(a, np.float64) for a in
            (U, V, S, dUdt if dUdt is not None else z0, dVdt if dVdt is not None else z0,
             dSdt if dSdt is not None else z0, z0, z0, z0, z0, deck.z)]
    fmt = (lambda v: f" {v:.17g}") if exact else (lambda v: f" {v:14.6e}")
    ncn = np.where(deck.nnd == 6, 3, 4)
    with open(path, "w") as f:
        f.write(f"# {time:>22s}   Release {release}   U,V,S,dUdt,dVdt,dSdt,K,D,C,qb,Zb\n")
        f.write(f"{n}  {deck.n_elems}\n")
        for i in range(n):
            f.write(f"{i + 1:7d}" + "".join(fmt(c[i]) for c in cols) + "\n")
        for e in range(deck.n_elems):
            corners = deck.elems[e, :ncn[e]]
            vals = (cols[0][corners].mean(), cols[1][corners].mean(), cols[2][corners].mean(), 0.0)
            f.write(f"{e + 1:7d}" + "".join(fmt(v) for v in vals) + "\n")


def _scan(tokens, defaults, types):
    """The fields of a keyword line the way sscanf fills them: in order, until the first
    token that does not convert; the rest keep their previous values."""
    out = list(defaults)
    for i, (tok, ty) in enumerate(zip(tokens, types)):
        try:
            out[i] = ty(tok)
        except ValueError:
            break
    return out


# PROJECT and DRYREW defaults (Project.cpp:71-92, Grid.h:109-114) for the project keywords
# read by PROJECT::Input (Project.cpp:1780-1852)
_PROJECT_KEYS = {
    # convUV convS convKD convQb
    "CONVERGENCE": ((1.0e-3, 1.0e-4, 1.0e-4, 1.0e-3), (float,) * 4),
    # relaxMethod relaxMin relaxMax maxDeltaUV maxDeltaS maxDeltaKD
    "RELAX": ((3, 0.001, 0.1, 0.1, 0.1, 0.1), (int,) + (float,) * 5),
    # method dryRewFreq dryLimit rewetLimit rewetPasses countDown
    "DRYREW": ((2, 10, 0.001, 0.005, 100, 5), (int, int, float, float, int, int)),
    # minVt minK minD maxUs minC
    "MINMAX": ((0.0, 0.0, 0.0, 100.0, 0.0), (float,) * 5),
}


def read_reference_deck(input_file: str) -> tuple[Deck, dict]:
    """Read a deck in the reference's keyword input format: the project file
    (``$REGIONFILE``, ``$CONTROLFILE``, ``$ASC_INITFILE``, ``$MATERIALFILE``,
    ``$TIMESTEPFILE``; Project.cpp:1296-1440, names relative to the project file's folder)
    and the files it names, read as the reference's readers read them.  Returns the deck
    (line 1 the inlet, line 2 the outlet, the solve path's convention) and the run settings
    (steps, dt, theta, stationary flag, Newton iterations, ``params``: the solve parameters
    of ``$CONVERGENCE``, ``$RELAX``, ``$DRYREW`` and ``$MINMAX`` with the reference's
    defaults, and the project file's other keyword lines as text).

    Supported: Manning roughness, emergent vegetation, constant / algebraic / anisotropic
    eddy viscosity, one QINLET or QTINLET line and one OUTLET line with a single water
    level, relaxation methods 0, 2, 3 and 4 (3 and 4: time-relaxed Newton steps, with
    relaxTimeFlow from ``$TM_INTERVAL``), drying and rewetting by method 2 or none
    (``$DRYREW``).  Anything else raises DeckFormatError."""
    base = os.path.dirname(os.path.abspath(input_file))
    files, other = {}, []
    proj = {k: list(v[0]) for k, v in _PROJECT_KEYS.items()}
    for key, t in _keyword_lines(input_file)[1]:
        if key in _PROJECT_KEYS:
            proj[key] = _scan(t, proj[key], _PROJECT_KEYS[key][1])
            other.append("$" + " ".join([key] + t))
            continue
        if key == "MAX_Us":
            # maxUs, also the fourth field of $MINMAX: the later line wins (Project.cpp:1786,
            # 1830).  ($MIN_VT*, $MIN_K, $MIN_D, $MIN_C only bound the k-epsilon models'
            # eddy viscosity, CoefsKD2D.cpp:374; they are kept as text.)
            proj["MINMAX"][3] = _scan(t[:1], [proj["MINMAX"][3]], (float,))[0]
            other.append("$" + " ".join([key] + t))
            continue
        if key in ("BIN_INITFILE", "BIN_INIFILE", "STA_INITFILE"):
            raise DeckFormatError(f"{input_file}: ${key}: only ASCII initial files "
                                  "($ASC_INITFILE) are supported")
        if key in ("REGIONFILE", "CONTROLFILE", "ASC_INITFILE", "ASC_INIFILE", "MATERIALFILE",
                   "TIMESTEPFILE"):
            files["ASC_INITFILE" if key == "ASC_INIFILE" else key] = os.path.join(base, t[0])
        else:
            other.append("$" + " ".join([key] + t))
    for key in ("REGIONFILE", "CONTROLFILE", "MATERIALFILE", "TIMESTEPFILE"):
        if key not in files:
            raise DeckFormatError(f"{input_file}: no ${key}")
    x, y, z, elems, nnd, mat = _read_region(files["REGIONFILE"])
    lines_in = _read_control(files["CONTROLFILE"])
    if "ASC_INITFILE" in files:
        U0, V0, S0 = _read_initial(files["ASC_INITFILE"], z)
    else:
        U0, V0, S0 = np.zeros(len(z)), np.zeros(len(z)), z.copy()
    mats = _read_materials(files["MATERIALFILE"])
    run, cond = _read_timesteps(files["TIMESTEPFILE"])
    missing = sorted(set(int(v) for v in np.unique(mat)) - set(mats))
    if missing:
        raise DeckFormatError(f"material(s) {missing} used by elements but not defined")
    law = {no: r["rtype"] for no, r in mats.items() if no in set(mat.tolist())}
    if any(v != _MANNING for v in law.values()):
        raise DeckFormatError(f"roughness laws {sorted(set(law.values()))}: only Manning (5) "
                              "is supported")

    def per_elem(k):
        return np.array([mats[int(v)][k] for v in mat], np.float64)
    inlets = [(k, no) for (k, no) in cond if k in ("QINLET", "QTINLET")]
    outlets = [(k, no) for (k, no) in cond if k == "OUTLET"]
    if len(inlets) != 1 or len(outlets) != 1:
        raise DeckFormatError(f"one inlet and one outlet line expected ({len(inlets)}, {len(outlets)})")
    (ik, ino), (_, ono) = inlets[0], outlets[0]
    if len(cond[("OUTLET", ono)]) != 1:
        raise DeckFormatError("OUTLET: a water-level profile along the line is not supported")
    for no in (ino, ono):
        if no not in lines_in:
            raise DeckFormatError(f"control line {no} has no line elements")
    q = cond[(ik, ino)]
    hydro = None
    if ik == "QTINLET":
        hydro = np.asarray([[r[2], r[0], r[1]] for r in q], np.float64)   # rows (t, Q, S)
        Q_in, S_in = float(hydro[0, 1]), float(hydro[0, 2])
    else:
        Q_in, S_in = float(q[0][0]), float(q[0][1])
    dp, sp = per_elem("dp"), per_elem("sp")
    deck = Deck(x=x, y=y, z=z, elems=elems, nnd=nnd, manning=per_elem("rcoef"),
                vt=per_elem("vt"), lines={INLET_LINE: lines_in[ino], OUTLET_LINE: lines_in[ono]},
                U0=U0, V0=V0, S0=S0, Q_inlet=Q_in, S_inlet=S_in,
                S_outlet=float(cond[("OUTLET", ono)][0][0]),
                name=os.path.basename(base),
                veg_dp=dp if (dp > 0).any() else None, veg_sp=sp if (dp > 0).any() else None,
                turb=run["turb"], turb_st=per_elem("st"), turb_kdest=per_elem("kdest"),
                turb_estx=per_elem("estx"), turb_esty=per_elem("esty"),
                hydrograph=hydro)
    run["project"] = other
    conv, relax, dry, minmax = (proj[k] for k in ("CONVERGENCE", "RELAX", "DRYREW", "MINMAX"))
    dry[3] = max(dry[3], dry[2])                 # rewetLimit not below dryLimit
    if relax[0] not in (0, 2, 3, 4):
        raise DeckFormatError(f"$RELAX: relaxation method {relax[0]} is not supported "
                              "(0, 2, 3, 4)")
    if relax[0] >= 3 and not run.get("dt"):
        raise DeckFormatError("$RELAX: relaxation methods 3 and 4 need $TM_INTERVAL")
    maxiter = run["maxiter"]                     # the flow cycle's $TM_MAXITER entry
    if dry[0] not in (0, 2) and dry[1] < maxiter:
        # EqsUVS2D.cpp:201: the flow cycle calls the drying and rewetting
        raise DeckFormatError(f"$DRYREW: drying and rewetting method {dry[0]} is not "
                              "supported (0, 2)")
    run["params"] = dict(relax_method=relax[0], max_delta_uv=relax[3], max_delta_s=relax[4],
                         conv_uv=conv[0], conv_s=conv[1], dry_method=dry[0], dry_limit=dry[2],
                         max_us=minmax[3], rewet_limit=dry[3], count_down=dry[5])
    if relax[0] >= 3:                            # time-relaxed Newton steps
        run["params"].update(relax_time=run["relax_time"], interval=run["dt"])
    run["dryrew"] = dict(method=dry[0], freq=dry[1], dry_limit=dry[2], rewet_limit=dry[3],
                         rewet_passes=dry[4], count_down=dry[5])
    return deck, run
