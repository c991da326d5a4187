"""Synthetic finite-volume systems in the reference's sparse layouts.

The reference assembles its systems cell by cell from a mesh deck
(src/Pform.f, src/Celllib*.f).  The solver path only sees the assembled
arrays, so this module builds the same kind of systems directly, vectorised
with numpy, and lays them out exactly as src/Datastruct.f does:

* ELLPACK (matriz 4): ``iax = [w, neq+1]``, ``ja(w, neq)`` and ``al(w, neq)``
  slot-fastest, padding ``ja = 1, a = 0`` (src/EllPack.f:25-36,80-89);
* CSRD (matriz 1): ``ia(neq+1)``, ``ja`` without the diagonal, sorted per row
  (src/Datastruct.f:40-77, src/Graph.f:45-70); symmetric storage keeps only
  the lower triangle;
* CSR (matriz 2): diagonal first, then neighbours in slot order
  (src/Csr.f:76-110 with diag), all values in ``ad``;
* CSRC (matriz 3): lower pattern sorted, ``al`` = row values, ``au(k)`` = the
  transposed upper value that row i scatters into ``y(ja(k))``
  (src/Matvec.f:102-135).

All indices are 1-based int32, all values float64, off-diagonals stored as
``-a_nb`` (src/EllPack.f:85, src/Csr.f:172).

Coefficient forms (interior face f between cells P and N, |S| face length,
xi = x_N - x_P, n the unit outward normal):
  diffusion   D_f = k_f |S| / ((n . xi/|xi|) |xi|)     src/CellibTrans.f:527
  harmonic    1/k_f = a/k_P + (1-a)/k_N,  a = d_P/|xi|  src/CellibTrans.f:510-511
  convection  F_f = psi(node2) - psi(node1)  (exactly conservative fluxes of
              u = curl psi), a_nb = D_f + max(-F_f, 0), a_P = sum a_nb + sum F_f
              src/CellibTrans.f:561,570-573, src/CellibSimple.f:789-816
  Dirichlet   a_P += k_P |S| / d_Pf, b += k_P |S| / d_Pf * T_wall
              src/CellibTrans.f:488-491
  pressure    a_nb = D_f, Neumann walls, cell 1 pinned with +1e60 on its
              diagonal (src/CellibPressure.f:414,557).

Random numbers come from SplitMix64 (Steele, Lea, Flood 2014) evaluated as a
counter-based stream, so every draw is reproducible from (seed, stream, i).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

PIN_PENALTY = 1.0e60
SEED_BASE = 20261202

_GAMMA = np.uint64(0x9E3779B97F4A7C15)


def _mix64(z: np.ndarray) -> np.ndarray:
    z = z.copy()
    z ^= z >> np.uint64(30)
    z *= np.uint64(0xBF58476D1CE4E5B9)
    z ^= z >> np.uint64(27)
    z *= np.uint64(0x94D049BB133111EB)
    z ^= z >> np.uint64(31)
    return z


def uniform(seed: int, stream: int, n: int) -> np.ndarray:
    """n doubles in [0, 1): SplitMix64 outputs of key + (i+1)*gamma, top 53 bits."""
    with np.errstate(over="ignore"):
        key = _mix64(np.array([(seed * 0x100000001B3 + stream * 0x9E3779B97F4A7C15)
                               & 0xFFFFFFFFFFFFFFFF], dtype=np.uint64))[0]
        ctr = np.arange(1, n + 1, dtype=np.uint64) * _GAMMA + key
        bits = _mix64(ctr) >> np.uint64(11)
    return bits.astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed: int, stream: int, n: int) -> np.ndarray:
    """Standard normals by Box-Muller from two uniform streams."""
    u1 = uniform(seed, 2 * stream + 1, n)
    u2 = uniform(seed, 2 * stream + 2, n)
    return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * math.pi * u2)


def lognormal_k(seed: int, n: int, sigma: float = 0.5) -> np.ndarray:
    return np.exp(sigma * normal(seed, 7, n))


# --------------------------------------------------------------------- mesh --

@dataclass
class Mesh:
    """Cell-centred 2-D mesh: per cell, ``w`` edges in the reference's edge
    order (src/Adjacency.f: edge s joins local nodes s and s+1)."""
    neq: int
    w: int
    nodes: np.ndarray          # (nnode, 2)
    cell_nodes: np.ndarray     # (neq, w) node ids, counter-clockwise
    nbr: np.ndarray            # (neq, w) 0-based neighbour cell or -1
    boundary_tag: np.ndarray   # (neq, w) int8: -1 interior, 0 S, 1 E, 2 N, 3 W wall
    _geo: dict = field(default_factory=dict)

    def geometry(self):
        if self._geo:
            return self._geo
        xy = self.nodes[self.cell_nodes]                      # (neq, w, 2)
        x, y = xy[..., 0], xy[..., 1]
        xn, yn = np.roll(x, -1, axis=1), np.roll(y, -1, axis=1)
        cross = x * yn - xn * y
        area = 0.5 * cross.sum(axis=1)
        cx = ((x + xn) * cross).sum(axis=1) / (6.0 * area)
        cy = ((y + yn) * cross).sum(axis=1) / (6.0 * area)
        tx, ty = xn - x, yn - y
        slen = np.hypot(tx, ty)                               # |S|
        nx_, ny_ = ty / slen, -tx / slen                      # outward unit normal
        mx, my = 0.5 * (x + xn), 0.5 * (y + yn)               # face midpoints
        dpf = (mx - cx[:, None]) * nx_ + (my - cy[:, None]) * ny_   # centroid-face distance
        self._geo = dict(area=area, cx=cx, cy=cy, slen=slen, nx=nx_, ny=ny_,
                         mx=mx, my=my, dpf=dpf)
        return self._geo


def quad_mesh(nx: int, ny: int | None = None, jitter: float = 0.0, seed: int = SEED_BASE) -> Mesh:
    """nx*ny quads on the unit square; interior nodes moved by U[-jitter*h, jitter*h]."""
    ny = nx if ny is None else ny
    hx, hy = 1.0 / nx, 1.0 / ny
    I, J = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), indexing="xy")
    X = (I * hx).astype(np.float64).ravel()
    Y = (J * hy).astype(np.float64).ravel()
    if jitter > 0.0:
        interior = ((I > 0) & (I < nx) & (J > 0) & (J < ny)).ravel()
        nn = X.size
        dx = (2.0 * uniform(seed, 1, nn) - 1.0) * jitter * hx
        dy = (2.0 * uniform(seed, 2, nn) - 1.0) * jitter * hy
        X = np.where(interior, X + dx, X)
        Y = np.where(interior, Y + dy, Y)
    nodes = np.stack([X, Y], axis=1)
    ci, cj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")
    ci, cj = ci.ravel(), cj.ravel()
    n0 = cj * (nx + 1) + ci
    cell_nodes = np.stack([n0, n0 + 1, n0 + nx + 2, n0 + nx + 1], axis=1)
    c = cj * nx + ci
    nbr = np.stack([
        np.where(cj > 0, c - nx, -1),          # edge 1 (south)
        np.where(ci < nx - 1, c + 1, -1),      # edge 2 (east)
        np.where(cj < ny - 1, c + nx, -1),     # edge 3 (north)
        np.where(ci > 0, c - 1, -1),           # edge 4 (west)
    ], axis=1)
    tag = np.where(nbr < 0, np.arange(4)[None, :], -1).astype(np.int8)
    return Mesh(nx * ny, 4, nodes, cell_nodes.astype(np.int64), nbr.astype(np.int64), tag)


def tri_mesh(n: int, jitter: float = 0.0, seed: int = SEED_BASE) -> Mesh:
    """n*n quads, each split along its A-C diagonal into triangles (A,B,C) and
    (A,C,D); cell id = 2*(i + n*j) + t."""
    q = quad_mesh(n, n, jitter, seed)
    ci, cj = np.meshgrid(np.arange(n), np.arange(n), indexing="xy")
    ci, cj = ci.ravel(), cj.ravel()
    A, B, C, D = (q.cell_nodes[:, k] for k in range(4))
    qid = cj * n + ci
    t0 = np.stack([A, B, C], axis=1)
    t1 = np.stack([A, C, D], axis=1)
    cell_nodes = np.empty((2 * n * n, 3), dtype=np.int64)
    cell_nodes[0::2] = t0
    cell_nodes[1::2] = t1
    nb0 = np.stack([
        np.where(cj > 0, 2 * (qid - n) + 1, -1),      # (A,B) -> lower quad, tri 1
        np.where(ci < n - 1, 2 * (qid + 1) + 1, -1),  # (B,C) -> east quad, tri 1
        2 * qid + 1,                                  # (C,A) -> same quad, tri 1
    ], axis=1)
    nb1 = np.stack([
        2 * qid,                                      # (A,C) -> same quad, tri 0
        np.where(cj < n - 1, 2 * (qid + n), -1),      # (C,D) -> north quad, tri 0
        np.where(ci > 0, 2 * (qid - 1), -1),          # (D,A) -> west quad, tri 0
    ], axis=1)
    nbr = np.empty((2 * n * n, 3), dtype=np.int64)
    nbr[0::2] = nb0
    nbr[1::2] = nb1
    tag = np.full((2 * n * n, 3), -1, dtype=np.int8)
    tag[0::2, 0] = np.where(cj > 0, -1, 0)
    tag[0::2, 1] = np.where(ci < n - 1, -1, 1)
    tag[1::2, 1] = np.where(cj < n - 1, -1, 2)
    tag[1::2, 2] = np.where(ci > 0, -1, 3)
    return Mesh(2 * n * n, 3, q.nodes, cell_nodes, nbr, tag)


def ref_mesh_arrays(mesh: Mesh, wall_bc=(3, 6, 1, 3)) -> dict:
    """The mesh in the reference's arrays, as C arrays with the first Fortran
    index fastest (src/Rdata_mvf.f, src/Adjacency.f): x(2, nnode) ->
    (nnode, 2); ix(nen+1, numel) -> (numel, nen+1), 1-based nodes then the
    material (1); nelcon(nen, numel) -> (numel, nen), 1-based neighbours, -1
    across a wall; pedge(nen+1, numel) -> (numel, nen+1), the boundary
    condition of each wall edge by wall (S, E, N, W): 3 prescribed pressure,
    6 zero gradient, anything else extrapolated (src/CellibPressure.f:
    443-458).  cell_type is the material's cell type: 65 (quad) or 63
    (triangle), the SIMPLE cells of src/Celllib.f:127-138."""
    w = mesh.w
    ix = np.empty((mesh.neq, w + 1), dtype=np.int32)
    ix[:, :w] = mesh.cell_nodes + 1
    ix[:, w] = 1
    nelcon = np.where(mesh.nbr >= 0, mesh.nbr + 1, -1).astype(np.int32)
    pedge = np.zeros((mesh.neq, w + 1), dtype=np.int32)
    bc = np.asarray(wall_bc, dtype=np.int32)
    pedge[:, :w] = np.where(mesh.boundary_tag >= 0, bc[np.maximum(mesh.boundary_tag, 0)], 0)
    return dict(x=np.ascontiguousarray(mesh.nodes, dtype=np.float64), ix=ix, nelcon=nelcon, pedge=pedge,
                cell_type=65 if w == 4 else 63)


# ------------------------------------------------------------------ systems --

@dataclass
class System:
    """Assembled cell-major system: row i = ad[i] x_i + sum_s off[i,s] x_nbr[i,s]."""
    neq: int
    w: int
    ad: np.ndarray        # (neq,)
    off: np.ndarray       # (neq, w) stored off-diagonal values (-a_nb), 0 at walls
    nbr: np.ndarray       # (neq, w) 0-based neighbour or -1
    b: np.ndarray         # (neq,)
    symmetric: bool = False
    meta: dict = field(default_factory=dict)

    @property
    def nad(self) -> int:
        return int((self.nbr >= 0).sum())


def _face_diffusion(mesh: Mesh, k: np.ndarray):
    g = mesh.geometry()
    nbr = mesh.nbr
    inner = nbr >= 0
    nb = np.where(inner, nbr, 0)
    xix = g["cx"][nb] - g["cx"][:, None]
    xiy = g["cy"][nb] - g["cy"][:, None]
    mxi = np.hypot(xix, xiy)
    mxi = np.where(inner, mxi, 1.0)
    nk = (g["nx"] * xix + g["ny"] * xiy) / mxi
    nk = np.where(inner, nk, 1.0)
    # share of the normal distance P->N on the P side; the N side sees 1 - a,
    # so the operator stays symmetric on distorted meshes
    a = np.clip(g["dpf"] / (nk * mxi), 0.0, 1.0)
    kP = k[:, None]
    kN = k[nb]
    kf = 1.0 / (a / kP + (1.0 - a) / kN)
    D = np.where(inner, kf * g["slen"] / (nk * mxi), 0.0)
    return D, dict(alpha=a, kf=kf, mxi=mxi, nk=nk)


def pressure_system(mesh: Mesh, k: np.ndarray | None = None, pin: str = "penalty",
                    seed: int = SEED_BASE) -> System:
    """Pressure-correction / diffusion operator with Neumann walls; b ~ U[-1,1]
    with its mean removed.  pin: "penalty" (+1e60 on cell 1's diagonal, the
    reference's method), "identity" (row 1 of the identity), or "none"."""
    k = np.ones(mesh.neq) if k is None else k
    D, _ = _face_diffusion(mesh, k)
    ad = D.sum(axis=1)
    off = -D
    b = 2.0 * uniform(seed, 11, mesh.neq) - 1.0
    b -= b.mean()
    if pin == "penalty":
        ad[0] = ad[0] + PIN_PENALTY
    elif pin == "identity":
        ad[0] = 1.0
        off[0, :] = 0.0
        b[0] = 0.0
    return System(mesh.neq, mesh.w, ad, off, mesh.nbr.copy(), b, symmetric=(pin != "identity"),
                  meta=dict(kind="pressure", pin=pin))


def stream_function(nodes: np.ndarray, seed: int, modes: int = 3) -> np.ndarray:
    """psi = sum_{m,n<=modes} c_mn sin(m pi x) sin(n pi y) / (m^2+n^2), psi = 0 on walls."""
    c = 2.0 * uniform(seed, 21, modes * modes) - 1.0
    x, y = nodes[:, 0], nodes[:, 1]
    psi = np.zeros(len(nodes))
    for m in range(1, modes + 1):
        sx = np.sin(m * math.pi * x)
        for n in range(1, modes + 1):
            psi += c[(m - 1) * modes + n - 1] * sx * np.sin(n * math.pi * y) / (m * m + n * n)
    on_wall = (x <= 0.0) | (x >= 1.0) | (y <= 0.0) | (y >= 1.0)
    psi[on_wall] = 0.0
    return psi


def advdiff_system(mesh: Mesh, k: np.ndarray | None = None, peclet_max: float = 100.0,
                   seed: int = SEED_BASE, walls=("dirichlet",) * 4,
                   wall_values=(0.0, 0.0, 0.0, 0.0), rhs: str = "random",
                   dt: float | None = None, under_relax: float = 1.0) -> System:
    """First-order upwind advection-diffusion (momentum / energy form).  The
    velocity is u = curl psi of a seeded smooth stream function, scaled so the
    largest cell Peclet number |F_f|/D_f equals peclet_max.  With ``dt`` the
    transient mass term rho*area/dt joins the diagonal, and ``under_relax``
    divides it as SIMPLE does (src/CellibSimple.f:347-355:
    a_P = (sum a_nb + sum F + m) / underU)."""
    k = np.ones(mesh.neq) if k is None else k
    D, _ = _face_diffusion(mesh, k)
    psi = stream_function(mesh.nodes, seed)
    cn = mesh.cell_nodes
    F = psi[np.roll(cn, -1, axis=1)] - psi[cn]              # outward flux per unit scale
    inner = mesh.nbr >= 0
    F = np.where(inner, F, 0.0)
    ratio = np.abs(F) / np.where(inner, D, np.inf)
    scale = peclet_max / ratio.max() if ratio.max() > 0 else 0.0
    F = F * scale
    anb = np.where(inner, D + np.maximum(-F, 0.0), 0.0)
    ad = anb.sum(axis=1) + F.sum(axis=1)
    b = np.zeros(mesh.neq)
    g = mesh.geometry()
    for side in range(4):
        on = mesh.boundary_tag == side
        if walls[side] == "dirichlet" and on.any():
            ap = np.where(on, k[:, None] * g["slen"] / g["dpf"], 0.0)
            ad += ap.sum(axis=1)
            b += ap.sum(axis=1) * wall_values[side]
    if dt is not None:
        ad = ad + g["area"] / dt
    ad = ad / under_relax
    if rhs == "random":
        b = b + (2.0 * uniform(seed, 31, mesh.neq) - 1.0)
    return System(mesh.neq, mesh.w, ad, -anb, mesh.nbr.copy(), b, symmetric=False,
                  meta=dict(kind="advdiff", peclet_max=peclet_max, flux_scale=scale,
                            face_flux=F, dt=dt, under_relax=under_relax))


# ------------------------------------------------------------------ layouts --

def to_ell(s: System) -> dict:
    """ELLPACK arrays of src/Datastruct.f:153-175 / src/EllPack.f."""
    inner = s.nbr >= 0
    ja = np.where(inner, s.nbr + 1, 1).astype(np.int32)
    al = np.where(inner, s.off, 0.0).astype(np.float64)
    return dict(matriz=4, neq=s.neq, nad=int(inner.sum()), iax=np.array([s.w, s.neq + 1], np.int32),
                ja=np.ascontiguousarray(ja.ravel()), ad=s.ad.copy(), al=np.ascontiguousarray(al.ravel()),
                au=None, b=s.b.copy(), w=s.w)


def _rows(s: System, keep):
    rows = np.repeat(np.arange(s.neq), s.w).reshape(s.neq, s.w)
    m = (s.nbr >= 0) & keep
    return rows[m], s.nbr[m], s.off[m]


def to_csrd(s: System, sym: bool = False) -> dict:
    """CSRD of src/Datastruct.f:40-77: no diagonal, columns sorted per row;
    sym keeps the lower triangle only."""
    keep = (s.nbr < np.arange(s.neq)[:, None]) if sym else np.ones_like(s.nbr, dtype=bool)
    r, c, v = _rows(s, keep)
    order = np.lexsort((c, r))
    r, c, v = r[order], c[order], v[order]
    ia = np.zeros(s.neq + 1, dtype=np.int64)
    np.add.at(ia, r + 1, 1)
    ia = np.cumsum(ia) + 1
    return dict(matriz=1, neq=s.neq, nad=int(len(c)), iax=ia.astype(np.int32),
                ja=(c + 1).astype(np.int32), ad=s.ad.copy(), al=v.astype(np.float64),
                au=None, b=s.b.copy(), sym=sym)


def to_csr(s: System) -> dict:
    """CSR of src/Datastruct.f:80-110: diagonal first then slot order; values in ad."""
    inner = s.nbr >= 0
    cnt = inner.sum(axis=1) + 1
    ia = np.concatenate([[1], 1 + np.cumsum(cnt)]).astype(np.int32)
    cols = np.concatenate([np.arange(s.neq)[:, None], s.nbr], axis=1)
    vals = np.concatenate([s.ad[:, None], s.off], axis=1)
    m = np.concatenate([np.ones((s.neq, 1), bool), inner], axis=1)
    return dict(matriz=2, neq=s.neq, nad=int(m.sum()), iax=ia, ja=(cols[m] + 1).astype(np.int32),
                ad=vals[m].astype(np.float64), al=None, au=None, b=s.b.copy())


def to_csrc(s: System, unsym: bool = True) -> dict:
    """CSRC: lower pattern sorted; al(k) = A(i, ja(k)), au(k) = A(ja(k), i)."""
    lower = s.nbr < np.arange(s.neq)[:, None]
    r, c, v = _rows(s, lower)
    order = np.lexsort((c, r))
    r, c, v = r[order], c[order], v[order]
    ia = np.zeros(s.neq + 1, dtype=np.int64)
    np.add.at(ia, r + 1, 1)
    ia = np.cumsum(ia) + 1
    d = dict(matriz=3, neq=s.neq, nad=int(len(c)), iax=ia.astype(np.int32),
             ja=(c + 1).astype(np.int32), ad=s.ad.copy(), al=v.astype(np.float64),
             au=None, b=s.b.copy(), sym=not unsym)
    if unsym:
        # value of row c, column r: find r among c's neighbours
        hit = s.nbr[c] == r[:, None]
        d["au"] = s.off[c][hit].astype(np.float64)
        assert d["au"].shape[0] == len(c)
    return d


def dense(s: System) -> np.ndarray:
    A = np.diag(s.ad.copy())
    for slot in range(s.w):
        m = s.nbr[:, slot] >= 0
        A[np.nonzero(m)[0], s.nbr[m, slot]] += s.off[m, slot]
    return A


def scipy_csr(s: System):
    import scipy.sparse as sp
    rows = np.repeat(np.arange(s.neq), s.w)
    m = s.nbr.ravel() >= 0
    A = sp.csr_matrix((s.off.ravel()[m], (rows[m], s.nbr.ravel()[m])), shape=(s.neq, s.neq))
    return (A + sp.diags(s.ad)).tocsr()


def transpose_system(s: System) -> System:
    """A^T in the same cell-major form (for the adjoint solve): entry (i, nbr)
    of A^T is A(nbr, i), found in the neighbour's slot that points back."""
    nb = np.where(s.nbr >= 0, s.nbr, 0)
    back = s.nbr[nb] == np.arange(s.neq)[:, None, None]       # (neq, w, w)
    offT = np.where(s.nbr >= 0, (s.off[nb] * back).sum(axis=2), 0.0)
    return System(s.neq, s.w, s.ad.copy(), offT, s.nbr.copy(), s.b.copy(), s.symmetric,
                  meta=dict(s.meta, transposed=True))


# ------------------------------------------------------------ named configs --

def config_poisson(n: int = 64, seed: int = SEED_BASE + 1, pin: str = "penalty") -> System:
    """Pressure-correction Poisson on an n*n uniform quad mesh, Neumann walls,
    cell 1 pinned (the 64x64 CPU-runnable case at n = 64)."""
    return pressure_system(quad_mesh(n), pin=pin, seed=seed)


def config_advdiff(n: int = 1024, seed: int = SEED_BASE + 2, peclet_max: float = 100.0,
                   dt: float | None = 0.005, under_relax: float = 0.8) -> System:
    """Nonsymmetric upwind advection-diffusion in the momentum form of a SIMPLE
    step on an n*n quad mesh with interior nodes jittered by +-10% of h:
    Dirichlet walls, transient term area/dt and under-relaxation (the decks'
    dt = 0.005, underU = 0.8), seeded RHS."""
    return advdiff_system(quad_mesh(n, jitter=0.1, seed=seed), peclet_max=peclet_max, seed=seed,
                          dt=dt, under_relax=under_relax)


def config_pressure_lognormal(n: int = 1024, seed: int = SEED_BASE + 3, triangles: bool = False,
                              sigma: float = 0.5) -> System:
    """Pressure system with lognormal(sigma) per-cell conductivity on an n*n
    jittered quad mesh, or on its 2*n*n triangles."""
    mesh = tri_mesh(n, jitter=0.1, seed=seed) if triangles else quad_mesh(n, jitter=0.1, seed=seed)
    k = lognormal_k(seed, mesh.neq, sigma)
    return pressure_system(mesh, k=k, seed=seed)


def config_batch_sides(count: int = 256, lo: int = 32, hi: int = 512, seed: int = SEED_BASE + 4):
    """Sides n = round(exp(U[ln lo, ln hi])) of the batched pressure systems."""
    u = uniform(seed, 41, count)
    return np.rint(np.exp(math.log(lo) + u * (math.log(hi) - math.log(lo)))).astype(np.int64)


def config_batch(count: int = 256, lo: int = 32, hi: int = 512, seed: int = SEED_BASE + 4):
    sides = config_batch_sides(count, lo, hi, seed)
    systems = []
    for i, n in enumerate(sides):
        sd = seed * 1000 + i
        mesh = quad_mesh(int(n))
        systems.append(pressure_system(mesh, k=lognormal_k(sd, mesh.neq), seed=sd))
    return systems


HEAT_WALLS = ("neumann", "dirichlet", "neumann", "dirichlet")
HEAT_WALL_T = (0.0, 0.0, 0.0, 1.0)


def config_heat(n: int = 512, seed: int = SEED_BASE + 5, peclet_max: float = 10.0):
    """Steady advection-diffusion heat problem: n*n jittered quads, lognormal k,
    hot (T=1) west wall, cold (T=0) east wall, adiabatic south/north walls,
    the stream-function velocity family of config_advdiff scaled to a largest
    cell Peclet number of 10 (at 100 the steady Jacobi-BiCGSTAB solve needs
    ~1e5 iterations at 256^2)."""
    mesh = quad_mesh(n, jitter=0.1, seed=seed)
    k = lognormal_k(seed, mesh.neq)
    s = advdiff_system(mesh, k=k, peclet_max=peclet_max, seed=seed, walls=HEAT_WALLS,
                       wall_values=HEAT_WALL_T, rhs="none")
    s.meta.update(kind="heat", k=k, mesh=mesh)
    return s


@dataclass
class HeatOperator:
    """The heat problem as a function of the per-cell conductivity k:
    A(k) T = b(k) + q.  Only the diffusion coefficients and the Dirichlet wall
    coefficients depend on k; the upwind convection part is fixed.

    Per (cell c, slot s):
      geo[c,s]    = |S| / ((n . xi^) |xi|)   (D_f = k_f * geo)
      alpha[c,s]  = c's share a of the normal distance (1/k_f = a/k_c + (1-a)/k_N)
      wall_g[c,s] = |S| / d_cf on Dirichlet faces, else 0 (a_P += k_c wall_g)
      wall_T[c,s] = the wall temperature there
      conv[c,s]   = max(-F_f, 0), convF[c] = sum_f F_f (fixed)
    """
    mesh: Mesh
    geo: np.ndarray
    alpha: np.ndarray
    wall_g: np.ndarray
    wall_T: np.ndarray
    conv: np.ndarray
    convF: np.ndarray

    @classmethod
    def build(cls, mesh: Mesh, k_for_flux_scale: np.ndarray, peclet_max: float, seed: int,
              walls=HEAT_WALLS, wall_values=HEAT_WALL_T) -> "HeatOperator":
        g = mesh.geometry()
        D, aux = _face_diffusion(mesh, k_for_flux_scale)
        inner = mesh.nbr >= 0
        geo = np.where(inner, g["slen"] / (aux["nk"] * aux["mxi"]), 0.0)
        psi = stream_function(mesh.nodes, seed)
        cn = mesh.cell_nodes
        F = np.where(inner, psi[np.roll(cn, -1, axis=1)] - psi[cn], 0.0)
        ratio = np.abs(F) / np.where(inner, D, np.inf)
        scale = peclet_max / ratio.max() if ratio.max() > 0 else 0.0
        F = F * scale
        wall_g = np.zeros_like(geo)
        wall_T = np.zeros_like(geo)
        for side in range(4):
            on = mesh.boundary_tag == side
            if walls[side] == "dirichlet":
                wall_g = np.where(on, g["slen"] / g["dpf"], wall_g)
                wall_T = np.where(on, wall_values[side], wall_T)
        return cls(mesh, geo, aux["alpha"], wall_g, wall_T, np.where(inner, np.maximum(-F, 0.0), 0.0),
                   F.sum(axis=1))

    def face_k(self, k: np.ndarray) -> np.ndarray:
        nb = np.where(self.mesh.nbr >= 0, self.mesh.nbr, 0)
        a = self.alpha
        return 1.0 / (a / k[:, None] + (1.0 - a) / k[nb])

    def assemble(self, k: np.ndarray, q: np.ndarray | None = None) -> System:
        inner = self.mesh.nbr >= 0
        D = np.where(inner, self.face_k(k) * self.geo, 0.0)
        anb = np.where(inner, D + self.conv, 0.0)
        ap_wall = k[:, None] * self.wall_g
        ad = anb.sum(axis=1) + self.convF + ap_wall.sum(axis=1)
        b = (ap_wall * self.wall_T).sum(axis=1)
        if q is not None:
            b = b + q
        return System(self.mesh.neq, self.mesh.w, ad, -anb, self.mesh.nbr.copy(), b,
                      meta=dict(kind="heat"))

    def grad_k(self, k: np.ndarray, T: np.ndarray, lam: np.ndarray) -> np.ndarray:
        """dL/dk = -lam^T (dA/dk T - db/dk) for the adjoint lam (A^T lam = dL/dT)."""
        inner = self.mesh.nbr >= 0
        nb = np.where(inner, self.mesh.nbr, 0)
        kf = self.face_k(k)
        dkf = kf * kf * self.alpha / (k[:, None] ** 2)
        face = -(lam[:, None] - lam[nb]) * (T[:, None] - T[nb]) * self.geo * dkf
        wall = -lam[:, None] * (T[:, None] - self.wall_T) * self.wall_g
        return np.where(inner, face, 0.0).sum(axis=1) + wall.sum(axis=1)


def ell_back_slots(nbr: np.ndarray) -> np.ndarray:
    """For slot (i, s) pointing at j, the slot t of row j that points back at
    i (int32; -1 for padding): A^T(i, s) = A(j, t)."""
    neq, w = nbr.shape
    nb = np.where(nbr >= 0, nbr, 0)
    back = np.full((neq, w), -1, dtype=np.int32)
    for t in range(w):
        hit = (nbr >= 0) & (nbr[nb, t] == np.arange(neq)[:, None])
        back = np.where(hit, t, back)
    return back


def heat_problem(n: int = 512, seed: int = SEED_BASE + 5, peclet_max: float = 10.0,
                 subset_frac: float = 0.01):
    """(operator, k, subset) of the differentiation config: the loss is
    mean(T[subset]^2) over a seeded subset of subset_frac of the cells."""
    mesh = quad_mesh(n, jitter=0.1, seed=seed)
    k = lognormal_k(seed, mesh.neq)
    op = HeatOperator.build(mesh, k, peclet_max, seed)
    m = max(1, int(round(subset_frac * mesh.neq)))
    subset = np.sort(np.argsort(uniform(seed, 51, mesh.neq))[:m])
    return op, k, subset
