"""The oracle for the post-solve step of a SIMPLE iteration: the pressure-
correction gradient of gform (src/Pform.f:466-676) for the cells that take
cell_si_p_i_ggl (types 63/65/73/75, src/Celllib.f:127-157), iws = 2: its
Green-Gauss reconstruction (src/CellibPressure.f:611-650), and simpleUpdate
(src/Simple.f:871-898).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

The restatement is vectorised over cells; every cell's arithmetic is the
reference's, operation by operation, so that results are bit-identical.
Arrays follow the reference's Fortran layout, held as C arrays with the
first Fortran index fastest: x(2, nnode) -> (nnode, 2), ix(nen+1, numel) ->
(numel, nen+1), nelcon(nen, numel) -> (numel, nen), pedge(nen+1, numel) ->
(numel, nen+1), grad(2, numel) -> (numel, 2); indices 1-based, a boundary
neighbour negative.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DROPIN_PATH = os.path.join(HERE, "_ref", "libmvf_dropin.so")

# local edge -> (first node, second node), 0-based (setcell, src/Celllib.f:541-583)
SN = {3: ((0, 1), (1, 2), (2, 0)), 4: ((0, 1), (1, 2), (2, 3), (3, 0))}


def _centroid(P, nsh):
    """cellGeom2D's centre: the vertex sum, then times 1/nshared
    (src/Celllib.f:725-733).  P: (..., nsh, 2)."""
    s = np.zeros(P.shape[:-2] + (2,))
    for l in range(nsh):
        s = s + P[..., l, :]
    return (1.0 / nsh) * s


def _cross_norm(a, b):
    """|a x b| for in-plane vectors: vet (src/Matvec.f:304-316) with the third
    components zero, then sqrt(dot_local(c, c, 3)) (src/Matvec.f:404-421)."""
    c1 = a[..., 1] * 0.0 - 0.0 * b[..., 1]
    c2 = 0.0 * b[..., 0] - a[..., 0] * 0.0
    c3 = a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]
    return np.sqrt(((0.0 + c1 * c1) + c2 * c2) + c3 * c3)


def _area(E, nsh):
    """areacell (src/Celllib.f:509-522): areatriacell (:282-297) or
    areaquadcell (:309-358), the mean of the two diagonal splits."""
    if nsh == 3:
        return 0.5 * _cross_norm(E[..., 0, :], E[..., 1, :])
    a = 0.5 * _cross_norm(E[..., 0, :], E[..., 1, :])
    a = a + 0.5 * _cross_norm(E[..., 2, :], E[..., 3, :])
    a = a + 0.5 * _cross_norm(E[..., 0, :], E[..., 3, :])
    a = a + 0.5 * _cross_norm(E[..., 1, :], E[..., 2, :])
    return 0.5 * a


def gform_p(x, ix, nelcon, pedge, u, grad):
    """New gradient and flux limiter of every cell (grad is the old gradient,
    used by the extrapolating boundary condition)."""
    numel, nsh = nelcon.shape
    sn = SN[nsh]
    X = x[ix[:, :nsh] - 1]                              # (numel, nsh, 2)
    xc = _centroid(X, nsh)
    # the central cell's edges: unit normals and centre-to-edge distances
    # (src/Celllib.f:737-808)
    n = np.empty((numel, nsh, 2))
    ca = np.empty((numel, nsh))
    for l, (a, b) in enumerate(sn):
        e = X[:, b] - X[:, a]
        meta = np.sqrt((0.0 + e[:, 0] * e[:, 0]) + e[:, 1] * e[:, 1])
        eu0, eu1 = e[:, 0] / meta, e[:, 1] / meta
        n[:, l, 0], n[:, l, 1] = eu1, -eu0
        xm = 0.5 * (X[:, b] + X[:, a])
        df = xm - xc
        ca[:, l] = df[:, 0] * n[:, l, 0] + df[:, 1] * n[:, l, 1]
    # neighbour centres, or ghost centres across boundary edges, and the
    # values there (src/CellibPressure.f:432-463, iws = 2)
    inner = nelcon > 0
    nb = np.where(inner, nelcon - 1, 0)
    xcn = _centroid(x[ix[nb, :nsh] - 1], nsh)           # (numel, nsh, 2)
    gama = 2.0 * ca
    ghost = xc[:, None, :] + gama[:, :, None] * n
    xcg = np.where(inner[:, :, None], xcn, ghost)
    uc = u[:, None]
    extrap = uc + gama * (grad[:, None, 0] * n[:, :, 0] + grad[:, None, 1] * n[:, :, 1])
    pe = pedge[:, :nsh]
    bc = np.where(pe == 3, 0.0, np.where(pe == 6, np.broadcast_to(uc, pe.shape), extrap))
    ug = np.where(inner, u[nb], bc)
    # Green-Gauss over the polygon of the neighbour centres
    # (src/CellibPressure.f:611-650)
    E = np.empty((numel, nsh, 2))
    um = np.empty((numel, nsh))
    for l, (a, b) in enumerate(sn):
        E[:, l] = xcg[:, b] - xcg[:, a]
        um[:, l] = 0.5 * (ug[:, b] + ug[:, a])
    area = _area(E, nsh)
    g0 = np.zeros(numel)
    g1 = np.zeros(numel)
    for l in range(nsh):
        g0 = g0 + um[:, l] * E[:, l, 1]
        g1 = g1 + um[:, l] * (-E[:, l, 0])
    return np.stack([g0 / area, g1 / area], axis=1), np.ones(numel)


def _cell_geometry(x, ix, nelcon):
    """cellGeom2D (src/Celllib.f:695-808) and vectorKm2d's intersection
    distance mKsiF (src/Celllib.f:596-656) for every cell and edge.  Returns
    a dict of (numel, nsh[, 2]) arrays: n (unit normal), meta (edge length),
    ca (centre-to-edge distance), ksi (unit centre-to-neighbour vector),
    mksi (its length), xi (where the centre line crosses the edge line),
    mksif (its distance from the centre), and xc (numel, 2)."""
    numel, nsh = nelcon.shape
    sn = SN[nsh]
    X = x[ix[:, :nsh] - 1]
    xc = _centroid(X, nsh)
    inner = nelcon > 0
    nb = np.where(inner, nelcon - 1, 0)
    xcn = _centroid(x[ix[nb, :nsh] - 1], nsh)
    n = np.empty((numel, nsh, 2))
    meta = np.empty((numel, nsh))
    ca = np.empty((numel, nsh))
    ksi = np.zeros((numel, nsh, 2))
    mksi = np.zeros((numel, nsh))
    mksif = np.zeros((numel, nsh))
    xis = np.zeros((numel, nsh, 2))
    with np.errstate(divide="ignore", invalid="ignore"):
        for l, (a, b) in enumerate(sn):
            e = X[:, b] - X[:, a]
            meta[:, l] = np.sqrt((0.0 + e[:, 0] * e[:, 0]) + e[:, 1] * e[:, 1])
            n[:, l, 0] = e[:, 1] / meta[:, l]
            n[:, l, 1] = -(e[:, 0] / meta[:, l])
            xm = 0.5 * (X[:, b] + X[:, a])
            ca[:, l] = (xm[:, 0] - xc[:, 0]) * n[:, l, 0] + (xm[:, 1] - xc[:, 1]) * n[:, l, 1]
            k = xcn[:, l] - xc
            mk = np.sqrt((0.0 + k[:, 0] * k[:, 0]) + k[:, 1] * k[:, 1])
            mksi[:, l] = np.where(inner[:, l], mk, 0.0)
            ksi[:, l, 0] = np.where(inner[:, l], k[:, 0] / mk, 0.0)
            ksi[:, l, 1] = np.where(inner[:, l], k[:, 1] / mk, 0.0)
            # vectorKm2d: the line through the two centres (a ghost centre
            # across a wall) against the line of edge l
            gama = 2.0 * ca[:, l]
            gx = np.where(inner[:, l], xcn[:, l, 0], xc[:, 0] + gama * n[:, l, 0])
            gy = np.where(inner[:, l], xcn[:, l, 1], xc[:, 1] + gama * n[:, l, 1])
            a1 = gy - xc[:, 1]
            b1 = xc[:, 0] - gx
            c1 = gx * xc[:, 1] - xc[:, 0] * gy
            a2 = X[:, a, 1] - X[:, b, 1]
            b2 = X[:, b, 0] - X[:, a, 0]
            c2 = X[:, a, 0] * X[:, b, 1] - X[:, b, 0] * X[:, a, 1]
            det = a1 * b2 - a2 * b1
            if np.any(det[inner[:, l]] == 0.0):
                raise ValueError("vectorKm2d: a centre line parallel to its edge")
            r = 1.0 / det
            f1, f2 = -c1, -c2
            xi0 = (r * b2) * f1 + (-r * b1) * f2
            xi1 = (-r * a2) * f1 + (r * a1) * f2
            s0, s1 = xi0 - xc[:, 0], xi1 - xc[:, 1]
            mksif[:, l] = np.sqrt(s0 * s0 + s1 * s1)
            xis[:, l, 0], xis[:, l, 1] = xi0, xi1
    return dict(n=n, meta=meta, ca=ca, ksi=ksi, mksi=mksi, mksif=mksif, xc=xc, xi=xis)


def pform_pc(x, ix, nelcon, pedge, num, e, sedge, w, d, gradp, p, int_vel=5):
    """The pressure-correction system as pform assembles it for SIMPLE
    (src/Simple.f:583-588: iws = 1, lib = 2, forces, matrix and residual),
    cell by cell through cell_si_p_i_ggl (src/CellibPressure.f:390-571) and
    into ELLPACK by ellPack (src/EllPack.f:68-93).

    int_vel selects the face velocity (the /simple_/ intVel, 5 in
    src/Simple.f:186): 1 linear, 2 Rhie-Chow, 5 Ferziger-Peric
    (cellVel, src/CellibSimple.f:1947-2070).  e (nmat, 10) holds the
    material properties (density in column 2), sedge (numel, nsh+1, 3)
    the prescribed edge values, num (numel) the 1-based equation numbers.
    Returns (ad (neq), al (neq, nsh), b (neq), rescell (numel)), neq =
    num.max(); al holds -a_nb per edge slot and 0 across walls."""
    numel, nsh = nelcon.shape
    G = _cell_geometry(x, ix, nelcon)
    n, meta, ca, ksi, mksi, mksif = G["n"], G["meta"], G["ca"], G["ksi"], G["mksi"], G["mksif"]
    inner = nelcon > 0
    nb = np.where(inner, nelcon - 1, 0)
    rho = e[ix[:, nsh] - 1, 1]                           # k(2, idCell)
    rhon = rho[nb]
    pe = pedge[:, :nsh]
    pz = np.zeros(numel)
    sp = np.zeros(numel)
    a = np.zeros((numel, nsh))
    with np.errstate(divide="ignore", invalid="ignore"):
        for i in range(nsh):
            n0, n1 = n[:, i, 0], n[:, i, 1]
            # walls: prescribed velocity (2) and locally parabolic (4) add the
            # prescribed flux; prescribed pressure (3) a diffusion term and
            # the cell's own flux (src/CellibPressure.f:476-507)
            wfn = sedge[:, i, 0] * n0 + sedge[:, i, 1] * n1
            cv_edge = (rho * wfn) * meta[:, i]
            ndd3 = (d[:, 0] * n0) * n0 + (d[:, 1] * n1) * n1
            dfd3 = ((rho * ndd3) * meta[:, i]) / ca[:, i]
            wfc = w[:, 0] * n0 + w[:, 1] * n1
            cv_cell = (rho * wfc) * meta[:, i]
            wall = ~inner[:, i]
            pz = np.where(wall & ((pe[:, i] == 2) | (pe[:, i] == 4)), pz - cv_edge, pz)
            sp = np.where(wall & (pe[:, i] == 3), sp + dfd3, sp)
            pz = np.where(wall & (pe[:, i] == 3), pz - cv_cell, pz)
            # inner faces (src/CellibPressure.f:511-553)
            alpha = mksif[:, i] / mksi[:, i]
            om = 1.0 - alpha
            dd0 = om * d[:, 0] + alpha * d[nb[:, i], 0]
            dd1 = om * d[:, 1] + alpha * d[nb[:, i], 1]
            nk = n0 * ksi[:, i, 0] + n1 * ksi[:, i, 1]
            if int_vel == 5:
                kf = om * rho + alpha * rhon[:, i]
                ndd = 0.5 * (dd0 + dd1)
            else:
                kf = rho
                ndd = (dd0 * n0) * n0 + (dd1 * n1) * n1
            dfd = ((kf * meta[:, i]) * ndd) / (nk * mksi[:, i])
            # cellVel: the face velocity (src/CellibSimple.f:1956-2036)
            v0 = om * w[:, 0] + alpha * w[nb[:, i], 0]
            v1 = om * w[:, 1] + alpha * w[nb[:, i], 1]
            wf = v0 * n0 + v1 * n1
            if int_vel in (2, 5):
                gf0 = om * gradp[:, 0] + alpha * gradp[nb[:, i], 0]
                gf1 = om * gradp[:, 1] + alpha * gradp[nb[:, i], 1]
                if int_vel == 2:
                    gfp1 = (p[nb[:, i]] - p) / mksi[:, i]
                    gfp2 = gf0 * ksi[:, i, 0] + gf1 * ksi[:, i, 1]
                    wf = wf + (0.5 * (dd0 + dd1)) * (gfp2 - gfp1)
                else:
                    gfp1 = p[nb[:, i]] - p
                    ks0, ks1 = mksi[:, i] * ksi[:, i, 0], mksi[:, i] * ksi[:, i, 1]
                    nks = ks0 * n0 + ks1 * n1
                    gfp2 = gf0 * ks0 + gf1 * ks1
                    gfp2 = (gfp2 - gfp1) / nks
                    wf = wf + (0.5 * (dd0 + dd1)) * gfp2
            elif int_vel != 1:
                raise ValueError("int_vel must be 1, 2 or 5")
            wf = wf * meta[:, i]
            a[:, i] = np.where(inner[:, i], dfd, 0.0)
            pz = np.where(inner[:, i], pz - kf * wf, pz)
    # the pin of the pressure level (src/CellibPressure.f:557)
    sp = np.where(pedge[:, nsh] == -1, sp + 1.0e60, sp)
    apc = sp
    for i in range(nsh):
        apc = apc + a[:, i]
    # ellPack: f(n) = p, ad(n) = a_P, a(j, n) = -a_j where the neighbour has
    # an equation (src/EllPack.f:78-90)
    neq = int(num.max())
    ad = np.zeros(neq)
    al = np.zeros((neq, nsh))
    b = np.zeros(neq)
    eq = num - 1
    ld = np.where(inner, num[nb], 0)
    ad[eq] = apc
    b[eq] = pz
    al[eq] = np.where(ld > 0, -a, 0.0)
    return ad, al, b, pz.copy()


def _tvd_limit(r, icod):
    """limit (src/Celllib.f:370-431): the TVD limiters of the momentum
    convection, by code (1 upwind ... 7 SuperBee)."""
    if icod == 1:
        return np.zeros_like(r)
    if icod == 2:
        return np.ones_like(r)
    if icod == 3:
        return (r + np.abs(r)) / (1.0 + np.abs(r))
    if icod == 4:
        return np.where(r > 0.0, (r + r * r) / (1.0 + r * r), 0.0)
    if icod == 5:
        return np.maximum(0.0, np.minimum(r, 1.0))
    if icod == 6:
        return np.maximum(0.0, np.minimum(2.0, r))
    if icod == 7:
        return np.maximum(0.0, np.maximum(np.minimum(2.0 * r, 1.0), np.minimum(r, 2.0)))
    raise ValueError("limiter code 1..7")


def pform_mom(x, ix, nelcon, pedge, num, e, sedge, u, pres, w, grad1, grad2, gradp, d, comp, dt,
              under_u=1.0, s_pressure=True, simple_c=False, bs=False):
    """A momentum system as pform assembles it for SIMPLE (src/Simple.f:
    425-431 for u1, :456-462 for u2: iws = 1, lib = 1, forces, matrix,
    momentum, D field and residual), cell by cell through cell_si_eb_i_ggl
    (src/CellibSimple.f:591-901) and into ELLPACK by ellPack (src/EllPack.f:
    68-93).

    comp is the velocity component (iws1 = 1 or 2); u (numel) that
    component, pres (numel) the pressure (pform's u0), w (numel, 2) the
    velocity, grad1 the gradient of u, grad2 that of the other component,
    gradp that of the pressure (numel, 2 each); e (nmat, 10) the material
    properties (viscosity, density, ..., column 6 the transient coefficient,
    column 7 the limiter code); sedge (numel, nsh+1, 3) the prescribed edge
    values; d (numel, 2) the D field (component comp is replaced).
    under_u, s_pressure, simple_c: the /simple_/ underU, sPressure, simpleC.
    Returns (ad, al, b, rescell, im (numel, 4) with columns comp and comp+2
    set, d) -- im and d start from zero and from the given d."""
    numel, nsh = nelcon.shape
    G = _cell_geometry(x, ix, nelcon)
    n, meta, ca, ksi, mksi, mksif = G["n"], G["meta"], G["ca"], G["ksi"], G["mksi"], G["mksif"]
    xc = G["xc"]
    sn = SN[nsh]
    X = x[ix[:, :nsh] - 1]
    inner = nelcon > 0
    nb = np.where(inner, nelcon - 1, 0)
    mat = ix[:, nsh] - 1
    visc, rho, cc = e[mat, 0], e[mat, 1], e[mat, 5]
    icod = e[mat, 6].astype(np.int64)          # integer = real*8 truncates
    c = comp - 1
    pe = pedge[:, :nsh]
    p = np.zeros(numel)
    pf = np.zeros(numel)
    sp = np.zeros(numel)
    a = np.zeros((numel, nsh))
    eps = 1.0e-14
    with np.errstate(divide="ignore", invalid="ignore"):
        # the central cell's area from its raw edge vectors (src/Celllib.f:
        # 736-744)
        E = np.stack([X[:, b] - X[:, a_] for a_, b in sn], axis=1)
        area = _area(E, nsh)
        for i in range(nsh):
            n0, n1 = n[:, i, 0], n[:, i, 1]
            ni = n[:, i, c]
            xm = 0.5 * (X[:, sn[i][1]] + X[:, sn[i][0]])
            df0, df1 = xm[:, 0] - xc[:, 0], xm[:, 1] - xc[:, 1]
            wall = ~inner[:, i]
            # ---- walls (src/CellibSimple.f:693-768)
            ap = visc * meta[:, i]
            if comp == 1:
                vw = ap * (grad1[:, 0] * n0 + grad2[:, 0] * n1)
            else:
                vw = ap * (grad1[:, 1] * n1 + grad2[:, 1] * n0)
            pw = p + vw
            pface_w = np.where(pe[:, i] == 3, sedge[:, i, 2],
                               pres + (gradp[:, 0] * df0 + gradp[:, 1] * df1) if s_pressure else pres)
            pfw = pf + (pface_w * meta[:, i]) * ni
            spw = sp.copy()
            apw = (visc * meta[:, i]) / ca[:, i]
            spw = np.where((pe[:, i] == 0) | (pe[:, i] == 1), spw + apw, spw)
            pw = np.where(pe[:, i] == 1, pw + apw * sedge[:, i, c], pw)
            wfn2 = sedge[:, i, 0] * n0 + sedge[:, i, 1] * n1
            pw = np.where(pe[:, i] == 2, pw - ((rho * wfn2) * meta[:, i]) * sedge[:, i, c], pw)
            wfc = w[:, 0] * n0 + w[:, 1] * n1
            cvc_ = (rho * wfc) * meta[:, i]
            big = np.abs(wfc) > 1.0e-7
            spw = np.where((pe[:, i] == 3) & big & (wfc > 0.0), spw + cvc_, spw)
            pw = np.where((pe[:, i] == 3) & big & (wfc < 0.0), pw + cvc_ * wfc, pw)
            spw = np.where(pe[:, i] == 4, spw + cvc_, spw)
            # ---- inner faces (src/CellibSimple.f:772-857)
            k = nb[:, i]
            alpha = mksif[:, i] / mksi[:, i]
            om = 1.0 - alpha
            kf = alpha / visc + om / visc[k]
            kf = 1.0 / kf
            wf0 = om * w[:, 0] + alpha * w[k, 0]
            wf1 = om * w[:, 1] + alpha * w[k, 1]
            gf0 = om * grad1[:, 0] + alpha * grad1[k, 0]
            gf1 = om * grad1[:, 1] + alpha * grad1[k, 1]
            k0, k1 = ksi[:, i, 0], ksi[:, i, 1]
            nk = n0 * k0 + n1 * k1
            wfn = wf0 * n0 + wf1 * n1
            gfk = gf0 * k0 + gf1 * k1
            corr = (u[k] - u) / mksi[:, i] - gfk
            gf0 = gf0 + corr * k0
            gf1 = gf1 + corr * k1
            gfn = gf0 * n0 + gf1 * n1
            gfk = gf0 * k0 + gf1 * k1
            dfd = (kf * meta[:, i]) / (nk * mksi[:, i])
            cd = (kf * meta[:, i]) * (gfn - gfk / nk)
            ai = dfd
            pi = p + cd
            cv = (rho * wfn) * meta[:, i]
            lim = np.empty(numel)
            neg = cv < 0.0
            du_n = (u - u[k]) + eps
            gpk_n = grad1[k, 0] * (xc[:, 0] - xc[k, 0]) + grad1[k, 1] * (xc[:, 1] - xc[k, 1])
            r_n = (2.0 * gpk_n) / du_n - 1.0
            du_p = (u[k] - u) + eps
            gpk_p = (grad1[:, 0] * k0 + grad1[:, 1] * k1) * mksi[:, i]
            r_p = (2.0 * gpk_p) / du_p - 1.0
            r = np.where(neg, r_n, r_p)
            for code in np.unique(icod):
                sel = icod == code
                lim[sel] = _tvd_limit(r[sel], int(code))
            cvc = np.where(neg, (om * lim) * du_n, (alpha * lim) * du_p)
            km0 = xm[:, 0] - G["xi"][:, i, 0]
            km1 = xm[:, 1] - G["xi"][:, i, 1]
            cvc = cvc + gf0 * km0 + gf1 * km1
            ai = ai + np.maximum(-cv, 0.0)
            spi = sp + cv
            pi = pi - cv * cvc
            if s_pressure:
                p1 = (pres + gradp[:, 0] * df0) + gradp[:, 1] * df1
                p2 = (pres[k] + gradp[k, 0] * (xm[:, 0] - xc[k, 0])) + gradp[k, 1] * (xm[:, 1] - xc[k, 1])
                pface = 0.5 * (p1 + p2)
            else:
                pface = om * pres + alpha * pres[k]
            pfi = pf + (pface * meta[:, i]) * ni
            g0 = om * grad1[:, c] + alpha * grad1[k, c]
            g1 = om * grad2[:, c] + alpha * grad2[k, c]
            if comp == 1:
                pi = pi + (kf * meta[:, i]) * (g0 * n0 + g1 * n1)
            else:
                pi = pi + (kf * meta[:, i]) * (g0 * n1 + g1 * n0)
            a[:, i] = np.where(wall, 0.0, ai)
            p = np.where(wall, pw, pi)
            pf = np.where(wall, pfw, pfi)
            sp = np.where(wall, spw, spi)
    # the pin (src/CellibSimple.f:861), a_P, the transient and the
    # under-relaxation (:866-883)
    sp = np.where(pedge[:, nsh] == -2, sp + 1.0e60, sp)
    apc = sp
    for i in range(nsh):
        apc = apc + a[:, i]
    m = ((rho * cc) * area) / dt
    if bs:
        m = 1.5 * m
    apc = apc + m
    apc = apc / under_u
    p = p + ((1.0 - under_u) * apc) * u
    rm = np.zeros(numel)
    anb = np.zeros(numel)
    for i in range(nsh):
        rm = rm + a[:, i] * np.where(inner[:, i], u[nb[:, i]], 0.0)
        anb = anb + a[:, i]
    im = np.zeros((numel, 4))
    im[:, c] = rm + p
    im[:, c + 2] = apc
    d = np.array(d, dtype=np.float64, copy=True)
    d[:, c] = area / (apc - anb) if simple_c else area / apc
    p = p - pf
    rm = (p + rm) - apc * u
    neq = int(num.max())
    ad = np.zeros(neq)
    al = np.zeros((neq, nsh))
    b = np.zeros(neq)
    eq = num - 1
    ld = np.where(inner, num[nb], 0)
    ad[eq] = apc
    b[eq] = p
    al[eq] = np.where(ld > 0, -a, 0.0)
    return ad, al, b, rm, im, d


def simple_update(w, pre, u1, u2, prec, gradpc, d, under_pc, simple_r):
    """simpleUpdate (src/Simple.f:871-898): w = u* - d grad p'; p += underPc p'
    (SIMPLE only).  Returns new (w, pre); w and d are (numel, 2)."""
    w = np.array(w, dtype=np.float64, copy=True)
    pre = np.array(pre, dtype=np.float64, copy=True)
    w[:, 0] = u1 - d[:, 0] * gradpc[:, 0]
    w[:, 1] = u2 - d[:, 1] * gradpc[:, 1]
    if not simple_r:
        pre = pre + under_pc * prec
    return w, pre


# ------------------------------------------------ the reference's own code --

def have_dropin() -> bool:
    return os.path.exists(DROPIN_PATH)


def _dropin():
    L = C.CDLL(DROPIN_PATH, mode=C.RTLD_LOCAL)
    return L


class _OmpControl(C.Structure):
    # COMMON /omp_control/ (include/openmp.fi:16-18)
    _fields_ = [("num_threads", C.c_int32), ("openmpCell", C.c_int32), ("openmpSolver", C.c_int32),
                ("nThreadsCell", C.c_int32), ("nThreadsSolver", C.c_int32), ("nThreadsDot", C.c_int32),
                ("nThreadsMatVec", C.c_int32)]


def ref_gform_p(x, ix, nelcon, pedge, u, grad, cell_type):
    """The reference's gform on the same arrays (sequential loop: openmpCell
    false), ndf = 1, iws = iws1 = 2, lib = 2: returns (grad, fluxl)."""
    L = _dropin()
    ctl = _OmpControl.in_dll(L, "omp_control_")
    ctl.openmpCell = 0
    numel, nsh = nelcon.shape
    nst = 3
    f64 = lambda a: np.asfortranarray(a, dtype=np.float64)
    i32 = lambda a: np.asfortranarray(a, dtype=np.int32)
    # C (numel, k) arrays are Fortran (k, numel): pass their transposes' memory
    xa = np.ascontiguousarray(x, dtype=np.float64)
    ixa = np.ascontiguousarray(ix, dtype=np.int32)
    nca = np.ascontiguousarray(nelcon, dtype=np.int32)
    pea = np.ascontiguousarray(pedge, dtype=np.int32)
    ua = np.ascontiguousarray(u, dtype=np.float64)
    g = np.ascontiguousarray(grad, dtype=np.float64).copy()
    fluxl = np.zeros(numel)
    sedge = np.zeros((numel, nsh + 1, nst))
    e = np.ones((1, 10))
    ls = np.zeros((numel, nsh, 2))
    ie = np.array([cell_type], np.int32)
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    ri = lambda v: C.byref(C.c_int32(v))
    L.gform_.restype = None
    L.gform_(p(ua), p(g), p(fluxl), p(xa), p(sedge), p(e), p(ls), p(ie), p(nca), p(pea), p(ixa),
             ri(numel), ri(2), ri(nsh), ri(nsh), ri(1), ri(nst), ri(2), ri(2), ri(2))
    del f64, i32
    return g, fluxl


def ref_simple_update(w, pre, u1, u2, prec, gradpc, d, under_u, under_pc, simple_r):
    """The reference's simpleUpdate on the same arrays: returns (w, pre)."""
    L = _dropin()
    numel = len(u1)
    wa = np.ascontiguousarray(w, dtype=np.float64).copy()
    pa = np.ascontiguousarray(pre, dtype=np.float64).copy()
    p = lambda a: np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(C.c_void_p)
    keep = [np.ascontiguousarray(a, dtype=np.float64) for a in (u1, u2, prec, gradpc, d)]
    L.simpleupdate_.restype = None
    L.simpleupdate_(wa.ctypes.data_as(C.c_void_p), pa.ctypes.data_as(C.c_void_p),
                    *[k.ctypes.data_as(C.c_void_p) for k in keep],
                    C.byref(C.c_int32(numel)), C.byref(C.c_int32(2)),
                    C.byref(C.c_double(under_u)), C.byref(C.c_double(under_pc)),
                    C.byref(C.c_int32(1 if simple_r else 0)))
    del p
    return wa, pa


def _simple_common(L):
    """COMMON /simple_/ (include/simple.fi): sixteen real*8 (underU, underP,
    underPc, underRo, g(3), ... fluxoM), intVel, maxItSimple, then the
    logicals in COMMON order."""
    class _Simple(C.Structure):
        _fields_ = [("reals", C.c_double * 16), ("intVel", C.c_int32), ("maxItSimple", C.c_int32),
                    ("simpleC", C.c_int32), ("simpleR", C.c_int32), ("singularPressure", C.c_int32),
                    ("vMass", C.c_int32), ("tDinamico", C.c_int32), ("sPressure", C.c_int32),
                    ("bBodyPressure", C.c_int32), ("closed", C.c_int32), ("skewnessCorrection", C.c_int32)]
    return _Simple.in_dll(L, "simple__")


def ref_pform_pc(x, ix, nelcon, pedge, num, e, sedge, w, d, gradp, p, cell_type, int_vel=5):
    """The reference's own pform on the same arrays, as src/Simple.f:583-588
    calls it for the pressure correction (iws = 1, iws1 = 2, lib = 2,
    forces, matrix and residual; ELLPACK, struc = 4), with openmpCell false
    and the given intVel.  Returns (ad, al, b, rescell) as pform_pc."""
    L = _dropin()
    ctl = _OmpControl.in_dll(L, "omp_control_")
    ctl.openmpCell = 0
    sim = _simple_common(L)
    saved = sim.intVel
    sim.intVel = int(int_vel)
    numel, nsh = nelcon.shape
    nnode = len(x)
    nst = 3
    neq = int(num.max())
    c64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    c32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
    ad = np.zeros(neq)
    al = np.zeros((neq, nsh))
    au = np.zeros(1)
    b = np.zeros(neq)
    rescell = np.zeros(numel)
    keep = dict(
        ia=np.zeros(2, np.int32), ja=np.zeros(1, np.int32), num=c32(num), un=np.zeros(nnode),
        u=np.zeros(numel), u1=c64(p), grad=np.zeros((numel, 2)), grad1=c64(gradp), div=np.zeros(numel),
        fluxl=np.zeros(numel), im=np.zeros((numel, 4)), ro=np.ones((numel, 3)), w=c64(w), x=c64(x),
        sedge=c64(sedge), e=c64(e), ddu=c64(d), eddy=np.zeros(numel),
        ie=np.full(len(e), cell_type, np.int32), nelcon=c32(nelcon), pedge=c32(pedge), ix=c32(ix))
    k = keep
    P = lambda a: a.ctypes.data_as(C.c_void_p)
    ri = lambda v: C.byref(C.c_int32(int(v)))
    rd = lambda v: C.byref(C.c_double(float(v)))
    L.pform_.restype = None
    try:
        L.pform_(P(ad), P(au), P(al), P(k["ia"]), P(k["ja"]), P(k["num"]), P(k["un"]),
                 P(k["u"]), P(k["u1"]), P(k["u1"]), P(k["grad"]), P(k["grad1"]), P(k["grad"]),
                 P(k["div"]), P(k["fluxl"]), P(k["im"]), P(k["ro"]), P(k["w"]), P(b), P(rescell),
                 P(k["x"]), P(k["sedge"]), P(k["e"]), P(k["ddu"]), P(k["eddy"]),
                 P(k["ie"]), P(k["nelcon"]), P(k["pedge"]), P(k["ix"]),
                 ri(numel), ri(2), ri(nsh), ri(nsh), ri(1), ri(nst), rd(0.0), rd(0.0),
                 ri(4), ri(1), ri(1), ri(2), ri(2), ri(1), ri(1), ri(0), ri(0), ri(1), ri(0))
    finally:
        sim.intVel = saved
    return ad, al, b, rescell


def ref_pform_mom(x, ix, nelcon, pedge, num, e, sedge, u, pres, w, grad1, grad2, gradp, d, comp, dt,
                  cell_type, under_u=1.0, s_pressure=True, simple_c=False, bs=False):
    """The reference's own pform for momentum component comp, as
    src/Simple.f:425-431,456-462 call it (iws = 1, iws1 = comp, lib = 1,
    forces, matrix, momentum, D field, residual; ELLPACK), with openmpCell
    false and /simple_/ underU, sPressure and simpleC set.  Returns
    (ad, al, b, rescell, im, d) as pform_mom (im starts from zero)."""
    L = _dropin()
    ctl = _OmpControl.in_dll(L, "omp_control_")
    ctl.openmpCell = 0
    sim = _simple_common(L)
    saved = (sim.reals[0], sim.sPressure, sim.simpleC)
    sim.reals[0] = float(under_u)
    sim.sPressure = 1 if s_pressure else 0
    sim.simpleC = 1 if simple_c else 0
    numel, nsh = nelcon.shape
    nnode = len(x)
    neq = int(num.max())
    c64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    c32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
    ad = np.zeros(neq)
    al = np.zeros((neq, nsh))
    au = np.zeros(1)
    b = np.zeros(neq)
    rescell = np.zeros(numel)
    im = np.zeros((numel, 4))
    dd = c64(d).copy()
    k = dict(
        ia=np.zeros(2, np.int32), ja=np.zeros(1, np.int32), num=c32(num), un=np.zeros(nnode), u=c64(u),
        pres=c64(pres), g1=c64(grad1), g2=c64(grad2), gp=c64(gradp), div=np.zeros(numel),
        fluxl=np.zeros(numel), ro=np.ones((numel, 3)), w=c64(w), x=c64(x), sedge=c64(sedge), e=c64(e),
        eddy=np.zeros(numel), ie=np.full(len(e), cell_type, np.int32), nelcon=c32(nelcon), pedge=c32(pedge),
        ix=c32(ix))
    P = lambda a: a.ctypes.data_as(C.c_void_p)
    ri = lambda v: C.byref(C.c_int32(int(v)))
    rd = lambda v: C.byref(C.c_double(float(v)))
    L.pform_.restype = None
    try:
        L.pform_(P(ad), P(au), P(al), P(k["ia"]), P(k["ja"]), P(k["num"]), P(k["un"]),
                 P(k["u"]), P(k["pres"]), P(k["pres"]), P(k["g1"]), P(k["g2"]), P(k["gp"]),
                 P(k["div"]), P(k["fluxl"]), P(im), P(k["ro"]), P(k["w"]), P(b), P(rescell),
                 P(k["x"]), P(k["sedge"]), P(k["e"]), P(dd), P(k["eddy"]),
                 P(k["ie"]), P(k["nelcon"]), P(k["pedge"]), P(k["ix"]),
                 ri(numel), ri(2), ri(nsh), ri(nsh), ri(1), ri(3), rd(dt), rd(0.0),
                 ri(4), ri(1), ri(1), ri(comp), ri(1), ri(1), ri(1), ri(1), ri(1), ri(1), ri(1 if bs else 0))
    finally:
        sim.reals[0], sim.sPressure, sim.simpleC = saved
    return ad, al, b, rescell, im, dd
