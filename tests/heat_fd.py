"""Central finite differences of the heat problem's loss, shared by the
adjoint tests and the fixture script.

L(k) = mean(T[S]^2) with A(k) T = b(k).  For a perturbation dk of cell c,
T(k + dk e_c) = T + d with d = A(k + dk e_c)^-1 r and
r = (b(k + dk e_c) - b(k)) - (A(k + dk e_c) - A(k)) T, formed face by face.
"""
import numpy as np

from fortranfvcell_amd import problems as P


def perturbation_rhs(op, k, T, c, dk):
    """r = (b(k + dk e_c) - b(k)) - (A(k + dk e_c) - A(k)) T, one face at a time.

    A face's diffusion coefficient is D = geo / u with u = a/k_P + (1-a)/k_N;
    raising k_c by dk changes 1/u by a dk / (k_c (k_c + dk) u u') on c's own
    faces (c is P, share a) and by (1-a) dk / (...) on the neighbours' faces
    that point back at c (c is N there).  Forming the change this way instead
    of subtracting assembled coefficients avoids the eps/dk relative rounding
    of that subtraction, which otherwise leaves a non-conservative residual
    that swamps the finite difference on cells whose gradient is a near
    cancellation (lambda_c ~ lambda_N far from the loss cells)."""
    nbr = op.mesh.nbr
    back = P.ell_back_slots(nbr)
    r = np.zeros(op.mesh.neq)
    kc, kp = k[c], k[c] + dk
    for s in range(nbr.shape[1]):
        n = nbr[c, s]
        if n >= 0:
            a = op.alpha[c, s]
            u, up = a / kc + (1 - a) / k[n], a / kp + (1 - a) / k[n]
            r[c] -= a * dk / (kc * kp * u * up) * op.geo[c, s] * (T[c] - T[n])
            t = back[c, s]
            a2 = op.alpha[n, t]
            u2, u2p = a2 / k[n] + (1 - a2) / kc, a2 / k[n] + (1 - a2) / kp
            r[n] -= (1 - a2) * dk / (kc * kp * u2 * u2p) * op.geo[n, t] * (T[n] - T[c])
        else:
            r[c] += dk * op.wall_g[c, s] * (op.wall_T[c, s] - T[c])
    return r


def loss_difference(T, subset, dp, dm, h):
    """(L(k+h) - L(k-h)) / 2h from the two solution corrections on S, as
    mean((d+ - d-) (2 T + d+ + d-)): equal to the difference of the two
    losses without cancelling against L itself."""
    return np.mean((dp - dm) * (2.0 * T[subset] + dp + dm)) / (2 * h)


def fd_gradient_lu(op, k, T, subset, cells, rel_step=1e-4, refine=6):
    """Central differences with the perturbed systems solved by sparse LU of
    A(k) plus iterative refinement on A(k +- h e_c)."""
    import scipy.sparse.linalg as sla
    A = P.scipy_csr(op.assemble(k)).tocsc()
    lu = sla.splu(A)
    out = []
    for c in cells:
        h = rel_step * k[c]
        d = []
        for sgn in (1.0, -1.0):
            kk = k.copy()
            kk[c] += sgn * h
            Ap = P.scipy_csr(op.assemble(kk))
            r = perturbation_rhs(op, k, T, c, sgn * h)
            x = lu.solve(r)
            for _ in range(refine):
                x = x + lu.solve(r - Ap @ x)
            d.append(x[subset])
        out.append(loss_difference(T, subset, d[0], d[1], h))
    return np.array(out)
