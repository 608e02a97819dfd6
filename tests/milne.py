"""Tensor calculus in Milne coordinates (tau, x, y, eta), written from the
definitions, for pinning the validity diagnostics by a route of their own.

Metric g = diag(1, -1, -1, -tau^2); Christoffel symbols Gamma^tau_{eta eta} =
tau, Gamma^eta_{tau eta} = Gamma^eta_{eta tau} = 1/tau.  Given a lattice's
velocity at two time levels, every quantity is a plain index contraction:
    D_mu u^nu = d_mu u^nu + Gamma^nu_{mu l} u^l     (covariant derivative)
    theta = D_mu u^mu,  Delta^{mu nu} = g^{mu nu} - u^mu u^nu
    grad^mu u^nu = Delta^{mu a} D_a u^nu
    sigma^{mu nu} = (grad^mu u^nu + grad^nu u^mu)/2 - Delta^{mu nu} theta/3
    z^mu = (tau u^eta, 0, 0, u^tau/tau) / sqrt(1 + u_x^2 + u_y^2)
    Xi^{mu nu} = Delta^{mu nu} + z^mu z^nu     (transverse projector)
    A = z_a z_b sigma^{ab},  B^mu = z_a sigma^{a mu}
    theta_T = 2 theta / 3 + A,  zDzu = theta - theta_T
    sigma_T^{mu nu} = sigma^{mu nu} + z^mu B^nu + B^mu z^nu + z^mu z^nu A - Xi^{mu nu} A / 2
The last two lines are Xi_{mu nu} grad^mu u^nu and
Xi^mu_a Xi^nu_b sigma^{ab} - Xi^{mu nu} (Xi_{ab} sigma^{ab})/2 for a sigma that is
transverse to u and traceless, as it is in the continuum; the lattice's
differences break those identities at O(dtau) and O(dx^2) (u_nu D u^nu != 0),
and the projections are taken in the form above, the one the VAH equations
(and the reference's diagnostics) use.  The time derivative is the
lattice's backward difference (u - u_prev)/dtau, the spatial ones are centred
differences with zero-gradient ghost cells -- the discretisation of the
reference's diagnostics -- so what is compared is the algebra."""
import numpy as np


def metric(tau):
    return np.diag([1.0, -1.0, -1.0, -tau * tau])


def gradients(u, up, shape, tau, dt, dx, dy, dz):
    """d_mu u^nu per interior cell: array (n, 4 [mu], 4 [nu]).  u, up: (4, n)
    interior arrays (x fastest), shape = (nz, ny, nx)."""
    nz, ny, nx = shape
    n = nx * ny * nz
    d = np.zeros((n, 4, 4))
    d[:, 0, :] = ((u - up) / dt).T
    for c in range(4):
        f = np.pad(u[c].reshape(shape), 1, mode="edge")
        d[:, 1, c] = ((f[1:-1, 1:-1, 2:] - f[1:-1, 1:-1, :-2]) / (2 * dx)).ravel()
        d[:, 2, c] = ((f[1:-1, 2:, 1:-1] - f[1:-1, :-2, 1:-1]) / (2 * dy)).ravel()
        if nz > 1:
            d[:, 3, c] = ((f[2:, 1:-1, 1:-1] - f[:-2, 1:-1, 1:-1]) / (2 * dz)).ravel()
    return d


def kinematics(u, up, shape, tau, dt, dx, dy, dz):
    """theta, theta_T, zDzu, sigma_T^{mu nu}, Xi^{mu nu} per cell."""
    g = metric(tau)
    gi = np.linalg.inv(g)
    U = u.T  # (n, 4)
    n = U.shape[0]
    Dm = gradients(u, up, shape, tau, dt, dx, dy, dz)
    Gam = np.zeros((4, 4, 4))  # Gamma^nu_{mu l}
    Gam[0, 3, 3] = tau
    Gam[3, 0, 3] = Gam[3, 3, 0] = 1.0 / tau
    D = Dm + np.einsum("vml,nl->nmv", Gam, U)  # D_mu u^nu
    theta = np.einsum("nmm->n", D)
    Delta = gi[None] - np.einsum("nm,nv->nmv", U, U)
    grad = np.einsum("nma,nav->nmv", Delta, D)
    sigma = 0.5 * (grad + grad.transpose(0, 2, 1)) - Delta * theta[:, None, None] / 3
    F = 1 + U[:, 1] ** 2 + U[:, 2] ** 2
    z = np.zeros((n, 4))
    z[:, 0] = tau * U[:, 3] / np.sqrt(F)
    z[:, 3] = U[:, 0] / tau / np.sqrt(F)
    Xi = Delta + np.einsum("nm,nv->nmv", z, z)
    z_lo = z @ g
    A = np.einsum("na,nb,nab->n", z_lo, z_lo, sigma)
    B = np.einsum("na,nam->nm", z_lo, sigma)
    theta_T = 2 * theta / 3 + A
    zz = np.einsum("nm,nv->nmv", z, z)
    sigma_T = (sigma + np.einsum("nm,nv->nmv", z, B) + np.einsum("nm,nv->nmv", B, z) + zz * A[:, None, None]
               - 0.5 * Xi * A[:, None, None])
    return theta, theta_T, theta - theta_T, sigma_T, Xi


def contract(a, b, tau):
    """a^{mu nu} b_{mu nu} per cell."""
    g = metric(tau)
    return np.einsum("nab,ac,bd,ncd->n", a, g, g, b)


def pi_tensor(q):
    """pi^{mu nu} per cell from the conserved fields (15, n): components 5..14."""
    n = q.shape[1]
    pi = np.zeros((n, 4, 4))
    comp = {(0, 0): 5, (0, 1): 6, (0, 2): 7, (0, 3): 8, (1, 1): 9, (1, 2): 10, (1, 3): 11, (2, 2): 12, (2, 3): 13,
            (3, 3): 14}
    for (a, b), f in comp.items():
        pi[:, a, b] = pi[:, b, a] = q[f]
    return pi
