"""Gauge-field constructions with known answers, for the tests.

All fields are float64 arrays of shape (V, 4, 3, 3, 2) in the reference order
(site x-fastest, mu, 3x3 row-major, re/im).
"""
import numpy as np

H = np.diag([1.0, -1.0, 0.0])
PLANES = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]


def coords(L):
    """(V, 4) integer coordinates (x, y, z, t) of every site, x fastest."""
    Lx, Ly, Lz, Lt = L
    t, z, y, x = np.meshgrid(np.arange(Lt), np.arange(Lz), np.arange(Ly), np.arange(Lx),
                             indexing="ij")
    return np.stack([x.ravel(), y.ravel(), z.ravel(), t.ravel()], axis=1)


def site_index(c, L):
    return ((c[..., 3] * L[2] + c[..., 2]) * L[1] + c[..., 1]) * L[0] + c[..., 0]


def to_real(U):
    return np.ascontiguousarray(np.stack([U.real, U.imag], axis=-1))


def to_cplx(f):
    return f[..., 0] + 1j * f[..., 1]


def abelian_flux(L, n):
    """Constant abelian field strength: plaquette angle phi_p = 2 pi n_p / (L_mu L_nu)
    in each plane p = (mu, nu) of PLANES, all links exp(i theta H).
    U_nu(x) carries phi x_mu; U_mu(x) carries -phi L_mu x_nu on the last mu-slice,
    which keeps every plaquette, including the wrapping ones, at angle phi.
    Returns (field, phis)."""
    c = coords(L)
    V = c.shape[0]
    theta = np.zeros((V, 4))
    phis = []
    for (mu, nu), nq in zip(PLANES, n):
        phi = 2 * np.pi * nq / (L[mu] * L[nu])
        phis.append(phi)
        theta[:, nu] += phi * c[:, mu]
        last = c[:, mu] == L[mu] - 1
        theta[last, mu] -= phi * L[mu] * c[last, nu]
    d = np.exp(1j * theta[..., None] * np.diag(H)[None, None, :])  # (V,4,3)
    U = np.zeros((V, 4, 3, 3), dtype=complex)
    for k in range(3):
        U[:, :, k, k] = d[:, :, k]
    return to_real(U), np.array(phis)


def abelian_expected(L, phis):
    """(P, E_ref, Q_ref) for abelian_flux: leaves all equal exp(i phi H), so
    F = i sin(phi) H, |F|^2 = 2 sin^2 phi, Re tr(F_a F_b) = -2 sin phi_a sin phi_b."""
    V = int(np.prod(L))
    s = np.sin(phis)
    P = np.mean((1 + 2 * np.cos(phis)) / 3)
    E = np.sum(s ** 2)
    q = -2 * (s[0] * s[5] - s[1] * s[4] + s[2] * s[3])
    Q = V * q / (4 * np.pi ** 2)
    return P, E, Q


def translate(f, L, shift):
    """U'_mu(x) = U_mu(x - shift) (periodic)."""
    c = coords(L)
    src = site_index((c - np.asarray(shift)) % np.asarray(L), L)
    return np.ascontiguousarray(f[src])


def reflect(f, L, d):
    """Lattice reflection x_d -> -x_d: U'_mu(x) = U_mu(Rx) for mu != d,
    U'_d(x) = U_d(Rx - d)^dag."""
    c = coords(L)
    Lv = np.asarray(L)
    r = c.copy()
    r[:, d] = (-c[:, d]) % L[d]
    out = f[site_index(r, L)].copy()
    rm = r.copy()
    rm[:, d] = (r[:, d] - 1) % L[d]
    U = to_cplx(f[site_index(rm % Lv, L), d])
    out[:, d] = to_real(np.conj(np.swapaxes(U, -1, -2)))
    return np.ascontiguousarray(out)


def gauge_apply(f, L, g):
    """U_mu(x) -> g(x) U_mu(x) g(x+mu)^dag for g of shape (V, 3, 3) complex."""
    c = coords(L)
    U = to_cplx(f)
    out = np.empty_like(U)
    for mu in range(4):
        cp = c.copy()
        cp[:, mu] = (cp[:, mu] + 1) % L[mu]
        gp = g[site_index(cp, L)]
        out[:, mu] = g @ U[:, mu] @ np.conj(np.swapaxes(gp, -1, -2))
    return to_real(out)


def random_su3(rng, n):
    Z = rng.normal(size=(n, 3, 3)) + 1j * rng.normal(size=(n, 3, 3))
    Q, R = np.linalg.qr(Z)
    ph = np.diagonal(R, axis1=1, axis2=2)
    Q = Q * (ph / np.abs(ph))[:, None, :]
    det = np.linalg.det(Q)
    return Q / (det ** (1.0 / 3.0))[:, None, None]


def random_algebra(rng, n=None):
    shape = (3, 3) if n is None else (n, 3, 3)
    A = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    X = 0.5 * (A - np.conj(np.swapaxes(A, -1, -2)))
    tr = np.trace(X, axis1=-2, axis2=-1)
    return X - (tr / 3.0)[..., None, None] * np.eye(3)
