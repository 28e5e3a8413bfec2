"""Shared helpers for the test suite: golden fixtures, random configurations,
and the comparison tolerances of the parity bar."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# FP64 parity bar: relative to the largest magnitude of the expected field
REL_TOL = 1e-12


def golden(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def golden_names(prefix):
    return sorted(f[:-4] for f in os.listdir(GOLDEN) if f.startswith(prefix) and f.endswith(".npz"))


def rand_conf(shape, seed, asig=0.7):
    rng = np.random.default_rng(seed)
    lx, ly, lz = shape
    n = lx * ly * lz // 2
    u = rng.normal(size=(3, 2, n, 4))
    u /= np.linalg.norm(u, axis=-1, keepdims=True)
    a = rng.normal(size=(2, n, 3)) * asig
    return u, a


def relerr(x, y):
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    scale = max(np.max(np.abs(y)), 1e-300)
    return float(np.max(np.abs(x - y)) / scale)


def assert_close(x, y, tol=REL_TOL, what=""):
    e = relerr(x, y)
    assert e <= tol, "%s: max relative deviation %.3e > %.1e" % (what, e, tol)


# ---- quaternion helpers for invariant tests (links as (c0,c3,c2,c1) in
# (Re alpha, Im alpha, Re beta, Im beta) order, f/INCLUDE/gaugefield.inc:89-101)

def to_mat(q):
    """(..., 4) (alpha, beta) -> (..., 2, 2) complex matrices [[a, b], [-b*, a*]]."""
    a = q[..., 0] + 1j * q[..., 1]
    b = q[..., 2] + 1j * q[..., 3]
    m = np.empty(q.shape[:-1] + (2, 2), dtype=complex)
    m[..., 0, 0] = a
    m[..., 0, 1] = b
    m[..., 1, 0] = -np.conj(b)
    m[..., 1, 1] = np.conj(a)
    return m


def from_mat(m):
    q = np.empty(m.shape[:-2] + (4,))
    q[..., 0] = m[..., 0, 0].real
    q[..., 1] = m[..., 0, 0].imag
    q[..., 2] = m[..., 0, 1].real
    q[..., 3] = m[..., 0, 1].imag
    return q


PAULI = np.array([[[0, 1], [1, 0]], [[0, -1j], [1j, 0]], [[1, 0], [0, -1]]], dtype=complex)


def higgs_mat(a):
    """X(a) = i a.sigma as a 2x2 matrix (f/matmul.f:150-154)."""
    return 1j * np.einsum("...i,ijk->...jk", a, PAULI)


def coords(shape):
    """per parity p, site s: 0-based (x, y, z) of the checkerboard numbering."""
    lx, ly, lz = shape
    h1, h2 = lx // 2, lx * ly // 2
    n = lx * ly * lz // 2
    s = np.arange(n)
    z = s // h2
    y = (s % h2) // h1
    xh = s % h1
    out = []
    for p in (0, 1):
        x = 2 * xh + ((p + y + z) & 1)
        out.append(np.stack([x, y, z], axis=1))
    return out


def lattice_fields(shape, u, a):
    """links (3,lx,ly,lz,2,2) matrices and higgs (lx,ly,lz,3) from the
    checkerboard arrays."""
    lx, ly, lz = shape
    U = np.zeros((3, lx, ly, lz, 2, 2), dtype=complex)
    A = np.zeros((lx, ly, lz, 3))
    cs = coords(shape)
    for p in (0, 1):
        x, y, z = cs[p].T
        for mu in range(3):
            U[mu, x, y, z] = to_mat(u[mu, p])
        A[x, y, z] = a[p]
    return U, A


def checkerboard(shape, U=None, A=None):
    lx, ly, lz = shape
    n = lx * ly * lz // 2
    cs = coords(shape)
    u = np.zeros((3, 2, n, 4)) if U is not None else None
    a = np.zeros((2, n, 3)) if A is not None else None
    for p in (0, 1):
        x, y, z = cs[p].T
        if U is not None:
            for mu in range(3):
                u[mu, p] = from_mat(U[mu, x, y, z])
        if A is not None:
            a[p] = A[x, y, z]
    return u, a


def plaquette_sum(shape, u):
    """sum over x and mu<nu of (1/2) Re Tr U_P, by explicit 2x2 products."""
    U, _ = lattice_fields(shape, u, np.zeros((2, u.shape[2], 3)))
    tot = 0.0
    for mu in range(3):
        for nu in range(mu + 1, 3):
            a = U[mu]
            b = np.roll(U[nu], -1, axis=mu)
            c = np.conj(np.swapaxes(np.roll(U[mu], -1, axis=nu), -1, -2))
            d = np.conj(np.swapaxes(U[nu], -1, -2))
            p = a @ b @ c @ d
            tot += 0.5 * np.sum(np.real(p[..., 0, 0] + p[..., 1, 1]))
    return tot
