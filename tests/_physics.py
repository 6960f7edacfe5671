"""Independent physics used to pin the oracle: Hamiltonians written as Pauli sums (not via
an MPO), dense MPO and MPS contractions, and the open-chain free-fermion closed form."""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

I2 = sp.identity(2, dtype=np.complex128, format="csr")
X = sp.csr_matrix(np.array([[0, 1], [1, 0]], dtype=np.complex128))
Y = sp.csr_matrix(np.array([[0, -1j], [1j, 0]], dtype=np.complex128))
Z = sp.csr_matrix(np.array([[1, 0], [0, -1]], dtype=np.complex128))


def site_op(op, k, L):
    """op acting on site k of L.  Site 0 is the FASTEST index of the state vector, which
    is how a column-major MPS contraction [s0, s1, ...] flattens."""
    out = sp.identity(1, dtype=np.complex128, format="csr")
    for j in range(L - 1, -1, -1):
        out = sp.kron(out, op if j == k else I2, format="csr")
    return out


def ising_hamiltonian(L, J, g, h=0.0, K=0.0):
    H = sp.csr_matrix((2 ** L, 2 ** L), dtype=np.complex128)
    for k in range(L - 1):
        H = H + J * site_op(Z, k, L) @ site_op(Z, k + 1, L)
    for k in range(L):
        H = H + g * site_op(X, k, L) + h * site_op(Z, k, L)
    return H + K * sp.identity(2 ** L, format="csr")


def heisenberg_hamiltonian(L, Jx, Jy, Jz, h=0.0):
    H = sp.csr_matrix((2 ** L, 2 ** L), dtype=np.complex128)
    for k in range(L - 1):
        for J, P in ((Jx, X), (Jy, Y), (Jz, Z)):
            H = H + (J / 4.0) * site_op(P, k, L) @ site_op(P, k + 1, L)
    for k in range(L):
        H = H + (h / 2.0) * site_op(Z, k, L)
    return H


def schwinger_hamiltonian(N, mu, x, l0=0.0, alpha=0.0, offset=0.0, nu=0.0, y=1.0):
    """Spin form of the lattice Schwinger model:
    x sum (s+s- + h.c.) + sum (mu/2)(1 + (-1)^(k+1) Z_k) + (nu/2)(1 + Z_k)
    + y sum_{n=0}^{N-2} [l0 + alpha + 1/2 sum_{k<=n} (Z_k + (-1)^(k+1))]^2 + offset."""
    dim = 2 ** N
    Id = sp.identity(dim, dtype=np.complex128, format="csr")
    H = sp.csr_matrix((dim, dim), dtype=np.complex128)
    for k in range(N - 1):
        H = H + (x / 2.0) * (site_op(X, k, N) @ site_op(X, k + 1, N) + site_op(Y, k, N) @ site_op(Y, k + 1, N))
    for k in range(N):
        sg = (-1.0) ** (k + 1)
        H = H + (mu / 2.0) * (Id + sg * site_op(Z, k, N)) + (nu / 2.0) * (Id + site_op(Z, k, N))
    Ln = (l0 + alpha) * Id
    for n in range(N - 1):
        Ln = Ln + 0.5 * (site_op(Z, n, N) + (-1.0) ** (n + 1) * Id)
        H = H + y * (Ln @ Ln)
    return H + offset * Id


def op_on(op, k, dims):
    """op acting on site k of a chain with per-site dimensions dims (site 0 fastest)."""
    out = sp.identity(1, dtype=np.complex128, format="csr")
    for j in range(len(dims) - 1, -1, -1):
        out = sp.kron(out, sp.csr_matrix(op) if j == k else sp.identity(dims[j], dtype=np.complex128,
                                                                          format="csr"), format="csr")
    return out


def chain_hamiltonian(bond_ops, field_ops):
    """sum_k sum_t A_t(k) B_t(k+1) + sum_k F(k) as a sparse Kronecker sum (no MPO)."""
    dims = [np.asarray(F).shape[0] for F in field_ops]
    n = int(np.prod(dims))
    H = sp.csr_matrix((n, n), dtype=np.complex128)
    for k, F in enumerate(field_ops):
        H = H + op_on(F, k, dims)
    for k, terms in enumerate(bond_ops):
        for A, B in terms:
            H = H + op_on(A, k, dims) @ op_on(B, k + 1, dims)
    return H


def ground_energy(H):
    if H.shape[0] <= 512:
        return float(np.linalg.eigvalsh(H.toarray())[0])
    return float(spla.eigsh(H, k=1, which="SA", tol=1e-14)[0][0])


def tfim_free_fermion(L, J, g):
    """Ground energy of H = -J sum Z Z - g sum X on an open chain: minus the sum of the
    singular values of the bidiagonal matrix with g on the diagonal and J above it."""
    B = np.diag(np.full(L, float(g))) + np.diag(np.full(L - 1, float(J)), 1)
    return -float(np.sum(np.linalg.svd(B, compute_uv=False)))


def mpo_to_dense(mpo):
    """Contract an MPO (list of W[du, Dl, dd, Dr]) into the dense operator on the
    column-major state vector (site 0 fastest)."""
    T = mpo[0][:, 0, :, :]                       # [u0, d0, r]
    for W in mpo[1:]:
        T = np.einsum("...r,urds->...uds", T, W)
    T = T[..., 0]
    L = len(mpo)
    # T indices: u0 d0 u1 d1 ... ; state index = sum s_k 2^k
    ups = list(range(0, 2 * L, 2))
    dns = list(range(1, 2 * L, 2))
    T = np.transpose(T, ups[::-1] + dns[::-1])
    dim = int(np.prod([W.shape[0] for W in mpo]))
    return T.reshape(dim, dim)


def mps_to_state(mps):
    """Column-major state vector of an MPS (site 0 fastest)."""
    psi = mps[0][:, 0, :]                        # [s0, r]
    for A in mps[1:]:
        psi = np.einsum("...r,srk->...sk", psi, A)
    psi = psi[..., 0]
    return np.reshape(psi, -1, order="F")


def site_isometry(mps, k):
    """P such that the state with site k replaced by x equals P @ x (x column-major flat)."""
    d, Dl, Dr = mps[k].shape
    n = d * Dl * Dr
    cols = []
    for j in range(n):
        e = np.zeros(n, dtype=np.complex128)
        e[j] = 1.0
        trial = list(mps)
        trial[k] = np.reshape(e, (d, Dl, Dr), order="F")
        cols.append(mps_to_state(trial))
    return np.stack(cols, axis=1)
