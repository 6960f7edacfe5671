"""The oracle's SVD truncation (wrapper::svd with a bond cut, MPS(const MPS&, D)) on the CPU,
pinned against the dense Schmidt decomposition of the state vector."""
import numpy as np
import pytest

from oracle import mpsd_oracle as orc
from tests import _physics as ph


def test_svd_keep_rule():
    M = orc.complex_gaussian(1, (6, 4))
    M[:, 3] = M[:, 0] * 2.0  # rank 3
    U, S, Vh = orc.svd_keep(M, 10)
    assert S.size == 4  # a rounding-level fourth value is still above zero and kept
    U, S, Vh = orc.svd_keep(M, 2)
    assert S.size == 2
    s_all = np.linalg.svd(M, compute_uv=False)
    assert np.allclose(S, s_all[:2], rtol=1e-14)
    assert np.allclose(U.conj().T @ U, np.eye(2), atol=1e-14)
    Z = np.zeros((3, 3), dtype=np.complex128)
    Z[0, 0] = 1.0
    _, S, _ = orc.svd_keep(Z, 3)
    assert S.size == 1  # exact zeros are dropped (tol = 0 keeps values above zero only)


@pytest.mark.parametrize("D,cut_bond", [(6, 3), (3, None)])
def test_truncation_is_schmidt_cut(D, cut_bond):
    """L = 6, d = 2, bonds [1,2,4,8,4,2,1]: from a left-canonical state, cutting to D = 6 touches
    only the middle bond, and the result is the rank-6 Schmidt truncation across it."""
    L = 6
    mps = orc.gauge_right_all(orc.random_mps(L, 8, seed=3))
    psi = ph.mps_to_state(mps)
    out, nf = orc.mps_truncate(mps, D)
    got = nf * ph.mps_to_state(out)
    if cut_bond is not None:
        M = psi.reshape(2 ** cut_bond, 2 ** (L - cut_bond), order="F")
        U, S, Vh = np.linalg.svd(M, full_matrices=False)
        want = ((U[:, :D] * S[:D]) @ Vh[:D, :]).reshape(-1, order="F")
        assert np.linalg.norm(got - want) <= 1e-12 * np.linalg.norm(want)
        assert abs(nf - np.sqrt(np.sum(S[:D] ** 2))) <= 1e-12
    else:
        # D = 3 cuts several bonds: the result has bond <= 3 and keeps most of the norm
        assert max(max(a.shape[1], a.shape[2]) for a in out) <= 3
        assert abs(np.vdot(got, psi)) / np.linalg.norm(got) ** 2 > 0.5


def test_truncation_without_cut_is_exact():
    mps = orc.random_mps(7, 8, seed=5)
    psi = ph.mps_to_state(mps)
    out, nf = orc.mps_truncate(mps, 64)
    got = nf * ph.mps_to_state(out)
    assert np.linalg.norm(got - psi) <= 1e-12 * np.linalg.norm(psi)
    # every site but the first is right-canonical after the sweep
    for A in out[1:]:
        d, Dl, Dr = A.shape
        M = orc.reshape_f(orc.permute(A, (2, 1, 3)), (Dl, d * Dr))
        assert np.allclose(M @ M.conj().T, np.eye(Dl), atol=1e-12)
