"""The oracle's start-state fitting of findGroundState on the CPU (mwArray::resize,
MPS::increaseBondDimension, MPS::adjustPhysDimensions, the truncating copy), the edge
normalisation of MPS::gaugeCond(pos, dir, next), and the residual ||(H - E) psi|| over the
MPO of (H - E)^2, each pinned by a closed form, the dense state vector or exact
diagonalisation."""
import numpy as np
import pytest

from oracle import mpsd_oracle as orc
from tests import _physics as ph


def test_resize_known_answer():
    """mwArray::resize keeps elements by multi-index (wrapper/mwArray.cpp:1645-1664)."""
    a = np.array([[1 + 1j, 2], [3, 4 - 2j]])  # a[0,0]=1+i, a[1,0]=3, a[0,1]=2, a[1,1]=4-2i
    grown = orc.resize(a, (3, 1))
    assert grown.shape == (3, 1)
    assert list(grown[:, 0]) == [1 + 1j, 3, 0]
    cut = orc.resize(a, (1, 3))
    assert list(cut[0]) == [1 + 1j, 2, 0]
    same = orc.resize(a, (2, 2))
    assert np.array_equal(same, a)


@pytest.mark.parametrize("seed", range(6))
def test_bond_profile_is_the_product_bound(seed):
    """The reference's running-product loop gives bond k = min(D, prod_{j<k} d_j,
    prod_{j>=k} d_j); the outer bonds are kept."""
    rng = np.random.default_rng(seed)
    L = int(rng.integers(2, 12))
    dvec = [int(x) for x in rng.integers(1, 4, size=L)]
    D = int(rng.integers(2, 40))
    mps = [np.ones((d, 1, 1), dtype=np.complex128) for d in dvec]
    out, bond = orc.increase_bond_dimension(mps, D, 1)
    left = [int(np.prod(dvec[:k])) for k in range(L + 1)]
    right = [int(np.prod(dvec[k:])) for k in range(L + 1)]
    want = [1] + [min(D, left[k], right[k]) for k in range(1, L)] + [1]
    got = [out[0].shape[1]] + [a.shape[2] for a in out]
    assert got == want
    assert bond == max(want)
    assert all(a.shape[0] == d for a, d in zip(out, dvec))


def test_padding_keeps_the_state():
    """Zero padding (noise 0) represents the same state vector; a D below the bond is a no-op."""
    mps = orc.gauge_right_all(orc.random_mps(7, 3, seed=5))
    psi = ph.mps_to_state(mps)
    grown, bond = orc.increase_bond_dimension(mps, 10, 3)
    assert bond == 8  # capped by 2^3 at the middle of 7 sites
    assert np.linalg.norm(ph.mps_to_state(grown) - psi) == 0.0
    same, b2 = orc.increase_bond_dimension(mps, 2, 3)
    assert b2 == 3 and all(np.array_equal(x, y) for x, y in zip(same, mps))


def test_phys_dimension_change_embeds_the_state():
    """adjustPhysDimensions 2 -> 3 on some sites: the old state sits in the d = 2 corner of the
    new Hilbert space (new components zero), and the bond profile is rebuilt for the new d."""
    L = 5
    mps = orc.gauge_right_all(orc.random_mps(L, 4, seed=2))
    psi = ph.mps_to_state(mps).reshape([2] * L, order="F")
    dvec = [2, 3, 2, 3, 2]
    out, bond = orc.adjust_phys_dimensions(mps, dvec, 4)
    assert [a.shape[0] for a in out] == dvec
    big = ph.mps_to_state(out).reshape(dvec, order="F")
    corner = big[tuple(slice(0, 2) for _ in dvec)]
    # bonds that grew are zero-padded and none is cut (4 <= every product bound here)
    assert np.linalg.norm(corner - psi) <= 1e-14 * np.linalg.norm(psi)
    assert np.linalg.norm(big) == pytest.approx(np.linalg.norm(psi), rel=1e-14)
    assert bond == 4


def test_fit_start_truncates_below_the_bond():
    """D below the start bond: the truncating copy MPS(init, D) (the Schmidt cut of
    test_truncate_oracle), D = 0 keeps the state."""
    mps = orc.gauge_right_all(orc.random_mps(6, 8, seed=3))
    mpo = orc.ising_mpo(6, -1.0, -1.0, -0.5)
    cut, bond = orc.fit_start(mpo, mps, 4)
    assert bond == 4 and max(max(a.shape[1], a.shape[2]) for a in cut) == 4
    want, _ = orc.mps_truncate(mps, 4)
    assert all(np.allclose(x, y, atol=1e-14) for x, y in zip(cut, want))
    kept, b0 = orc.fit_start(mpo, mps, 0)
    assert b0 == 8 and all(np.array_equal(x, y) for x, y in zip(kept, mps))


def test_growth_lets_the_sweep_reach_the_exact_energy():
    """TFIM L = 8 started at D = 2: asked for D = 16 (exact for 8 sites) the sweep reaches the
    ED energy; without D it stays at the D = 2 variational energy, 2.5e-6 above it (2e-7
    relative, far outside the 1e-10 bar)."""
    L = 8
    mpo = orc.ising_mpo(L, -1.0, -1.0, -0.5)
    e_ed = ph.ground_energy(ph.ising_hamiltonian(L, -1.0, -1.0, -0.5))
    start = orc.random_mps(L, 2, seed=11)
    e_grow = orc.find_ground_state(mpo, start, convtol=1e-12, D=16)[0]
    e_small = orc.find_ground_state(mpo, start, convtol=1e-12)[0]
    assert abs(e_grow - e_ed) <= 1e-10 * abs(e_ed)
    assert e_small - e_ed > 1e-6


@pytest.mark.parametrize("direction,pos", [("L", 0), ("R", 5)])
def test_edge_gauge_moves_the_norm(direction, pos):
    """MPS::gaugeCond(pos, dir, true) at the chain's end: the unit-norm canonical state with the
    edge site scaled by c has norm |c|, and that is what enters the norm factor."""
    L = 6
    mps = orc.random_mps(L, 4, seed=9)
    mps = orc.gauge_left_all(mps) if direction == "L" else orc.gauge_right_all(mps)
    c = 3.7 - 1.1j
    mps[pos] = mps[pos] * c
    out, nf = orc.gauge_cond_site(mps, pos, direction, True, normfact=0.5)
    assert nf == pytest.approx(0.5 * abs(c), rel=1e-14)
    assert np.linalg.norm(ph.mps_to_state(out)) == pytest.approx(1.0, rel=1e-14)
    # not at the edge, or without next: the norm factor is untouched
    _, nf2 = orc.gauge_cond_site(mps, pos, direction, False, normfact=0.5)
    assert nf2 == 0.5


def test_mpo_product_is_the_matrix_product():
    L = 4
    A = orc.ising_mpo(L, -1.0, -0.7, 0.3)
    B = orc.heisenberg_mpo(L, 1.0, 0.8, 1.2, 0.1)
    P = orc.mpo_product(A, B)
    assert [W.shape[1] for W in P][1:] == [15] * (L - 1)
    dense = ph.mpo_to_dense(P)
    want = ph.mpo_to_dense(A) @ ph.mpo_to_dense(B)
    assert np.linalg.norm(dense - want) <= 1e-13 * np.linalg.norm(want)


@pytest.mark.parametrize("model", ["ising", "heisenberg", "schwinger"])
def test_residual_norm_matches_the_dense_residual(model):
    """||(H - E) psi|| over the MPO of (H - E)^2, with -E spread over the local terms (the
    models' constant / offset arguments), against the dense residual of the same state."""
    L = 6
    if model == "ising":
        mpo = lambda E: orc.ising_mpo(L, -1.0, -1.0, -0.5, K=-E)
        H = ph.ising_hamiltonian(L, -1.0, -1.0, -0.5)
    elif model == "heisenberg":
        mpo = lambda E: orc.heisenberg_mpo(L, 1.0, 1.0, 1.0, 0.0, offset=-E)
        H = ph.heisenberg_hamiltonian(L, 1.0, 1.0, 1.0, 0.0)
    else:
        mpo = lambda E: orc.schwinger_mpo(L, 2.5, 4.0, offset=-E)
        H = ph.schwinger_hamiltonian(L, 2.5, 4.0)
    mps = orc.gauge_left_all(orc.random_mps(L, 4, seed=4))
    psi = ph.mps_to_state(mps)
    E = float(np.real(np.vdot(psi, H @ psi)))
    want = np.linalg.norm(H @ psi - E * psi)
    got = orc.residual_norm(mps, mpo(E))
    assert got == pytest.approx(want, rel=1e-9)
    # the exact ground state written as an MPS: its residual is zero up to the floor of any
    # variance form, sqrt of the rounding of <(H-E)^2> (about 3e-8 here); the shifted MPO keeps
    # that rounding at the scale of the local terms instead of E^2
    w, v = np.linalg.eigh(H.toarray())
    gs = _state_to_mps(v[:, 0], L)
    assert orc.residual_norm(gs, mpo(w[0])) ** 2 < 1e-14 * w[0] ** 2


def _state_to_mps(psi, L, d=2):
    """Exact MPS of a state vector (site 0 fastest) by successive SVDs."""
    sites, rest, Dl = [], psi.reshape(-1, order="F"), 1
    for k in range(L - 1):
        M = rest.reshape(Dl * d, -1, order="F")
        U, S, Vh = np.linalg.svd(M, full_matrices=False)
        sites.append(U.reshape(Dl, d, -1, order="F").transpose(1, 0, 2))
        rest = (S[:, None] * Vh).reshape(-1, order="F")
        Dl = S.size
    sites.append(rest.reshape(Dl, d, 1, order="F").transpose(1, 0, 2))
    return sites
