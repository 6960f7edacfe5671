"""Pins the CPU oracle (oracle/mpsd_oracle.py) against the reference's own known answers and
golden values, and against independent mathematics (dense Pauli-sum Hamiltonians, explicit
projections, closed forms).  CPU only."""
import json
import os

import numpy as np
import pytest

from oracle import mpsd_oracle as orc
from tests import _physics as ph

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _golden():
    with open(os.path.join(GOLDEN, "reference_values.json")) as f:
        return json.load(f)


# ---------------------------------------------------------------- layout known answers

def test_colmajor_layout_known_answer():
    # src/tests/testmwArray_new.cpp:92-110: element (j,k,i) of a [2,4,3] array holds
    # dims0*dims1*i + dims0*k + j, i.e. the buffer is column-major.
    dims = (2, 4, 3)
    buf = np.arange(24, dtype=float) * (1 + 1j)
    A = orc.from_colmajor(buf, dims)
    for i in range(3):
        for k in range(4):
            for j in range(2):
                assert A[j, k, i] == (2 * 4 * i + 2 * k + j) * (1 + 1j)
    assert np.array_equal(orc.to_colmajor(A), buf)


def test_permute_known_answer():
    # testmwArray_new.cpp:159-170: after permute(2,3,1), A4(k,i,j) == A3(j,k,i)
    g = _golden()["mwArray_permute_231"]
    A3 = orc.from_colmajor(np.arange(24, dtype=float) * (1 + 1j), (2, 4, 3))
    A4 = orc.permute(A3, (2, 3, 1))
    assert A4.shape == tuple(g["dims_after"])
    for k in range(4):
        for j in range(2):
            for i in range(3):
                assert A4[k, i, j] == (2 * 4 * i + 2 * k + j) * (1 + 1j)


def test_product_known_answer():
    # testmwArray_new.cpp:176-246: A1*A2 == A3 for the 2x2 complex matrices given there.
    g = _golden()["mwArray_product_2x2"]
    A1 = orc.from_colmajor(np.array(g["A1"][0]) + 1j * np.array(g["A1"][1]), (2, 2))
    A2 = orc.from_colmajor(np.array(g["A2"][0]) + 1j * np.array(g["A2"][1]), (2, 2))
    A3 = orc.from_colmajor(np.array(g["A3"][0]) + 1j * np.array(g["A3"][1]), (2, 2))
    assert np.array_equal(A1 @ A2, A3)


# ---------------------------------------------------------------- MPO builders

@pytest.mark.parametrize("L,J,g,h", [(5, -1.0, -1.0, -0.5), (6, 0.7, 0.3, 0.2), (4, -2.0, 1.5, 0.0)])
def test_ising_mpo_dense(L, J, g, h):
    H = ph.ising_hamiltonian(L, J, g, h).toarray()
    assert np.allclose(ph.mpo_to_dense(orc.ising_mpo(L, J, g, h)), H, atol=1e-13)


@pytest.mark.parametrize("L,Jx,Jy,Jz,h", [(5, 1.0, 1.0, 1.0, 0.0), (6, 1.0, 1.0, 0.6, 0.3)])
def test_heisenberg_mpo_dense(L, Jx, Jy, Jz, h):
    H = ph.heisenberg_hamiltonian(L, Jx, Jy, Jz, h).toarray()
    assert np.allclose(ph.mpo_to_dense(orc.heisenberg_mpo(L, Jx, Jy, Jz, h)), H, atol=1e-13)


@pytest.mark.parametrize("N,mg,x,l0", [(6, 0.125, 100.0, 0.0), (7, 0.3, 25.0, 0.0), (6, 0.0, 4.0, 0.5)])
def test_schwinger_mpo_dense(N, mg, x, l0):
    mu = 2 * mg * np.sqrt(x)
    H = ph.schwinger_hamiltonian(N, mu, x, l0).toarray()
    assert np.allclose(ph.mpo_to_dense(orc.schwinger_mpo(N, mu, x, l0)), H, atol=1e-10 * np.abs(H).max())


# ---------------------------------------------------------------- Heff and environments

def _mixed_canonical(L, D, seed, center):
    mps = orc.gauge_left_all(orc.gauge_right_all(orc.random_mps(L, D, seed=seed)))
    for k in range(center):
        q, r = orc._qr_site(mps[k])
        mps[k] = q
        mps[k + 1] = orc._absorb_right(mps[k + 1], r)
    return mps


def _envs(mps, mpo, k):
    Lenv = np.ones((1, 1, 1), dtype=complex)
    for j in range(k):
        Lenv = orc.contract_left(Lenv, mps[j], mpo[j], mps[j])
    Renv = np.ones((1, 1, 1), dtype=complex)
    for j in range(len(mps) - 1, k, -1):
        Renv = orc.contract_right(Renv, mps[j], mpo[j], mps[j])
    return Lenv, Renv


@pytest.mark.parametrize("model", ["ising", "heis", "schw"])
@pytest.mark.parametrize("k", [0, 2, 5])
def test_heff_is_projected_hamiltonian(model, k):
    L, D = 6, 4
    if model == "ising":
        mpo, H = orc.ising_mpo(L, -1, -1, -0.5), ph.ising_hamiltonian(L, -1, -1, -0.5)
    elif model == "heis":
        mpo, H = orc.heisenberg_mpo(L, 1, 1, 0.8, 0.1), ph.heisenberg_hamiltonian(L, 1, 1, 0.8, 0.1)
    else:
        mpo, H = orc.schwinger_mpo(L, 1.0, 4.0), ph.schwinger_hamiltonian(L, 1.0, 4.0)
    mps = _mixed_canonical(L, D, seed=11 + k, center=k)
    Lenv, Renv = _envs(mps, mpo, k)
    P = ph.site_isometry(mps, k)
    Hp = P.conj().T @ (H.toarray() @ P)
    Hf = orc.heff_full(Lenv, mpo[k], Renv)
    assert np.allclose(Hf, Hp, atol=1e-11 * np.abs(Hp).max())
    x = orc.complex_gaussian(5, (P.shape[1],))
    y = orc.heff_product(Lenv, mpo[k], Renv, x, offset=0.25)
    assert np.allclose(y, Hp @ x + 0.25 * x, atol=1e-11 * np.abs(Hp).max() * np.linalg.norm(x))
    # P is an isometry in the mixed-canonical gauge
    assert np.allclose(P.conj().T @ P, np.eye(P.shape[1]), atol=1e-12)


def test_expectation_matches_dense():
    L = 7
    mpo = orc.heisenberg_mpo(L, 1.0, 0.5, 1.3, 0.2)
    H = ph.heisenberg_hamiltonian(L, 1.0, 0.5, 1.3, 0.2)
    mps = orc.random_mps(L, 5, seed=3)
    psi = ph.mps_to_state(mps)
    assert np.isclose(orc.expectation(mps, mpo), psi.conj() @ (H @ psi), rtol=1e-12)


def _explicit_heff(L, W, R, x):
    """Seven nested loops straight from the definition y = sum L W R x."""
    Db, wl, Dk = L.shape
    du, _, dd, wr = W.shape
    Dbr, _, Dkr = R.shape
    X = orc.from_colmajor(x, (dd, Dk, Dkr))
    Y = np.zeros((du, Db, Dbr), dtype=complex)
    for sp_ in range(du):
        for b in range(Db):
            for c in range(Dbr):
                acc = 0j
                for al in range(wl):
                    for s in range(dd):
                        for be in range(wr):
                            w = W[sp_, al, s, be]
                            if w == 0:
                                continue
                            for a in range(Dk):
                                for a2 in range(Dkr):
                                    acc += L[b, al, a] * w * R[c, be, a2] * X[s, a, a2]
                Y[sp_, b, c] = acc
    return orc.to_colmajor(Y)


@pytest.mark.parametrize("Dl,Dr,wl,wr", [(3, 4, 3, 3), (1, 2, 1, 5), (4, 1, 5, 1), (2, 3, 2, 4)])
def test_heff_explicit_loops(Dl, Dr, wl, wr):
    d = 2
    L = orc.complex_gaussian(1, (Dl, wl, Dl))
    W = orc.complex_gaussian(2, (d, wl, d, wr))
    R = orc.complex_gaussian(3, (Dr, wr, Dr))
    x = orc.complex_gaussian(4, (d * Dl * Dr,))
    assert np.allclose(orc.heff_product(L, W, R, x), _explicit_heff(L, W, R, x), rtol=1e-12, atol=1e-12)
    Hf = orc.heff_full(L, W, R)
    for j in range(d * Dl * Dr):
        e = np.zeros(d * Dl * Dr, dtype=complex)
        e[j] = 1
        assert np.allclose(Hf[:, j], orc.heff_product(L, W, R, e), atol=1e-13)


def _explicit_left(L, ket, W, bra):
    Db, wl, Dk = L.shape
    d, _, Dr = bra.shape
    _, _, Drk = ket.shape
    wr = W.shape[3]
    out = np.zeros((Dr, wr, Drk), dtype=complex)
    for br in range(Dr):
        for be in range(wr):
            for ar in range(Drk):
                acc = 0j
                for sp_ in range(d):
                    for s in range(d):
                        for b in range(Db):
                            for al in range(wl):
                                for a in range(Dk):
                                    acc += np.conj(bra[sp_, b, br]) * L[b, al, a] * W[sp_, al, s, be] * ket[s, a, ar]
                out[br, be, ar] = acc
    return out


def _explicit_right(R, ket, W, bra):
    Db, wr, Dk = R.shape
    d, Dl, _ = bra.shape
    _, Dlk, _ = ket.shape
    wl = W.shape[1]
    out = np.zeros((Dl, wl, Dlk), dtype=complex)
    for bl in range(Dl):
        for al in range(wl):
            for alk in range(Dlk):
                acc = 0j
                for sp_ in range(d):
                    for s in range(d):
                        for b in range(Db):
                            for be in range(wr):
                                for a in range(Dk):
                                    acc += np.conj(bra[sp_, bl, b]) * W[sp_, al, s, be] * R[b, be, a] * ket[s, alk, a]
                out[bl, al, alk] = acc
    return out


@pytest.mark.parametrize("Dl,Dr,wl,wr", [(3, 4, 3, 3), (1, 2, 1, 5), (2, 1, 5, 1)])
def test_env_explicit_loops(Dl, Dr, wl, wr):
    d = 2
    L = orc.complex_gaussian(11, (Dl, wl, Dl))
    R = orc.complex_gaussian(12, (Dr, wr, Dr))
    W = orc.complex_gaussian(13, (d, wl, d, wr))
    ket = orc.complex_gaussian(14, (d, Dl, Dr))
    bra = orc.complex_gaussian(15, (d, Dl, Dr))
    assert np.allclose(orc.contract_left(L, ket, W, bra), _explicit_left(L, ket, W, bra), rtol=1e-12, atol=1e-12)
    assert np.allclose(orc.contract_right(R, ket, W, bra), _explicit_right(R, ket, W, bra), rtol=1e-12, atol=1e-12)


def test_heff_hermitian_and_linear():
    L_, D = 8, 6
    mpo = orc.heisenberg_mpo(L_, 1, 1, 1)
    mps = _mixed_canonical(L_, D, seed=4, center=3)
    Lenv, Renv = _envs(mps, mpo, 3)
    n = mps[3].size
    u = orc.complex_gaussian(21, (n,))
    v = orc.complex_gaussian(22, (n,))
    Hu = orc.heff_product(Lenv, mpo[3], Renv, u)
    Hv = orc.heff_product(Lenv, mpo[3], Renv, v)
    assert np.isclose(np.vdot(u, Hv), np.vdot(Hu, v), rtol=1e-12)
    a, b = 0.3 - 1.1j, 2.0 + 0.5j
    assert np.allclose(orc.heff_product(Lenv, mpo[3], Renv, a * u + b * v), a * Hu + b * Hv, rtol=1e-12, atol=1e-12)


# ---------------------------------------------------------------- RNG

@pytest.mark.parametrize("Dl,Dr,wl,wr,offset", [(3, 4, 2, 3, 0.0), (5, 5, 3, 3, 0.75), (1, 6, 1, 3, -1.5)])
def test_heff_product_hermitian_is_dense_average(Dl, Dr, wl, wr, offset):
    # non-Hermitian operands (F != F^H), so the conjugate-permuted copies must reproduce the
    # dense adjoint exactly and the average must differ from the plain product
    L = orc.complex_gaussian(81, (Dl, wl, Dl))
    W = orc.complex_gaussian(82, (2, wl, 2, wr))
    R = orc.complex_gaussian(83, (Dr, wr, Dr))
    x = orc.complex_gaussian(84, (2 * Dl * Dr,))
    F = orc.heff_full(L, W, R)
    assert np.linalg.norm(F - F.conj().T) > 1e-3 * np.linalg.norm(F)
    want = 0.5 * (F + F.conj().T) @ x + offset * x
    got = orc.heff_product_hermitian(L, W, R, x, offset)
    assert np.linalg.norm(got - want) <= 1e-13 * np.linalg.norm(want)
    assert np.linalg.norm(got - (orc.heff_product(L, W, R, x) + offset * x)) > 1e-3 * np.linalg.norm(want)


def test_rng_reproducible_and_normal():
    a = orc.complex_gaussian(7, (4, 5))
    b = orc.complex_gaussian(7, (4, 5))
    assert np.array_equal(a, b)
    big = orc.complex_gaussian(9, (200000,))
    assert abs(big.real.mean()) < 0.01 and abs(big.imag.std() - 1) < 0.01
    # SplitMix64 known value: first output for seed 0 is 0xE220A8397B1DCDAF
    assert int(orc.splitmix64(0, 1)[0]) == 0xE220A8397B1DCDAF


# ---------------------------------------------------------------- ground states

def test_tfim_closed_form_matches_ed():
    for L in (4, 7, 9):
        e_ed = ph.ground_energy(ph.ising_hamiltonian(L, -1.0, -1.0))
        assert np.isclose(ph.tfim_free_fermion(L, 1.0, 1.0), e_ed, rtol=1e-12)


def test_gs_tfim_L8_reference_value():
    g = _golden()["tfim_L8_critical"]
    mpo = orc.ising_mpo(8, -1.0, -1.0, 0.0)
    e, _, _, _ = orc.find_ground_state(mpo, orc.random_mps(8, g["D"], seed=2), convtol=1e-12)
    assert abs(e - g["reference_energy"]) < 1e-9
    assert abs(e - ph.tfim_free_fermion(8, 1.0, 1.0)) < 1e-9


def test_gs_heisenberg_ed():
    L = 10
    e_ed = ph.ground_energy(ph.heisenberg_hamiltonian(L, 1, 1, 1))
    e, _, _, _ = orc.find_ground_state(orc.heisenberg_mpo(L, 1, 1, 1), orc.random_mps(L, 32, seed=5), convtol=1e-13)
    assert abs(e - e_ed) < 1e-10 * abs(e_ed)


def test_gs_schwinger_ed():
    N, x, mg = 10, 100.0, 0.125
    mu = 2 * mg * np.sqrt(x)
    e_ed = ph.ground_energy(ph.schwinger_hamiltonian(N, mu, x))
    e, _, _, _ = orc.find_ground_state(orc.schwinger_mpo(N, mu, x), orc.random_mps(N, 32, seed=6), convtol=1e-13)
    assert abs(e - e_ed) < 1e-10 * abs(e_ed)


@pytest.mark.slow
def test_gs_config1_reference_value():
    """TFIM -sum ZZ - sum X - 0.5 sum Z, L=20, D=20 -> reference value."""
    g = _golden()["tfim_L20_D20"]
    mpo = orc.ising_mpo(20, -1.0, -1.0, -0.5)
    e, _, _, _ = orc.find_ground_state(mpo, orc.random_mps(20, 20, seed=1), convtol=1e-12)
    assert abs(e - g["reference_energy"]) < 1e-10 * abs(g["reference_energy"])
