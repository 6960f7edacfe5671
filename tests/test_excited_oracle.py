"""The oracle's partial sweeps, projector multipliers and excited states (CPU), pinned against
quantities fixed independently of the contraction code:
  * the effective vector of a level equals P^H |level>, P the isometry that places a site
    tensor into the state (brute force over unit vectors);
  * the projector multiplier equals P^H (H + sum_k p_k |k><k|) P with H the dense MPO matrix;
  * overlaps equal inner products of the dense state vectors;
  * excited-state energies equal exact diagonalisation's next eigenvalues;
  * alternating partial sweeps reach the exact ground energy;
  * the MPS file reader/writer against a hand-assembled file."""
import struct

import numpy as np
import pytest

from oracle import mpsd_oracle as orc
from tests import _physics as ph


def _normalise(mps):
    mps = [a.copy() for a in mps]
    n = np.sqrt(abs(orc.overlap(mps, mps)))
    mps[0] = mps[0] / n
    return mps


def _mixed_canonical(mps, pos):
    mps = [a.copy() for a in mps]
    for k in range(pos):
        orc._gauge_site(mps, k, "R", True)
    for k in range(len(mps) - 1, pos, -1):
        orc._gauge_site(mps, k, "L", True)
    return mps


def test_overlap_matches_state_vectors():
    a = orc.random_mps(6, 4, seed=1)
    b = orc.random_mps(6, 3, seed=2)
    want = np.vdot(ph.mps_to_state(b), ph.mps_to_state(a))
    assert abs(orc.overlap(a, b) - want) <= 1e-12 * abs(want)


@pytest.mark.parametrize("pos", [0, 2, 5])
def test_effective_vector_is_projection(pos):
    L = 6
    state = orc.random_mps(L, 4, seed=3)
    level = orc.random_mps(L, 3, seed=4)
    NL, NR = orc._norm_envs_at(state, level, pos)
    v = orc.effective_bra(NL, level[pos], NR)
    P = ph.site_isometry(state, pos)
    want = P.conj().T @ ph.mps_to_state(level)
    assert np.linalg.norm(v.conj() - want) <= 1e-12 * np.linalg.norm(want)


@pytest.mark.parametrize("pos", [1, 3])
def test_projector_multiplier_is_projected_hamiltonian(pos):
    L = 5
    mpo = orc.heisenberg_mpo(L, 1.0, 0.7, 0.9, 0.3)
    H = ph.mpo_to_dense(mpo)
    state = _mixed_canonical(orc.random_mps(L, 4, seed=5), pos)
    levels = [_normalise(orc.random_mps(L, 4, seed=6)), _normalise(orc.random_mps(L, 2, seed=7))]
    pens = [2.5, 0.75]
    EL, ER = orc._envs_at(state, mpo, pos)
    vb = []
    for lev in levels:
        NL, NR = orc._norm_envs_at(state, lev, pos)
        vb.append(orc.effective_bra(NL, lev[pos], NR))
    F = orc.proj_multi_full(EL, mpo[pos], ER, [-p for p in pens], vb, True, offset=0.4)
    P = ph.site_isometry(state, pos)
    assert np.linalg.norm(P.conj().T @ P - np.eye(P.shape[1])) <= 1e-12
    Hp = H + sum(p * np.outer(ph.mps_to_state(l), ph.mps_to_state(l).conj()) for p, l in zip(pens, levels))
    want = P.conj().T @ Hp @ P + 0.4 * np.eye(P.shape[1])
    assert np.linalg.norm(F - want) <= 1e-12 * np.linalg.norm(want)
    x = orc.complex_gaussian(8, (P.shape[1],))
    y = orc.proj_multi_product(EL, mpo[pos], ER, x, [-p for p in pens], vb, True, offset=0.4)
    assert np.linalg.norm(y - want @ x) <= 1e-12 * np.linalg.norm(want @ x)


def test_excited_states_match_exact_diagonalisation():
    L, D = 7, 8
    mpo = orc.ising_mpo(L, -1.0, -1.3, 0.2)
    exact = np.linalg.eigvalsh(ph.mpo_to_dense(mpo))[:3]
    e0, gs, _, _ = orc.find_ground_state(mpo, orc.random_mps(L, D, seed=1))
    e1, ex1, _ = orc.find_next_excited_state(mpo, orc.random_mps(L, D, seed=2), [gs])
    e2, ex2, _ = orc.find_next_excited_state(mpo, orc.random_mps(L, D, seed=3), [gs, ex1])
    assert np.allclose([e0, e1, e2], exact, rtol=1e-9, atol=0)
    assert abs(orc.overlap(ex1, gs)) < 1e-6 and abs(orc.overlap(ex2, gs)) < 1e-6 and abs(orc.overlap(ex2, ex1)) < 1e-6


def test_alternating_partial_sweeps_reach_ground_state():
    L, D = 6, 8
    mpo = orc.heisenberg_mpo(L, 1.0, 1.0, 0.8, 0.1)
    exact = np.linalg.eigvalsh(ph.mpo_to_dense(mpo))[0]
    mps = orc.random_mps(L, D, seed=9)
    last = np.inf
    for it in range(12):
        mps, _ = orc.sweep_part(mpo, mps, 0, L) if it % 2 == 0 else orc.sweep_part(mpo, mps, L - 1, -1)
        e = orc.expectation(mps, mpo).real / orc.overlap(mps, mps).real
        assert e <= last + 1e-12
        last = e
    assert abs(last - exact) <= 1e-10 * abs(exact)


def test_penalty_sweep_part_lowers_penalised_energy():
    L, D = 6, 4
    mpo = orc.ising_mpo(L, -1.0, -0.9)
    gs = _normalise(orc.random_mps(L, 4, seed=11))
    mps = orc.random_mps(L, D, seed=12)

    def pen_energy(m):
        n = orc.overlap(m, m).real
        return (orc.expectation(m, mpo).real + 3.0 * abs(orc.overlap(m, gs)) ** 2 / n) / n

    before = pen_energy(orc.gauge_left_all(orc.gauge_right_all(mps)))
    out, val = orc.penalty_sweep_part(mpo, mps, [-3.0], [gs], 0, L)
    assert abs(pen_energy(out) - val) <= 1e-10 * abs(val)
    assert val <= before + 1e-12


def test_mps_file_known_answer(tmp_path):
    A0 = orc.complex_gaussian(1, (2, 1, 3))
    A1 = orc.complex_gaussian(2, (2, 3, 1))
    raw = struct.pack("<iidi", 2, 3, 0.5, 1)
    for k, A in enumerate((A0, A1)):
        raw += struct.pack("<ii", k + 1, 3) + struct.pack("<3i", *A.shape)
        flat = [A[i, j, l] for l in range(A.shape[2]) for j in range(A.shape[1]) for i in range(A.shape[0])]
        raw += b"".join(struct.pack("<dd", c.real, c.imag) for c in flat)
    p = tmp_path / "m.mps"
    p.write_bytes(raw)
    sites, nf = orc.mps_read(str(p))
    assert nf == 0.5 and np.array_equal(sites[0], A0) and np.array_equal(sites[1], A1)
    q = tmp_path / "w.mps"
    orc.mps_write(str(q), [A0, A1], bond=3, normfact=0.5, gauge=1)
    assert q.read_bytes() == raw
