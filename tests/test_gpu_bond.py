"""findGroundState's start-state handling on the device, through the C-ABI: the device resize
(mwArray::resize), bond growth by zero padding (MPS::increaseBondDimension), truncation of a
larger start state (MPS(init, D)), physical dimensions fitted to the MPO
(MPS::adjustPhysDimensions), the edge normalisation of MPS::gaugeCond(pos, dir, next), and the
residual ||H psi - E psi|| of a solve against the oracle's.  Energies: 1e-10 relative."""
import json
import os

import numpy as np
import pytest

from oracle import mpsd_oracle as orc
from tests import _physics as ph

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "reference_values.json")
ETOL = 1e-10


@pytest.fixture(scope="module")
def ctx():
    from mpsdyn_amd import device as dev
    c = dev.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("old,new", [((3, 4, 5), (3, 6, 7)), ((3, 4, 5), (2, 2, 9)), ((2, 8, 8), (2, 8, 8)),
                                     ((5, 1, 7, 2), (6, 2, 3, 2)), ((64, 33, 40), (70, 40, 20))])
def test_tensor_resize_matches_oracle(ctx, old, new):
    """Bit-exact against mwArray::resize with zero noise: kept, zero-padded and cut indices."""
    print(f"resize {old} -> {new}", flush=True)
    a = orc.complex_gaussian(7, old)
    t = ctx.tensor(a)
    t.resize(new)
    assert t.dims == tuple(new)
    got = orc.from_colmajor(t.download(), tuple(new))
    assert np.array_equal(got, orc.resize(a, new))


def test_resize_rejects_a_rank_change(ctx):
    from mpsdyn_amd import device as dev
    t = ctx.tensor(orc.complex_gaussian(1, (2, 3, 4)))
    with pytest.raises(dev.MpsdError) as e:
        t.resize((2, 12))
    assert e.value.code == -2


def test_grown_start_reaches_the_reference_energy(ctx, tmp_path):
    """TFIM L = 20 (J = g = 1, h = 0.5) started from a D = 2 state and solved with D = 20: the
    bond grows by zero padding and the sweep reaches the reference's D = 20 energy."""
    from mpsdyn_amd import device as dev
    ref = json.load(open(GOLDEN))["tfim_L20_D20"]["reference_energy"]
    start = orc.gauge_right_all(orc.random_mps(20, 2, seed=3))
    path = str(tmp_path / "start.mps")
    orc.mps_write(path, start)
    cfg = dev.gs_config(dev.MODEL_ISING, 20, 20, [-1.0, -1.0, -0.5], convtol=1e-12)
    out = str(tmp_path / "gs.mps")
    r = dev.gs_solve_from(ctx, cfg, path, export_path=out)
    sites, _ = orc.mps_read(out)
    bonds = [a.shape[2] for a in sites[:-1]]
    print(f"grown E={r.energy:.15f} ref={ref:.15f} max bond {max(bonds)}", flush=True)
    assert max(bonds) == 20 and bonds == orc.bond_dims(20, 20)[1:-1]
    assert abs(r.energy - ref) <= ETOL * abs(ref)
    # without D (0: keep the start bond) the solve stays at D = 2, measurably higher (a D = 8
    # state would already be within 1e-13 of the D = 20 energy for this gapped chain)
    r2 = dev.gs_solve_from(ctx, dev.gs_config(dev.MODEL_ISING, 20, 0, [-1.0, -1.0, -0.5], convtol=1e-12), path)
    print(f"kept D=2: E={r2.energy:.15f}", flush=True)
    assert r2.energy > ref + 1e-7


def test_truncated_start_matches_the_oracle(ctx, tmp_path):
    """TFIM L = 10 started from a D = 32 state and solved with D = 16: the start state is cut
    by MPS(init, D) and swept; energy equal to the oracle's cut-then-swept state."""
    from mpsdyn_amd import device as dev
    L = 10
    start = orc.gauge_right_all(orc.random_mps(L, 32, seed=5))
    path = str(tmp_path / "start.mps")
    orc.mps_write(path, start)
    mpo = orc.ising_mpo(L, -1.0, -1.0, -0.5)
    e_orc = orc.find_ground_state(mpo, start, convtol=1e-12, D=16)[0]
    r = dev.gs_solve_from(ctx, dev.gs_config(dev.MODEL_ISING, L, 16, [-1.0, -1.0, -0.5], convtol=1e-12), path)
    print(f"truncated E={r.energy:.15f} oracle={e_orc:.15f}", flush=True)
    assert abs(r.energy - e_orc) <= ETOL * abs(e_orc)
    # D = 32 is exact for 10 sites: kept as it is, the solve reaches exact diagonalisation
    e_ed = ph.ground_energy(ph.ising_hamiltonian(L, -1.0, -1.0, -0.5))
    r32 = dev.gs_solve_from(ctx, dev.gs_config(dev.MODEL_ISING, L, 0, [-1.0, -1.0, -0.5], convtol=1e-12), path)
    assert abs(r32.energy - e_ed) <= ETOL * abs(e_ed)
    assert r.energy >= r32.energy - 1e-12 * abs(e_ed)


def test_start_state_fitted_to_the_mpo_dimensions(ctx, tmp_path):
    """A spin-1/2 (d = 2) start state solved against a spin-1 (d = 3) chain from an MPO file:
    the physical dimensions are raised to the MPO's (adjustPhysDimensions), the bond profile is
    rebuilt, and with D = 27 (exact for 6 sites) the energy equals exact diagonalisation."""
    from mpsdyn_amd import device as dev
    sx, sy, sz = orc.spin_ops(2)
    L = 6
    bonds = [[(sx, sx), (sy, sy), (sz, sz)]] * (L - 1)
    fields = [0.3 * sz @ sz + 0.1 * sz] * L
    mpo_path = str(tmp_path / "h.mpo")
    orc.mpo_write(mpo_path, orc.chain_mpo(bonds, fields), "binary")
    e_ed = ph.ground_energy(ph.chain_hamiltonian(bonds, fields))
    start = orc.gauge_right_all(orc.random_mps(L, 4, seed=8))
    path = str(tmp_path / "start.mps")
    orc.mps_write(path, start)
    out = str(tmp_path / "gs.mps")
    r = dev.gs_solve_from(ctx, dev.gs_config(0, 0, 27, [], convtol=1e-12), path, mpo_path=mpo_path,
                          fmt=dev.MPO_BINARY, export_path=out)
    sites, _ = orc.mps_read(out)
    print(f"spin-1 from d=2 start: E={r.energy:.15f} ED={e_ed:.15f}", flush=True)
    assert [a.shape[0] for a in sites] == [3] * L
    assert abs(r.energy - e_ed) <= ETOL * abs(e_ed)


@pytest.mark.parametrize("direction,pos", [("L", 0), ("R", 5), ("L", 3), ("R", 2)])
def test_gauge_site_edge_normalisation(ctx, tmp_path, direction, pos):
    """MPS::gaugeCond(pos, dir, true): at the chain's end the edge factor's magnitude enters the
    norm factor (MPS.cpp:380-387, 402-409); inside, the factor goes to the neighbour.  Sites
    and norm factor against the oracle."""
    from mpsdyn_amd import device as dev
    L = 6
    mps = orc.random_mps(L, 4, seed=9)
    mps = orc.gauge_left_all(mps) if direction == "L" else orc.gauge_right_all(mps)
    mps[pos] = mps[pos] * (3.7 - 1.1j)
    src, dst = str(tmp_path / "in.mps"), str(tmp_path / "out.mps")
    orc.mps_write(src, mps, normfact=0.5)
    dev.mps_gauge_site(ctx, src, pos, direction, True, dst)
    got, nf = orc.mps_read(dst)
    want, nf_want = orc.gauge_cond_site(mps, pos, direction, True, normfact=0.5)
    print(f"gauge {direction}@{pos}: normfact {nf!r} oracle {nf_want!r}", flush=True)
    assert nf == pytest.approx(nf_want, rel=1e-13)
    for a, b in zip(got, want):
        assert a.shape == b.shape
        assert np.linalg.norm(a - b) <= 1e-12 * max(np.linalg.norm(b), 1.0)


@pytest.mark.parametrize("model,L,D,p", [
    (0, 12, 4, [-1.0, -1.0, -0.5]),
    (1, 12, 6, [1.0, 1.0, 1.0, 0.0]),
    (2, 12, 6, [2.5, 100.0, 0.0, 0.0]),
])
def test_residual_norm_matches_the_oracle(ctx, tmp_path, model, L, D, p):
    """||H psi - E psi|| of a solve's state (device, over the MPO of (H - E)^2) against the
    oracle's on the exported state, away from the rounding floor (one half-sweep at small D,
    residuals 1e-3..1e-1); and against the dense residual of the same state."""
    from mpsdyn_amd import device as dev
    out = str(tmp_path / "gs.mps")
    cfg = dev.gs_config(model, L, D, p, convtol=1e-12, max_halfsweeps=1, want_residual=True)
    r = dev.gs_solve(ctx, cfg, export_path=out)
    sites, nf = orc.mps_read(out)
    E = r.energy
    if model == 0:
        shifted = orc.ising_mpo(L, p[0], p[1], p[2], K=-E)
        H = ph.ising_hamiltonian(L, p[0], p[1], p[2])
    elif model == 1:
        shifted = orc.heisenberg_mpo(L, p[0], p[1], p[2], p[3], offset=-E)
        H = ph.heisenberg_hamiltonian(L, p[0], p[1], p[2], p[3])
    else:
        shifted = orc.schwinger_mpo(L, p[0], p[1], offset=-E)
        H = ph.schwinger_hamiltonian(L, p[0], p[1])
    want = orc.residual_norm(sites, shifted)
    psi = ph.mps_to_state(sites)
    dense = np.linalg.norm(H @ psi - E * psi) / np.linalg.norm(psi)
    print(f"model {model}: residual device {r.residual_norm:.6e} oracle {want:.6e} dense {dense:.6e}", flush=True)
    assert r.residual_norm > 1e-6
    assert r.residual_norm == pytest.approx(want, rel=1e-8)
    assert r.residual_norm == pytest.approx(dense, rel=1e-8)
