"""Partial sweeps, projector multipliers and excited states on the device, through the C-ABI:
  * mpsd_vec_proj_update (the projector terms of TensorMultiplierProjMulti) against numpy;
  * Contractor::sweepPart and sweepPartWithProjectorPenalty from a state the oracle wrote, against
    the oracle's restatement from the same state: energy to 1e-10 relative, and the result
    states equal up to a global phase (|<dev|oracle>| = 1 to 1e-9);
  * findGroundState + findNextExcitedState (mpsd_gs_solve_levels) against exact
    diagonalisation: the lowest levels to 1e-9 relative, mutual overlaps below 1e-6;
  * bad arguments are errors."""
import numpy as np
import pytest

from oracle import mpsd_oracle as orc
from tests import _physics as ph

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from mpsdyn_amd import device as dev
    c = dev.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("n,m", [(1, 1), (1000, 3), (4097, 8), (70000, 32)])
def test_proj_update(ctx, n, m):
    import ctypes as C
    print(f"proj_update n={n} m={m}", flush=True)
    Phi = orc.complex_gaussian(n + m, (n, m))
    x = orc.complex_gaussian(1, (n,))
    y = orc.complex_gaussian(2, (n,))
    w = np.linspace(-2.0, 3.0, m)
    tP, tx, ty = ctx.tensor(Phi), ctx.tensor(x), ctx.tensor(y)
    ctx.check(ctx.lib.mpsd_vec_proj_update(ctx.handle, tP.devptr(), n, m, (C.c_double * m)(*w), tx.devptr(), ty.devptr(), n))
    got = ty.download()
    want = y - Phi @ (w * (Phi.conj().T @ x))
    assert np.linalg.norm(got - want) <= 1e-13 * np.linalg.norm(want)


def _state_file(tmp_path, name, mps):
    p = str(tmp_path / name)
    orc.mps_write(p, mps)
    return p


def _norm_energy(mps, mpo):
    return orc.expectation(mps, mpo).real / orc.overlap(mps, mps).real


def _fidelity(a, b):
    return abs(orc.overlap(a, b)) / np.sqrt(orc.overlap(a, a).real * orc.overlap(b, b).real)


PARTS = [(0, 8), (7, -1), (2, 5), (5, 1), (3, 3)]


@pytest.mark.parametrize("pos0,lim", PARTS)
def test_sweep_part_matches_oracle(ctx, tmp_path, pos0, lim):
    from mpsdyn_amd import device as dev
    print(f"sweep_part {pos0}->{lim}", flush=True)
    L, D = 8, 8
    mpo = orc.heisenberg_mpo(L, 1.0, 0.9, 0.8, 0.15)
    start = orc.random_mps(L, D, seed=21)
    cfg = dev.gs_config(dev.MODEL_HEISENBERG, L, D, [1.0, 0.9, 0.8, 0.15], eigtol=1e-13)
    out = str(tmp_path / "out.mps")
    e = dev.gs_sweep_part(ctx, cfg, _state_file(tmp_path, "in.mps", start), pos0, lim, mps_out=out)
    want, _ = orc.sweep_part(mpo, start, pos0, lim)
    got, _ = orc.mps_read(out)
    assert abs(e - _norm_energy(want, mpo)) <= 1e-10 * abs(e)
    assert abs(_norm_energy(got, mpo) - e) <= 1e-10 * abs(e)
    assert abs(_fidelity(got, want) - 1.0) <= 1e-9


@pytest.mark.parametrize("pos0,lim", [(0, 6), (5, -1), (1, 4)])
def test_penalty_sweep_part_matches_oracle(ctx, tmp_path, pos0, lim):
    from mpsdyn_amd import device as dev
    print(f"penalty sweep_part {pos0}->{lim}", flush=True)
    L, D = 6, 8
    p = [-1.0, -1.2, 0.3]
    mpo = orc.ising_mpo(L, *p)
    start = orc.random_mps(L, D, seed=31)
    levels = [orc.gauge_left_all(orc.gauge_right_all(orc.random_mps(L, 4, seed=32 + k))) for k in range(2)]
    weights = [-2.0, -0.75]
    paths = [_state_file(tmp_path, f"lev{k}.mps", lev) for k, lev in enumerate(levels)]
    cfg = dev.gs_config(dev.MODEL_ISING, L, D, p, eigtol=1e-13)
    out = str(tmp_path / "out.mps")
    dev.gs_sweep_part(ctx, cfg, _state_file(tmp_path, "in.mps", start), pos0, lim, mps_out=out, proj=paths,
                      weights=weights)
    want, val = orc.penalty_sweep_part(mpo, start, weights, levels, pos0, lim)
    got, _ = orc.mps_read(out)
    assert abs(_fidelity(got, want) - 1.0) <= 1e-9
    assert abs(_norm_energy(got, mpo) - _norm_energy(want, mpo)) <= 1e-10 * abs(val)


LEVELS = [
    ("ising_L10", dict(model=0, L=10, p=[-1.0, -1.2, 0.1]), 32, 3),
    ("heisenberg_L10", dict(model=1, L=10, p=[1.0, 1.0, 0.7, 0.23]), 32, 3),
]


@pytest.mark.parametrize("name,spec,D,nlev", LEVELS, ids=[x[0] for x in LEVELS])
def test_levels_match_exact(ctx, name, spec, D, nlev):
    from mpsdyn_amd import device as dev
    print(f"levels {name} D={D}", flush=True)
    L, p = spec["L"], spec["p"]
    if spec["model"] == 0:
        H = ph.ising_hamiltonian(L, p[0], p[1], p[2])
    else:
        H = ph.heisenberg_hamiltonian(L, p[0], p[1], p[2], p[3])
    exact = np.linalg.eigvalsh(H.toarray())[:nlev]
    cfg = dev.gs_config(spec["model"], L, D, p, convtol=1e-13, seed=4)
    e, ov, res = dev.gs_solve_levels(ctx, cfg, nlev)
    assert np.all(np.abs(e - exact) <= 1e-9 * np.abs(exact)), (e, exact)
    assert np.allclose(np.diag(ov), 1.0, atol=1e-10)
    off = ov - np.diag(np.diag(ov))
    assert off.max() <= 1e-6
    assert all(r.matvecs > 0 for r in res)


def test_levels_from_mpo_file_spin1(ctx, tmp_path):
    """Spin-1 chain from an MPO file: lowest three levels against sparse exact diagonalisation."""
    from mpsdyn_amd import device as dev
    sx, sy, sz = orc.spin_ops(2)
    L = 6
    bonds = [[(sx, sx), (sy, sy), (sz, 1.1 * sz)]] * (L - 1)
    fields = [0.2 * sz @ sz + 0.05 * sz + 0.03 * sx] * L
    path = str(tmp_path / "h.mpo")
    orc.mpo_write(path, orc.chain_mpo(bonds, fields), "binary")
    exact = np.linalg.eigvalsh(ph.chain_hamiltonian(bonds, fields).toarray())[:3]
    e, ov, _ = dev.gs_solve_levels(ctx, dev.gs_config(0, 0, 27, [], convtol=1e-13, seed=8), 3, mpo_path=path)
    assert np.all(np.abs(e - exact) <= 1e-9 * np.abs(exact)), (e, exact)


def test_orthogonality_retry_runs(ctx):
    """setExcOrthTol: the check runs and the result stays orthogonal to the ground state."""
    from mpsdyn_amd import device as dev
    cfg = dev.gs_config(0, 8, 16, [-1.0, -1.5, 0.05], convtol=1e-12, seed=2)
    e, ov, _ = dev.gs_solve_levels(ctx, cfg, 2, exc_orth_tol=1e-4)
    exact = np.linalg.eigvalsh(ph.ising_hamiltonian(8, -1.0, -1.5, 0.05).toarray())[:2]
    assert np.all(np.abs(e - exact) <= 1e-9 * np.abs(exact))
    assert ov[0, 1] <= 1e-4


def test_bad_arguments(ctx, tmp_path):
    from mpsdyn_amd import device as dev
    cfg = dev.gs_config(0, 6, 4, [-1.0, -1.0])
    start = _state_file(tmp_path, "s.mps", orc.random_mps(6, 4, seed=1))
    with pytest.raises(dev.MpsdError):
        dev.gs_sweep_part(ctx, cfg, start, 6, 0)       # pos0 past the end
    with pytest.raises(dev.MpsdError):
        dev.gs_sweep_part(ctx, cfg, start, 0, 8)       # lim past the end
    short = _state_file(tmp_path, "short.mps", orc.random_mps(5, 4, seed=2))
    with pytest.raises(dev.MpsdError):
        dev.gs_sweep_part(ctx, cfg, start, 0, 6, proj=[short], weights=[-1.0])
    with pytest.raises(dev.MpsdError):
        dev.gs_solve_levels(ctx, cfg, 0)
    r = dev.gs_solve(ctx, dev.gs_config(0, 6, 4, [-1.0, -1.0], convtol=1e-10))
    assert np.isfinite(r.energy)
