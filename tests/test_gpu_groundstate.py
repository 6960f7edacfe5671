"""Ground-state energies of the device sweep (C++ Contractor over the HIP kernels) against
the reference's own energies, exact diagonalisation, free-fermion closed forms and the
CPU oracle sweep.  Energy tolerance: 1e-10 relative (FP64)."""
import json
import os

import numpy as np
import pytest

from oracle import mpsd_oracle as orc
from tests import _physics as ph

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "reference_values.json")


@pytest.fixture(scope="module")
def ctx():
    from mpsdyn_amd import device as dev
    c = dev.Context(0)
    yield c
    c.close()


def _solve(ctx, model, L, D, p, **kw):
    from mpsdyn_amd import device as dev
    print(f"solve model={model} L={L} D={D} p={p}", flush=True)
    return dev.gs_solve(ctx, dev.gs_config(model, L, D, p, **kw))


def test_tfim_L8_reference_value(ctx):
    from mpsdyn_amd import device as dev
    g = json.load(open(GOLDEN))["tfim_L8_critical"]
    r = _solve(ctx, dev.MODEL_ISING, 8, g["D"], [-1.0, -1.0, 0.0], convtol=1e-12)
    assert abs(r.energy - g["reference_energy"]) < 1e-9
    assert abs(r.energy - ph.tfim_free_fermion(8, 1.0, 1.0)) < 1e-9


def test_config1_reference_value(ctx):
    """TFIM -sum ZZ - sum X - 0.5 sum Z, L=20, D=20, convtol 1e-12."""
    from mpsdyn_amd import device as dev
    g = json.load(open(GOLDEN))["tfim_L20_D20"]
    r = _solve(ctx, dev.MODEL_ISING, 20, 20, [-1.0, -1.0, -0.5], convtol=1e-12)
    assert abs(r.energy - g["reference_energy"]) < 1e-10 * abs(g["reference_energy"])
    assert r.halfsweeps >= 2 and r.matvecs > 0


def test_heisenberg_ed(ctx):
    from mpsdyn_amd import device as dev
    L = 10
    e_ed = ph.ground_energy(ph.heisenberg_hamiltonian(L, 1, 1, 1))
    r = _solve(ctx, dev.MODEL_HEISENBERG, L, 32, [1.0, 1.0, 1.0, 0.0], convtol=1e-13)
    assert abs(r.energy - e_ed) < 1e-10 * abs(e_ed)


def test_schwinger_ed(ctx):
    from mpsdyn_amd import device as dev
    N, x, mg = 10, 100.0, 0.125
    mu = 2 * mg * np.sqrt(x)
    e_ed = ph.ground_energy(ph.schwinger_hamiltonian(N, mu, x))
    r = _solve(ctx, dev.MODEL_SCHWINGER, N, 32, [mu, x, 0.0, 0.0], convtol=1e-13)
    assert abs(r.energy - e_ed) < 1e-10 * abs(e_ed)


def test_tfim_L60_free_fermions(ctx):
    """Open critical TFIM at L=60, D=48 against the free-fermion closed form."""
    from mpsdyn_amd import device as dev
    r = _solve(ctx, dev.MODEL_ISING, 60, 48, [-1.0, -1.0, 0.0], convtol=1e-12)
    e = ph.tfim_free_fermion(60, 1.0, 1.0)
    assert abs(r.energy - e) < 1e-9 * abs(e)


def test_matches_oracle_sweep(ctx):
    """Device sweep and the CPU oracle sweep agree on a Heisenberg XXZ chain (Delta=0.8)."""
    from mpsdyn_amd import device as dev
    L, D = 16, 16
    r = _solve(ctx, dev.MODEL_HEISENBERG, L, D, [1.0, 1.0, 0.8, 0.0], convtol=1e-12)
    e, _, _, _ = orc.find_ground_state(orc.heisenberg_mpo(L, 1, 1, 0.8), orc.random_mps(L, D, seed=3),
                                       convtol=1e-12)
    assert abs(r.energy - e) < 1e-10 * abs(e)


def test_export_import_checkpoint(ctx, tmp_path):
    """exportMPS layout round trip: the exported ground state's <H> equals the energy."""
    import ctypes as C
    from mpsdyn_amd import device as dev
    cfg = dev.gs_config(dev.MODEL_ISING, 12, 16, [-1.0, -0.7, 0.0], convtol=1e-12)
    path = str(tmp_path / "gs.mps")
    r = dev.gs_solve(ctx, cfg, export_path=path)
    re, im = C.c_double(), C.c_double()
    ctx.check(ctx.lib.mpsd_gs_expectation(ctx.handle, C.byref(cfg), path.encode(), C.byref(re), C.byref(im)))
    assert abs(re.value - r.energy) < 1e-10 * abs(r.energy)
    assert abs(im.value) < 1e-10
    # header of MPS::exportMPS: int L, int bond, double normfact, int gauge
    raw = open(path, "rb").read()
    assert np.frombuffer(raw[:8], dtype=np.int32).tolist() == [12, 16]


@pytest.mark.parametrize("model,L,D,p,convtol", [
    (0, 100, 128, [-1.0, -1.0, 0.0], 1e-10),
    (0, 100, 256, [-1.0, -1.0, 0.0], 1e-10),
    (2, 40, 48, [2.5, 100.0, 0.0, 0.0], 1e-12),
    (1, 32, 64, [1.0, 1.0, 1.0, 0.0], 1e-12),
    (0, 20, 20, [-1.0, -1.0, -0.5], 1e-12),
])
def test_tolerance_schedule_reaches_the_fixed_tolerance_energy(ctx, model, L, D, p, convtol):
    """The local-tolerance schedule (loose FP64 local solves while the sweep is far from
    converged, eigtol for the last two half-sweeps) ends at the energy of the sweep that solves
    every local problem to eigtol, to 1e-10 relative, with fewer matvecs on the large sites."""
    from mpsdyn_amd import device as dev
    fixed = dev.gs_solve(ctx, dev.gs_config(model, L, D, p, convtol=convtol))
    sched = dev.gs_solve(ctx, dev.gs_config(model, L, D, p, convtol=convtol, tol_schedule=True))
    print(f"model {model} L={L} D={D}: fixed E={fixed.energy:.15f} mv={fixed.matvecs} {fixed.seconds:.2f}s | "
          f"schedule E={sched.energy:.15f} mv={sched.matvecs} {sched.seconds:.2f}s", flush=True)
    assert abs(sched.energy - fixed.energy) <= 1e-10 * abs(fixed.energy)
    if L == 100:
        assert sched.matvecs < fixed.matvecs


@pytest.mark.parametrize("prec", [0, 1])
def test_schedule_tightens_every_half_sweep(ctx, prec):
    """Per-half-sweep record of a scheduled solve (Schwinger L=40, D=48): each half-sweep's
    local tolerance is at least 100 times tighter than the one before until it reaches eigtol
    (1e-10), the last two run at eigtol, the matvec counts add up to the total and the seconds fit
    in the wall time."""
    from mpsdyn_amd import device as dev
    r = dev.gs_solve(ctx, dev.gs_config(dev.MODEL_SCHWINGER, 40, 48, [2.5, 100.0, 0.0, 0.0], convtol=1e-12,
                                        tol_schedule=True, precision=prec))
    n = r.n_history
    tols = [r.hs_tol[i] for i in range(n)]
    mvs = [r.hs_matvecs[i] for i in range(n)]
    print(f"prec={prec}: tols {tols} matvecs {mvs} seconds {[round(r.hs_seconds[i], 3) for i in range(n)]}",
          flush=True)
    assert n == r.halfsweeps and n >= 3
    assert tols[0] == 1e-3
    for a, b in zip(tols, tols[1:]):
        assert b == 1e-10 or b <= 1e-2 * a * (1 + 1e-12)
    assert tols[-1] == tols[-2] == 1e-10
    assert sum(mvs) == r.matvecs and all(m > 0 for m in mvs)
    assert 0 < sum(r.hs_seconds[i] for i in range(n)) <= r.seconds


@pytest.mark.parametrize("L,D,g", [(180, 64, 1.5), (120, 256, 1.5)])
def test_long_chain_reaches_the_closed_form(ctx, L, D, g):
    """Long chains (the served mix draws L up to 200): the seeded start state stays inside the
    double range through its right gauge (its norm would otherwise grow like prod sqrt(2 d D) to
    about 1e160 here), and the open TFIM at g = 1.5 reaches its free-fermion energy."""
    from mpsdyn_amd import device as dev
    r = dev.gs_solve(ctx, dev.gs_config(dev.MODEL_ISING, L, D, [-1.0, -g, 0.0], convtol=1e-10, tol_schedule=True))
    e = ph.tfim_free_fermion(L, 1.0, g)
    print(f"L={L} D={D} g={g}: E={r.energy:.12f} closed form {e:.12f} ({r.seconds:.2f}s, {r.matvecs} matvecs)",
          flush=True)
    assert abs(r.energy - e) <= 1e-10 * abs(e)
