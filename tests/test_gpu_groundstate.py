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
