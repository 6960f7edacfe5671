"""The graph-replayed Lanczos path (device control block, one captured hipGraph per chunk)
against the stream-launched path: same eigenpairs and the same ground-state energies."""
import json
import os

import numpy as np
import pytest

from oracle import mpsd_oracle as orc
from tests import _physics as ph

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "reference_values.json")


@pytest.fixture
def ctx():
    from mpsdyn_amd import device as dev
    c = dev.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("graph,fused", [(0, 0), (1, 0), (1, 1)])
def test_eigs_graph_and_stream_match_dense(ctx, graph, fused):
    from mpsdyn_amd import device as dev
    ctx.set_option(dev.OPT_GRAPH, graph)
    ctx.set_option(dev.OPT_FUSED_LANCZOS, fused)
    L_, k = 10, 5
    mpo = orc.heisenberg_mpo(L_, 1.0, 1.0, 0.9)
    mps = orc.gauge_left_all(orc.gauge_right_all(orc.random_mps(L_, 16, seed=8)))
    for j in range(k):
        q, r = orc._qr_site(mps[j])
        mps[j] = q
        mps[j + 1] = orc._absorb_right(mps[j + 1], r)
    envL = np.ones((1, 1, 1), dtype=complex)
    for j in range(k):
        envL = orc.contract_left(envL, mps[j], mpo[j], mps[j])
    envR = np.ones((1, 1, 1), dtype=complex)
    for j in range(L_ - 1, k, -1):
        envR = orc.contract_right(envR, mps[j], mpo[j], mps[j])
    H = orc.heff_full(envL, mpo[k], envR, offset=0.5)
    e0 = np.linalg.eigvalsh(0.5 * (H + H.conj().T))[0]
    h = dev.Heff(ctx, ctx.tensor(envL), ctx.tensor(mpo[k]), ctx.tensor(envR), offset=0.5)
    x = ctx.tensor(mps[k])
    val, res, nmv = h.eigs_lowest(x, tol=1e-11)
    assert abs(val - e0) <= 1e-10 * max(1, abs(e0))
    # the returned vector is the eigenvector: ||H v - val v|| small, |v| = 1
    v = x.download().ravel(order="F")
    A = 0.5 * (H + H.conj().T)
    assert abs(np.linalg.norm(v) - 1.0) <= 1e-12
    assert np.linalg.norm(A @ v - val * v) <= 1e-8 * max(1, abs(e0))
    h.close()
    ctx.set_option(dev.OPT_FUSED_LANCZOS, 1)


def test_solve_graph_equals_stream():
    """Whole sweeps with graph replay on and off give the same energies (the two paths share
    their arithmetic; tolerance 1e-12 relative)."""
    from mpsdyn_amd import device as dev
    cfgs = [dev.gs_config(dev.MODEL_ISING, 20, 20, [-1.0, -1.0, -0.5], convtol=1e-12),
            dev.gs_config(dev.MODEL_SCHWINGER, 12, 64, [2.5, 100.0, 0.0, 0.0], convtol=1e-12),
            dev.gs_config(dev.MODEL_HEISENBERG, 16, 24, [1.0, 1.0, 0.8, 0.0], convtol=1e-12, precision=1)]
    graph = []
    for cfg in cfgs:
        out = []
        for g in (0, 1):
            c = dev.Context(0)
            c.set_option(dev.OPT_GRAPH, g)
            c.set_option(dev.OPT_FUSED_LANCZOS, 0)  # small sites would take the one-workgroup solver
            out.append(dev.gs_solve(c, cfg).energy)
            c.close()
        print(f"stream={out[0]:.16f} graph={out[1]:.16f} bitwise={out[0] == out[1]}", flush=True)
        assert abs(out[0] - out[1]) <= 1e-12 * abs(out[0])
        graph.append(out[1])
    # the graph path's Schwinger energy against exact diagonalisation (D = 64 is exact for 12 sites)
    e_ed = ph.ground_energy(ph.schwinger_hamiltonian(12, 2.5, 100.0))
    print(f"schwinger graph {graph[1]:.15f} ED {e_ed:.15f}", flush=True)
    assert abs(graph[1] - e_ed) <= 1e-9 * abs(e_ed)


def test_fused_solver_sweeps_match():
    """Sweeps with the one-workgroup eigensolver (small sites) and with the general solver
    converge to the same energies (1e-10 relative), and the fused path is the faster one."""
    import time
    from mpsdyn_amd import device as dev
    cfgs = [dev.gs_config(dev.MODEL_ISING, 20, 20, [-1.0, -1.0, -0.5], convtol=1e-12, seed=4),
            dev.gs_config(dev.MODEL_HEISENBERG, 16, 24, [1.0, 1.0, 0.8, 0.0], convtol=1e-12, seed=5),
            dev.gs_config(dev.MODEL_SCHWINGER, 12, 32, [2.5, 100.0, 0.0, 0.0], convtol=1e-12, seed=6)]
    for cfg in cfgs:
        e, t = [], []
        for f in (0, 1):
            c = dev.Context(0)
            c.set_option(dev.OPT_FUSED_LANCZOS, f)
            dev.gs_solve(c, cfg)  # warm
            t0 = time.perf_counter()
            e.append(dev.gs_solve(c, cfg).energy)
            t.append(time.perf_counter() - t0)
            c.close()
        print(f"general={e[0]:.15f} ({t[0]:.3f}s) fused={e[1]:.15f} ({t[1]:.3f}s)", flush=True)
        assert abs(e[0] - e[1]) <= 1e-10 * abs(e[0])
    ref = json.load(open(GOLDEN))["tfim_L20_D20"]["reference_energy"]
    c = dev.Context(0)
    r = dev.gs_solve(c, cfgs[0])
    c.close()
    assert abs(r.energy - ref) <= 1e-10 * abs(ref)
