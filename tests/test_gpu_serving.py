"""Batched and served forms: grouped matvecs are bit-identical to single matvecs, served
solves give the same energies alone and in the mix, admission respects the HBM budget."""
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


@pytest.mark.parametrize("prec", [0, 1])
def test_grouped_matvec_bit_identical(ctx, prec):
    """Uneven shapes (edge sites, w=3 and 5, D from 1 to 200) in one grouped launch equal the
    single-problem matvec bit for bit, and the oracle to the form's tolerance."""
    from mpsdyn_amd import device as dev
    shapes = [(16, 16, 3, 3), (1, 2, 1, 3), (64, 64, 5, 5), (37, 50, 5, 5), (200, 128, 3, 3), (2, 1, 3, 1),
              (128, 128, 5, 5)]
    hs, xs, ys, ref, keep = [], [], [], [], []
    for k, (Db, Dbr, wl, wr) in enumerate(shapes):
        L = orc.complex_gaussian(10 + k, (Db, wl, Db))
        W = orc.complex_gaussian(20 + k, (2, wl, 2, wr))
        R = orc.complex_gaussian(30 + k, (Dbr, wr, Dbr))
        x = orc.complex_gaussian(40 + k, (2, Db, Dbr))
        tL, tW, tR = ctx.tensor(L), ctx.tensor(W), ctx.tensor(R)
        h = dev.Heff(ctx, tL, tW, tR, offset=0.25 * k, precision=prec)
        hs.append(h)
        xs.append(ctx.tensor(x))
        ys.append(ctx.tensor(dims=(2, Db, Dbr)))
        ref.append(orc.heff_product(L, W, R, x, offset=0.25 * k))
        keep += [tL, tW, tR]
    single = []
    for h, x in zip(hs, xs):
        y1 = ctx.tensor(dims=x.dims)
        h.apply_device(x.devptr(), y1.devptr())
        single.append(y1.download())
    dev.heff_apply_grouped(ctx, hs, xs, ys)
    tol = 1e-12 if prec == 0 else 1e-6
    for k in range(len(hs)):
        yg = ys[k].download()
        assert np.array_equal(yg, single[k]), f"problem {k} differs from its single matvec"
        r = ref[k].reshape(-1, order="F")
        assert np.linalg.norm(yg.reshape(-1, order="F") - r) <= tol * np.linalg.norm(r)
    for h in hs:
        h.close()


def _cfgs(dev):
    return [
        dev.gs_config(dev.MODEL_ISING, 20, 20, [-1.0, -1.0, -0.5], convtol=1e-12, seed=1),
        dev.gs_config(dev.MODEL_HEISENBERG, 16, 24, [1.0, 1.0, 0.8, 0.0], convtol=1e-12, seed=2),
        dev.gs_config(dev.MODEL_SCHWINGER, 12, 32, [2.5, 100.0, 0.0, 0.0], convtol=1e-12, seed=3),
        dev.gs_config(dev.MODEL_ISING, 32, 48, [-1.0, -0.7, -0.2], convtol=1e-12, seed=4),
        dev.gs_config(dev.MODEL_HEISENBERG, 24, 64, [1.0, 1.0, 1.0, 0.0], convtol=1e-12, seed=7),
        dev.gs_config(dev.MODEL_ISING, 24, 32, [-1.0, -1.3, 0.0], convtol=1e-12, seed=5, precision=1),
        dev.gs_config(dev.MODEL_HEISENBERG, 20, 16, [1.0, 1.0, 1.2, 0.0], convtol=1e-12, seed=6, precision=2),
    ]


def test_served_results_equal_alone(ctx):
    """Each request solved alone (own context, no batching) and inside a served mix
    (concurrent, matvecs grouped) gives the same energy (tolerance 1e-12 relative; the paths
    are built to be bit-identical)."""
    from mpsdyn_amd import device as dev
    cfgs = _cfgs(dev)
    alone = [dev.gs_solve(ctx, c).energy for c in cfgs]
    srv = dev.Server(slots=len(cfgs), batching=True, window_us=300)
    res = srv.run_trace(cfgs, [0.0] * len(cfgs))
    st = srv.stats()
    srv.close()
    print(f"batches={st.batches} batched_matvecs={st.batched_matvecs} max_batch={st.max_batch} "
          f"peak_running={st.peak_running}", flush=True)
    for a, r in zip(alone, res):
        assert r.status == 0
        print(f"alone={a:.16f} served={r.energy:.16f} bitwise={a == r.energy}", flush=True)
        assert abs(a - r.energy) <= 1e-12 * abs(a)
    assert st.max_batch >= 2  # matvecs of concurrent solves were grouped
    # physics check on one of them
    e_ed = ph.ground_energy(ph.schwinger_hamiltonian(12, 2.5, 100.0))
    assert abs(res[2].energy - e_ed) <= 1e-10 * abs(e_ed)


@pytest.mark.parametrize("batching", [False, True])
def test_large_request_served_equals_alone(ctx, batching):
    """A config-sized request (Ising L = 100, D = 256) served next to a Schwinger L = 100,
    D = 128 one and two small ones, with and without grouped matvecs: each result equals the
    same solve alone (1e-12 relative; the paths are built to be bit-identical)."""
    from mpsdyn_amd import device as dev
    cfgs = [dev.gs_config(dev.MODEL_ISING, 100, 256, [-1.0, -0.8, -0.2], convtol=1e-10, seed=21),
            dev.gs_config(dev.MODEL_SCHWINGER, 100, 128, [2.5, 100.0, 0.0, 0.0], convtol=1e-10, seed=22),
            dev.gs_config(dev.MODEL_HEISENBERG, 24, 32, [1.0, 1.0, 0.8, 0.0], convtol=1e-10, seed=23),
            dev.gs_config(dev.MODEL_ISING, 40, 64, [-1.0, -1.1, 0.0], convtol=1e-10, seed=24, precision=1)]
    alone = [dev.gs_solve(ctx, c).energy for c in cfgs]
    srv = dev.Server(slots=len(cfgs), batching=batching, window_us=300)
    res = srv.run_trace(cfgs, [0.0, 0.05, 0.1, 0.15])
    st = srv.stats()
    srv.close()
    print(f"batching={batching} batches={st.batches} max_batch={st.max_batch} peak_running={st.peak_running}",
          flush=True)
    for a, r in zip(alone, res):
        assert r.status == 0
        print(f"alone={a:.16f} served={r.energy:.16f} bitwise={a == r.energy} "
              f"latency={r.finish_s - r.arrival_s:.2f}s", flush=True)
        assert abs(a - r.energy) <= 1e-12 * abs(a)
    assert st.peak_running >= 2  # the solves overlapped on the device


def test_served_graph_steps_equal_alone(ctx):
    """Served with graph = 1 (each worker's Lanczos steps replayed from captured hipGraphs, no
    batcher), every request gives the energy of the same solve alone on stream launches (1e-12
    relative) and the workers did launch captured graphs."""
    from mpsdyn_amd import device as dev
    cfgs = _cfgs(dev)[:5]
    alone = [dev.gs_solve(ctx, c).energy for c in cfgs]
    srv = dev.Server(slots=len(cfgs), batching=False, graph=True)
    res = srv.run_trace(cfgs, [0.0] * len(cfgs))
    st = srv.stats()
    srv.close()
    print(f"graph_replays={st.graph_replays} peak_running={st.peak_running}", flush=True)
    for a, r in zip(alone, res):
        assert r.status == 0
        print(f"alone={a:.16f} served={r.energy:.16f} bitwise={a == r.energy}", flush=True)
        assert abs(a - r.energy) <= 1e-12 * abs(a)
    assert st.graph_replays > 0


def test_admission_respects_budget():
    """With a budget that holds two of the requests, never more than two run at once and
    the admitted footprint stays within the budget; later arrivals wait, then finish."""
    from mpsdyn_amd import device as dev
    cfgs = [dev.gs_config(dev.MODEL_ISING, 24, 32, [-1.0, -1.0, -0.1 * k], convtol=1e-10, seed=k + 1)
            for k in range(5)]
    fp = max(dev.gs_footprint(c) for c in cfgs)
    srv = dev.Server(slots=5, hbm_budget=2 * fp + fp // 2, batching=True)
    res = srv.run_trace(cfgs, [0.0, 0.0, 0.0, 0.01, 0.02])
    st = srv.stats()
    srv.close()
    assert all(r.status == 0 for r in res)
    assert st.peak_running <= 2
    assert st.peak_admitted_bytes <= 2 * fp + fp // 2
    waits = [r.admit_s - r.arrival_s for r in res]
    assert max(waits) > 0  # someone queued for memory
    for k, r in enumerate(res):
        e = ph.tfim_free_fermion(24, 1.0, 1.0) if k == 0 else None
        if e is not None:
            assert abs(r.energy - e) <= 1e-9 * abs(e)


def test_request_too_large_is_rejected():
    from mpsdyn_amd import device as dev
    cfg = dev.gs_config(dev.MODEL_ISING, 20, 20, [-1.0, -1.0, 0.0])
    srv = dev.Server(slots=1, hbm_budget=1024, batching=False)
    with pytest.raises(dev.MpsdError) as e:
        srv.run_trace([cfg], [0.0])
    assert e.value.code == -4
    srv.close()
