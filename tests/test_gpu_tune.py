"""Tile autotuner (mpsd_tune_*): every tile configuration it may choose gives bit-identical
results, the tuned table changes only timing, and the table survives save / load.
Parity of the tuned matvec against the CPU oracle: relative 2-norm error <= 1e-12."""
import numpy as np
import pytest

from oracle import mpsd_oracle as orc

pytestmark = pytest.mark.gpu

TOL = 1e-12
CONFIGS = [(0, 0, 16, 0), (32, 32, 16, 1), (32, 32, 32, 0), (32, 32, 32, 1), (64, 32, 16, 0),
           (64, 32, 16, 1), (32, 64, 16, 0), (32, 64, 16, 1), (64, 64, 16, 0), (64, 64, 32, 2),
           (16, 16, 16, 0), (16, 16, 32, 0), (16, 16, 16, 1), (16, 16, 32, 2)]


@pytest.fixture(scope="module")
def ctx():
    from mpsdyn_amd import device as dev
    c = dev.Context(0)
    yield c
    c.close()


def _rel(a, b):
    return np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300)


def _set_cfg(ctx, cfg):
    from mpsdyn_amd import device as dev
    tm, tn, tk, nbw = cfg
    ctx.set_tile(tm, tn)
    ctx.set_option(dev.OPT_TILE_K, tk)
    ctx.set_option(dev.OPT_BATCH_GROUP, nbw)


def _heff(ctx, Db, Dbr, wl, wr, seed, offset=0.0):
    from mpsdyn_amd import device as dev
    L = orc.complex_gaussian(seed + 1, (Db, wl, Db))
    W = orc.complex_gaussian(seed + 2, (2, wl, 2, wr))
    R = orc.complex_gaussian(seed + 3, (Dbr, wr, Dbr))
    x = orc.complex_gaussian(seed + 4, (2, Db, Dbr))
    h = dev.Heff(ctx, ctx.tensor(L), ctx.tensor(W), ctx.tensor(R), offset=offset)
    return h, (L, W, R, x)


@pytest.mark.parametrize("Db,Dbr,wl,wr", [(37, 50, 5, 5), (96, 96, 3, 3), (130, 70, 3, 5), (256, 256, 3, 3)])
def test_every_tile_config_is_bit_identical(ctx, Db, Dbr, wl, wr):
    print(f"tile configs Db={Db} Dbr={Dbr} wl={wl} wr={wr}", flush=True)
    h, (L, W, R, x) = _heff(ctx, Db, Dbr, wl, wr, seed=Db + Dbr, offset=0.5)
    xt = ctx.tensor(x)
    yt = ctx.tensor(dims=(2, Db, Dbr))
    ys = []
    for cfg in CONFIGS:
        _set_cfg(ctx, cfg)
        h.apply_device(xt.devptr(), yt.devptr())
        ys.append(yt.download())
    _set_cfg(ctx, (0, 0, 16, 0))
    for cfg, y in zip(CONFIGS, ys):
        assert np.array_equal(y, ys[0]), f"config {cfg} changed the result"
    ref = orc.heff_product(L, W, R, x.ravel(order="F")) + 0.5 * x.ravel(order="F")
    assert _rel(ys[0].ravel(order="F"), ref) <= TOL
    h.close()


def test_tune_heff_keeps_results_and_parity(ctx):
    print("tune heff D=192 w=3", flush=True)
    ctx.tune_clear()
    h, (L, W, R, x) = _heff(ctx, 192, 192, 3, 3, seed=11)
    xt = ctx.tensor(x)
    yt = ctx.tensor(dims=(2, 192, 192))
    h.apply_device(xt.devptr(), yt.devptr())
    before = yt.download()
    rows = h.tune(xt.devptr(), yt.devptr(), reps=3)
    assert ctx.tune_count() >= 2 and len(rows) >= 2
    for r in rows:
        assert (r["tm"], r["tn"]) in [(0, 0), (32, 32), (64, 32), (32, 64), (64, 64)]
        assert r["tk"] in (16, 32) and r["nbw"] in (0, 1, 2)
        assert r["us_best"] > 0 and r["us_best"] <= r["us_builtin"] * 1.0001
    assert np.array_equal(yt.download(), before)  # tune_heff leaves y = Heff x
    h.apply_device(xt.devptr(), yt.devptr())
    assert np.array_equal(yt.download(), before)
    assert _rel(before.ravel(order="F"), orc.heff_product(L, W, R, x.ravel(order="F"))) <= TOL
    h.close()


def test_tune_env_and_save_load(ctx, tmp_path):
    from mpsdyn_amd import device as dev
    print("tune env D=160 w=5", flush=True)
    ctx.tune_clear()
    D, w, d = 160, 5, 2
    E = orc.complex_gaussian(21, (D, w, D))
    A = orc.complex_gaussian(22, (d, D, D))
    W = orc.complex_gaussian(23, (d, w, d, w))
    tE, tA, tW = ctx.tensor(E), ctx.tensor(A), ctx.tensor(W)
    out = ctx.tensor(dims=(D, w, D))
    results = {}
    for side, fn in ((0, dev.env_left), (1, dev.env_right)):
        fn(ctx, tE, tA, tW, tA, out)
        results[side] = out.download()
        rows = ctx.tune_env(side, tE, tA, tW, tA, out, reps=3)
        assert len(rows) >= 1
        fn(ctx, tE, tA, tW, tA, out)
        assert np.array_equal(out.download(), results[side])
    ref_left = orc.contract_left(E, A, W, A)
    assert _rel(results[0], ref_left) <= TOL
    n = ctx.tune_count()
    path = tmp_path / "tiles.txt"
    ctx.tune_save(path)
    c2 = dev.Context(0)
    try:
        assert c2.tune_load(path) == n
        assert c2.tune_count() == n
        tE2, tA2, tW2 = c2.tensor(E), c2.tensor(A), c2.tensor(W)
        out2 = c2.tensor(dims=(D, w, D))
        dev.env_left(c2, tE2, tA2, tW2, tA2, out2)
        assert np.array_equal(out2.download(), results[0])
    finally:
        c2.close()


def test_tune_load_rejects_bad_table(ctx, tmp_path):
    from mpsdyn_amd import device as dev
    bad = tmp_path / "bad.txt"
    bad.write_text("1 2 3 4 5 6 48 32 16 0\n")
    with pytest.raises(dev.MpsdError):
        ctx.tune_load(bad)
    with pytest.raises(dev.MpsdError):
        ctx.tune_load(tmp_path / "missing.txt")
