"""Precision form: complex-FP32 and split-BF16 MFMA matvecs against the FP64 oracle, and
ground-state energies of the mixed-precision sweep against the FP64 device sweep and exact
diagonalisation.  Tolerances: matvec relative error 1e-6 (FP32) / 1e-4 (split BF16);
energies 1e-10 relative to the FP64 result (absolute differences are printed; the
reference's stopping rule |dE| < convtol |E| is relative)."""
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


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("prec,tol", [(1, 1e-6), (2, 1e-4)])
@pytest.mark.parametrize("Db,Dbr,wl,wr", [(20, 20, 3, 3), (37, 50, 5, 5), (1, 4, 1, 3), (128, 128, 5, 5),
                                          (512, 512, 3, 3), (256, 256, 5, 5), (400, 400, 5, 5)])
def test_lowp_matvec_accuracy(ctx, prec, tol, Db, Dbr, wl, wr):
    from mpsdyn_amd import device as dev
    print(f"prec={prec} Db={Db} Dbr={Dbr} wl={wl} wr={wr}", flush=True)
    L = orc.complex_gaussian(1, (Db, wl, Db))
    W = orc.complex_gaussian(2, (2, wl, 2, wr))
    R = orc.complex_gaussian(3, (Dbr, wr, Dbr))
    x = orc.complex_gaussian(4, (2 * Db * Dbr,))
    h = dev.Heff(ctx, ctx.tensor(L), ctx.tensor(W), ctx.tensor(R), offset=0.3, precision=prec)
    err = _rel(h.apply(x), orc.heff_product(L, W, R, x, offset=0.3))
    assert err <= tol, err
    assert err > 1e-15  # really the reduced-precision path
    h.set_precision(0)
    assert _rel(h.apply(x), orc.heff_product(L, W, R, x, offset=0.3)) <= 1e-12
    h.close()


@pytest.mark.parametrize("prec,tol", [(0, 1e-12), (1, 1e-6), (2, 1e-4)])
@pytest.mark.parametrize("d,Db,Dbr,wl,wr", [(3, 24, 17, 3, 5), (1, 40, 33, 4, 4), (4, 16, 16, 2, 3),
                                            (3, 96, 96, 5, 5)])
def test_matvec_other_physical_dimensions(ctx, prec, tol, d, Db, Dbr, wl, wr):
    """Physical dimension d != 2 (spin-1 chains, d = 1 and 4 as well) in every arithmetic form:
    the packed precision forms take any d, the FP64 form its in-place path, all against the
    oracle's product."""
    from mpsdyn_amd import device as dev
    print(f"prec={prec} d={d} Db={Db} Dbr={Dbr} wl={wl} wr={wr}", flush=True)
    L = orc.complex_gaussian(61, (Db, wl, Db))
    W = orc.complex_gaussian(62, (d, wl, d, wr))
    R = orc.complex_gaussian(63, (Dbr, wr, Dbr))
    x = orc.complex_gaussian(64, (d * Db * Dbr,))
    h = dev.Heff(ctx, ctx.tensor(L), ctx.tensor(W), ctx.tensor(R), offset=-0.4, precision=prec)
    err = _rel(h.apply(x), orc.heff_product(L, W, R, x, offset=-0.4))
    assert err <= tol, err
    h.close()


@pytest.mark.parametrize("prec", [1, 2])
def test_mixed_precision_schwinger_energy(ctx, prec):
    """Schwinger N=12, x=100, m/g=0.125: mixed-precision energy within 1e-10 of the FP64
    sweep and of exact diagonalisation, with most matvecs in the low-precision form."""
    from mpsdyn_amd import device as dev
    N, x, mg = 12, 100.0, 0.125
    mu = 2 * mg * np.sqrt(x)
    e_ed = ph.ground_energy(ph.schwinger_hamiltonian(N, mu, x))
    p = [mu, x, 0.0, 0.0]
    r64 = dev.gs_solve(ctx, dev.gs_config(dev.MODEL_SCHWINGER, N, 48, p, convtol=1e-12))
    rlp = dev.gs_solve(ctx, dev.gs_config(dev.MODEL_SCHWINGER, N, 48, p, convtol=1e-12, precision=prec))
    print(f"prec={prec} E64={r64.energy:.15f} Elp={rlp.energy:.15f} ED={e_ed:.15f} "
          f"mv={rlp.matvecs} lowp={rlp.matvecs_lowp} refine_from={rlp.refine_from}", flush=True)
    assert abs(rlp.energy - r64.energy) <= 1e-10 * abs(r64.energy)
    assert abs(rlp.energy - e_ed) <= 1e-10 * abs(e_ed)
    assert rlp.refine_from >= 1


@pytest.mark.parametrize("prec", [1, 2])
def test_mixed_precision_bulk_in_low_precision(ctx, prec):
    """Schwinger N=40, x=100, m/g=0.125, D=48: energy within 1e-10 of the FP64 sweep; the early,
    loose half-sweeps run the low-precision matvec, and at least the last two half-sweeps run the
    FP64 one (the precision follows the local tolerance)."""
    from mpsdyn_amd import device as dev
    N, x, mg = 40, 100.0, 0.125
    p = [2 * mg * np.sqrt(x), x, 0.0, 0.0]
    r64 = dev.gs_solve(ctx, dev.gs_config(dev.MODEL_SCHWINGER, N, 48, p, convtol=1e-12))
    rlp = dev.gs_solve(ctx, dev.gs_config(dev.MODEL_SCHWINGER, N, 48, p, convtol=1e-12, precision=prec))
    print(f"prec={prec} E64={r64.energy:.15f} Elp={rlp.energy:.15f} mv64={r64.matvecs} "
          f"mv={rlp.matvecs} lowp={rlp.matvecs_lowp} refine_from={rlp.refine_from} hs={rlp.halfsweeps}", flush=True)
    assert abs(rlp.energy - r64.energy) <= 1e-10 * abs(r64.energy)
    assert rlp.matvecs_lowp >= 0.15 * rlp.matvecs
    assert 1 <= rlp.refine_from <= rlp.halfsweeps - 2


def test_mixed_precision_ising_closed_form(ctx):
    from mpsdyn_amd import device as dev
    L = 40
    r = dev.gs_solve(ctx, dev.gs_config(dev.MODEL_ISING, L, 32, [-1.0, -1.0, 0.0], convtol=1e-12, precision=1))
    e = ph.tfim_free_fermion(L, 1.0, 1.0)
    assert abs(r.energy - e) <= 1e-10 * abs(e)


def test_schwinger_L100_precision_forms():
    """Config 3 at its size: Schwinger L = 100, x = 100, m/g = 0.125 (mu = 2 (m/g) sqrt(x) = 2.5,
    schwingerPD.cpp:51), zero background field, D = 128.  The complex-FP32 and split-BF16 forms
    (FP64 orthogonalisation, Rayleigh-Ritz, gauge and refinement) reach the FP64 energy to 1e-10
    relative, and their states' residuals ||H psi - E psi|| are those of the FP64 state (the D =
    128 truncation, not the arithmetic, sets them: within 50%)."""
    from mpsdyn_amd import device as dev
    p = [2 * 0.125 * np.sqrt(100.0), 100.0, 0.0, 0.0]
    c = dev.Context(0)
    out = {}
    for prec in (0, 1, 2):
        r = dev.gs_solve(c, dev.gs_config(dev.MODEL_SCHWINGER, 100, 128, p, convtol=1e-10, precision=prec,
                                          want_residual=True))
        out[prec] = r
        print(f"prec={prec} E={r.energy:.12f} residual={r.residual_norm:.3e} (floor {r.residual_floor:.1e}) "
              f"halfsweeps={r.halfsweeps} mv={r.matvecs} lowp={r.matvecs_lowp} {r.seconds:.2f}s", flush=True)
    c.close()
    e64, res64 = out[0].energy, out[0].residual_norm
    assert out[0].residual_floor < 0.1 * res64 < 1e-3  # resolved: the D = 128 truncation sets it
    for prec in (1, 2):
        r = out[prec]
        assert abs(r.energy - e64) <= 1e-10 * abs(e64)
        assert abs(r.residual_norm - res64) <= 0.5 * res64
        assert r.matvecs_lowp > 0


@pytest.mark.parametrize("prec,tol", [(1, 1e-6), (2, 1e-4)])
@pytest.mark.parametrize("Db,Dbr,wl,wr", [(20, 20, 3, 3), (37, 50, 5, 5), (1, 4, 1, 3), (3, 1, 3, 1),
                                          (100, 70, 3, 5), (256, 256, 3, 3)])
def test_packed_tiles_bit_identical(ctx, prec, tol, Db, Dbr, wl, wr):
    """Packed-operand precision forms: every BRGEMM tile (32x32, 64x32, 32x64, 64x64) sums each
    output in the same k order, so all give the same bits; they match the oracle to the form's
    tolerance, as does the form on FP64 operands converted per tile."""
    from mpsdyn_amd import device as dev
    print(f"prec={prec} Db={Db} Dbr={Dbr} wl={wl} wr={wr}", flush=True)
    L = orc.complex_gaussian(5, (Db, wl, Db))
    W = orc.complex_gaussian(6, (2, wl, 2, wr))
    R = orc.complex_gaussian(7, (Dbr, wr, Dbr))
    x = orc.complex_gaussian(8, (2 * Db * Dbr,))
    ref = orc.heff_product(L, W, R, x, offset=-0.7)
    h = dev.Heff(ctx, ctx.tensor(L), ctx.tensor(W), ctx.tensor(R), offset=-0.7, precision=prec)
    outs = []
    try:
        for tile in (0, 1, 2, 3, 4):
            ctx.set_option(dev.OPT_LP_TILE, tile)
            outs.append(h.apply(x))
        ctx.set_option(dev.OPT_LP_TILE, 0)
        ctx.set_option(dev.OPT_LP_PACKED, 0)
        staged = h.apply(x)
    finally:
        ctx.set_option(dev.OPT_LP_TILE, 0)
        ctx.set_option(dev.OPT_LP_PACKED, 1)
    for k, y in enumerate(outs[1:]):
        assert np.array_equal(y, outs[0]), f"tile option {k + 1} differs"
    assert _rel(outs[0], ref) <= tol
    assert _rel(staged, ref) <= tol
    h.close()


@pytest.mark.parametrize("prec", [1, 2])
def test_packed_operands_per_handle(ctx, prec):
    """Two handles on one context, applied alternately, each use their own packed environments;
    set_precision takes a fresh copy after the environment tensor was rewritten."""
    from mpsdyn_amd import device as dev
    tol = 1e-6 if prec == 1 else 1e-4
    hs, data = [], []
    for k, (D, w) in enumerate([(48, 3), (33, 5)]):
        L = orc.complex_gaussian(50 + k, (D, w, D))
        W = orc.complex_gaussian(60 + k, (2, w, 2, w))
        R = orc.complex_gaussian(70 + k, (D, w, D))
        tL = ctx.tensor(L)
        keep = (tL, ctx.tensor(W), ctx.tensor(R))
        hs.append(dev.Heff(ctx, *keep, offset=0.1, precision=prec))
        data.append([L, W, R, keep])
    for it in range(3):
        for k, h in enumerate(hs):
            L, W, R, _ = data[k]
            x = orc.complex_gaussian(80 + 10 * it + k, (h.size,))
            assert _rel(h.apply(x), orc.heff_product(L, W, R, x, offset=0.1)) <= tol
    L, W, R, keep = data[0]
    L2 = orc.complex_gaussian(99, L.shape)
    keep[0].upload(L2)
    hs[0].set_precision(prec)
    x = orc.complex_gaussian(98, (hs[0].size,))
    assert _rel(hs[0].apply(x), orc.heff_product(L2, W, R, x, offset=0.1)) <= tol
    for h in hs:
        h.close()
