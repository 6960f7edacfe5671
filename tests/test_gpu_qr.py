"""The library's blocked Householder QR behind the gauge steps (Site::gaugeR / gaugeL), against
rocSOLVER's zgeqrf/zungqr (zgelqf/zunglq) on the same inputs and against the defining
properties: Q^H Q = I, Q R = A, R upper triangular with a real diagonal.  Tolerances:
orthogonality and reconstruction 1e-13 relative (x sqrt of the size), agreement with
rocSOLVER 1e-12 relative."""
import numpy as np
import pytest

from oracle import mpsd_oracle as orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from mpsdyn_amd import device as dev
    c = dev.Context(0)
    yield c
    c.set_option(dev.OPT_GAUGE_QR, 0)
    c.close()


def _gauge_r(ctx, A, path):
    from mpsdyn_amd import device as dev
    ctx.set_option(dev.OPT_GAUGE_QR, path)
    tA, tR = ctx.tensor(A), ctx.tensor(dims=(1, 1))
    ctx.check(ctx.lib.mpsd_site_gauge_r(ctx.handle, tA.handle, tR.handle))
    out = tA.download(), tR.download()
    ctx.set_option(dev.OPT_GAUGE_QR, 0)
    return out


def _gauge_l(ctx, A, path):
    from mpsdyn_amd import device as dev
    ctx.set_option(dev.OPT_GAUGE_QR, path)
    tA, tL = ctx.tensor(A), ctx.tensor(dims=(1, 1))
    ctx.check(ctx.lib.mpsd_site_gauge_l(ctx.handle, tA.handle, tL.handle))
    out = tA.download(), tL.download()
    ctx.set_option(dev.OPT_GAUGE_QR, 0)
    return out


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


# (d, Dl, Dr): the single-workgroup form (m = d Dl <= 64, n = Dr <= 64: single rows and columns,
# one, two and four rows per thread, wide and the largest cases; with m + n <= 64 Q is accumulated
# in the spare lanes, up to the boundary m + n = 64), panel edges (16), one- to
# four-row-per-thread panels (m up to 1024), wide factors
SHAPES = [(2, 1, 1), (2, 1, 2), (2, 8, 1), (2, 8, 16), (2, 20, 20), (2, 32, 32), (2, 32, 64), (2, 16, 64),
          (3, 20, 20), (2, 33, 17), (2, 50, 64), (2, 64, 64), (2, 65, 64), (2, 64, 65), (2, 128, 16),
          (2, 128, 128), (2, 200, 150), (2, 200, 4), (2, 256, 256), (2, 300, 512), (2, 512, 512), (3, 10, 40),
          (3, 12, 20), (2, 30, 4), (2, 16, 33)]


@pytest.mark.parametrize("d,Dl,Dr", SHAPES)
def test_gauge_r_householder(ctx, d, Dl, Dr):
    print(f"gauge_r d={d} Dl={Dl} Dr={Dr}", flush=True)
    A = orc.complex_gaussian(100 + Dl + Dr, (d, Dl, Dr))
    Q, R = _gauge_r(ctx, A, 0)
    m, k = d * Dl, min(d * Dl, Dr)
    Qm = orc.reshape_f(Q, (m, k))
    Am = orc.reshape_f(A, (m, Dr))
    assert np.linalg.norm(Qm.conj().T @ Qm - np.eye(k)) <= 1e-13 * np.sqrt(k) * 10
    assert _rel(Qm @ R, Am) <= 1e-13 * np.sqrt(m)
    assert np.abs(np.tril(R, -1)).max() == 0.0
    assert np.abs(np.diag(R).imag).max() == 0.0
    # same factorisation as rocSOLVER (LAPACK conventions) up to rounding
    Q2, R2 = _gauge_r(ctx, A, 1)
    assert _rel(R, R2) <= 1e-12 and _rel(Q, Q2) <= 1e-12


@pytest.mark.parametrize("d,Dl,Dr", [(2, 1, 1), (2, 16, 8), (2, 40, 20), (2, 64, 64), (2, 64, 32), (2, 256, 200),
                                     (2, 512, 512), (3, 40, 10), (2, 20, 20), (3, 8, 12), (2, 32, 16)])
def test_gauge_l_householder(ctx, d, Dl, Dr):
    print(f"gauge_l d={d} Dl={Dl} Dr={Dr}", flush=True)
    A = orc.complex_gaussian(200 + Dl + Dr, (d, Dl, Dr))
    Q, Lm = _gauge_l(ctx, A, 0)
    k = min(Dl, d * Dr)
    Qm = orc.reshape_f(orc.permute(Q, (2, 1, 3)), (k, d * Dr))
    Mm = orc.reshape_f(orc.permute(A, (2, 1, 3)), (Dl, d * Dr))
    assert np.linalg.norm(Qm @ Qm.conj().T - np.eye(k)) <= 1e-13 * np.sqrt(k) * 10
    assert _rel(Lm @ Qm, Mm) <= 1e-13 * np.sqrt(d * Dr)
    assert np.abs(np.triu(Lm, 1)).max() == 0.0
    Q2, L2 = _gauge_l(ctx, A, 1)
    assert _rel(Lm, L2) <= 1e-12 and _rel(Q, Q2) <= 1e-12


@pytest.mark.parametrize("d,Dl,Dr", [(2, 30, 48), (2, 100, 96)])
def test_rank_deficient_and_zero_columns(ctx, d, Dl, Dr):
    """Exactly zero and linearly dependent columns (tau = 0 reflectors, tiny pivots) still give
    an isometry that reconstructs A, in the single-workgroup and the panel form."""
    A = orc.complex_gaussian(7, (d, Dl, Dr))
    Am = orc.reshape_f(A, (d * Dl, Dr)).copy()
    Am[:, 5] = 0.0
    Am[:, 17] = 0.0
    Am[:, 30] = 2.0 * Am[:, 3] - 1j * Am[:, 9]
    Am[:, 40] = 1e-200 * Am[:, 2]
    A = orc.reshape_f(Am, (d, Dl, Dr))
    Q, R = _gauge_r(ctx, A, 0)
    k = min(d * Dl, Dr)
    Qm = orc.reshape_f(Q, (d * Dl, k))
    assert np.all(np.isfinite(Qm)) and np.all(np.isfinite(R))
    assert np.linalg.norm(Qm.conj().T @ Qm - np.eye(k)) <= 1e-12
    assert _rel(Qm @ R, Am) <= 1e-13 * np.sqrt(d * Dl)


def test_rank_deficient_fused_q(ctx):
    """The same degenerate columns where Q is accumulated in the spare lanes (m + n <= 64)."""
    d, Dl, Dr = 2, 12, 30
    A = orc.complex_gaussian(8, (d, Dl, Dr))
    Am = orc.reshape_f(A, (d * Dl, Dr)).copy()
    Am[:, 4] = 0.0
    Am[:, 11] = 0.0
    Am[:, 20] = 2.0 * Am[:, 3] - 1j * Am[:, 9]
    Am[:, 25] = 1e-200 * Am[:, 2]
    A = orc.reshape_f(Am, (d, Dl, Dr))
    Q, R = _gauge_r(ctx, A, 0)
    k = min(d * Dl, Dr)
    Qm = orc.reshape_f(Q, (d * Dl, k))
    assert np.all(np.isfinite(Qm)) and np.all(np.isfinite(R))
    assert np.linalg.norm(Qm.conj().T @ Qm - np.eye(k)) <= 1e-12
    assert _rel(Qm @ R, Am) <= 1e-13 * np.sqrt(d * Dl)


def test_ground_state_unchanged_by_qr_path(ctx):
    """A full sweep gives the same energy with either QR path (parity of the gauge steps)."""
    from mpsdyn_amd import device as dev
    cfg = dev.gs_config(dev.MODEL_HEISENBERG, 16, 24, [1.0, 1.0, 1.0, 0.0], convtol=1e-12, seed=3)
    e0 = dev.gs_solve(ctx, cfg).energy
    ctx.set_option(dev.OPT_GAUGE_QR, 1)
    try:
        e1 = dev.gs_solve(ctx, cfg).energy
    finally:
        ctx.set_option(dev.OPT_GAUGE_QR, 0)
    assert abs(e0 - e1) <= 1e-10 * abs(e1)


# (d, Dl, Dr, Dn): site [d, Dl, Dr] and its neighbour's outer bond Dn.  The single-workgroup
# form in both Q layouts (m + n <= 64 and above), d = 3, a growing bond (k < n), the edge, and
# panel sizes where the fused entries fall back to the two calls.
ABSORB = [(2, 20, 20, 20), (2, 8, 16, 16), (2, 32, 32, 32), (3, 12, 20, 7), (2, 1, 2, 4), (2, 4, 16, 3),
          (2, 64, 64, 64), (2, 100, 40, 9), (2, 1, 1, 2)]


@pytest.mark.parametrize("d,Dl,Dr,Dn", ABSORB)
def test_gauge_absorb_matches_two_calls(ctx, d, Dl, Dr, Dn):
    """mpsd_site_gauge_r_absorb / gauge_l_absorb against mpsd_site_gauge_r + reduce_dim_r and
    mpsd_site_gauge_l + reduce_dim_l: the site and the factor bit for bit (the same QR kernel),
    the neighbour to 1e-13 relative (a different summation order)."""
    print(f"gauge+absorb d={d} Dl={Dl} Dr={Dr} Dn={Dn}", flush=True)
    A = orc.complex_gaussian(300 + Dl + Dr, (d, Dl, Dr))
    Bn = orc.complex_gaussian(301 + Dn, (d, Dr, Dn))   # right neighbour of a gauge_r step
    Bp = orc.complex_gaussian(302 + Dn, (d, Dn, Dl))   # left neighbour of a gauge_l step
    for side, B in (("r", Bn), ("l", Bp)):
        outs = []
        for fused in (False, True):
            tA, tF, tB = ctx.tensor(A), ctx.tensor(dims=(1, 1)), ctx.tensor(B)
            if fused:
                f = ctx.lib.mpsd_site_gauge_r_absorb if side == "r" else ctx.lib.mpsd_site_gauge_l_absorb
                ctx.check(f(ctx.handle, tA.handle, tF.handle, tB.handle))
            else:
                g = ctx.lib.mpsd_site_gauge_r if side == "r" else ctx.lib.mpsd_site_gauge_l
                rdim = ctx.lib.mpsd_site_reduce_dim_r if side == "r" else ctx.lib.mpsd_site_reduce_dim_l
                ctx.check(g(ctx.handle, tA.handle, tF.handle))
                ctx.check(rdim(ctx.handle, tB.handle, tF.handle))
            outs.append([t.download() for t in (tA, tF, tB)])
        (a0, f0, b0), (a1, f1, b1) = outs
        assert a0.shape == a1.shape and f0.shape == f1.shape and b0.shape == b1.shape
        assert np.array_equal(a0, a1) and np.array_equal(f0, f1)
        assert _rel(b1, b0) <= 1e-13, side
        # the neighbour is the factor applied to the old neighbour
        k = f0.shape[0] if side == "r" else f0.shape[1]
        if side == "r":
            want = np.einsum("ij,sjc->sic", f0, B)
        else:
            want = np.einsum("sac,cj->saj", B, f0)
        assert b0.shape[1 if side == "r" else 2] == k
        assert _rel(b1, want) <= 1e-13, side


@pytest.mark.parametrize("d,Dl,Dr", [(2, 16, 12), (2, 20, 20), (2, 128, 96), (2, 256, 200)])
@pytest.mark.parametrize("scale", [1e200, 1e-200])
def test_huge_and_tiny_columns(ctx, d, Dl, Dr, scale):
    """Columns whose squared norms overflow (entries 1e200) or underflow (1e-200) are rescaled
    before the reflector (as zlarfg / dznrm2 do): the factorisation of s A is the one of A, with R
    scaled by s, in the single-workgroup and the panel form."""
    print(f"scaled gauge_r d={d} Dl={Dl} Dr={Dr} scale={scale}", flush=True)
    A = orc.complex_gaussian(11 + Dl, (d, Dl, Dr))
    Q1, R1 = _gauge_r(ctx, A, 0)
    Qs, Rs = _gauge_r(ctx, scale * A, 0)
    assert np.all(np.isfinite(Qs)) and np.all(np.isfinite(Rs))
    assert _rel(Qs, Q1) <= 1e-13 * np.sqrt(d * Dl)
    assert _rel(Rs / scale, R1) <= 1e-13 * np.sqrt(d * Dl)
