"""GPU checks of the eigensolver support (multi-dot, GEMV update, axpby), the gauge steps
(QR/LQ on the device, neighbour absorption) and the device Lanczos solver, through the
C-ABI, against numpy."""
import ctypes as C

import numpy as np
import pytest

from oracle import mpsd_oracle as orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from mpsdyn_amd import device as dev
    c = dev.Context(0)
    yield c
    c.close()


def test_vec_dots_gemv_axpby(ctx):
    n, m = 10007, 7
    V = orc.complex_gaussian(1, (n, m))
    w = orc.complex_gaussian(2, (n,))
    tV, tw = ctx.tensor(V), ctx.tensor(w)
    out = np.zeros(m, dtype=np.complex128)
    ctx.check(ctx.lib.mpsd_vec_dots(ctx.handle, tV.devptr(), n, m, tw.devptr(), n, out.ctypes.data))
    assert np.allclose(out, V.conj().T @ w, rtol=1e-13, atol=1e-10)
    c = orc.complex_gaussian(3, (m,))
    ctx.check(ctx.lib.mpsd_vec_gemv(ctx.handle, tV.devptr(), n, m, c.ctypes.data, tw.devptr(), n))
    assert np.allclose(tw.download(), w + V @ c, rtol=1e-13, atol=1e-12)
    a = np.array([0.5, -1.0])
    b = np.array([2.0, 0.25])
    x = orc.complex_gaussian(4, (n,))
    tx, ty = ctx.tensor(x), ctx.tensor(w)
    ctx.check(ctx.lib.mpsd_vec_axpby(ctx.handle, a.ctypes.data, tx.devptr(), b.ctypes.data, ty.devptr(), n))
    assert np.allclose(ty.download(), (0.5 - 1j) * x + (2 + 0.25j) * w, rtol=1e-14, atol=1e-13)
    nr = C.c_double()
    ctx.check(ctx.lib.mpsd_vec_norm2(ctx.handle, tx.devptr(), n, C.byref(nr)))
    assert np.isclose(nr.value, np.vdot(x, x).real, rtol=1e-13)


def test_dots_are_bit_reproducible(ctx):
    n = 300001
    V = orc.complex_gaussian(5, (n, 3))
    w = orc.complex_gaussian(6, (n,))
    tV, tw = ctx.tensor(V), ctx.tensor(w)
    outs = []
    for _ in range(3):
        out = np.zeros(3, dtype=np.complex128)
        ctx.check(ctx.lib.mpsd_vec_dots(ctx.handle, tV.devptr(), n, 3, tw.devptr(), n, out.ctypes.data))
        outs.append(out)
    assert all(np.array_equal(outs[0], o) for o in outs)


@pytest.mark.parametrize("n,j", [(1, 0), (300, 2), (2049, 5), (32768, 9), (8000, 40), (32769, 3), (70000, 12)])
def test_gram_schmidt_passes_bit_identical(ctx, n, j):
    """The two Gram-Schmidt passes of a Lanczos step in their folded form (the dots' last
    reduction inside the consumer, n <= 32768) leave the same bits as dots + gemv and
    dots + coef_update; w is column j+1 of V, as in the solver.  Above n = 32768 the entries
    take the unfolded launches."""
    print(f"gram-schmidt n={n} j={j}", flush=True)
    V = orc.complex_gaussian(7 + j, (n, j + 2))
    V[:, : j + 1] /= np.sqrt(n)
    runs = []
    for folded in (False, True):
        tV = ctx.tensor(V)
        h1, h2, ab = (ctx.tensor(np.zeros(k, dtype=np.complex128)) for k in (j + 1, j + 2, j + 2))
        base = tV.devptr()
        w = base + (j + 1) * n * 16
        if folded:
            ctx.check(ctx.lib.mpsd_lanczos_gs1_dev(ctx.handle, base, n, j + 1, w, n, h1.devptr()))
            ctx.check(ctx.lib.mpsd_lanczos_gs2_dev(ctx.handle, base, n, j, h1.devptr(), h2.devptr(), ab.devptr(),
                                                   w, n))
        else:
            ctx.check(ctx.lib.mpsd_vec_dots_dev(ctx.handle, base, n, j + 1, w, n, h1.devptr()))
            ctx.check(ctx.lib.mpsd_vec_gemv_dev(ctx.handle, base, n, j + 1, h1.devptr(), -1.0, w, n))
            ctx.check(ctx.lib.mpsd_vec_dots_dev(ctx.handle, base, n, j + 2, w, n, h2.devptr()))
            ctx.check(ctx.lib.mpsd_lanczos_coef_update(ctx.handle, base, n, h1.devptr(), h2.devptr(), j,
                                                       ab.devptr(), w, n))
        runs.append([t.download() for t in (tV, h1, h2, ab)])
    for a, b in zip(*runs):
        assert np.array_equal(np.ravel(a, order="F").view(np.float64), np.ravel(b, order="F").view(np.float64))
    h1d = runs[1][1]
    assert np.allclose(h1d, V[:, : j + 1].conj().T @ V[:, j + 1], rtol=1e-12, atol=1e-12 * np.sqrt(n))


def _cgs2_numpy(V, w):
    """Classical Gram-Schmidt applied twice, the Lanczos step's coefficients (numpy)."""
    h1 = V.conj().T @ w
    w1 = w - V @ h1
    h2 = V.conj().T @ w1
    nw1 = np.vdot(w1, w1).real
    beta = np.sqrt(nw1 - np.vdot(h2, h2).real)
    return h1, h2, nw1, (w1 - V @ h2) / beta, (h1[-1] + h2[-1]).real, beta


# n chosen for both ways of summing the second pass (n <= 32768: folded into the update
# kernel; above: a separate launch), j for each register width (8, 16, 24, 32 columns)
@pytest.mark.parametrize("n,j,pad", [(1, 0, 0), (300, 2, 0), (2049, 7, 0), (32768, 15, 0), (32769, 16, 3),
                                     (131072, 23, 0), (70000, 24, 0), (300001, 29, 5), (5000, 31, 0)])
def test_gram_schmidt_step_fused(ctx, n, j, pad):
    """mpsd_lanczos_gs_step_dev against numpy's CGS2 and against the two unfused passes: a
    random w (well conditioned), and a w that is the basis plus 1e-7 noise (as H v_j is late in
    a Lanczos run: the second pass carries the orthogonality)."""
    print(f"fused gram-schmidt n={n} j={j} ldv={n + pad}", flush=True)
    m1 = j + 1
    ldv = n + pad
    Q, _ = np.linalg.qr(orc.complex_gaussian(11 + j, (n, m1)))
    exhausted = n <= m1  # w lies in the span of V: beta is rounding noise (the solver stops there)
    for case in ("random",) if exhausted else ("random", "near_span"):
        w = orc.complex_gaussian(12 + j, (n,))
        if case == "near_span":
            w = Q @ orc.complex_gaussian(13 + j, (m1,)) + 1e-7 * w
        tight = case == "random"
        if exhausted:
            h1r = Q.conj().T @ w
        else:
            h1r, h2r, nw1, w2r, alpha, beta = _cgs2_numpy(Q, w)
        Vp = np.zeros((ldv, m1), dtype=np.complex128)
        Vp[:n] = Q
        tV, tw = ctx.tensor(Vp), ctx.tensor(w)
        h1, h2, ab = (ctx.tensor(np.zeros(k, dtype=np.complex128)) for k in (m1, m1 + 1, m1 + 1))
        ctx.check(ctx.lib.mpsd_lanczos_gs_step_dev(ctx.handle, tV.devptr(), ldv, j, tw.devptr(), n, h1.devptr(),
                                                   h2.devptr(), ab.devptr()))
        wo, g1, g2 = tw.download(), h1.download(), h2.download()
        abv = ab.download().view(np.float64)
        scale = np.linalg.norm(w)
        assert np.allclose(g1, h1r, rtol=0, atol=1e-12 * scale)
        if exhausted:
            assert abv[2 * j + 1] <= 1e-14 * scale
            continue
        assert np.allclose(g2[:m1], h2r, rtol=0, atol=1e-12 * scale)
        assert abs(g2[m1].real - nw1) <= (1e-12 if tight else 1e-8) * nw1
        assert abs(abv[2 * j] - alpha) <= 1e-12 * scale
        assert abs(abv[2 * j + 1] - beta) <= (1e-12 if tight else 1e-7) * beta
        assert np.abs(Q.conj().T @ wo).max() <= 1e-12
        assert abs(np.linalg.norm(wo) - 1.0) <= 1e-12
        assert np.allclose(wo, w2r, rtol=0, atol=1e-13 if tight else 1e-8)
        # the unfused passes (j + 1 basis vectors, w in column j + 1)
        Vu = np.zeros((ldv, m1 + 1), dtype=np.complex128)
        Vu[:n, :m1], Vu[:n, m1] = Q, w
        tU = ctx.tensor(Vu)
        u1, u2, uab = (ctx.tensor(np.zeros(k, dtype=np.complex128)) for k in (m1, m1 + 1, m1 + 1))
        base = tU.devptr()
        ctx.check(ctx.lib.mpsd_lanczos_gs1_dev(ctx.handle, base, ldv, m1, base + m1 * ldv * 16, n, u1.devptr()))
        ctx.check(ctx.lib.mpsd_lanczos_gs2_dev(ctx.handle, base, ldv, j, u1.devptr(), u2.devptr(), uab.devptr(),
                                               base + m1 * ldv * 16, n))
        wu = np.asfortranarray(tU.download())[:n, m1]
        assert np.allclose(wo, wu, rtol=0, atol=1e-13 if tight else 1e-8)
        ua = uab.download().view(np.float64)
        assert abs(abv[2 * j] - ua[2 * j]) <= 1e-12 * scale
        assert abs(abv[2 * j + 1] - ua[2 * j + 1]) <= (1e-12 if tight else 1e-7) * beta
    # the same call twice gives the same bits
    outs = []
    for _ in range(2):
        tV, tw = ctx.tensor(Vp), ctx.tensor(w)
        h1, h2, ab = (ctx.tensor(np.zeros(k, dtype=np.complex128)) for k in (m1, m1 + 1, m1 + 1))
        ctx.check(ctx.lib.mpsd_lanczos_gs_step_dev(ctx.handle, tV.devptr(), ldv, j, tw.devptr(), n, h1.devptr(),
                                                   h2.devptr(), ab.devptr()))
        outs.append(np.concatenate([tw.download(), h1.download(), h2.download(), ab.download()]))
    assert np.array_equal(outs[0].view(np.float64), outs[1].view(np.float64))


def test_gram_schmidt_step_fused_limit(ctx):
    """Above 32 basis vectors the fused step refuses (the solver then takes the two passes)."""
    n, j = 64, 32
    tV = ctx.tensor(np.zeros((n, j + 2), dtype=np.complex128))
    h = ctx.tensor(np.zeros(j + 2, dtype=np.complex128))
    rc = ctx.lib.mpsd_lanczos_gs_step_dev(ctx.handle, tV.devptr(), n, j, tV.devptr() + (j + 1) * n * 16, n,
                                          h.devptr(), h.devptr(), h.devptr())
    assert rc == -6 and "32" in ctx.lib.mpsd_last_error(ctx.handle).decode()


@pytest.mark.parametrize("d,Dl,Dr", [(2, 8, 16), (2, 16, 8), (2, 1, 2), (2, 64, 64), (2, 4, 1), (3, 5, 7)])
def test_gauge_r_and_reduce(ctx, d, Dl, Dr):
    A = orc.complex_gaussian(7 + Dl, (d, Dl, Dr))
    tA = ctx.tensor(A)
    tR = ctx.tensor(dims=(1, 1))
    ctx.check(ctx.lib.mpsd_site_gauge_r(ctx.handle, tA.handle, tR.handle))
    Q = tA.download()
    R = tR.download()
    k = min(d * Dl, Dr)
    assert Q.shape == (d, Dl, k) and R.shape == (k, Dr)
    Qm = orc.reshape_f(Q, (d * Dl, k))
    assert np.allclose(Qm.conj().T @ Qm, np.eye(k), atol=1e-13)
    assert np.allclose(Qm @ R, orc.reshape_f(A, (d * Dl, Dr)), atol=1e-12)
    assert np.allclose(np.tril(R, -1), 0)
    # absorb R into a neighbour: B'[s, :, c] = R . B[s, :, c]
    B = orc.complex_gaussian(9, (d, Dr, 5))
    tB = ctx.tensor(B)
    ctx.check(ctx.lib.mpsd_site_reduce_dim_r(ctx.handle, tB.handle, tR.handle))
    assert np.allclose(tB.download(), orc._absorb_right(B, R), atol=1e-12)


@pytest.mark.parametrize("d,Dl,Dr", [(2, 8, 16), (2, 16, 8), (2, 2, 1), (2, 64, 64), (2, 1, 4), (3, 7, 5)])
def test_gauge_l_and_reduce(ctx, d, Dl, Dr):
    A = orc.complex_gaussian(17 + Dr, (d, Dl, Dr))
    tA = ctx.tensor(A)
    tL = ctx.tensor(dims=(1, 1))
    ctx.check(ctx.lib.mpsd_site_gauge_l(ctx.handle, tA.handle, tL.handle))
    Q = tA.download()
    Lm = tL.download()
    k = min(Dl, d * Dr)
    assert Q.shape == (d, k, Dr) and Lm.shape == (Dl, k)
    Qm = orc.reshape_f(orc.permute(Q, (2, 1, 3)), (k, d * Dr))
    assert np.allclose(Qm @ Qm.conj().T, np.eye(k), atol=1e-13)
    assert np.allclose(Lm @ Qm, orc.reshape_f(orc.permute(A, (2, 1, 3)), (Dl, d * Dr)), atol=1e-12)
    B = orc.complex_gaussian(19, (d, 3, Dl))
    tB = ctx.tensor(B)
    ctx.check(ctx.lib.mpsd_site_reduce_dim_l(ctx.handle, tB.handle, tL.handle))
    assert np.allclose(tB.download(), orc._absorb_left(B, Lm), atol=1e-12)


def test_edge_phase_convention(ctx):
    """Site::gaugeR on a last site (Dr = 1) leaves a real positive factor (Site.cpp:202-207)."""
    A = orc.complex_gaussian(21, (2, 4, 1))
    tA, tR = ctx.tensor(A), ctx.tensor(dims=(1, 1))
    ctx.check(ctx.lib.mpsd_site_gauge_r(ctx.handle, tA.handle, tR.handle))
    r = tR.download().ravel()[0]
    assert abs(r.imag) < 1e-14 and r.real > 0
    assert np.isclose(r.real, np.linalg.norm(A))
    assert np.allclose(tA.download() * r, A, atol=1e-13)


@pytest.mark.parametrize("D,w", [(6, 3), (20, 3), (32, 5)])
def test_lanczos_lowest_matches_dense(ctx, D, w):
    """Device Lanczos on a Hermitian Heff against numpy's eigh of the dense oracle Heff."""
    from mpsdyn_amd import device as dev
    L_ = 8
    mpo = orc.heisenberg_mpo(L_, 1.0, 1.0, 0.7, 0.1) if w == 5 else orc.ising_mpo(L_, -1.0, -1.0, -0.5)
    mps = orc.gauge_left_all(orc.gauge_right_all(orc.random_mps(L_, D, seed=D)))
    k = 4
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
    H = orc.heff_full(envL, mpo[k], envR)
    evals = np.linalg.eigvalsh(0.5 * (H + H.conj().T))
    h = dev.Heff(ctx, ctx.tensor(envL), ctx.tensor(mpo[k]), ctx.tensor(envR))
    x = ctx.tensor(mps[k])
    val, res, nmv = h.eigs_lowest(x, tol=1e-11)
    assert abs(val - evals[0]) <= 1e-10 * max(1, abs(evals[0]))
    v = x.download().reshape(-1, order="F")
    assert np.isclose(np.linalg.norm(v), 1.0)
    assert np.linalg.norm(H @ v - val * v) <= 1e-9 * max(1, abs(val))
    h.close()


def test_lanczos_reports_nan(ctx):
    """A NaN in an environment is reported as MPSD_ERR_NUMERIC, not a crash or hang."""
    from mpsdyn_amd import device as dev
    L = orc.complex_gaussian(1, (4, 3, 4))
    L[1, 1, 1] = np.nan
    W = orc.ising_mpo(6, -1, -1, 0)[2]
    R = orc.complex_gaussian(2, (4, 3, 4))
    h = dev.Heff(ctx, ctx.tensor(L), ctx.tensor(W), ctx.tensor(R))
    x = ctx.tensor(orc.complex_gaussian(3, (2, 4, 4)))
    with pytest.raises(dev.MpsdError) as e:
        h.eigs_lowest(x)
    assert e.value.code == -5
    h.close()
