"""SVD-truncating gauge steps on the device (Site::gaugeR/gaugeL with cutD, rocSOLVER zgesvd) and
MPS(const MPS&, D) against numpy and the oracle.  Bars: singular values and the truncated
reconstruction to 1e-12 relative, isometries to 1e-13, truncated states equal to the oracle's
(dense vectors, 1e-11 relative) with the same norm factor."""
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


@pytest.mark.parametrize("d,Dl,Dr,cut", [(2, 8, 16, 5), (2, 20, 20, 40), (3, 10, 33, 12), (2, 64, 100, 64),
                                         (2, 1, 2, 1), (2, 128, 128, 100)])
def test_gauge_r_svd(ctx, d, Dl, Dr, cut):
    print(f"gauge_r_svd {d} {Dl} {Dr} cut={cut}", flush=True)
    A = orc.complex_gaussian(Dl + Dr + cut, (d, Dl, Dr))
    tA, tR = ctx.tensor(A), ctx.tensor(dims=(1, 1))
    ctx.check(ctx.lib.mpsd_site_gauge_r_svd(ctx.handle, tA.handle, tR.handle, cut))
    Q, R = tA.download(), tR.download()
    M = orc.reshape_f(A, (d * Dl, Dr))
    U, S, Vh = orc.svd_keep(M, cut)
    X = S.size
    assert Q.shape == (d, Dl, X) and R.shape == (X, Dr)
    Qm = orc.reshape_f(Q, (d * Dl, X))
    assert np.linalg.norm(Qm.conj().T @ Qm - np.eye(X)) <= 1e-13 * np.sqrt(X) * 10
    best = (U * S) @ Vh
    assert np.linalg.norm(Qm @ R - best) <= 1e-12 * np.linalg.norm(best)
    assert np.allclose(np.linalg.svd(R, compute_uv=False), S, rtol=1e-12)


@pytest.mark.parametrize("d,Dl,Dr,cut", [(2, 16, 8, 5), (2, 20, 20, 40), (3, 33, 10, 12), (2, 100, 64, 64),
                                         (2, 2, 1, 1)])
def test_gauge_l_svd(ctx, d, Dl, Dr, cut):
    print(f"gauge_l_svd {d} {Dl} {Dr} cut={cut}", flush=True)
    A = orc.complex_gaussian(Dl * Dr + cut, (d, Dl, Dr))
    tA, tL = ctx.tensor(A), ctx.tensor(dims=(1, 1))
    ctx.check(ctx.lib.mpsd_site_gauge_l_svd(ctx.handle, tA.handle, tL.handle, cut))
    Q, Lm = tA.download(), tL.download()
    M = orc.reshape_f(orc.permute(A, (2, 1, 3)), (Dl, d * Dr))
    U, S, Vh = orc.svd_keep(M, cut)
    X = S.size
    assert Q.shape == (d, X, Dr) and Lm.shape == (Dl, X)
    Qm = orc.reshape_f(orc.permute(Q, (2, 1, 3)), (X, d * Dr))
    assert np.linalg.norm(Qm @ Qm.conj().T - np.eye(X)) <= 1e-13 * np.sqrt(X) * 10
    best = (U * S) @ Vh
    assert np.linalg.norm(Lm @ Qm - best) <= 1e-12 * np.linalg.norm(best)


def test_zero_matrix_is_an_error(ctx):
    from mpsdyn_amd import device as dev
    tA, tR = ctx.tensor(np.zeros((2, 3, 4), dtype=np.complex128)), ctx.tensor(dims=(1, 1))
    with pytest.raises(dev.MpsdError):
        ctx.check(ctx.lib.mpsd_site_gauge_r_svd(ctx.handle, tA.handle, tR.handle, 2))


@pytest.mark.parametrize("L,Dstart,D", [(8, 16, 6), (8, 16, 16), (10, 32, 9), (6, 8, 1)])
def test_mps_truncate_matches_oracle(ctx, tmp_path, L, Dstart, D):
    from mpsdyn_amd import device as dev
    print(f"mps_truncate L={L} {Dstart}->{D}", flush=True)
    mps = orc.random_mps(L, Dstart, seed=L + D)
    pin, pout = str(tmp_path / "in.mps"), str(tmp_path / "out.mps")
    orc.mps_write(pin, mps, normfact=0.75)
    dev.mps_truncate(ctx, pin, D, pout)
    got, nf = orc.mps_read(pout)
    want, nf_want = orc.mps_truncate(mps, D, normfact=0.75)
    assert [a.shape for a in got] == [a.shape for a in want]
    assert abs(nf - nf_want) <= 1e-12 * nf_want
    g, w = ph.mps_to_state(got), ph.mps_to_state(want)
    assert np.linalg.norm(g - w) <= 1e-11 * np.linalg.norm(w)
