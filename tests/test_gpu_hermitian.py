"""Hermitian form of the matvec (TensorMultiplierHermitian(Offset)), the dense Heff of that
form and the dense eigensolver (fulleig / small-size switch), against the CPU oracle.
Tolerances: matvec and dense matrix 1e-12 relative 2-norm; eigenvalues 1e-12 relative;
ground-state energies 1e-10 relative."""
import json
import os

import numpy as np
import pytest

from oracle import mpsd_oracle as orc

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "reference_values.json")


@pytest.fixture(scope="module")
def ctx():
    from mpsdyn_amd import device as dev
    c = dev.Context(0)
    yield c
    c.close()


def _rel(a, b):
    return np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300)


def _ops(Dl, Dr, wl, wr, seed):
    L = orc.complex_gaussian(seed + 1, (Dl, wl, Dl))
    W = orc.complex_gaussian(seed + 2, (2, wl, 2, wr))
    R = orc.complex_gaussian(seed + 3, (Dr, wr, Dr))
    x = orc.complex_gaussian(seed + 4, (2 * Dl * Dr,))
    return L, W, R, x


@pytest.mark.parametrize("Dl,Dr,wl,wr,offset", [(1, 2, 1, 3, 0.0), (7, 9, 3, 3, 0.5), (40, 33, 5, 5, -2.0),
                                                (130, 128, 3, 3, 1.25)])
def test_hermitian_matvec(ctx, Dl, Dr, wl, wr, offset):
    from mpsdyn_amd import device as dev
    print(f"hermitian matvec Dl={Dl} Dr={Dr} wl={wl} wr={wr} offset={offset}", flush=True)
    L, W, R, x = _ops(Dl, Dr, wl, wr, seed=Dl + 3 * Dr)
    h = dev.Heff(ctx, ctx.tensor(L), ctx.tensor(W), ctx.tensor(R), offset=offset)
    plain = h.apply(x)
    h.set_hermitian(True)
    y = h.apply(x)
    want = orc.heff_product_hermitian(L, W, R, x, offset)
    assert _rel(y, want) <= 1e-12
    assert _rel(y, plain) > 1e-6  # the operands are not Hermitian: the average differs
    h.set_hermitian(False)
    assert np.array_equal(h.apply(x), plain)
    h.close()


def test_hermitian_full_and_grouped(ctx):
    from mpsdyn_amd import device as dev
    print("hermitian full + grouped", flush=True)
    L, W, R, x = _ops(6, 5, 3, 3, seed=50)
    h = dev.Heff(ctx, ctx.tensor(L), ctx.tensor(W), ctx.tensor(R), offset=0.3)
    h.set_hermitian(True)
    F = orc.heff_full(L, W, R)
    want = 0.5 * (F + F.conj().T) + 0.3 * np.eye(F.shape[0])
    assert _rel(h.full(), want) <= 1e-12
    # a Hermitian handle inside a grouped call gives the same bits as alone
    xt = ctx.tensor(x.reshape(2, 6, 5, order="F"))
    y1, y2 = ctx.tensor(dims=(2, 6, 5)), ctx.tensor(dims=(2, 6, 5))
    h.apply_device(xt.devptr(), y1.devptr())
    dev.heff_apply_grouped(ctx, [h], [xt], [y2])
    assert np.array_equal(y1.download(), y2.download())
    h.close()


def test_hermitian_needs_square(ctx):
    from mpsdyn_amd import device as dev
    L = orc.complex_gaussian(1, (4, 3, 5))  # bra and ket bonds differ
    W = orc.complex_gaussian(2, (2, 3, 2, 3))
    R = orc.complex_gaussian(3, (4, 3, 4))
    h = dev.Heff(ctx, ctx.tensor(L), ctx.tensor(W), ctx.tensor(R))
    with pytest.raises(dev.MpsdError) as e:
        h.set_hermitian(True)
    assert e.value.code == -2
    h.close()


@pytest.mark.parametrize("Dl,Dr,w,herm", [(5, 6, 3, False), (12, 12, 5, True), (20, 20, 3, False)])
def test_dense_eigensolver(ctx, Dl, Dr, w, herm):
    from mpsdyn_amd import device as dev
    print(f"dense eig Dl={Dl} Dr={Dr} w={w} herm={herm}", flush=True)
    L, W, R, x = _ops(Dl, Dr, w, w, seed=7 * Dl + Dr)
    h = dev.Heff(ctx, ctx.tensor(L), ctx.tensor(W), ctx.tensor(R), offset=0.25)
    if herm:
        h.set_hermitian(True)
    F = orc.heff_full(L, W, R)
    A = 0.5 * (F + F.conj().T) + 0.25 * np.eye(F.shape[0])
    evals = np.linalg.eigvalsh(A)
    xt = ctx.tensor(x.reshape(2, Dl, Dr, order="F"))
    val = h.eigs_dense(xt)
    assert abs(val - evals[0]) <= 1e-12 * max(1.0, abs(evals).max())
    v = xt.download().ravel(order="F")
    assert abs(np.linalg.norm(v) - 1.0) <= 1e-12
    assert np.linalg.norm(A @ v - val * v) <= 1e-10 * abs(evals).max()
    h.close()


def test_fulleig_ground_state_matches_reference_value(ctx):
    from mpsdyn_amd import device as dev
    print("fulleig TFIM L=8 D=8", flush=True)
    with open(GOLDEN) as f:
        ref = json.load(f)["tfim_L8_critical"]["reference_energy"]
    cfg = dev.gs_config(dev.MODEL_ISING, 8, 8, [-1.0, -1.0, 0.0], convtol=1e-12, seed=5,
                        solver=dev.SOLVER_FULLEIG)
    r = dev.gs_solve(ctx, cfg)
    assert abs(r.energy - ref) <= 1e-10 * abs(ref)
    lz = dev.gs_solve(ctx, dev.gs_config(dev.MODEL_ISING, 8, 8, [-1.0, -1.0, 0.0], convtol=1e-12, seed=5))
    assert abs(r.energy - lz.energy) <= 1e-10 * abs(ref)
