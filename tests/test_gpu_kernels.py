"""GPU parity of the device matvec, environment updates and dense Heff against the CPU
oracle, through the C-ABI.  Tolerance (FP64 drop-in): relative 2-norm error <= 1e-12."""
import numpy as np
import pytest

from oracle import mpsd_oracle as orc

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module")
def ctx():
    from mpsdyn_amd import device as dev
    c = dev.Context(0)
    yield c
    c.close()


def _rel(a, b):
    return np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300)


def _operands(Db, Dbr, wl, wr, d=2, seed=0, Dk=None, Dkr=None):
    Dk = Db if Dk is None else Dk
    Dkr = Dbr if Dkr is None else Dkr
    L = orc.complex_gaussian(seed + 1, (Db, wl, Dk))
    W = orc.complex_gaussian(seed + 2, (d, wl, d, wr))
    R = orc.complex_gaussian(seed + 3, (Dbr, wr, Dkr))
    x = orc.complex_gaussian(seed + 4, (d * Dk * Dkr,))
    return L, W, R, x


SHAPES = [
    # (Db, Dbr, wl, wr): edges, the L=20 D=20 shape, ragged tiles, real DMRG shapes
    (1, 2, 1, 3), (2, 1, 3, 1), (1, 1, 1, 1), (20, 20, 3, 3), (2, 4, 3, 3), (37, 50, 5, 5),
    (64, 64, 3, 3), (128, 128, 3, 3), (100, 128, 5, 5), (16, 33, 5, 3),
]


@pytest.mark.parametrize("Db,Dbr,wl,wr", SHAPES)
def test_heff_matches_oracle(ctx, Db, Dbr, wl, wr):
    from mpsdyn_amd import device as dev
    print(f"heff Db={Db} Dbr={Dbr} wl={wl} wr={wr}", flush=True)
    L, W, R, x = _operands(Db, Dbr, wl, wr, seed=Db + 7 * Dbr)
    tL, tW, tR = ctx.tensor(L), ctx.tensor(W), ctx.tensor(R)
    h = dev.Heff(ctx, tL, tW, tR)
    y = h.apply(x)
    assert _rel(y, orc.heff_product(L, W, R, x)) <= TOL
    h.close()


def test_heff_offset_and_block(ctx):
    from mpsdyn_amd import device as dev
    L, W, R, _ = _operands(24, 24, 3, 3, seed=5)
    X = np.stack([orc.complex_gaussian(40 + j, (2 * 24 * 24,)) for j in range(3)], axis=1)
    h = dev.Heff(ctx, ctx.tensor(L), ctx.tensor(W), ctx.tensor(R), offset=1.75)
    Y = h.apply(X)
    for j in range(3):
        assert _rel(Y[:, j], orc.heff_product(L, W, R, X[:, j], offset=1.75)) <= TOL
    h.close()


def test_heff_known_answer_2x2(ctx):
    """The reference's 2x2 complex product (testmwArray_new.cpp:176-246) as an Heff with
    d = 1, W = 1, R = 1: y = L x, exact in FP64."""
    from mpsdyn_amd import device as dev
    A1 = orc.from_colmajor(np.array([1, 3, 2, 4]) * (1 + 1j), (2, 2))
    A2 = 2 * A1
    L = A1.reshape(2, 1, 2)
    W = np.ones((1, 1, 1, 1), dtype=complex)
    R = np.ones((1, 1, 1), dtype=complex)
    h = dev.Heff(ctx, ctx.tensor(L), ctx.tensor(W), ctx.tensor(R))
    Y = h.apply(A2)
    assert np.array_equal(Y, orc.from_colmajor(np.array([28j, 60j, 40j, 88j]), (2, 2)))
    h.close()


@pytest.mark.parametrize("Dl,Dr,wl,wr", [(1, 2, 1, 3), (20, 20, 3, 3), (37, 50, 5, 5), (2, 1, 3, 1),
                                         (128, 96, 3, 3), (64, 64, 5, 5)])
def test_env_updates_match_oracle(ctx, Dl, Dr, wl, wr):
    from mpsdyn_amd import device as dev
    print(f"env Dl={Dl} Dr={Dr} wl={wl} wr={wr}", flush=True)
    d = 2
    L = orc.complex_gaussian(1, (Dl, wl, Dl))
    R = orc.complex_gaussian(2, (Dr, wr, Dr))
    W = orc.complex_gaussian(3, (d, wl, d, wr))
    ket = orc.complex_gaussian(4, (d, Dl, Dr))
    bra = orc.complex_gaussian(5, (d, Dl, Dr))
    tk, tb, tW = ctx.tensor(ket), ctx.tensor(bra), ctx.tensor(W)
    out = ctx.tensor(dims=(1,))
    dev.env_left(ctx, ctx.tensor(L), tk, tW, tb, out)
    assert out.dims == (Dr, wr, Dr)
    assert _rel(out.download(), orc.contract_left(L, ket, W, bra)) <= TOL
    dev.env_right(ctx, ctx.tensor(R), tk, tW, tb, out)
    assert out.dims == (Dl, wl, Dl)
    assert _rel(out.download(), orc.contract_right(R, ket, W, bra)) <= TOL


@pytest.mark.parametrize("D,w", [(256, 3), (512, 3), (256, 5), (512, 5), (400, 5)])
def test_env_updates_full_size(ctx, D, w):
    """Environment updates at the sweep shapes the bench times (D=512 w=3, Heisenberg D=512
    w=5, Schwinger D=400 w=5) against Operator::contractleftop / contractrightop."""
    from mpsdyn_amd import device as dev
    print(f"env full-size D={D} w={w}", flush=True)
    d = 2
    L = orc.complex_gaussian(21, (D, w, D))
    R = orc.complex_gaussian(22, (D, w, D))
    W = orc.complex_gaussian(23, (d, w, d, w))
    ket = orc.complex_gaussian(24, (d, D, D))
    bra = orc.complex_gaussian(25, (d, D, D))
    tk, tb, tW = ctx.tensor(ket), ctx.tensor(bra), ctx.tensor(W)
    out = ctx.tensor(dims=(1,))
    dev.env_left(ctx, ctx.tensor(L), tk, tW, tb, out)
    assert out.dims == (D, w, D)
    assert _rel(out.download(), orc.contract_left(L, ket, W, bra)) <= TOL
    dev.env_right(ctx, ctx.tensor(R), tk, tW, tb, out)
    assert out.dims == (D, w, D)
    assert _rel(out.download(), orc.contract_right(R, ket, W, bra)) <= TOL


@pytest.mark.parametrize("on_device", [0, 1])
def test_heff_apply_leading_dimensions(ctx, on_device):
    """A block of 3 vectors with ldx = n + 5, ldy = ny + 7, as PRIMME's matrixMatvec hands them
    over (eigs.cpp:296-302): each column equals the oracle's product and the padding between
    columns of y is left untouched."""
    from mpsdyn_amd import device as dev
    L, W, R, _ = _operands(24, 30, 3, 5, seed=13, Dk=20, Dkr=26)
    h = dev.Heff(ctx, ctx.tensor(L), ctx.tensor(W), ctx.tensor(R))
    n, ny = 2 * 20 * 26, 2 * 24 * 30
    ldx, ldy, nvec = n + 5, ny + 7, 3
    X = np.full((ldx, nvec), 9.0 - 4.0j, order="F")
    for j in range(nvec):
        X[:n, j] = orc.complex_gaussian(60 + j, (n,))
    sentinel = -123.25 + 7.5j
    Y = np.full((ldy, nvec), sentinel, dtype=np.complex128, order="F")
    if on_device:
        tx = ctx.tensor(np.reshape(X, -1, order="F"))
        ty = ctx.tensor(np.reshape(Y, -1, order="F"))
        h.apply_device(tx.devptr(), ty.devptr(), nvec=nvec, ldx=ldx, ldy=ldy)
        Y = np.reshape(ty.download(), (ldy, nvec), order="F")
    else:
        xb = np.asfortranarray(X)
        ctx.check(ctx.lib.mpsd_heff_apply(h._h, xb.ctypes.data, Y.ctypes.data, nvec, ldx, ldy, 0))
    for j in range(nvec):
        assert _rel(Y[:ny, j], orc.heff_product(L, W, R, X[:n, j])) <= TOL
        assert np.all(Y[ny:, j] == sentinel)
    h.close()


def test_heff_full_matches_oracle(ctx):
    from mpsdyn_amd import device as dev
    L, W, R, _ = _operands(10, 20, 3, 3, seed=9)
    h = dev.Heff(ctx, ctx.tensor(L), ctx.tensor(W), ctx.tensor(R), offset=-0.5)
    H = h.full()
    assert _rel(H, orc.heff_full(L, W, R, offset=-0.5)) <= TOL
    h.close()


def test_shape_errors_are_reported(ctx):
    from mpsdyn_amd import device as dev
    L, W, R, _ = _operands(8, 8, 3, 3)
    bad_R = orc.complex_gaussian(1, (8, 5, 8))
    with pytest.raises(dev.MpsdError) as e:
        dev.Heff(ctx, ctx.tensor(L), ctx.tensor(W), ctx.tensor(bad_R))
    assert e.value.code == -2
    ket = orc.complex_gaussian(2, (2, 7, 8))
    with pytest.raises(dev.MpsdError):
        dev.env_left(ctx, ctx.tensor(L), ctx.tensor(ket), ctx.tensor(W), ctx.tensor(ket), ctx.tensor(dims=(1,)))


@pytest.mark.parametrize("D,w", [(256, 3), (512, 3), (512, 5), (256, 5), (400, 5)])
def test_heff_full_size(ctx, D, w):
    """Real DMRG shapes (D=256/512 Ising w=3, Heisenberg w=5, Schwinger D=256/400 w=5) against
    the oracle,
    plus Hermiticity <u, H v> = <H u, v> for a Hermitian-structured operator."""
    from mpsdyn_amd import device as dev
    print(f"full-size D={D} w={w}", flush=True)
    L, W, R, x = _operands(D, D, w, w, seed=D + w)
    h = dev.Heff(ctx, ctx.tensor(L), ctx.tensor(W), ctx.tensor(R))
    assert _rel(h.apply(x), orc.heff_product(L, W, R, x)) <= TOL
    h.close()
    # Hermitian operands: L, R Hermitian in (bra, ket) per MPO index, W Hermitian in (s', s)
    Lh = 0.5 * (L + np.conj(np.transpose(L, (2, 1, 0))))
    Rh = 0.5 * (R + np.conj(np.transpose(R, (2, 1, 0))))
    Wh = 0.5 * (W + np.conj(np.transpose(W, (2, 1, 0, 3))))
    h = dev.Heff(ctx, ctx.tensor(Lh), ctx.tensor(Wh), ctx.tensor(Rh))
    u = orc.complex_gaussian(77, (x.size,))
    Hu, Hx = h.apply(u), h.apply(x)
    assert abs(np.vdot(u, Hx) - np.vdot(Hu, x)) <= 1e-11 * np.linalg.norm(Hu) * np.linalg.norm(x)
    h.close()


@pytest.mark.parametrize("d,Dl,Dr,wl,wr", [(2, 1, 2, 1, 3), (2, 20, 20, 3, 3), (2, 2, 1, 3, 1), (3, 9, 14, 5, 5),
                                           (2, 17, 23, 4, 2), (2, 24, 24, 3, 3), (2, 8, 8, 9, 9)])
def test_env_chain_bit_identical(ctx, d, Dl, Dr, wl, wr):
    """Small environment updates run as one single-workgroup launch (MPSD_OPT_ENV_CHAIN): the
    same bits as the three separate launches, and the oracle's result to 1e-12."""
    from mpsdyn_amd import device as dev
    print(f"env chain d={d} Dl={Dl} Dr={Dr} wl={wl} wr={wr}", flush=True)
    L = orc.complex_gaussian(11, (Dl, wl, Dl))
    R = orc.complex_gaussian(12, (Dr, wr, Dr))
    W = orc.complex_gaussian(13, (d, wl, d, wr))
    ket = orc.complex_gaussian(14, (d, Dl, Dr))
    bra = orc.complex_gaussian(15, (d, Dl, Dr))
    tL, tR, tk, tb, tW = ctx.tensor(L), ctx.tensor(R), ctx.tensor(ket), ctx.tensor(bra), ctx.tensor(W)
    got = {}
    try:
        for chain in (1, 0):
            ctx.set_option(dev.OPT_ENV_CHAIN, chain)
            out = ctx.tensor(dims=(1,))
            dev.env_left(ctx, tL, tk, tW, tb, out)
            left = out.download()
            dev.env_right(ctx, tR, tk, tW, tb, out)
            got[chain] = (left, out.download())
    finally:
        ctx.set_option(dev.OPT_ENV_CHAIN, 1)
    assert np.array_equal(got[1][0], got[0][0]) and np.array_equal(got[1][1], got[0][1])
    assert _rel(got[1][0], orc.contract_left(L, ket, W, bra)) <= TOL
    assert _rel(got[1][1], orc.contract_right(R, ket, W, bra)) <= TOL


@pytest.mark.parametrize("Db,Dbr,wl,wr", [(20, 20, 3, 3), (37, 50, 5, 5), (1, 4, 1, 3), (4, 1, 3, 1), (3, 2, 3, 2),
                                          (100, 70, 3, 5), (128, 128, 5, 5), (512, 512, 3, 3), (400, 400, 5, 5)])
def test_packed_fp64_matvec_same_bits(ctx, Db, Dbr, wl, wr):
    """The FP64 matvec on packed operands (L, R packed once per handle; x packed per call or read
    in place by the first stage) equals the matvec on operands read in place bit for bit, for
    every BRGEMM tile, with the offset term; both match the oracle to 1e-12."""
    from mpsdyn_amd import device as dev
    print(f"Db={Db} Dbr={Dbr} wl={wl} wr={wr}", flush=True)
    L = orc.complex_gaussian(11, (Db, wl, Db))
    W = orc.complex_gaussian(12, (2, wl, 2, wr))
    R = orc.complex_gaussian(13, (Dbr, wr, Dbr))
    x = orc.complex_gaussian(14, (2 * Db * Dbr,))
    h = dev.Heff(ctx, ctx.tensor(L), ctx.tensor(W), ctx.tensor(R), offset=0.45)
    outs = []
    try:
        ctx.set_option(dev.OPT_PACKED_FP64, 0)
        plain = h.apply(x)
        ctx.set_option(dev.OPT_PACKED_FP64, 1)
        for xin in (0, 1):  # x read in place by the first stage, or packed first
            ctx.set_option(dev.OPT_PACKED_FP64_X, xin)
            for tile in (0, 1, 2, 3, 4):
                ctx.set_option(dev.OPT_PACKED_FP64_TILE, tile)
                outs.append(h.apply(x))
    finally:
        ctx.set_option(dev.OPT_PACKED_FP64, 1)
        ctx.set_option(dev.OPT_PACKED_FP64_TILE, 0)
        ctx.set_option(dev.OPT_PACKED_FP64_X, -1)
    for k, y in enumerate(outs):
        assert np.array_equal(y, plain), f"tile option {k} differs from the in-place matvec"
    ref = orc.heff_product(L, W, R, x, offset=0.45)
    assert np.linalg.norm(plain - ref) <= 1e-12 * np.linalg.norm(ref)
    h.close()
