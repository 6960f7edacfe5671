"""CPU oracle for the DMRG effective-Hamiltonian hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``
may import this module, and only as the checker (or as the timed CPU baseline).  The
product path (``mpsdyn_amd``) never calls it and fails loudly when its HIP library is
missing.

It restates, in numpy, the algorithm of the reference library (mcbapo/mpsdyn) for the
path that the device library implements.  Tensors are numpy arrays indexed exactly like
the reference's ``mwArray`` (first index first); every reshape is column-major
(``order='F'``), as ``mwArray`` stores its data (wrapper/mwArray.cpp:452-466), and
``permute`` takes 1-based orders like ``Indices`` (wrapper/mwArray.cpp:646-724).

Conventions (reference layout, column-major):
  site tensor   A[s, a_l, a_r]            (core/headers/Site.h:49)
  MPO tensor    W[s_up, w_l, s_down, w_r] (core/headers/Operator.h:37-43)
  environments  L[b, alpha, a], R[b, beta, a]  (bra, MPO, ket)

Pinning (how each function is checked; tests/test_oracle.py):
  heff_product      -- against heff_full applied to unit vectors, against the explicit
                       projection P^H H P of a dense Hamiltonian built from Pauli
                       Kronecker products (independent of the MPO path), Hermiticity,
                       linearity, and 7-fold explicit loops on tiny shapes.
  contract_left/right -- env of the full chain against <psi|H|psi> from the dense
                       Hamiltonian; explicit loops on tiny shapes.
  heff_full         -- against heff_product on unit vectors and the dense projection.
  heff_product_hermitian -- against (F + F^H) / 2 applied to x, F = heff_full (the dense
                       adjoint, independent of the conjugate-permuted copies), on
                       non-Hermitian operands where F != F^H.
  *_mpo builders    -- MPO contracted to a dense matrix against the Pauli-sum Hamiltonian.
  chain_mpo, spin_ops -- dense contraction against Kronecker-product sums with mixed site
                       dimensions; spin_ops against the spin-1/2 Pauli matrices and the
                       commutation relations / S^2 = S(S+1) for S = 1, 3/2.
  overlap_left/right, effective_bra, proj_multi_* -- overlaps against dense state vectors; the
                       effective vector against P^H |level> and the projector multiplier
                       against P^H (H + sum p|k><k|) P, P the brute-force site isometry.
  sweep_part, penalty_sweep_part, find_*_excited_state -- ED excited energies; alternating
                       partial sweeps reach the ED ground energy; penalised energy decreases.
  mps_write/mps_read -- a hand-assembled exportMPS file (known answer).
  svd_keep, mps_truncate -- the truncated state equals the dense Schmidt truncation across the
                       one bond that exceeds D (left-canonical start), and is exact when no
                       bond exceeds D.
  mpo_write/mpo_read -- a hand-written text file and a hand-assembled binary file laid out
                       as the reference's save / savetext code writes them (known answers),
                       and against the device library's own writer.
  find_ground_state -- reference energies recorded in tests/golden (the TFIM L=20, D=20 value from
                       the compiled reference), exact diagonalisation and the open-chain
                       free-fermion closed form.
  permute/reshape   -- the known answers of the reference's own unit test
                       (src/tests/testmwArray_new.cpp:92-170, 176-246).
  random_mps (seeded)-- parity unpinned: the reference draws libc random(); energies do not
                       depend on the start state, which is what the tests compare.
"""
from __future__ import annotations

import math
import numpy as np

__all__ = [
    "reshape_f", "permute", "heff_product", "heff_full", "contract_left", "contract_right",
    "ising_mpo", "heisenberg_mpo", "schwinger_mpo", "random_mps", "gauge_right_all",
    "gauge_left_all", "find_ground_state", "complex_gaussian", "splitmix64",
    "matvec_flops", "to_colmajor", "from_colmajor", "mpo_write", "mpo_read", "spin_ops",
    "chain_mpo", "mps_write", "mps_read", "overlap_left", "overlap_right", "effective_bra",
    "proj_multi_product", "proj_multi_full", "sweep_part", "penalty_sweep_part",
    "find_ground_state_with_projector_penalty", "find_next_excited_state", "overlap",
    "svd_keep", "mps_truncate", "resize", "increase_bond_dimension", "adjust_phys_dimensions",
    "fit_start", "gauge_cond_site", "mpo_product", "residual_norm",
]


# --------------------------------------------------------------------------- layout helpers

def reshape_f(a, dims):
    """Column-major reshape, as mwArray::reshape (wrapper/mwArray.cpp:556-583)."""
    return np.reshape(a, dims, order="F")


def permute(a, order):
    """mwArray::permute with a 1-based order (wrapper/mwArray.cpp:646-724):
    new index i is old index order[i]-1."""
    return np.transpose(a, [o - 1 for o in order])


def to_colmajor(a):
    """Flat complex128 buffer in mwArray storage order (interleaved re/im, column-major)."""
    return np.ascontiguousarray(np.reshape(np.asarray(a, dtype=np.complex128), -1, order="F"))


def from_colmajor(buf, dims):
    return reshape_f(np.asarray(buf, dtype=np.complex128), tuple(dims))


# --------------------------------------------------------------------------- RNG (project's)

_MASK = (1 << 64) - 1


def splitmix64(seed, n):
    """n successive SplitMix64 outputs (uint64) for a seed.  Pure integer arithmetic, so
    the C++ generator in the device library reproduces it bit for bit."""
    out = np.empty(n, dtype=np.uint64)
    state = np.uint64(seed & _MASK)
    with np.errstate(over="ignore"):
        idx = np.arange(1, n + 1, dtype=np.uint64)
        z = state + idx * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        out[:] = z ^ (z >> np.uint64(31))
    return out


def complex_gaussian(seed, shape):
    """Complex tensor whose re and im parts are N(0,1): SplitMix64 uniforms (53-bit) fed
    to Box--Muller.  Values are laid out column-major over ``shape`` (the i-th draw is the
    i-th element of the mwArray buffer)."""
    n = int(np.prod(shape))
    u = splitmix64(seed, 2 * n)
    u1 = ((u[0::2] >> np.uint64(11)).astype(np.float64) + 1.0) * (1.0 / 9007199254740992.0)
    u2 = (u[1::2] >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    r = np.sqrt(-2.0 * np.log(u1))
    re = r * np.cos(2.0 * np.pi * u2)
    im = r * np.sin(2.0 * np.pi * u2)
    return reshape_f(re + 1j * im, shape)


# --------------------------------------------------------------------------- the hot path

def heff_product(L, W, R, x, offset=0.0):
    """y = Heff x (+ offset x).  Restates TensorMultiplier::TensorMultiplier
    (src/wrapper/multipliers/TensorMultiplier.cpp:8-21) and ::product (:61-93), and the
    offset of TensorMultiplierOffset::product (TensorMultiplierOffset.cpp:14-17).
    ``x`` is the flat column-major vector of the site tensor [dd, betak, betakr]; the
    result is flat over [du, betab, betabr]."""
    betab, alpha, betak = L.shape
    du, al1, dd, al2 = W.shape
    betabr, alphar, betakr = R.shape
    AL = reshape_f(L, (betab * alpha, betak))
    Aop = reshape_f(permute(W, (1, 4, 2, 3)), (du * al2, al1 * dd))
    AR = permute(reshape_f(R, (betabr, alphar * betakr)), (2, 1))
    x = np.asarray(x, dtype=np.complex128).reshape(-1, order="F")
    aux = reshape_f(x, (dd, betak, betakr))
    aux = reshape_f(permute(aux, (2, 1, 3)), (betak, dd * betakr))
    aux = AL @ aux                                    # stage 1: L . x
    aux = reshape_f(aux, (betab, alpha * dd, betakr))
    aux = reshape_f(permute(aux, (2, 1, 3)), (alpha * dd, betab * betakr))
    aux = Aop @ aux                                   # stage 2: W
    aux = reshape_f(aux, (du, al2, betab, betakr))
    aux = reshape_f(permute(aux, (1, 3, 2, 4)), (du * betab, al2 * betakr))
    aux = aux @ AR                                    # stage 3: . R
    y = reshape_f(aux, (-1,))
    if offset != 0.0:
        y = y + offset * x
    return y


def heff_full(L, W, R, offset=0.0):
    """Dense Heff with rows (du, betab, betabr) and columns (dd, betak, betakr).
    Restates TensorMultiplier::getFullTensor (TensorMultiplier.cpp:102-117) and the offset
    of TensorMultiplierOffset::getFullTensor (TensorMultiplierOffset.cpp:21-25)."""
    betab, alpha, betak = L.shape
    du, al1, dd, al2 = W.shape
    betabr, alphar, betakr = R.shape
    AL = reshape_f(L, (betab * alpha, betak))
    Aop = reshape_f(permute(W, (1, 4, 2, 3)), (du * al2, al1 * dd))
    AR = permute(reshape_f(R, (betabr, alphar * betakr)), (2, 1))
    res = reshape_f(permute(reshape_f(AL, (betab, alpha, betak)), (1, 3, 2)), (betab * betak, alpha))
    tmp = reshape_f(permute(reshape_f(Aop, (du, al2, al1, dd)), (3, 1, 4, 2)), (al1, du * dd * al2))
    res = reshape_f(res @ tmp, (betab * betak * du * dd, al2))
    tmp = reshape_f(AR, (alphar, betakr * betabr))
    res = reshape_f(res @ tmp, (betab, betak, du, dd, betakr, betabr))
    res = reshape_f(permute(res, (3, 1, 6, 4, 2, 5)), (du * betab * betabr, dd * betak * betakr))
    if offset != 0.0:
        res = res + offset * np.eye(res.shape[0])
    return res


def heff_product_hermitian(L, W, R, x, offset=0.0):
    """(Heff + Heff^H) x / 2 + offset x.  Restates TensorMultiplierHermitian: the constructor's
    conjugate-permuted copies ALh = conj(permute(L,(3,2,1))), Aoph = conj(permute(W,(3,2,1,4))),
    ARh = conj(permute(R,(3,2,1))) (TensorMultiplierHermitian.cpp:21-26), ::product's second
    contraction over them and result = .5*(result+aux) (:36-66), and the offset of
    TensorMultiplierHermitianOffset::product (TensorMultiplierHermitianOffset.cpp:13-16)."""
    Lh = np.conj(permute(L, (3, 2, 1)))
    Wh = np.conj(permute(W, (3, 2, 1, 4)))
    Rh = np.conj(permute(R, (3, 2, 1)))
    y = 0.5 * (heff_product(L, W, R, x) + heff_product(Lh, Wh, Rh, x))
    return y + offset * np.asarray(x).reshape(y.shape, order="F") if offset != 0.0 else y


def contract_left(L, ket, W, bra):
    """New left environment L'[b_r, w_r, a_r].  Restates Operator::contractleftop
    (src/core/cpp/Operator.cpp:273-322), the order after 'better order of contractions'."""
    d, Dl, Dr = bra.shape
    dp, Dlp, Drp = ket.shape
    betab, alpha, betak = L.shape
    du, al1, dd, al2 = W.shape
    res = reshape_f(permute(ket, (2, 1, 3)), (Dlp, dp * Drp))
    res = reshape_f(L, (betab * alpha, betak)) @ res
    res = reshape_f(res, (betab, alpha * dp, Drp))
    res = reshape_f(permute(res, (2, 1, 3)), (alpha * dp, betab * Drp))
    wm = reshape_f(permute(reshape_f(W, (du, al1 * dd, al2)), (1, 3, 2)), (du * al2, al1 * dd))
    res = wm @ res
    res = reshape_f(res, (du, al2, betab, Drp))
    res = reshape_f(permute(res, (1, 3, 2, 4)), (du * betab, al2 * Drp))
    res = reshape_f(bra, (d * Dl, Dr)).conj().T @ res
    return reshape_f(res, (Dr, al2, Drp))


def contract_right(R, ket, W, bra):
    """New right environment R'[b_l, w_l, a_l].  Restates Operator::contractrightop
    (src/core/cpp/Operator.cpp:336-387)."""
    d, Dl, Dr = bra.shape
    dp, Dlp, Drp = ket.shape
    betab, alpha, betak = R.shape
    d1, al1, d2, al2 = W.shape
    res = reshape_f(permute(ket, (3, 1, 2)), (Drp, dp * Dlp))
    res = reshape_f(R, (betab * alpha, betak)) @ res
    res = reshape_f(res, (betab, alpha, dp, Dlp))
    res = reshape_f(permute(res, (3, 2, 1, 4)), (dp * alpha, betab * Dlp))
    res = reshape_f(W, (d1 * al1, d2 * al2)) @ res
    res = reshape_f(res, (d1, al1, betab, Dlp))
    res = reshape_f(permute(res, (1, 3, 2, 4)), (d1 * betab, al1 * Dlp))
    res = reshape_f(permute(bra.conj(), (2, 1, 3)), (Dl, d * Dr)) @ res
    return reshape_f(res, (Dl, al1, Dlp))


def matvec_flops(Dl, Dr, wl, wr, d=2):
    """Algorithmic real flops of one matvec (or env update), complex MAC = 8 flops."""
    return 8 * (d * wl * Dl * Dl * Dr + d * d * wl * wr * Dl * Dr + d * wr * Dl * Dr * Dr)


# --------------------------------------------------------------------------- model MPOs

_I2 = np.eye(2, dtype=np.complex128)
_SX = np.array([[0, 1], [1, 0]], dtype=np.complex128)
_SY = np.array([[0, -1j], [1j, 0]], dtype=np.complex128)
_SZ = np.array([[1, 0], [0, -1]], dtype=np.complex128)


def _w_from_c(C, ops):
    """W[du, Dl, dd, Dr] = sum_o C[Dl, Dr, o] op_o[du, dd]: the reshape C*Z then
    permute(3,1,4,2) of IsingHamiltonian.cpp:110-115 (same step in the other models)."""
    return np.einsum("lro,oud->uldr", C, np.stack(ops))


def ising_mpo(N, J, g, h=0.0, K=0.0):
    """H = sum J sz sz + g sx + h sz (+K).  Restates IsingHamiltonian::setHMPO
    (src/models/cpp/IsingHamiltonian.cpp:70-121) with Z = {1, sz, sx} (:51-68); a negative
    J enters as i*sqrt|J| on both sides (:84)."""
    ops = [_I2, _SZ, _SX]
    effJ = 1j * math.sqrt(abs(J)) if J < 0 else math.sqrt(J)
    mpo = []
    for k in range(N):
        Dl = 1 if k == 0 else 3
        Dr = 1 if k == N - 1 else 3
        C = np.zeros((Dl, Dr, 3), dtype=np.complex128)
        if k != N - 1:
            C[0, 0, 0] = 1
        if k != 0:
            C[Dl - 1, Dr - 1, 0] = 1
        C[0, Dr - 1, 2] = g
        C[0, Dr - 1, 1] = h
        C[0, Dr - 1, 0] = K / N
        if k != N - 1:
            C[0, 1, 1] = effJ
        if k != 0:
            C[1, Dr - 1, 1] = effJ
        mpo.append(_w_from_c(C, ops))
    return mpo


def heisenberg_mpo(L, Jx, Jy, Jz, h=0.0, offset=0.0):
    """H = sum Jx SxSx + Jy SySy + Jz SzSz + h Sz with S = sigma/2 (+offset).  Restates
    HeisenbergHamiltonian::initZ/initHMPO (src/models/cpp/HeisenbergHamiltonian.cpp:266-331)
    without the penalty term; couplings may be scalars or per-bond sequences."""
    ops = [_I2, 0.5 * _SX, 0.5 * _SY, 0.5 * _SZ]
    jx = np.broadcast_to(np.asarray(Jx, dtype=float), (max(L - 1, 1),))
    jy = np.broadcast_to(np.asarray(Jy, dtype=float), (max(L - 1, 1),))
    jz = np.broadcast_to(np.asarray(Jz, dtype=float), (max(L - 1, 1),))
    hh = np.broadcast_to(np.asarray(h, dtype=float), (L,))
    mpo = []
    for k in range(L):
        Dl = 1 if k == 0 else 5
        Dr = 1 if k == L - 1 else 5
        C = np.zeros((Dl, Dr, 4), dtype=np.complex128)
        if k != L - 1:
            C[0, 0, 0] = 1
            C[0, 1, 1] = jx[k]
            C[0, 2, 2] = jy[k]
            C[0, 3, 3] = jz[k]
        if k != 0:
            C[Dl - 1, Dr - 1, 0] = 1
            C[1, Dr - 1, 1] = 1
            C[2, Dr - 1, 2] = 1
            C[3, Dr - 1, 3] = 1
        C[0, Dr - 1, 3] = hh[k]
        if offset != 0:
            C[0, Dr - 1, 0] = offset / L
        mpo.append(_w_from_c(C, ops))
    return mpo


def schwinger_mpo(N, mu, x, l0=0.0, alpha=0.0, offset=0.0, nu=0.0, y=1.0):
    """Schwinger model (staggered fermions, gauge field integrated out).  Restates
    SchwingerHamiltonian::initZ/initHMPO (src/models/cpp/SchwingerHamiltonian.cpp:23-121);
    mu = 2 (m/g) sqrt(x) as in src/programs/banulsm/schwinger/schwingerPD.cpp:51."""
    ops = [_I2, _SX, _SY, _SZ]
    a_ = alpha + l0
    alphas = np.array([a_ - 0.25 * (1.0 - (1.0 if (k + 1) % 2 == 0 else -1.0)) for k in range(N)])
    sumA = np.zeros(N)
    tot = 0.0
    for k in range(N - 2, -1, -1):
        tot += alphas[k]
        sumA[k] = tot
    mpo = []
    for k in range(N):
        Dl = 1 if k == 0 else 5
        Dr = 1 if k == N - 1 else 5
        C = np.zeros((Dl, Dr, 4), dtype=np.complex128)
        if k != N - 1:
            C[0, 0, 0] = 1
        if k != 0:
            C[4, Dr - 1, 0] = 1
        if 0 < k < N - 2:
            C[3, 3, 0] = 1
        ak = alphas[k]
        sg = 1.0 if (k + 1) % 2 == 0 else -1.0
        if k != N - 1:
            C[0, Dr - 1, 0] = y * ak * ak + y * (k + 1) / 4.0 + mu / 2.0 + nu / 2.0 + offset / N
            C[0, Dr - 1, 3] = sg * mu / 2.0 + nu / 2.0 + y * sumA[k]
        else:
            C[0, Dr - 1, 0] = mu / 2.0 + nu / 2.0 + offset / N
            C[0, Dr - 1, 3] = sg * mu / 2.0 + nu / 2.0
        if k != N - 1:
            C[0, 1, 1] = math.sqrt(x / 2)
            C[0, 2, 2] = math.sqrt(x / 2)
            if k < N - 2:
                C[0, 3, 3] = y
            if k != 0:
                C[1, 4, 1] = math.sqrt(x / 2)
                C[2, 4, 2] = math.sqrt(x / 2)
                C[3, 4, 3] = (N - k - 1) / 2.0
        else:
            C[1, 0, 1] = math.sqrt(x / 2)
            C[2, 0, 2] = math.sqrt(x / 2)
        mpo.append(_w_from_c(C, ops))
    return mpo


def spin_ops(S2):
    """Spin-S matrices (Sx, Sy, Sz) for 2S = S2, basis m = S, S-1, ..., -S (d = S2 + 1):
    <m+1|S+|m> = sqrt(S(S+1) - m(m+1))."""
    S = S2 / 2.0
    m = S - np.arange(S2 + 1)
    sp_ = np.zeros((S2 + 1, S2 + 1), dtype=np.complex128)
    for i in range(1, S2 + 1):
        sp_[i - 1, i] = math.sqrt(S * (S + 1) - m[i] * (m[i] + 1))
    sx = 0.5 * (sp_ + sp_.conj().T)
    sy = -0.5j * (sp_ - sp_.conj().T)
    return sx, sy, np.diag(m).astype(np.complex128)


def chain_mpo(bond_ops, field_ops):
    """Nearest-neighbour chain H = sum_k sum_t A_t(k) B_t(k+1) + sum_k F(k) as an MPO in the
    reference's [du, Dl, dd, Dr] layout with the lower-triangular finite-automaton form the
    reference's models use (IsingHamiltonian.cpp:70-121): bond index 0 = 'nothing yet',
    1..T = 'A_t placed', T+1 = 'done'.  bond_ops[k] = list of (A_t, B_t) pairs for bond
    (k, k+1), field_ops[k] = F on site k; site k's dimension is F's.  Sites may have
    different dimensions."""
    L = len(field_ops)
    T = len(bond_ops[0]) if L > 1 else 0
    w = T + 2
    mpo = []
    for k in range(L):
        F = np.asarray(field_ops[k], dtype=np.complex128)
        d = F.shape[0]
        Dl = 1 if k == 0 else w
        Dr = 1 if k == L - 1 else w
        W = np.zeros((d, Dl, d, Dr), dtype=np.complex128)
        last_r = Dr - 1
        eye = np.eye(d, dtype=np.complex128)
        if k != L - 1:
            W[:, 0, :, 0] = eye
            for t, (A, _) in enumerate(bond_ops[k]):
                W[:, 0, :, 1 + t] = A
        if k != 0:
            W[:, Dl - 1, :, last_r] = eye
            for t, (_, B) in enumerate(bond_ops[k - 1]):
                W[:, 1 + t, :, last_r] = B
        W[:, 0, :, last_r] += F
        mpo.append(W)
    return mpo


# --------------------------------------------------------------------------- MPO files

def mpo_write(path, mpo, fmt, digits=None):
    """Write an MPO as MPO::saveMPO (fmt 'binary', src/core/cpp/MPO.cpp:214-232) or
    MPO::exportMPOtext (fmt 'text', MPO.cpp:234-248).  Binary: int32 length, then per operator
    mwArray::save (wrapper/mwArray.cpp:430-443): int32 rank, int32 dims[rank], then the
    column-major data as interleaved (re, im) float64.  Text: "length " then per operator
    mwArray::savetext (mwArray.cpp:1672-1688): "rank dims... re im ..." each followed by one
    space; digits = significant digits (None: 6, the default stream precision the reference
    writes with)."""
    if fmt == "binary":
        with open(path, "wb") as f:
            f.write(np.int32(len(mpo)).tobytes())
            for W in mpo:
                W = np.asarray(W, dtype=np.complex128)
                f.write(np.int32(W.ndim).tobytes())
                f.write(np.asarray(W.shape, dtype=np.int32).tobytes())
                f.write(to_colmajor(W).view(np.float64).tobytes())
    elif fmt == "text":
        p = 6 if digits is None else digits
        parts = [f"{len(mpo)} "]
        for W in mpo:
            W = np.asarray(W, dtype=np.complex128)
            parts.append(f"{W.ndim} " + "".join(f"{n} " for n in W.shape))
            parts.append("".join(f"{v:.{p}g} " for v in to_colmajor(W).view(np.float64)))
        with open(path, "w") as f:
            f.write("".join(parts))
    else:
        raise ValueError(fmt)


def mpo_read(path, fmt):
    """Read a file written as by mpo_write (MPO::loadMPO, MPO.cpp:293-316, and
    MPO::importMPOtext, MPO.cpp:250-271)."""
    mpo = []
    if fmt == "binary":
        raw = open(path, "rb").read()
        off = 0

        def ints(n):
            nonlocal off
            v = np.frombuffer(raw, dtype=np.int32, count=n, offset=off)
            off += 4 * n
            return [int(x) for x in v]

        (L,) = ints(1)
        for _ in range(L):
            (rank,) = ints(1)
            dims = ints(rank)
            n = int(np.prod(dims))
            v = np.frombuffer(raw, dtype=np.float64, count=2 * n, offset=off).view(np.complex128)
            off += 16 * n
            mpo.append(from_colmajor(v.copy(), tuple(dims)))
        if off != len(raw):
            raise ValueError("trailing bytes")
    elif fmt == "text":
        tok = open(path).read().split()
        pos = 0
        L = int(tok[pos]); pos += 1
        for _ in range(L):
            rank = int(tok[pos]); pos += 1
            dims = [int(t) for t in tok[pos:pos + rank]]; pos += rank
            n = int(np.prod(dims))
            v = np.array([float(t) for t in tok[pos:pos + 2 * n]], dtype=np.float64).view(np.complex128)
            pos += 2 * n
            mpo.append(from_colmajor(v, tuple(dims)))
    else:
        raise ValueError(fmt)
    return mpo


# --------------------------------------------------------------------------- MPS and gauge

def bond_dims(L, D, d=2):
    """Bond profile after MPS(L,D,d) and gaugeCond('R'),('L') (MPS.cpp:15-35, 306-364):
    min(d^k, D, d^(L-k)) for k = 0..L."""
    return [min(d ** min(k, 62), D, d ** min(L - k, 62)) for k in range(L + 1)]


def random_mps(L, D, d=2, seed=1):
    """Seeded complex-Gaussian MPS with the canonical bond profile; site k draws from seed
    seed*1000003 + k.  The reference fills uniform [0,1) values from libc random()
    (wrapper/mwArray.cpp:1595-1602): parity unpinned, only energies are compared."""
    dims = bond_dims(L, D, d)
    return [complex_gaussian(seed * 1000003 + k, (d, dims[k], dims[k + 1])) for k in range(L)]


def _qr_site(A):
    """Site::gaugeR (Site.cpp:179-208): QR of A as (d*Dl, Dr); returns (Q-site, R)."""
    d, Dl, Dr = A.shape
    q, r = np.linalg.qr(reshape_f(A, (d * Dl, Dr)))
    return reshape_f(q, (d, Dl, q.shape[1])), r


def _lq_site(A):
    """Site::gaugeL (Site.cpp:210-238): LQ of A permuted to (Dl, d*Dr)."""
    d, Dl, Dr = A.shape
    m = reshape_f(permute(A, (2, 1, 3)), (Dl, d * Dr))
    q, r = np.linalg.qr(m.conj().T)             # m = r^H q^H = L Q
    Lm, Q = r.conj().T, q.conj().T
    return permute(reshape_f(Q, (Q.shape[0], d, Dr)), (2, 1, 3)), Lm


def _absorb_right(A, R):
    """Site::reduceDimR (Site.cpp:240-261): A'[s, :, a_r] = R . A[s, :, a_r]."""
    return np.einsum("ij,sjk->sik", R, A)


def _absorb_left(A, Lm):
    """Site::reduceDimL (Site.cpp:263-283): A'[s, a_l, :] = A[s, a_l, :] . L."""
    return np.einsum("sij,jk->sik", A, Lm)


def gauge_right_all(mps):
    """MPS::gaugeCond('R', normal) (MPS.cpp:306-337): left-canonical sites 0..L-1, norm
    dropped at the end."""
    mps = [a.copy() for a in mps]
    for k in range(len(mps)):
        q, r = _qr_site(mps[k])
        mps[k] = q
        if k + 1 < len(mps):
            mps[k + 1] = _absorb_right(mps[k + 1], r)
        else:
            mps[k] = mps[k] * (r[0, 0] / abs(r[0, 0]))
    return mps


def gauge_left_all(mps):
    """MPS::gaugeCond('L', normal) (MPS.cpp:310-321): right-canonical sites L-1..0."""
    mps = [a.copy() for a in mps]
    for k in range(len(mps) - 1, -1, -1):
        q, lm = _lq_site(mps[k])
        mps[k] = q
        if k > 0:
            mps[k - 1] = _absorb_left(mps[k - 1], lm)
        else:
            mps[k] = mps[k] * (lm[0, 0] / abs(lm[0, 0]))
    return mps


# --------------------------------------------------------------------------- the sweep

def _local_ground(L, W, R, x0, offset, dense_below=801):
    n = x0.size
    if n < dense_below:
        # Contractor::solveEigs small-size switch (Contractor.cpp:3126-3133)
        H = heff_full(L, W, R, offset)
        H = 0.5 * (H + H.conj().T)
        vals, vecs = np.linalg.eigh(H)
        return float(vals[0]), vecs[:, 0]
    from scipy.sparse.linalg import LinearOperator, eigsh
    op = LinearOperator((n, n), matvec=lambda v: heff_product(L, W, R, v, offset), dtype=np.complex128)
    vals, vecs = eigsh(op, k=1, which="SA", v0=x0, tol=0.0)
    return float(vals[0]), vecs[:, 0]


def environments(mps, mpo):
    """All right environments of a right-canonical MPS plus the trivial left edge:
    TmpManager edges are the scalar 1 (TmpManager.cpp:44-54), the rest follow
    Contractor::calculateR (Contractor.cpp:744-764)."""
    L = len(mps)
    one = np.ones((1, 1, 1), dtype=np.complex128)
    envL = [None] * L
    envR = [None] * L
    envL[0] = one
    envR[L - 1] = one
    for k in range(L - 2, -1, -1):
        envR[k] = contract_right(envR[k + 1], mps[k + 1], mpo[k + 1], mps[k + 1])
    return envL, envR


def expectation(mps, mpo):
    """<psi|H|psi> via a full left-environment sweep (Contractor::contract, Contractor.cpp:528)."""
    e = np.ones((1, 1, 1), dtype=np.complex128)
    for k in range(len(mps)):
        e = contract_left(e, mps[k], mpo[k], mps[k])
    return complex(e.reshape(-1)[0])


def mpo_product(a, b):
    """The MPO of the operator product A B: P[u, a1 + Dl_A a2, v, b1 + Dr_A b2] =
    sum_s A[u, a1, s, b1] B[s, a2, v, b2] per site (column-major merged bonds)."""
    out = []
    for A, B in zip(a, b):
        du, al, _dm, ar = A.shape
        _, bl, dd, br = B.shape
        P = np.einsum("uasb,scvd->uacvbd", A, B)        # [u, a1, a2, v, b1, b2]
        out.append(P.reshape((du, al * bl, dd, ar * br), order="F"))
    return out


def residual_norm(mps, shifted_mpo):
    """||(H - E) psi|| for a unit-norm psi, given the MPO of H - E: sqrt(<psi|(H-E)^2|psi>)."""
    v = expectation(mps, mpo_product(shifted_mpo, shifted_mpo)).real
    return math.sqrt(max(v, 0.0))


def find_ground_state(mpo, mps, convtol=1e-12, offset=0.0, max_halfsweeps=200, dense_below=801, D=0):
    """Single-site ground-state sweep.  Restates Contractor::findGroundState
    (src/core/cpp/Contractor.cpp:1352-1533): fit the start state to the MPO and to D
    (fit_start), right-canonicalise, sweep 0..L-1, L-2..0, ... solving the lowest eigenpair of
    Heff at each site (warm start from the current site), gauge-moving the centre with QR/LQ
    (MPS::gaugeCond(pos,dir,true), MPS.cpp:366-419) and stopping when
    |E_prev - E| < convtol |E + offset| at the end of a half-sweep.
    Returns (energy, mps, n_halfsweeps, energies_per_halfsweep)."""
    Ls = len(mps)
    mps, _ = fit_start(mpo, mps, D)
    mps = gauge_left_all(gauge_right_all(mps))
    envL, envR = environments(mps, mpo)
    energy = expectation(mps, mpo).real
    val = energy
    pos, right, done, rounds = 0, True, False, 0
    history = []
    while not done:
        if pos > 0 and envL[pos] is None:
            for k in range(1, pos + 1):
                if envL[k] is None:
                    envL[k] = contract_left(envL[k - 1], mps[k - 1], mpo[k - 1], mps[k - 1])
        if pos < Ls - 1 and envR[pos] is None:
            for k in range(Ls - 2, pos - 1, -1):
                if envR[k] is None:
                    envR[k] = contract_right(envR[k + 1], mps[k + 1], mpo[k + 1], mps[k + 1])
        A = mps[pos]
        val, vec = _local_ground(envL[pos], mpo[pos], envR[pos], to_colmajor(A), offset, dense_below)
        val -= offset
        mps[pos] = from_colmajor(vec, A.shape)
        if right:
            q, r = _qr_site(mps[pos])
            mps[pos] = q
            if pos + 1 < Ls:
                mps[pos + 1] = _absorb_right(mps[pos + 1], r)
            else:
                mps[pos] = q * (r[0, 0] / abs(r[0, 0]))
        else:
            q, lm = _lq_site(mps[pos])
            mps[pos] = q
            if pos > 0:
                mps[pos - 1] = _absorb_left(mps[pos - 1], lm)
            else:
                mps[pos] = q * (lm[0, 0] / abs(lm[0, 0]))
        for l in range(pos + 1, Ls):
            envL[l] = None
        for l in range(pos - 1, -1, -1):
            envR[l] = None
        if right:
            if pos == Ls - 1:
                right = False
                done = abs(energy - val) < convtol * abs(energy + offset) or rounds + 1 >= max_halfsweeps
                energy = val
                history.append(val)
                if not done:
                    q, lm = _lq_site(mps[pos])
                    mps[pos] = q
                    mps[pos - 1] = _absorb_left(mps[pos - 1], lm)
                    pos -= 1
                rounds += 1
            else:
                pos += 1
        else:
            if pos == 0:
                right = True
                done = abs(energy - val) < convtol * abs(energy + offset) or rounds + 1 >= max_halfsweeps
                energy = val
                history.append(val)
                if not done:
                    q, r = _qr_site(mps[pos])
                    mps[pos] = q
                    mps[pos + 1] = _absorb_right(mps[pos + 1], r)
                    pos += 1
                rounds += 1
            else:
                pos -= 1
    return energy, mps, rounds, history


# --------------------------------------------------------------------------- MPS files

def mps_write(path, mps, bond=None, normfact=1.0, gauge=0):
    """MPS::exportMPS (src/core/cpp/MPS.cpp:440-459): int32 length, int32 bond, float64
    normfact, int32 gauge, then per site Site::save (Site.cpp:570-578): int32 position
    (1-based) and mwArray::save of A[d, Dl, Dr]."""
    bond = max(max(a.shape[1], a.shape[2]) for a in mps) if bond is None else bond
    with open(path, "wb") as f:
        f.write(np.int32(len(mps)).tobytes() + np.int32(bond).tobytes() + np.float64(normfact).tobytes()
                + np.int32(gauge).tobytes())
        for k, A in enumerate(mps):
            A = np.asarray(A, dtype=np.complex128)
            f.write(np.int32(k + 1).tobytes() + np.int32(3).tobytes() + np.asarray(A.shape, dtype=np.int32).tobytes())
            f.write(to_colmajor(A).view(np.float64).tobytes())


def mps_read(path):
    """MPS::importMPS (MPS.cpp:487-509).  Returns (sites, normfact)."""
    raw = open(path, "rb").read()
    L, _bond = np.frombuffer(raw, dtype=np.int32, count=2, offset=0)
    normfact = float(np.frombuffer(raw, dtype=np.float64, count=1, offset=8)[0])
    off = 20
    sites = []
    for _ in range(int(L)):
        _pos, rank = np.frombuffer(raw, dtype=np.int32, count=2, offset=off)
        dims = [int(x) for x in np.frombuffer(raw, dtype=np.int32, count=int(rank), offset=off + 8)]
        off += 8 + 4 * int(rank)
        n = int(np.prod(dims))
        v = np.frombuffer(raw, dtype=np.float64, count=2 * n, offset=off).view(np.complex128)
        off += 16 * n
        sites.append(from_colmajor(v.copy(), tuple(dims)))
    return sites, normfact


# --------------------------------------------------------------------------- partial sweeps

def _gauge_site(mps, pos, direction, nxt):
    """MPS::gaugeCond(pos, dir, next) (MPS.cpp:366-419) on a list of sites, in place: the
    site becomes an isometry; its factor goes to the neighbour when next is set or the bond
    shrank.  A 1x1 factor's phase stays in the site (Site.cpp:202-207, 232-237)."""
    L = len(mps)
    if direction == "R":
        old = mps[pos].shape[2]
        q, r = _qr_site(mps[pos])
        if r.shape == (1, 1) and r[0, 0] != 0:
            q, r = q * (r[0, 0] / abs(r[0, 0])), np.abs(r)
        mps[pos] = q
        if pos < L - 1 and (q.shape[2] != old or nxt):
            mps[pos + 1] = _absorb_right(mps[pos + 1], r)
    else:
        old = mps[pos].shape[1]
        q, lm = _lq_site(mps[pos])
        if lm.shape == (1, 1) and lm[0, 0] != 0:
            q, lm = q * (lm[0, 0] / abs(lm[0, 0])), np.abs(lm)
        mps[pos] = q
        if pos > 0 and (q.shape[1] != old or nxt):
            mps[pos - 1] = _absorb_left(mps[pos - 1], lm)


def gauge_cond_site(mps, pos, direction, nxt, normfact=1.0):
    """MPS::gaugeCond(pos, dir, next) (MPS.cpp:366-419) with its edge normalisation: after
    gaugeL at site 0 or gaugeR at the last site with next set, the 1x1 factor's magnitude
    multiplies the norm factor (MPS.cpp:380-387, 402-409).  Returns (sites, normfact)."""
    mps = [a.copy() for a in mps]
    L = len(mps)
    A = mps[pos]
    if direction == "R":
        q, f = _qr_site(A)
        edge = nxt and pos == L - 1
    else:
        q, f = _lq_site(A)
        edge = nxt and pos == 0
    if edge:
        if f.size != 1:
            raise ValueError("edge factor is not a scalar")
        normfact = normfact * abs(f[0, 0])
    _gauge_site(mps, pos, direction, nxt)
    return mps, normfact


# --------------------------------------------------------------------------- bond and physical dimensions

def resize(a, dims):
    """mwArray::resize with zero noise (wrapper/mwArray.cpp:1604-1670): the element at each
    multi-index inside both shapes is kept, new ones are zero, the rest are cut."""
    a = np.asarray(a, dtype=np.complex128)
    out = np.zeros(tuple(dims), dtype=np.complex128)
    box = tuple(slice(0, min(o, n)) for o, n in zip(a.shape, dims))
    out[box] = a[box]
    return out


def increase_bond_dimension(mps, D, bond):
    """MPS::increaseBondDimension(D) with zero noise (MPS.cpp:650-719), restated loop for
    loop: boundR is the first site from the right where the running right product reaches D;
    cumd / cumdr are the running left / right products of the physical dimensions; every site
    is resized (Site::increaseBondDimWithNoise -> mwArray::resize).  Returns (sites, bond)."""
    if D <= bond:
        return [a.copy() for a in mps], bond
    L = len(mps)
    cumd, cumdr_last, bound_r = 1, 1, L - 1
    while bound_r > 0:
        dim = mps[bound_r].shape[0]
        if cumdr_last * dim < D:
            bound_r -= 1
            cumdr_last *= dim
        else:
            break
    cumdr = D if bound_r > 0 else cumdr_last * mps[0].shape[0]
    out, maxbond = [], 1
    for k in range(L):
        d, oDl, oDr = mps[k].shape
        if k == 0:
            nDl = oDl
        else:
            nDl = min(cumd, cumdr) if (cumd < D or cumdr < D) else D
        if cumd < D:
            cumd *= d
        if k == bound_r:
            cumdr = cumdr_last
        elif k > bound_r:
            cumdr //= d
        if k == L - 1:
            nDr = oDr
        else:
            nDr = min(cumd, cumdr) if (cumd < D or cumdr < D) else D
        out.append(resize(mps[k], (d, nDl, nDr)))
        maxbond = max(maxbond, nDl, nDr)
    return out, maxbond


def adjust_phys_dimensions(mps, dvec, bond):
    """MPS::adjustPhysDimensions (MPS.cpp:722-756): site k resized to physical dimension
    dvec[k] (Site::increasePhysDim / decreasePhysDim, Site.cpp:545-559); when any site changed,
    the bond profile is rebuilt by increaseBondDimension(bond) from bond 1.  Returns
    (sites, bond)."""
    if len(dvec) != len(mps):
        raise ValueError("one physical dimension per site")
    out, changed = [], False
    for A, dk in zip(mps, dvec):
        if dk != A.shape[0]:
            A = resize(A, (dk, A.shape[1], A.shape[2]))
            changed = True
        out.append(A.copy())
    if changed:
        return increase_bond_dimension(out, bond, 1)
    return out, bond


def fit_start(mpo, mps, D, bond=None):
    """The start state of findGroundState (Contractor.cpp:1360-1370): physical dimensions from
    the MPO, then for D > 0 grown by zero padding or cut by mps_truncate (MPS(init, D)).
    Returns (sites, bond)."""
    bond = max(max(a.shape[1], a.shape[2]) for a in mps) if bond is None else bond
    mps, bond = adjust_phys_dimensions(mps, [W.shape[0] for W in mpo], bond)
    if D > 0:
        if D > bond:
            mps, bond = increase_bond_dimension(mps, D, bond)
        elif D < bond:
            mps, _ = mps_truncate(mps, D)
            bond = min(bond, max(max(a.shape[1], a.shape[2]) for a in mps))
    return mps, bond


def _envs_at(mps, mpo, pos):
    """Operator environments of site pos (Contractor::calculateL/R, Contractor.cpp:723-764)."""
    one = np.ones((1, 1, 1), dtype=np.complex128)
    EL = one
    for k in range(pos):
        EL = contract_left(EL, mps[k], mpo[k], mps[k])
    ER = one
    for k in range(len(mps) - 1, pos, -1):
        ER = contract_right(ER, mps[k], mpo[k], mps[k])
    return EL, ER


def sweep_part(mpo, mps, pos0, lim, offset=0.0):
    """Contractor::sweepPart (Contractor.cpp:1260-1350): sites before pos0 right-gauged, after
    it left-gauged, then solve sites pos0 .. lim (excluded), passing the gauge factor on only at
    the last step.  Local problems by dense diagonalisation.  Returns (mps, last value)."""
    mps = [a.copy() for a in mps]
    L = len(mps)
    right = lim >= pos0
    for k in range(pos0):
        _gauge_site(mps, k, "R", True)
    for k in range(L - 1, pos0, -1):
        _gauge_site(mps, k, "L", True)
    pos, val = pos0, None
    while True:
        EL, ER = _envs_at(mps, mpo, pos)
        val, vec = _local_ground(EL, mpo[pos], ER, to_colmajor(mps[pos]), offset, dense_below=10 ** 9)
        val -= offset
        mps[pos] = from_colmajor(vec, mps[pos].shape)
        pass_on = (right and pos == lim - 1) or (not right and pos == lim + 1) or pos == lim
        _gauge_site(mps, pos, "R" if right else "L", pass_on)
        pos += 1 if right else -1
        if (pos >= lim) if right else (pos <= lim):
            return mps, val


# --------------------------------------------------------------------------- projector penalties

def overlap_left(N, ket, bra):
    """Contractor::contractL (Contractor.cpp:819-838): N'[r_bra, r_ket] =
    sum conj(bra[s, l, r]) N[l, l'] ket[s, l', r']."""
    d, Dl, Dr = bra.shape
    dp, Dlp, Drp = ket.shape
    res = N @ reshape_f(permute(ket, (2, 1, 3)), (Dlp, dp * Drp))
    res = reshape_f(res, (Dl * d, Drp))
    aux = reshape_f(permute(bra, (3, 2, 1)).conj(), (Dr, Dl * d))
    return aux @ res


def overlap_right(N, ket, bra):
    """Contractor::contractR (Contractor.cpp:840-858): N'[l_bra, l_ket] =
    sum conj(bra[s, l, r]) N[r, r'] ket[s, l', r']."""
    d, Dl, Dr = bra.shape
    dp, Dlp, Drp = ket.shape
    res = reshape_f(permute(bra, (2, 1, 3)).conj(), (Dl * d, Dr)) @ N
    res = reshape_f(res, (Dl, d * Drp))
    return res @ reshape_f(permute(ket, (1, 3, 2)), (dp * Drp, Dlp))


def effective_bra(NL, bra, NR):
    """Contractor::contractMbraId (Contractor.cpp:880-903): the row vector
    v[s, a, c] = sum conj(bra[s, l, r]) NL[l, a] NR[r, c], flat column-major."""
    dp, Dlp, Drp = bra.shape
    res = reshape_f(bra.conj(), (dp * Dlp, Drp)) @ NR
    res = reshape_f(permute(reshape_f(res, (dp, Dlp, NR.shape[1])), (1, 3, 2)), (dp * NR.shape[1], Dlp))
    res = res @ NL
    res = permute(reshape_f(res, (dp, NR.shape[1], NL.shape[1])), (1, 3, 2))
    return reshape_f(res, (-1,))


def proj_multi_product(L, W, R, x, weights, vbras, hermitian=True, offset=0.0):
    """TensorMultiplierProjMulti::product (TensorMultiplierProjMulti.cpp:87-96): the basic
    (Hermitian-averaged when hermitian) product, minus weight_k (v_k . x) v_k^H, plus offset x."""
    x = np.asarray(x, dtype=np.complex128)
    y = heff_product_hermitian(L, W, R, x) if hermitian else heff_product(L, W, R, x)
    for w, v in zip(weights, vbras):
        y = y - w * (v @ x) * v.conj()
    return y + offset * x


def proj_multi_full(L, W, R, weights, vbras, hermitian=True, offset=0.0):
    """TensorMultiplierProjMulti::getFullTensor (TensorMultiplierProjMulti.cpp:98-106)."""
    F = heff_full(L, W, R)
    if hermitian:
        F = 0.5 * (F + F.conj().T)
    for w, v in zip(weights, vbras):
        F = F - w * np.outer(v.conj(), v)
    F = F + offset * np.eye(F.shape[0])
    return 0.5 * (F + F.conj().T)


def _norm_envs_at(mps, level, pos):
    """normL[k][pos], normR[k][pos] of TmpManagerExc (Contractor.cpp:2958-2989): overlaps of
    the level (bra) with the state (ket) up to site pos."""
    NL = np.ones((1, 1), dtype=np.complex128)
    for k in range(pos):
        NL = overlap_left(NL, mps[k], level[k])
    NR = np.ones((1, 1), dtype=np.complex128)
    for k in range(len(mps) - 1, pos, -1):
        NR = overlap_right(NR, mps[k], level[k])
    return NL, NR


def overlap(ket, bra):
    """<bra|ket> (Contractor::contract(ket, bra), Contractor.cpp:467-499), unit norm factors."""
    N = np.ones((1, 1), dtype=np.complex128)
    for k in range(len(ket)):
        N = overlap_left(N, ket[k], bra[k])
    return complex(N[0, 0])


def _penalty_site(mpo, mps, levels, weights, pos, offset):
    EL, ER = _envs_at(mps, mpo, pos)
    vb = []
    for lev in levels:
        NL, NR = _norm_envs_at(mps, lev, pos)
        vb.append(effective_bra(NL, lev[pos], NR))
    F = proj_multi_full(EL, mpo[pos], ER, weights, vb, True, offset)
    vals, vecs = np.linalg.eigh(F)
    return float(vals[0]) - offset, vecs[:, 0]


def penalty_sweep_part(mpo, mps, weights, levels, pos0, lim, offset=0.0, pass_always=False):
    """Contractor::sweepPartWithProjectorPenalty (Contractor.cpp:2575-2666): sweepPart on
    H - sum_k weights[k] |k><k| (the weights enter the multiplier as given);
    pass_always gives sweepPartWithProjPenalty (Contractor.cpp:3640-3734)."""
    mps = [a.copy() for a in mps]
    L = len(mps)
    right = lim >= pos0
    for k in range(pos0):
        _gauge_site(mps, k, "R", True)
    for k in range(L - 1, pos0, -1):
        _gauge_site(mps, k, "L", True)
    pos, val = pos0, None
    while True:
        val, vec = _penalty_site(mpo, mps, levels, weights, pos, offset)
        mps[pos] = from_colmajor(vec, mps[pos].shape)
        pass_on = pass_always or (right and pos == lim - 1) or (not right and pos == lim + 1) or pos == lim
        _gauge_site(mps, pos, "R" if right else "L", pass_on)
        pos += 1 if right else -1
        if (pos >= lim) if right else (pos <= lim):
            return mps, val


def find_ground_state_with_projector_penalty(mpo, mps, penalties, levels, convtol=1e-12, offset=0.0,
                                             max_halfsweeps=400):
    """Contractor::findGroundStateWithProjectorPenalty (Contractor.cpp:2739-2956): ground state
    of H + sum_k penalties[k] |k><k| by full sweeps, each site's problem the Hermitian
    projector multiplier with weights -penalties.  Returns (energy, mps, halfsweeps)."""
    L = len(mps)
    mps = gauge_left_all(gauge_right_all(mps))
    energy = expectation(mps, mpo).real + offset
    for p, lev in zip(penalties, levels):
        energy += p * abs(overlap(mps, lev)) ** 2
    energy -= offset
    weights = [-p for p in penalties]
    pos, right, rounds = 0, True, 0
    while True:
        val, vec = _penalty_site(mpo, mps, levels, weights, pos, offset)
        mps[pos] = from_colmajor(vec, mps[pos].shape)
        _gauge_site(mps, pos, "R" if right else "L", True)
        end = (right and pos == L - 1) or (not right and pos == 0)
        if end:
            done = abs(energy - val) < convtol * abs(energy + offset) or abs(energy - val) < 1e-12
            done = done or rounds + 1 >= max_halfsweeps
            energy = val
            rounds += 1
            if done:
                return energy, mps, rounds
            _gauge_site(mps, pos, "L" if right else "R", True)
            pos += -1 if right else 1
            right = not right
        else:
            pos += 1 if right else -1


def find_next_excited_state(mpo, mps, levels, convtol=1e-12, offset=0.0):
    """Contractor::findNextExcitedState (Contractor.cpp:2676-2727) without the orthogonality
    retry: penalty max(20 (maxE - minE), |offset|), or L/2 for one (or degenerate) level."""
    es = [expectation(lev, mpo).real + offset for lev in levels]
    pen = max(20 * (max(es) - min(es)), abs(offset))
    if len(levels) == 1 or pen < 1e-6:
        pen = len(levels[0]) * 0.5
    return find_ground_state_with_projector_penalty(mpo, mps, [pen] * len(levels), levels, convtol, offset)


# --------------------------------------------------------------------------- SVD truncation

def svd_keep(M, cutD):
    """wrapper::svd with tol = 0 and Xi = cutD (src/wrapper/svd.cpp:23-118): thin SVD, keeping
    the leading singular values above zero among the first min(cutD, k).  Returns (U, S, Vh)."""
    U, S, Vh = np.linalg.svd(M, full_matrices=False)
    X = 0
    while X < min(cutD, S.size) and S[X] > 0:
        X += 1
    if X == 0:
        raise ValueError("no singular value above zero")
    return U[:, :X], S[:X], Vh[:X, :]


def mps_truncate(mps, D, normfact=1.0):
    """MPS(const MPS&, D) (MPS.cpp:761-773) = gaugeCond('L', false, D) (MPS.cpp:306-364) with
    Site::gaugeL's SVD branch (Site.cpp:210-238): from the right, M[a, s + d c] = U S V^H, the
    site becomes V^H, U S goes into the left neighbour; the last 1x1 factor's phase stays in
    site 0 and its magnitude multiplies the norm factor.  Returns (sites, normfact)."""
    mps = [a.copy() for a in mps]
    for pos in range(len(mps) - 1, -1, -1):
        d, Dl, Dr = mps[pos].shape
        M = reshape_f(permute(mps[pos], (2, 1, 3)), (Dl, d * Dr))
        U, S, Vh = svd_keep(M, D)
        X = S.size
        lt = U * S[None, :]
        site = permute(reshape_f(Vh, (X, d, Dr)), (2, 1, 3))
        if X == 1 and Dl == 1:
            site = site * (lt[0, 0] / abs(lt[0, 0]))
            lt = np.abs(lt)
        mps[pos] = site
        if pos > 0:
            mps[pos - 1] = _absorb_left(mps[pos - 1], lt)
        else:
            normfact = normfact * abs(lt[0, 0])
    return mps, normfact

