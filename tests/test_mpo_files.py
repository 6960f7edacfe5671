"""MPO files on the CPU: the oracle's reader/writer against hand-assembled files laid out as the
reference's MPO::saveMPO / exportMPOtext write them (src/core/cpp/MPO.cpp:214-248,
wrapper/mwArray.cpp:430-443, 1672-1688), the general chain MPO against a Kronecker sum, and the
device library's host-only writer (mpsd_mpo_save, no GPU needed) against the oracle's model
MPOs.  Integer/byte layout is compared exactly; values written by the library bit-exactly."""
import os
import struct

import numpy as np
import pytest

from oracle import mpsd_oracle as orc
from tests import _physics as ph


def _seq_mpo_text(shapes):
    """Text file whose numbers are 1, 2, 3, ... in file order, and the tensors that implies
    (first index fastest, re then im)."""
    parts, mpo, v = [f"{len(shapes)} "], [], 1
    for shp in shapes:
        n = int(np.prod(shp))
        parts.append(f"{len(shp)} " + "".join(f"{x} " for x in shp))
        vals = np.arange(v, v + 2 * n, dtype=np.float64)
        parts.append("".join(f"{int(x)} " for x in vals))
        v += 2 * n
        flat = vals[0::2] + 1j * vals[1::2]
        W = np.zeros(shp, dtype=np.complex128)
        for idx in np.ndindex(*shp):
            lin, stride = 0, 1
            for i, s in zip(idx, shp):
                lin += i * stride
                stride *= s
            W[idx] = flat[lin]
        mpo.append(W)
    return "".join(parts), mpo


def test_text_known_answer(tmp_path):
    text, want = _seq_mpo_text([(2, 1, 2, 3), (2, 3, 2, 1)])
    p = tmp_path / "m.txt"
    p.write_text(text)
    got = orc.mpo_read(str(p), "text")
    assert len(got) == 2
    for g, w in zip(got, want):
        assert g.shape == w.shape and np.array_equal(g, w)
    # element [1, 0, 1, 2] of the first operator: linear index 1 + 2*1 + 4*2 = 11 -> numbers 23, 24
    assert got[0][1, 0, 1, 2] == 23 + 24j
    # the writer reproduces the file (integers print the same at 6 digits)
    q = tmp_path / "w.txt"
    orc.mpo_write(str(q), want, "text")
    assert q.read_text() == text


def test_binary_known_answer(tmp_path):
    W0 = orc.complex_gaussian(1, (3, 1, 3, 2))
    W1 = orc.complex_gaussian(2, (3, 2, 3, 1))
    raw = struct.pack("<i", 2)
    for W in (W0, W1):
        raw += struct.pack("<i", 4) + struct.pack("<4i", *W.shape)
        flat = [W[i, j, k, l] for l in range(W.shape[3]) for k in range(W.shape[2])
                for j in range(W.shape[1]) for i in range(W.shape[0])]
        raw += b"".join(struct.pack("<dd", c.real, c.imag) for c in flat)
    p = tmp_path / "m.bin"
    p.write_bytes(raw)
    got = orc.mpo_read(str(p), "binary")
    assert np.array_equal(got[0], W0) and np.array_equal(got[1], W1)
    q = tmp_path / "w.bin"
    orc.mpo_write(str(q), [W0, W1], "binary")
    assert q.read_bytes() == raw


def test_text_digits_round_trip(tmp_path):
    W = [orc.complex_gaussian(3, (2, 1, 2, 4)), orc.complex_gaussian(4, (2, 4, 2, 1))]
    p = tmp_path / "m.txt"
    orc.mpo_write(str(p), W, "text", digits=17)
    assert all(np.array_equal(a, b) for a, b in zip(orc.mpo_read(str(p), "text"), W))
    orc.mpo_write(str(p), W, "text")  # the reference's 6 significant digits
    got = orc.mpo_read(str(p), "text")
    err = max(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)) for a, b in zip(got, W))
    assert 0 < err <= 1e-5


@pytest.mark.parametrize("S2,want", [(1, None), (2, 2.0), (3, 3.75)])
def test_spin_ops(S2, want):
    sx, sy, sz = orc.spin_ops(S2)
    if S2 == 1:
        assert np.allclose(sx, 0.5 * np.array([[0, 1], [1, 0]]), atol=0, rtol=0)
        assert np.allclose(sy, 0.5 * np.array([[0, -1j], [1j, 0]]), atol=1e-16)
        assert np.allclose(sz, 0.5 * np.diag([1, -1]), atol=0)
        return
    assert np.allclose(sx @ sy - sy @ sx, 1j * sz, atol=1e-14)
    assert np.allclose(sy @ sz - sz @ sy, 1j * sx, atol=1e-14)
    assert np.allclose(sx @ sx + sy @ sy + sz @ sz, want * np.eye(S2 + 1), atol=1e-14)


def _mixed_chain(dims, seed=0):
    rng = np.random.default_rng(seed)
    ops = {d: orc.spin_ops(d - 1) for d in set(dims)}
    fields = [0.3 * ops[d][2] + 0.2 * rng.standard_normal() * ops[d][0] for d in dims]
    bonds = []
    for k in range(len(dims) - 1):
        a, b = ops[dims[k]], ops[dims[k + 1]]
        bonds.append([(a[0], b[0]), (a[1], (0.5 + 0.1 * k) * b[1]), (a[2], -0.7 * b[2])])
    return bonds, fields


@pytest.mark.parametrize("dims", [[2, 2, 2, 2], [3, 3, 3], [2, 3, 2, 3, 3], [4, 2, 3]])
def test_chain_mpo_against_kronecker_sum(dims):
    bonds, fields = _mixed_chain(dims, seed=len(dims))
    mpo = orc.chain_mpo(bonds, fields)
    H = ph.chain_hamiltonian(bonds, fields).toarray()
    assert np.linalg.norm(ph.mpo_to_dense(mpo) - H) <= 1e-13 * np.linalg.norm(H)


def test_chain_mpo_reproduces_heisenberg():
    """The general builder and the Heisenberg restatement describe the same operator."""
    sx, sy, sz = orc.spin_ops(1)
    L = 5
    mpo = orc.chain_mpo([[(sx, 1.0 * sx), (sy, 0.8 * sy), (sz, 0.6 * sz)]] * (L - 1), [0.3 * sz] * L)
    ref = orc.heisenberg_mpo(L, 1.0, 0.8, 0.6, 0.3)
    assert np.allclose(ph.mpo_to_dense(mpo), ph.mpo_to_dense(ref), atol=1e-14)


MODELS = [
    ("ising", 0, 7, [-1.0, -0.7, 0.25, 0.4], lambda L, p: orc.ising_mpo(L, *p)),
    ("ising_posJ", 0, 4, [1.0, 0.5, 0.0, 0.0], lambda L, p: orc.ising_mpo(L, *p)),
    ("heisenberg", 1, 6, [1.0, 0.9, 1.1, 0.2], lambda L, p: orc.heisenberg_mpo(L, *p)),
    ("schwinger", 2, 8, [0.3, 4.0, 0.1, 0.0], lambda L, p: orc.schwinger_mpo(L, *p)),
]


@pytest.mark.parametrize("name,model,L,p,build", MODELS, ids=[m[0] for m in MODELS])
@pytest.mark.parametrize("fmt", ["binary", "text"])
def test_library_writer_matches_oracle(tmp_path, name, model, L, p, build, fmt):
    from mpsdyn_amd import device as dev
    cfg = dev.gs_config(model, L, 4, p)
    path = str(tmp_path / f"{name}.{fmt}")
    dev.mpo_save(cfg, path, dev.MPO_BINARY if fmt == "binary" else dev.MPO_TEXT)
    got = orc.mpo_read(path, fmt)
    want = build(L, p)
    assert len(got) == L
    for k, (g, w) in enumerate(zip(got, want)):
        assert g.shape == w.shape, k
        assert np.max(np.abs(g - w)) <= 1e-15 * max(1.0, np.max(np.abs(w))), k
    if fmt == "binary":
        assert os.path.getsize(path) == 4 + sum(4 + 16 + 16 * w.size for w in want)


def test_library_writer_errors(tmp_path):
    from mpsdyn_amd import device as dev
    cfg = dev.gs_config(dev.MODEL_ISING, 6, 4, [1.0, 1.0])
    with pytest.raises(dev.MpsdError):
        dev.mpo_save(cfg, str(tmp_path / "x.bin"), 7)
    with pytest.raises(dev.MpsdError):
        dev.mpo_save(cfg, str(tmp_path / "no_such_dir" / "x.bin"), dev.MPO_BINARY)
    bad = dev.gs_config(9, 6, 4, [1.0])
    with pytest.raises(dev.MpsdError):
        dev.mpo_save(bad, str(tmp_path / "y.bin"), dev.MPO_BINARY)
