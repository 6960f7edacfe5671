"""Ground states from MPO files (MPO::loadMPO / importMPOtext, src/core/cpp/MPO.cpp:250-316,
then Contractor::findGroundState) through mpsd_gs_solve_mpo.  Checks:
  * a model's MPO written by the library and read back gives the same energy, bit for bit, as
    the solve that builds the model in memory (same operator bits, same seed);
  * chains the built-in models do not cover -- spin-1 (d = 3) and mixed d = 2 / 3 sites -- written
    by the oracle in either layout, against exact diagonalisation of the Kronecker-sum
    Hamiltonian (energies to 1e-10 relative, D at the exact bond dimension);
  * a file with the reference's 6-digit text precision reads and solves;
  * malformed files are rejected with an error, before any device work."""
import numpy as np
import pytest

from oracle import mpsd_oracle as orc
from tests import _physics as ph

pytestmark = pytest.mark.gpu

ETOL = 1e-10


@pytest.fixture(scope="module")
def ctx():
    from mpsdyn_amd import device as dev
    c = dev.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("model,L,D,p", [(0, 16, 16, [-1.0, -1.0, 0.5]), (1, 14, 24, [1.0, 1.0, 0.8, 0.1]),
                                         (2, 12, 20, [0.4, 3.0, 0.2, 0.0])])
@pytest.mark.parametrize("fmt", [0, 1])
def test_file_solve_matches_model_solve(ctx, tmp_path, model, L, D, p, fmt):
    from mpsdyn_amd import device as dev
    print(f"mpo file model={model} L={L} D={D} fmt={fmt}", flush=True)
    cfg = dev.gs_config(model, L, D, p, convtol=1e-12, seed=5)
    path = str(tmp_path / "h.mpo")
    dev.mpo_save(cfg, path, fmt, ctx=ctx)
    want = dev.gs_solve(ctx, cfg)
    got = dev.gs_solve_mpo(ctx, cfg, path, fmt)
    assert got.energy == want.energy
    assert got.halfsweeps == want.halfsweeps and got.matvecs == want.matvecs


def _spin1_chain(L, J=(1.0, 1.0, 1.0), D2=0.0, h=0.0):
    sx, sy, sz = orc.spin_ops(2)
    bonds = [[(sx, J[0] * sx), (sy, J[1] * sy), (sz, J[2] * sz)]] * (L - 1)
    fields = [D2 * sz @ sz + h * sz] * L
    return bonds, fields


def _mixed_chain(dims):
    ops = {d: orc.spin_ops(d - 1) for d in set(dims)}
    bonds = [[(ops[dims[k]][0], ops[dims[k + 1]][0]), (ops[dims[k]][1], ops[dims[k + 1]][1]),
              (ops[dims[k]][2], 1.3 * ops[dims[k + 1]][2])] for k in range(len(dims) - 1)]
    fields = [0.15 * ops[d][2] + 0.1 * ops[d][0] for d in dims]
    return bonds, fields


CHAINS = [
    ("spin1_L6", lambda: _spin1_chain(6, D2=0.3, h=0.1), 27),
    ("spin1_L8", lambda: _spin1_chain(8, J=(1.0, 0.9, 1.2)), 81),
    ("mixed_2323", lambda: _mixed_chain([2, 3, 2, 3, 2, 3]), 18),
    ("mixed_33222", lambda: _mixed_chain([3, 3, 2, 2, 2, 3, 2]), 24),
]


@pytest.mark.parametrize("name,make,D", CHAINS, ids=[c[0] for c in CHAINS])
@pytest.mark.parametrize("fmt", ["binary", "text"])
def test_general_chain_against_exact(ctx, tmp_path, name, make, D, fmt):
    from mpsdyn_amd import device as dev
    print(f"chain {name} D={D} fmt={fmt}", flush=True)
    bonds, fields = make()
    mpo = orc.chain_mpo(bonds, fields)
    path = str(tmp_path / "h.mpo")
    orc.mpo_write(path, mpo, fmt, digits=17)
    e_exact = ph.ground_energy(ph.chain_hamiltonian(bonds, fields))
    cfg = dev.gs_config(0, 0, D, [], convtol=1e-12, seed=11)
    r = dev.gs_solve_mpo(ctx, cfg, path, dev.MPO_BINARY if fmt == "binary" else dev.MPO_TEXT,
                         export_path=str(tmp_path / "gs.mps"))
    assert abs(r.energy - e_exact) <= ETOL * abs(e_exact)
    # the exported state evaluated on the same file gives the same energy
    e = dev.mpo_expectation(ctx, path, dev.MPO_BINARY if fmt == "binary" else dev.MPO_TEXT, str(tmp_path / "gs.mps"))
    assert abs(e.real - r.energy) <= ETOL * abs(e_exact) and abs(e.imag) <= 1e-10


def test_reference_text_precision(ctx, tmp_path):
    """A 6-significant-digit text file (the reference's default stream precision) solves to the
    exact ground energy of the rounded operators it holds."""
    from mpsdyn_amd import device as dev
    bonds, fields = _spin1_chain(6, J=(1.0, 1.0, 1.0), h=0.123456789)
    path = str(tmp_path / "h.txt")
    orc.mpo_write(path, orc.chain_mpo(bonds, fields), "text")
    rounded = orc.mpo_read(path, "text")
    e_exact = ph.ground_energy(ph.sp.csr_matrix(ph.mpo_to_dense(rounded)))
    r = dev.gs_solve_mpo(ctx, dev.gs_config(0, 0, 27, [], convtol=1e-12, seed=2), path, dev.MPO_TEXT)
    assert abs(r.energy - e_exact) <= ETOL * abs(e_exact)


def _write_raw(path, ops):
    orc.mpo_write(path, ops, "binary")


@pytest.mark.parametrize("case", ["bond_mismatch", "not_square", "open_end", "one_site", "truncated",
                                  "missing", "bad_format"])
def test_malformed_files_rejected(ctx, tmp_path, case):
    from mpsdyn_amd import device as dev
    path = str(tmp_path / "bad.mpo")
    g = orc.complex_gaussian
    fmt = dev.MPO_BINARY
    if case == "bond_mismatch":
        _write_raw(path, [g(1, (2, 1, 2, 3)), g(2, (2, 4, 2, 1))])
    elif case == "not_square":
        _write_raw(path, [g(1, (2, 1, 3, 3)), g(2, (2, 3, 2, 1))])
    elif case == "open_end":
        _write_raw(path, [g(1, (2, 1, 2, 3)), g(2, (2, 3, 2, 2))])
    elif case == "one_site":
        _write_raw(path, [g(1, (2, 1, 2, 1))])
    elif case == "truncated":
        orc.mpo_write(path, [g(1, (2, 1, 2, 3)), g(2, (2, 3, 2, 1))], "text")
        txt = open(path).read()
        open(path, "w").write(txt[: len(txt) // 2])
        fmt = dev.MPO_TEXT
    elif case == "missing":
        path = str(tmp_path / "does_not_exist.mpo")
    elif case == "bad_format":
        _write_raw(path, [g(1, (2, 1, 2, 3)), g(2, (2, 3, 2, 1))])
        fmt = 5
    with pytest.raises(dev.MpsdError):
        dev.gs_solve_mpo(ctx, dev.gs_config(0, 0, 8, [], convtol=1e-10), path, fmt)
    # the context still works afterwards
    r = dev.gs_solve(ctx, dev.gs_config(0, 8, 8, [-1.0, -1.0], convtol=1e-10))
    assert np.isfinite(r.energy)
