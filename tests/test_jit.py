"""On-the-fly B and D blocks (BD_JIT = 1), on the host.

A BD_JIT representation stores U only; every matvec evaluates the coupling
(B) and near-field (D) blocks from the shell pairs
(src/H2ERI_matvec.c:411-568,761-879).  Those blocks are the uncompressed ERI
blocks, so the reference value of an on-the-fly matvec is the oracle's matvec
on the same molecule built with B and D stored dense
(H2ERI_set_block_compression(h, 0)): same partition, same U (seeded), same
skeleton rows.  The device row-task code runs here through the host emulator
(H2ERI_jit_emulate: the same task function on one lane), so this pins the task
lists, the three coupling cases, the partner side rules and the
McMurchie-Davidson integrals against the host integral engine, which the
reference's J_alk32 / K_alk26 fixtures pin in turn.
"""
import numpy as np
import pytest

import oracle_bind
from conftest import relerr


def _pair(kind, size, seed, tmp_path):
    import h2eri
    from h2eri import molecules as M
    rng = M.rng_for(seed)
    atoms = M.alkane(size, rng) if kind == "alkane" else M.water_cluster(size, rng)
    txt = M.mol_text(atoms)
    hd = h2eri.H2ERI(1e-10, 1e-10, 1e-6)
    hd.set_block_compression(0)
    hd.build_from_mol(text=txt)
    p = str(tmp_path / f"{kind}{size}_dense.rep")
    hd.save(p)
    hj = h2eri.H2ERI(1e-10, 1e-10, 1e-6)
    hj.read_mol_text(txt).process_shells().partition().build_h2(1)
    return hd, hj, p


def _level_cases(h):
    s = h.s
    lv = s.node_level
    pairs = s.r_adm_pairs
    same = diff = 0
    for i in range(s.n_r_adm_pair):
        a, b = lv[pairs[2 * i]], lv[pairs[2 * i + 1]]
        if a == b:
            same += 1
        else:
            diff += 1
    return same, diff


@pytest.fixture(scope="module")
def water6(tmp_path_factory):
    return _pair("water", 6, 906, tmp_path_factory.mktemp("jit_w6"))


def test_jit_emulation_matches_dense_oracle(water6):
    """(H2O)6: coupling blocks of both kinds (equal levels, and a leaf at a
    coarser level) and near-field blocks; the emulated on-the-fly matvec
    equals the oracle on the dense-stored representation to rounding."""
    hd, hj, p = water6
    same, diff = _level_cases(hj)
    assert same > 0 and diff > 0, (same, diff)
    assert hj.s.n_r_inadm_pair > 0
    o = oracle_bind.Oracle(p)
    rng = np.random.default_rng(11)
    for _ in range(2):
        x = rng.standard_normal(hj.n)
        e = relerr(hj.jit_emulate(x), o.matvec(x))
        assert e < 1e-13, e


def test_jit_linear_and_symmetric(water6):
    """The on-the-fly operator is linear and symmetric (u^T A v = v^T A u),
    as the ERI tensor is: a dropped transpose-side contribution, a wrong
    partner side or a skeleton index off by one breaks one of these."""
    hd, hj, p = water6
    rng = np.random.default_rng(12)
    u = rng.standard_normal(hj.n)
    v = rng.standard_normal(hj.n)
    Au, Av = hj.jit_emulate(u), hj.jit_emulate(v)
    assert abs(u @ Av - v @ Au) <= 1e-12 * abs(u @ Av)
    Auv = hj.jit_emulate(2.0 * u - 3.0 * v)
    assert relerr(Auv, 2.0 * Au - 3.0 * Av) < 1e-13


def test_jit_representation_stores_no_blocks(water6, tmp_path):
    """BD_JIT = 1 keeps U and the B/D shapes but no block data: the block
    arrays hold null entries, the B and D storage statistics are zero, and
    saving such a representation is refused."""
    import h2eri
    hd, hj, p = water6
    s = hj.s
    assert s.BD_JIT == 1 and hd.s.BD_JIT == 0
    assert not any(bool(s.c_B_blks[i]) for i in range(s.n_B))
    assert not any(bool(s.c_D_blks[i]) for i in range(s.n_D))
    assert s.mat_size[1] == 0 and s.mat_size[2] == 0   # B and D elements (B_SIZE_IDX, D_SIZE_IDX)
    assert hd.s.mat_size[1] > 0 and hd.s.mat_size[2] > 0
    with pytest.raises(h2eri.H2ERIError):
        hj.save(str(tmp_path / "jit.rep"))
    st = hj.jit_stats()
    assert st["tasks"] > 0 and st["quartets"] > 0 and st["eri_values"] >= st["quartets"]
    # the stored-block plan helpers refuse it; the dense-stored handle has no on-the-fly plan
    with pytest.raises(h2eri.H2ERIError, match="BD_JIT"):
        hj.plan_check()
    with pytest.raises(h2eri.H2ERIError, match="BD_JIT"):
        hj.plan_emulate(np.zeros(hj.n))
    with pytest.raises(h2eri.H2ERIError):
        hd.jit_stats()


def test_dense_blocks_option(water6):
    """H2ERI_set_block_compression(h, 0) stores every B and D block dense
    (ld >= 0); the default compresses where it pays."""
    import h2eri
    hd, hj, p = water6
    s = hd.s
    assert all(s.c_B_blks[i].contents.ld >= 0 for i in range(s.n_B))
    assert all(s.c_D_blks[i].contents.ld >= 0 for i in range(s.n_D))
    from h2eri import molecules as M
    rng = M.rng_for(906)
    hc = h2eri.H2ERI(1e-10, 1e-10, 1e-6)
    hc.build_from_mol(text=M.mol_text(M.water_cluster(6, rng)))
    sc = hc.s
    assert any(sc.c_D_blks[i].contents.ld < 0 for i in range(sc.n_D))


def test_jit_refuses_shells_above_d():
    """BD_JIT = 1 evaluates shells up to d on the device; a basis with an f
    shell is refused at H2ERI_build_H2 with an error (nothing built), and the
    same molecule still builds with stored blocks."""
    import h2eri
    mol = """2
H 2 4 4
0.0 0.0 0.0
S 3 1
13.01    0.0196850
1.962    0.1379770
0.4446    0.4781480
F 1 1
0.8    1.0
H 2 4 4
0.0 0.0 1.4
S 3 1
13.01    0.0196850
1.962    0.1379770
0.4446    0.4781480
F 1 1
0.8    1.0
"""
    h = h2eri.H2ERI(1e-10, 1e-10, 1e-6)
    h.read_mol_text(mol).process_shells().partition()
    with pytest.raises(h2eri.H2ERIError, match="up to d"):
        h.build_h2(1)
    assert h.s.BD_JIT == 0
    h.build_h2(0)
    assert h.s.BD_JIT == 0 and h.s.n_D > 0


def test_single_leaf_molecule(tmp_path):
    """A molecule small enough to be one leaf (no far field, one diagonal
    near-field block): the stored and the on-the-fly representation both
    build, and the device plan (host emulation) and the on-the-fly task code
    (host emulation) equal the oracle."""
    import h2eri
    mol = """2
H 2 4 4
0.0 0.0 0.0
S 3 1
13.01    0.0196850
1.962    0.1379770
0.4446    0.4781480
P 1 1
0.727    1.0
H 2 4 4
0.0 0.0 1.4
S 3 1
13.01    0.0196850
1.962    0.1379770
0.4446    0.4781480
P 1 1
0.727    1.0
"""
    hd = h2eri.H2ERI(1e-10, 1e-10, 1e-6)
    hd.set_block_compression(0)
    hd.read_mol_text(mol).process_shells().partition().build_h2(0)
    assert hd.s.n_node == 1 and hd.s.n_r_adm_pair == 0
    p = str(tmp_path / "h2.rep")
    hd.save(p)
    o = oracle_bind.Oracle(p)
    x = np.random.default_rng(41).standard_normal(hd.n)
    yo = o.matvec(x)
    y, conflicts = hd.plan_emulate(x)
    assert conflicts == 0 and relerr(y, yo) < 1e-13
    hj = h2eri.H2ERI(1e-10, 1e-10, 1e-6)
    hj.read_mol_text(mol).process_shells().partition().build_h2(1)
    assert relerr(hj.jit_emulate(x), yo) < 1e-13
