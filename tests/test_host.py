"""CPU tests of the library boundary and host logic (no GPU calls)."""
import os
import re

import numpy as np
import pytest

from conftest import ROOT, make_synthetic


def _declared_functions():
    text = open(os.path.join(ROOT, "include", "h2eri.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:[A-Za-z_][\w ]*\*?\s*\**)\b([A-Za-z_]\w*)\s*\(", text, flags=re.M))
                  - {"NCART"})


def test_library_exports_every_declared_symbol():
    import ctypes
    import h2eri
    lib = h2eri.capi.load()
    names = _declared_functions()
    assert "H2ERI_matvec" in names and "H2ERI_build_Coulomb" in names and "H2ERI_build_exchange" in names
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) <= set(h2eri.capi.SIGNATURES), set(names) - set(h2eri.capi.SIGNATURES)
    assert lib.h2eri_struct_size() == ctypes.sizeof(h2eri.capi.H2ERIStruct)


def test_struct_leading_fields_match_reference_order():
    # callers of the reference write natom/nshell/shells and read num_bf, n_matvec, timers
    import h2eri
    S = h2eri.capi.H2ERIStruct
    order = [f[0] for f in S._fields_]
    assert order[:9] == ["natom", "nshell", "max_am", "max_shell_nbf", "num_bf", "num_sp", "num_sp_bfp",
                         "pp_npts_layer", "pp_nlayer_ext"]
    assert order[-4:] == ["n_matvec", "mat_size", "timers", "dev"]


@pytest.mark.parametrize("seed,n,leaf,mr,shape", [(1, 900, 30, 12, "chain"), (2, 1200, 150, 40, "chain"),
                                                  (3, 2500, 200, 700, "chain"), (4, 1500, 40, 20, "cube"),
                                                  (5, 40, 300, 8, "chain")])
def test_device_plan_bounds(tmp_path, seed, n, leaf, mr, shape):
    h = make_synthetic(tmp_path / "r.rep", n, seed, leaf=leaf, max_rank=mr, lr_frac=0.4, diag_lr_frac=0.3,
                       shape=shape)
    for bits in (64, 32):
        assert h.plan_check(bits) > 0


def test_save_load_roundtrip(tmp_path):
    import h2eri
    h = make_synthetic(tmp_path / "a.rep", 700, 9, leaf=25, max_rank=10, lr_frac=0.5, diag_lr_frac=0.5)
    h2 = h2eri.H2ERI.load(str(tmp_path / "a.rep"))
    h2.save(str(tmp_path / "b.rep"))
    assert open(tmp_path / "a.rep", "rb").read() == open(tmp_path / "b.rep", "rb").read()
    t1, t2 = h.tree(), h2.tree()
    for k in ("parent", "children", "mat_cluster", "r_adm_pairs", "r_inadm_pairs", "level_nodes"):
        assert np.array_equal(t1[k], t2[k]), k


def test_errors_are_reported_not_aborting(tmp_path):
    import h2eri
    h = h2eri.H2ERI()
    with pytest.raises(h2eri.H2ERIError):
        h.read_mol(str(tmp_path / "missing.mol"))
    with pytest.raises(h2eri.H2ERIError):
        h.partition()
    p = tmp_path / "trunc.rep"
    make_synthetic(tmp_path / "full.rep", 300, 1, leaf=20)
    p.write_bytes(open(tmp_path / "full.rep", "rb").read()[:5000])
    with pytest.raises(h2eri.H2ERIError):
        h2eri.H2ERI.load(str(p))
    h3 = make_synthetic(tmp_path / "x.rep", 300, 2, leaf=20)
    with pytest.raises(h2eri.H2ERIError):
        h3.build_h2(bd_jit=1)
    # tuning setters refuse out-of-range values (no GPU involved)
    with pytest.raises(h2eri.H2ERIError):
        h3.set_near_field_grid(-1)
    with pytest.raises(h2eri.H2ERIError):
        h3.set_exchange_scratch(100)
    with pytest.raises(h2eri.H2ERIError):
        h3.set_schedule(258)
    with pytest.raises(h2eri.H2ERIError):
        h3.set_wave_tile_bytes(-1)
    h3.set_near_field_grid(0)
    h3.set_exchange_scratch(2 << 30)
    h3.set_schedule(257)


def _rep_offsets(data):
    """Byte offsets of the sections of a representation file without
    chemistry metadata (the layout written by H2ERI_save_rep)."""
    hdr = np.frombuffer(data, dtype="<i4", count=11, offset=16)
    num_sp, n_node, n_leaf, max_level = int(hdr[3]), int(hdr[5]), int(hdr[6]), int(hdr[7])
    n_adm, n_inadm = int(hdr[9]), int(hdr[10])
    L = max_level + 1
    off = {"hdr": 16, "sp_bfp_sidx": 60}
    o = 60 + 4 * (num_sp + 1)
    for name, cnt in (("parent", n_node), ("children", 8 * n_node), ("n_child", n_node), ("node_level", n_node),
                      ("node_height", n_node), ("pt_cluster", 2 * n_node), ("mat_cluster", 2 * n_node),
                      ("level_n_node", L), ("level_nodes", L * n_leaf), ("height_n_node", L),
                      ("height_nodes", L * n_leaf), ("r_adm_pairs", 2 * n_adm), ("r_inadm_pairs", 2 * n_inadm)):
        off[name] = o
        o += 4 * cnt
    off["blocks"] = o
    return off, hdr


def test_load_rejects_corrupt_files(tmp_path):
    """A representation file with counts, indices or block shapes outside
    what its tree allows is refused with an error, before anything reads
    through them."""
    import h2eri
    make_synthetic(tmp_path / "ok.rep", 600, 3, leaf=25, max_rank=10, lr_frac=0.5, diag_lr_frac=0.5)
    data = bytearray(open(tmp_path / "ok.rep", "rb").read())
    off, hdr = _rep_offsets(data)
    n_node = int(hdr[5])
    root_children = off["children"] + 4 * 8 * (n_node - 1)
    cases = {
        "negative n_node": (off["hdr"] + 4 * 5, -5),
        "huge n_node": (off["hdr"] + 4 * 5, 1 << 30),
        "negative num_sp": (off["hdr"] + 4 * 3, -1),
        "x offset past N": (off["sp_bfp_sidx"] + 4, 1 << 27),
        "child index past the tree": (root_children, n_node + 10),
        "admissible pair names no node": (off["r_adm_pairs"], n_node),
        "leaf range past the pairs": (off["pt_cluster"] + 4, int(hdr[3]) + 5),
        "first block taller than its leaf": (off["blocks"], 100000),
    }
    for name, (pos, val) in cases.items():
        bad = bytearray(data)
        bad[pos:pos + 4] = int(val).to_bytes(4, "little", signed=True)
        p = tmp_path / "bad.rep"
        p.write_bytes(bytes(bad))
        with pytest.raises(h2eri.H2ERIError):
            h2eri.H2ERI.load(str(p))
            pytest.fail(name)
    for cut in (20, off["parent"] + 7, off["blocks"] + 10, len(data) - 4):
        p = tmp_path / "cut.rep"
        p.write_bytes(bytes(data[:cut]))
        with pytest.raises(h2eri.H2ERIError):
            h2eri.H2ERI.load(str(p))
    # the untouched file still loads
    h2eri.H2ERI.load(str(tmp_path / "ok.rep"))


@pytest.mark.parametrize("seed,n,leaf,mr,shape", [(1, 900, 30, 12, "chain"), (2, 1200, 150, 40, "chain"),
                                                  (4, 1500, 40, 20, "cube"), (5, 40, 300, 8, "chain"),
                                                  (6, 2000, 120, 600, "chain")])
def test_device_plan_semantics_match_oracle(tmp_path, seed, n, leaf, mr, shape):
    """The launch plan executed on the host with the kernels' semantics equals
    the oracle, and no launch has a read/write conflict between its tasks."""
    import oracle_bind
    from conftest import relerr
    h = make_synthetic(tmp_path / "r.rep", n, seed, leaf=leaf, max_rank=mr, lr_frac=0.5, diag_lr_frac=0.4,
                       shape=shape)
    o = oracle_bind.Oracle(str(tmp_path / "r.rep"))
    x = np.random.default_rng(seed).standard_normal(h.n)
    yo = o.matvec(x)
    y, conflicts = h.plan_emulate(x, 64)
    assert conflicts == 0
    assert relerr(y, yo) < 1e-13
    y32, conflicts = h.plan_emulate(x, 32)
    assert conflicts == 0
    assert relerr(y32, yo) < 2e-6
    # near field in the sweep launches / own launches / 1, 3, 8 or 39 groups
    # pipelined over launches of a second stream (more groups than blocks
    # leave launches empty)
    for sched in (0, 1, 2, 4, 9, 40):
        h.set_schedule(sched)
        y1, conflicts = h.plan_emulate(x, 64)
        assert conflicts == 0
        assert relerr(y1, yo) < 1e-13


@pytest.mark.parametrize("nv", [2, 4, 8, 16])
@pytest.mark.parametrize("seed,n,leaf,mr,shape", [(2, 1200, 150, 40, "chain"), (4, 1500, 40, 20, "cube"),
                                                  (6, 2000, 120, 600, "chain")])
def test_device_plan_semantics_batched(tmp_path, nv, seed, n, leaf, mr, shape):
    """Batched plans (nv vectors per pass over the matrix data) reproduce the
    oracle for every vector, with no intra-launch conflicts."""
    import oracle_bind
    from conftest import relerr
    h = make_synthetic(tmp_path / "r.rep", n, seed, leaf=leaf, max_rank=mr, lr_frac=0.5, diag_lr_frac=0.4,
                       shape=shape)
    o = oracle_bind.Oracle(str(tmp_path / "r.rep"))
    X = np.random.default_rng(seed + 10 * nv).standard_normal((nv, h.n))
    for sched in (3, 9):
        h.set_schedule(sched)
        Y, conflicts = h.plan_emulate(X, 64)
        assert conflicts == 0
        for v in range(nv):
            assert relerr(Y[v], o.matvec(X[v])) < 1e-13


@pytest.mark.parametrize("seed,n,leaf,mr,shape", [(1, 900, 30, 12, "chain"), (4, 1500, 40, 20, "cube"),
                                                  (5, 40, 300, 8, "chain"), (6, 2000, 120, 600, "chain")])
def test_queue_order_covers_every_dependency(tmp_path, seed, n, leaf, mr, shape):
    """The persistent executor's queue, executed in ticket order, reads no
    workspace element that is not produced by a task counted in the list it
    reads through, and reproduces the oracle; for several workgroup counts
    (which change the order), schedules and batch widths."""
    import oracle_bind
    from conftest import relerr
    h = make_synthetic(tmp_path / "r.rep", n, seed, leaf=leaf, max_rank=mr, lr_frac=0.5, diag_lr_frac=0.4,
                       shape=shape)
    o = oracle_bind.Oracle(str(tmp_path / "r.rep"))
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(h.n)
    yo = o.matvec(x)
    for sched in (3, 0, 12):
        h.set_schedule(sched)
        for wg in (1, 7, 1024):
            y, bad = h.queue_emulate(x, 64, workgroups=wg)
            assert bad == 0
            assert relerr(y, yo) < 1e-13
    for nv in (4, 8, 16):
        X = rng.standard_normal((nv, h.n))
        Y, bad = h.queue_emulate(X, 32, workgroups=256)
        assert bad == 0
        for v in range(nv):
            assert relerr(Y[v], o.matvec(X[v])) < 2e-6


def test_exchange_needs_molecule(tmp_path):
    """A synthetic (chemistry-free) representation has no shell pairs: the
    exchange build reports an error before touching the device."""
    import h2eri
    h = make_synthetic(tmp_path / "s.rep", 300, 3, leaf=20)
    buf = np.zeros(16)
    h.lib.H2ERI_build_exchange(h.h, buf.ctypes.data_as(h2eri.capi.c_dbl_p), buf.ctypes.data_as(h2eri.capi.c_dbl_p))
    with pytest.raises(h2eri.H2ERIError, match="shell-pair metadata"):
        h2eri._check(h.lib)


@pytest.mark.parametrize("kind,size", [("alkane", 4), ("alkane", 10), ("water", 6)])
def test_exchange_plan_checks_on_host(kind, size):
    """The exchange plan is built and every kernel access bounds-checked on
    the host (no GPU): entry lists, the row splits of the bra (C10H22 has
    260 basis functions, more than one split's 192), G offsets and the
    basis-product GEMMs.  Every D and B block of these
    representations becomes one term, and basis products exist exactly when
    there are admissible (B) pairs (C4H10 has none)."""
    import h2eri
    from h2eri import molecules as M
    rng = M.rng_for(700 + size)
    atoms = M.alkane(size, rng) if kind == "alkane" else M.water_cluster(size, rng)
    h = h2eri.H2ERI(1e-10, 1e-10, 1e-6)
    h.build_from_mol(text=M.mol_text(atoms))
    r = h.exchange_plan_check()
    s = h.s
    assert r["terms"] == s.n_leaf_node + s.n_r_inadm_pair + s.n_r_adm_pair
    assert r["terms_f"] == r["terms"]
    assert (r["gemm_flops"] > 0) == (s.n_r_adm_pair > 0)
    assert r["scratch_doubles"] > 0
    # resident bytes include the density, its screened copy and the three K buffers
    assert r["resident_bytes"] > 8 * 5 * h.nbf ** 2
    # G in chunks of at most 64 KB (nearly a chunk per term): the plan still
    # passes every check (chunk ranges, one chunk per term, G offsets within
    # the shared region) and its scratch shrinks
    h.set_exchange_scratch(65536)
    r2 = h.exchange_plan_check()
    assert r2["terms"] == r["terms"] and r2["scratch_doubles"] < r["scratch_doubles"]


def test_traffic_records_are_readable():
    """Every committed HBM-traffic record carries the fields the benchmark
    reads (kernel, launches per matvec, traffic bytes per matvec), so a
    record never breaks the benchmark line it feeds."""
    import glob
    import json
    import os
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    files = glob.glob(os.path.join(root, "profiles", "traffic_*_j_matvec.json"))
    assert files
    for f in files:
        with open(f) as fh:
            rec = json.load(fh)
        assert isinstance(rec["kernel"], str) and rec["launches_per_matvec"] >= 1, f
        assert rec["traffic_bytes_per_matvec"] > 0, f
