"""On-the-fly B and D blocks (BD_JIT = 1) on the GPU.

The device evaluates every coupling and near-field block from the shell pairs
in each matvec (src/H2ERI_matvec.c:411-568,761-879).  Checks:

* (H2O)6, whose coupling blocks include both level cases: the GPU matvec
  equals the oracle's matvec on the dense-stored representation (the same
  blocks, stored) to <= 1e-13, and the host emulation of the same task code;
  repeated calls are bitwise equal (every output row has one writer, fixed
  summation order); several vectors in one call equal single calls;
* the reference's fixture J_alk32 (tests/J_alk32_ccpvdz.bin of H2P-ERI) from
  the on-the-fly Coulomb build, bar 5e-7 as for the stored form, and that J
  against the oracle's J on the dense-stored C10H22 to <= 1e-13;
* entry points that need stored blocks refuse a BD_JIT handle;
* at the benchmark size (C30H62): symmetry and linearity to rounding, and
  agreement with the stored (compressed) representation within the
  compression tolerance.
"""
import os

import numpy as np
import pytest

import oracle_bind
from conftest import GOLDEN, relerr

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def water6(tmp_path_factory):
    from test_jit import _level_cases, _pair
    hd, hj, p = _pair("water", 6, 906, tmp_path_factory.mktemp("gjit_w6"))
    same, diff = _level_cases(hj)
    assert same > 0 and diff > 0
    return hd, hj, p


def test_jit_matvec_matches_dense_oracle(water6):
    hd, hj, p = water6
    o = oracle_bind.Oracle(p)
    x = np.random.default_rng(21).standard_normal(hj.n)
    print(f"jit matvec (H2O)6 N={hj.n}", flush=True)
    y = hj.matvec(x)
    assert relerr(y, o.matvec(x)) < 1e-13
    assert relerr(y, hj.jit_emulate(x)) < 1e-13
    assert np.array_equal(hj.matvec(x), y)   # bit-reproducible


def test_jit_multi_equals_single(water6):
    hd, hj, p = water6
    X = np.random.default_rng(22).standard_normal((3, hj.n))
    Y = hj.matvec_multi(X)
    for k in range(3):
        assert np.array_equal(Y[k], hj.matvec(X[k]))


def test_jit_device_pointers(water6):
    import torch
    hd, hj, p = water6
    x = np.random.default_rng(23).standard_normal(hj.n)
    xd = torch.from_numpy(x).to("cuda:0")
    yd = torch.empty_like(xd)
    torch.cuda.synchronize()
    hj.matvec_dev(xd.data_ptr(), yd.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert np.array_equal(yd.cpu().numpy(), hj.matvec(x))


def test_jit_refuses_plan_entry_points(water6):
    import torch
    import h2eri
    hd, hj, p = water6
    with pytest.raises(h2eri.H2ERIError, match="BD_JIT"):
        hj.autotune(nv=1, reps=1)
    xd = torch.zeros(hj.n, dtype=torch.float64, device="cuda:0")
    yd = torch.empty_like(xd)
    with pytest.raises(h2eri.H2ERIError, match="BD_JIT"):
        hj.graph_capture("matvec", 1, xd.data_ptr(), yd.data_ptr())
    # the handle still works afterwards
    x = np.random.default_rng(24).standard_normal(hj.n)
    assert relerr(hj.matvec(x), oracle_bind.Oracle(p).matvec(x)) < 1e-13


def test_jit_coulomb_alk32_fixture(tmp_path):
    import h2eri
    mol = os.path.join(GOLDEN, "alk32_ccpvdz.mol")
    hj = h2eri.H2ERI(1e-10, 1e-10, 1e-6)
    hj.read_mol(mol).process_shells().partition().build_h2(1)
    nbf = hj.nbf
    D = np.fromfile(os.path.join(GOLDEN, "D_alk32_ccpvdz.bin")).reshape(nbf, nbf)
    Jref = np.fromfile(os.path.join(GOLDEN, "J_alk32_ccpvdz.bin")).reshape(nbf, nbf)
    print(f"jit Coulomb alk32 N={hj.n}", flush=True)
    J = hj.build_coulomb(D)
    assert relerr(J, Jref) < 5e-7
    assert np.array_equal(J, J.T)
    hd = h2eri.H2ERI(1e-10, 1e-10, 1e-6)
    hd.set_block_compression(0)
    hd.build_from_mol(mol)
    p = str(tmp_path / "alk32_dense.rep")
    hd.save(p)
    assert relerr(J, oracle_bind.Oracle(p).coulomb(D)) < 1e-13


def test_jit_c30h62_full_size():
    """C30H62 at reltol 1e-6 (the benchmark molecule, 6.1e7 shell quartets
    per matvec): the on-the-fly operator is symmetric and linear to rounding
    at full size, and it differs from the stored (compressed) representation
    of the same molecule only by the compression tolerance (reltol 1e-6 on
    the blocks; 2e-6 on C10H22)."""
    import h2eri
    from h2eri import molecules as M
    atoms, txt, _, reltol = M.workload(1)
    hj = h2eri.H2ERI(1e-10, 1e-10, reltol)
    hj.read_mol_text(txt).process_shells().partition().build_h2(1)
    rng = np.random.default_rng(31)
    u, v = rng.standard_normal(hj.n), rng.standard_normal(hj.n)
    print(f"jit C30H62 N={hj.n}", flush=True)
    Au, Av = hj.matvec(u), hj.matvec(v)
    assert abs(u @ Av - v @ Au) <= 1e-12 * abs(u @ Av)
    assert relerr(hj.matvec(2.0 * u - 3.0 * v), 2.0 * Au - 3.0 * Av) < 1e-13
    hs = h2eri.H2ERI(1e-10, 1e-10, reltol)
    hs.build_from_mol(text=txt)
    e = relerr(hs.matvec(u), Au)
    assert 0.0 < e < 1e-5, e


def test_jit_phase_timers(water6):
    """A host-pointer matvec of a BD_JIT handle fills the reference's phase
    timer slots (src/H2ERI_matvec.c:890-954): forward, intermediate,
    backward and dense (the task time split by planned B and D cost), and
    the vector operations."""
    hd, hj, p = water6
    t0 = [hj.s.timers[i] for i in range(4, 9)]
    hj.matvec(np.random.default_rng(25).standard_normal(hj.n))
    t1 = [hj.s.timers[i] for i in range(4, 9)]
    d = [b - a for a, b in zip(t0, t1)]
    assert d[0] > 0 and d[1] > 0 and d[2] > 0 and d[3] > 0 and d[4] > 0, d
