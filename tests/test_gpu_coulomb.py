"""GPU parity of the Coulomb (J) build: uncontract -> H2 matvec -> contract.

* an octane molecule built here, GPU J vs the oracle's J on the same
  representation (FP64: <= 1e-13; FP32-stored U/B: <= 1e-7 vs FP64);
* the reference's fixture J_alk32 (tests/J_alk32_ccpvdz.bin of H2P-ERI) from
  the GPU path, bar 5e-7 as in the CPU fixture test;
* several densities in one call equal single-density calls.
"""
import os

import numpy as np
import pytest

import oracle_bind
from conftest import GOLDEN, relerr

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def butane(tmp_path_factory):
    import h2eri
    from h2eri import molecules as M
    rng = M.rng_for(100)
    atoms = M.alkane(8, rng)
    D = M.density(atoms, 33, rng)
    h = h2eri.H2ERI(1e-10, 1e-10, 1e-6)
    h.build_from_mol(text=M.mol_text(atoms))
    p = tmp_path_factory.mktemp("c4") / "c4.rep"
    h.save(str(p))
    assert h.s.n_B > 0 and h.s.n_r_inadm_pair > 0   # far and near field both present
    return h, str(p), D


def test_coulomb_fp64_matches_oracle(butane):
    h, path, D = butane
    o = oracle_bind.Oracle(path)
    J = h.build_coulomb(D)
    Jo = o.coulomb(D)
    assert relerr(J, Jo) < 1e-13
    assert np.array_equal(J, J.T)


def test_coulomb_fp32_storage(butane):
    h, path, D = butane
    J64 = h.build_coulomb(D)
    h.set_precision(32)
    try:
        J32 = h.build_coulomb(D)
    finally:
        h.set_precision(64)
    assert 0.0 < relerr(J32, J64) < 1e-7   # precision gate of the FP32-stored form


def test_coulomb_multi_equals_single(butane):
    h, path, D = butane
    rng = np.random.default_rng(3)
    Ds = []
    for _ in range(5):
        A = rng.standard_normal(D.shape) * 1e-2
        Ds.append(D + A + A.T)
    Ds = np.array(Ds)
    Js = h.build_coulomb_multi(Ds)
    o = oracle_bind.Oracle(path)
    for k in range(5):
        assert relerr(Js[k], o.coulomb(Ds[k])) < 1e-13


def test_coulomb_device_pointers(butane):
    import torch
    h, path, D = butane
    Dd = torch.from_numpy(D).to("cuda:0")
    Jd = torch.empty_like(Dd)
    torch.cuda.synchronize()
    h.build_coulomb_dev(Dd.data_ptr(), Jd.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert relerr(Jd.cpu().numpy(), oracle_bind.Oracle(path).coulomb(D)) < 1e-13


def test_coulomb_alk32_fixture():
    import h2eri
    h = h2eri.H2ERI(1e-10, 1e-10, 1e-6)
    h.build_from_mol(os.path.join(GOLDEN, "alk32_ccpvdz.mol"))
    nbf = h.nbf
    D = np.fromfile(os.path.join(GOLDEN, "D_alk32_ccpvdz.bin")).reshape(nbf, nbf)
    Jref = np.fromfile(os.path.join(GOLDEN, "J_alk32_ccpvdz.bin")).reshape(nbf, nbf)
    assert relerr(h.build_coulomb(D), Jref) < 5e-7
    h.set_precision(32)
    assert relerr(h.build_coulomb(D), Jref) < 5e-7


def test_coulomb_graph_replay(butane):
    import torch
    h, path, D = butane
    o = oracle_bind.Oracle(path)
    Dd = torch.from_numpy(D).to("cuda:0")
    Jd = torch.empty_like(Dd)
    s = torch.cuda.current_stream().cuda_stream
    h.graph_capture("coulomb", 1, Dd.data_ptr(), Jd.data_ptr())
    for k in range(3):
        Dk = D * (1.0 + 0.1 * k)
        Dd.copy_(torch.from_numpy(Dk))
        Jd.zero_()
        h.graph_replay(s)
        torch.cuda.synchronize()
        assert relerr(Jd.cpu().numpy(), o.coulomb(Dk)) < 1e-13


def test_coulomb_graph_keeps_its_buffers(butane):
    """A Coulomb graph captured on a fresh handle stays valid when later eager
    calls need larger buffers than the graph's (several densities, several
    vectors), and a change that releases it makes the next replay fail with
    an error instead of running on freed memory."""
    import torch
    import h2eri
    _, path, D = butane
    h = h2eri.H2ERI.load(path)
    o = oracle_bind.Oracle(path)
    Jo = o.coulomb(D)
    Dd = torch.from_numpy(D).to("cuda:0")
    Jd = torch.empty_like(Dd)
    s = torch.cuda.current_stream().cuda_stream
    h.graph_capture("coulomb", 1, Dd.data_ptr(), Jd.data_ptr())
    Ds = np.array([D * (1.0 + 0.1 * k) for k in range(4)])
    Js = h.build_coulomb_multi(Ds)
    X = np.random.default_rng(9).standard_normal((6, h.n))
    Y = h.matvec_multi(X)
    for k in range(2):
        Jd.zero_()
        h.graph_replay(s)
        torch.cuda.synchronize()
        assert relerr(Jd.cpu().numpy(), Jo) < 1e-13
        assert relerr(Js[k], o.coulomb(Ds[k])) < 1e-13
    assert relerr(Y[5], o.matvec(X[5])) < 1e-13
    h.set_precision(32)
    with pytest.raises(h2eri.H2ERIError):
        h.graph_replay(s)
