"""GPU parity of the HIP H2 matvec against the CPU oracle.

Synthetic (chemistry-free) representations with every coupling case, dense
and low-rank B/D blocks, low-rank diagonal blocks, blocks wider than one
column tile, and ranks above one column tile.  Bar: relative error <= 1e-13
in FP64 (summation order differs from the oracle, values do not).
"""
import numpy as np
import pytest

import oracle_bind
from conftest import make_synthetic, relerr

pytestmark = pytest.mark.gpu

CASES = [
    # name, n, seed, leaf, max_rank, lr_frac, diag_lr, shape
    ("chain", 900, 11, 30, 12, 0.4, 0.0, "chain"),
    ("chain_lrdiag", 900, 12, 30, 16, 0.5, 0.6, "chain"),
    ("cube", 1500, 13, 40, 20, 0.3, 0.2, "cube"),
    ("wide_leaves", 1200, 14, 150, 40, 0.3, 0.3, "chain"),
    ("big_rank", 2500, 15, 200, 700, 0.0, 0.0, "chain"),
    ("tiny", 40, 16, 300, 8, 0.0, 0.0, "chain"),
]


@pytest.fixture(scope="module")
def reps(tmp_path_factory):
    d = tmp_path_factory.mktemp("reps")
    out = {}
    for name, n, seed, leaf, mr, lr, dlr, shape in CASES:
        p = d / f"{name}.rep"
        h = make_synthetic(p, n, seed, leaf=leaf, max_rank=mr, lr_frac=lr, diag_lr_frac=dlr, shape=shape)
        out[name] = (h, str(p))
    return out


@pytest.mark.parametrize("name", [c[0] for c in CASES])
def test_matvec_fp64_matches_oracle(reps, name):
    h, path = reps[name]
    o = oracle_bind.Oracle(path)
    rng = np.random.default_rng(5)
    for _ in range(2):
        x = rng.standard_normal(h.n)
        assert relerr(h.matvec(x), o.matvec(x)) < 1e-13


def test_matvec_bitwise_reproducible(reps):
    h, _ = reps["cube"]
    x = np.random.default_rng(1).standard_normal(h.n)
    y1, y2 = h.matvec(x), h.matvec(x)
    assert np.array_equal(y1, y2)


def test_matvec_loaded_equals_built(reps):
    import h2eri
    h, path = reps["chain_lrdiag"]
    h2 = h2eri.H2ERI.load(path)
    x = np.random.default_rng(2).standard_normal(h.n)
    assert np.array_equal(h.matvec(x), h2.matvec(x))


def test_matvec_fp32_storage(reps):
    h, path = reps["chain"]
    o = oracle_bind.Oracle(path)
    x = np.random.default_rng(3).standard_normal(h.n)
    y64 = o.matvec(x)
    h.set_precision(32)
    try:
        y32 = h.matvec(x)
    finally:
        h.set_precision(64)
    e = relerr(y32, y64)
    assert 0.0 < e < 2e-6


def test_matvec_device_pointers(reps):
    import torch
    h, path = reps["wide_leaves"]
    o = oracle_bind.Oracle(path)
    x = np.random.default_rng(4).standard_normal(h.n)
    xd = torch.from_numpy(x).to("cuda:0")
    yd = torch.empty_like(xd)
    torch.cuda.synchronize()
    h.matvec_dev(xd.data_ptr(), yd.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert relerr(yd.cpu().numpy(), o.matvec(x)) < 1e-13


@pytest.mark.parametrize("nvec", [2, 3, 4, 8, 15])
@pytest.mark.parametrize("name", ["cube", "chain_lrdiag", "big_rank"])
def test_matvec_multi(reps, name, nvec):
    """Batched passes (8, 4, 2, 1 vectors per pass over the matrix data)."""
    h, path = reps[name]
    o = oracle_bind.Oracle(path)
    X = np.random.default_rng(6 + nvec).standard_normal((nvec, h.n))
    Y = h.matvec_multi(X)
    for k in range(nvec):
        assert relerr(Y[k], o.matvec(X[k])) < 1e-13


@pytest.mark.parametrize("nvec", [16, 17, 33])
@pytest.mark.parametrize("name", ["cube", "chain_lrdiag", "big_rank", "wide_leaves", "tiny"])
def test_matvec_sixteen_per_pass(reps, name, nvec):
    """16 vectors per pass (FP64 MFMA tile products), with the rest in
    narrower passes; both executors, FP64 and FP32-stored U/B."""
    h, path = reps[name]
    o = oracle_bind.Oracle(path)
    X = np.random.default_rng(70 + nvec).standard_normal((nvec, h.n))
    Yo = [o.matvec(X[k]) for k in range(nvec)]
    try:
        for ex in (0, 1):
            h.set_executor(ex)
            for bits in (64, 32):
                h.set_precision(bits)
                Y = h.matvec_multi(X)
                for k in range(nvec):
                    e = relerr(Y[k], Yo[k])
                    assert e < (1e-13 if bits == 64 else 2e-6), (ex, bits, k, e)
                if ex == 1:
                    assert h.device_status() == 0
    finally:
        h.set_executor(0)
        h.set_precision(64)


@pytest.mark.parametrize("nvec", [1, 8])
def test_matvec_multi_fp32_storage(reps, nvec):
    h, path = reps["cube"]
    o = oracle_bind.Oracle(path)
    X = np.random.default_rng(7).standard_normal((nvec, h.n))
    h.set_precision(32)
    try:
        Y = h.matvec_multi(X)
        Y2 = h.matvec_multi(X)
    finally:
        h.set_precision(64)
    assert np.array_equal(Y, Y2)
    for k in range(nvec):
        assert 0.0 < relerr(Y[k], o.matvec(X[k])) < 2e-6


def test_matvec_multi_device_pointers(reps):
    import torch
    h, path = reps["chain"]
    o = oracle_bind.Oracle(path)
    X = np.random.default_rng(8).standard_normal((4, h.n))
    Xd = torch.from_numpy(X).to("cuda:0")
    Yd = torch.empty_like(Xd)
    torch.cuda.synchronize()
    h.matvec_multi_dev(4, Xd.data_ptr(), Yd.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    Y = Yd.cpu().numpy()
    for k in range(4):
        assert relerr(Y[k], o.matvec(X[k])) < 1e-13


@pytest.mark.parametrize("nvec", [1, 4])
def test_graph_replay_matches_eager(reps, nvec):
    """A captured graph replayed several times, with eager work and new input
    values between the replays, equals the eager call every time."""
    import torch
    h, path = reps["chain_lrdiag"]
    rng = np.random.default_rng(20 + nvec)
    Xd = torch.empty((nvec, h.n), dtype=torch.float64, device="cuda:0")
    Yd = torch.empty_like(Xd)
    Ye = torch.empty_like(Xd)
    s = torch.cuda.current_stream().cuda_stream
    h.graph_capture("matvec", nvec, Xd.data_ptr(), Yd.data_ptr())
    for _ in range(3):
        Xd.copy_(torch.from_numpy(rng.standard_normal((nvec, h.n))))
        Yd.zero_()
        h.graph_replay(s)
        h.matvec_multi_dev(nvec, Xd.data_ptr(), Ye.data_ptr(), s)
        torch.cuda.synchronize()
        assert torch.equal(Yd, Ye)


@pytest.mark.parametrize("name", [c[0] for c in CASES])
def test_persistent_executor_equals_staged(reps, name):
    """One persistent launch per pass gives the staged executor's results
    (FP64 and FP32-stored U/B, 1 to 15 vectors; equal to rounding: the staged
    plan runs small matrix parts as wave tiles) and reports no dependency wait
    past its bound; its own results are bitwise reproducible."""
    h, path = reps[name]
    X = np.random.default_rng(30).standard_normal((15, h.n))
    try:
        for bits in (64, 32):
            h.set_precision(bits)
            h.set_executor(0)
            Y0, y0 = h.matvec_multi(X), h.matvec(X[0])
            h.set_executor(1)
            Y1, y1 = h.matvec_multi(X), h.matvec(X[0])
            assert h.device_status() == 0
            for k in range(15):
                assert relerr(Y1[k], Y0[k]) < 1e-13
            assert relerr(y1, y0) < 1e-13
            assert np.array_equal(y1, h.matvec(X[0]))
    finally:
        h.set_executor(0)
        h.set_precision(64)
    o = oracle_bind.Oracle(path)
    h.set_executor(1)
    try:
        assert relerr(h.matvec(X[1]), o.matvec(X[1])) < 1e-13
    finally:
        h.set_executor(0)


@pytest.mark.parametrize("nvec", [1, 8])
def test_persistent_executor_graph_replay(reps, nvec):
    """The persistent pass captured once in a graph: every replay (new inputs
    and eager work between replays) equals bitwise the eager persistent call
    of a second handle on the same representation."""
    import torch
    import h2eri
    h, path = reps["cube"]
    ref = h2eri.H2ERI.load(path)
    ref.set_executor(1)
    rng = np.random.default_rng(40 + nvec)
    Xd = torch.empty((nvec, h.n), dtype=torch.float64, device="cuda:0")
    Yd = torch.empty_like(Xd)
    Ye = torch.empty_like(Xd)
    s = torch.cuda.current_stream().cuda_stream
    try:
        h.set_executor(1)
        h.graph_capture("matvec", nvec, Xd.data_ptr(), Yd.data_ptr())
        for _ in range(3):
            Xd.copy_(torch.from_numpy(rng.standard_normal((nvec, h.n))))
            Yd.zero_()
            h.graph_replay(s)
            ref.matvec_multi_dev(nvec, Xd.data_ptr(), Ye.data_ptr(), s)
            torch.cuda.synchronize()
            assert torch.equal(Yd, Ye)
        assert h.device_status() == 0
    finally:
        h.set_executor(0)


@pytest.mark.parametrize("wave_bytes", [0, 4096, 1 << 20])
def test_wave_tile_limits(reps, wave_bytes):
    """The staged executor with no, few and (nearly) all matrix parts as wave
    tiles (one wave each, four per workgroup) equals the oracle."""
    h, path = reps["cube"]
    o = oracle_bind.Oracle(path)
    X = np.random.default_rng(50).standard_normal((3, h.n))
    try:
        h.set_wave_tile_bytes(wave_bytes)
        Y = h.matvec_multi(X)
    finally:
        h.set_wave_tile_bytes(32768)
    for k in range(3):
        assert relerr(Y[k], o.matvec(X[k])) < 1e-13


@pytest.mark.parametrize("grid", [1, 7, 300])
def test_near_field_grid_cap(reps, grid):
    """Near-field launches with fewer workgroups than tasks (each workgroup
    loops over tasks g, g + grid, ...) run the same tasks with the same sums:
    results bitwise equal to one workgroup per task, at 1, 4 and 16 vectors
    per pass, eager and as a graph replay."""
    import torch
    h, path = reps["chain_lrdiag"]
    X = np.random.default_rng(60).standard_normal((21, h.n))   # passes of 16, 4, 1
    h.set_executor(0)
    h.set_schedule(3)
    Y0 = h.matvec_multi(X)
    try:
        h.set_near_field_grid(grid)
        Y1 = h.matvec_multi(X)
        Xd = torch.from_numpy(X).to("cuda:0")
        Yd = torch.empty_like(Xd)
        h.graph_capture("matvec", 21, Xd.data_ptr(), Yd.data_ptr())
        s = torch.cuda.current_stream().cuda_stream
        for _ in range(2):
            Yd.zero_()
            h.graph_replay(s)
            torch.cuda.synchronize()
            assert np.array_equal(Yd.cpu().numpy(), Y0)
    finally:
        h.set_near_field_grid(0)
    assert np.array_equal(Y1, Y0)
    o = oracle_bind.Oracle(path)
    assert relerr(Y1[20], o.matvec(X[20])) < 1e-13


def test_persistent_wait_bound_reports_error(tmp_path):
    """The persistent executor's dependency waits given up at once (a bound
    of zero polls, set from the host): every kernel still exits, and each
    synchronous call either reports the failed pass with an error or, when
    no wait had to give up, returns the correct product; on a tree of
    thousands of dependent tasks at least one call reports it.  With the
    default bound the same handle is correct again."""
    import h2eri
    h = make_synthetic(tmp_path / "w.rep", 8000, 31, leaf=60, max_rank=30, lr_frac=0.4, diag_lr_frac=0.3,
                       shape="cube")
    o = oracle_bind.Oracle(str(tmp_path / "w.rep"))
    X = np.random.default_rng(60).standard_normal((3, h.n))
    yo = o.matvec(X[0])
    h.set_executor(1)
    failed = 0
    try:
        h.set_spin_limit(0)
        for _ in range(3):
            try:
                y = h.matvec(X[0])
            except h2eri.H2ERIError as e:
                failed += 1
                assert "bound" in str(e)
            else:
                assert relerr(y, yo) < 1e-13
        try:
            Y = h.matvec_multi(X)
        except h2eri.H2ERIError:
            failed += 1
        else:
            assert relerr(Y[0], yo) < 1e-13
    finally:
        h.set_spin_limit(1 << 22)
    assert failed >= 1
    assert relerr(h.matvec(X[0]), yo) < 1e-13
    assert h.device_status() == 0
    h.set_executor(0)
