"""The allocation-free matvec (H2ERI_matvec_dev_noalloc) against
H2ERI_matvec_dev, both called through the C-ABI.

Two representations on one GPU, the second smaller than the first, called
one after the other so the second call finds the first one's contents in
the shared scratch array; with both executors.  Same plan, same kernels,
other workspace address: the results are bitwise equal.  A handle that is
not prepared, or whose device state changed since, is refused with an error.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_bind
from conftest import make_synthetic, relerr

pytestmark = pytest.mark.gpu


def _call(h, fn, x, y, stream):
    getattr(h.lib, fn)(h.h, C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr()), C.c_void_p(stream))
    err = h.lib.h2eri_last_error()
    msg = h.lib.h2eri_last_error_msg().decode() if err else ""
    h.lib.h2eri_clear_error()
    return err, msg


def test_noalloc_matches_matvec_dev(tmp_path):
    import torch
    big = make_synthetic(tmp_path / "big.rep", 3000, 71, leaf=60, max_rank=30, lr_frac=0.4, diag_lr_frac=0.3,
                         shape="cube")
    small = make_synthetic(tmp_path / "small.rep", 900, 72, leaf=30, max_rank=12, lr_frac=0.4, diag_lr_frac=0.3)
    oracles = {"big": oracle_bind.Oracle(str(tmp_path / "big.rep")),
               "small": oracle_bind.Oracle(str(tmp_path / "small.rep"))}
    s = torch.cuda.current_stream().cuda_stream
    rng = np.random.default_rng(70)
    for ex in (0, 1):
        for bits in (64, 32):
            for h in (big, small):
                h.set_executor(ex)
                h.set_precision(bits)
                assert h.upload() > 0
            for name, h in (("big", big), ("small", small)):   # the second call: a smaller input
                x = rng.standard_normal(h.n)
                xd = torch.from_numpy(x).to("cuda:0")
                y_ref = torch.full_like(xd, np.nan)
                y_na = torch.full_like(xd, np.nan)
                assert _call(h, "H2ERI_matvec_dev", xd, y_ref, s) == (0, "")
                assert _call(h, "H2ERI_matvec_dev_noalloc", xd, y_na, s) == (0, "")
                torch.cuda.synchronize()
                assert torch.equal(y_na, y_ref), (name, ex, bits)
                assert h.device_status() == 0
                e = relerr(y_na.cpu().numpy(), oracles[name].matvec(x))
                assert e < (1e-13 if bits == 64 else 2e-6), (name, ex, bits, e)
    # a change of precision releases the prepared state: refused, not run
    small.set_precision(64)
    xd = torch.zeros(small.n, dtype=torch.float64, device="cuda:0")
    err, msg = _call(small, "H2ERI_matvec_dev_noalloc", xd, torch.empty_like(xd), s)
    assert err != 0 and "prepared" in msg
    for h in (big, small):
        h.set_executor(0)
        h.set_precision(64)
