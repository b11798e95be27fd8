"""The drop-in API called from C: tests/c_driver/h2eri_c_driver.c, compiled
with gcc against include/H2ERI.h (the struct layout and signatures as a C
compiler sees them, not as a binding mirrors them), runs the reference
caller's sequence init -> CMS_read_mol_file -> process_shells -> partition ->
build_H2 -> build_Coulomb / build_exchange -> print_statistic -> destroy on
the reference's fixtures:

* J of D_alk32 for alk32_ccpvdz.mol against J_alk32 (bar 5e-7, as the CPU
  fixture test), n_matvec counting the calls;
* K of D_alk26 for alk26_ccpvdz.mol against K_alk26 (bar 5e-6);
* J of D_alk32 with H2ERI_build_H2(h2eri, 1): B and D evaluated on the
  device in every call (bar 5e-7), print_statistic reporting the mode.
"""
import os
import re
import subprocess

import pytest

from conftest import GOLDEN, ROOT

pytestmark = pytest.mark.gpu


def _driver():
    path = os.path.join(ROOT, "_build", "h2eri_c_driver")
    if not os.path.exists(path):   # normally built by __graft_entry__.build()
        import sys
        sys.path.insert(0, ROOT)
        import __graft_entry__
        __graft_entry__.build_c_driver()
    return path


def _run(*args):
    r = subprocess.run([_driver(), *args], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    m = re.search(r"result=(\w) nbf=(\d+) relerr=(\S+) n_matvec=(\d+) ms_per_call=(\S+)", r.stdout)
    assert m, r.stdout[-2000:]
    print(m.group(0), flush=True)
    return m.group(1), int(m.group(2)), float(m.group(3)), int(m.group(4)), r.stdout


def test_c_caller_coulomb_alk32():
    kind, nbf, err, nmv, out = _run(os.path.join(GOLDEN, "alk32_ccpvdz.mol"), os.path.join(GOLDEN, "D_alk32_ccpvdz.bin"),
                                    "1e-6", "J", os.path.join(GOLDEN, "J_alk32_ccpvdz.bin"), "3")
    assert kind == "J" and nbf == 260 and nmv == 3
    assert err < 5e-7
    assert "H2ERI" in out   # print_statistic ran


def test_c_caller_exchange_alk26():
    kind, nbf, err, nmv, _ = _run(os.path.join(GOLDEN, "alk26_ccpvdz.mol"), os.path.join(GOLDEN, "D_alk26_ccpvdz.bin"),
                                  "1e-6", "K", os.path.join(GOLDEN, "K_alk26_ccpvdz.bin"), "2")
    assert kind == "K" and nbf == 210 and nmv == 2
    assert err < 5e-6


def test_c_caller_coulomb_alk32_bd_jit():
    kind, nbf, err, nmv, out = _run(os.path.join(GOLDEN, "alk32_ccpvdz.mol"), os.path.join(GOLDEN, "D_alk32_ccpvdz.bin"),
                                    "1e-6", "J", os.path.join(GOLDEN, "J_alk32_ccpvdz.bin"), "2", "1")
    assert kind == "J" and nbf == 260 and nmv == 2
    assert err < 5e-7
    assert "Yes (B & D not allocated)" in out
