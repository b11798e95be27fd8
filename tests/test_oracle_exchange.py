"""Pins of the CPU oracle's exchange matrix (CPU only).

orc_build_exchange forms the blocks of A_H2 from expanded nested bases and
contracts them with D.  It must equal an independent contraction of the dense
brute-force matrix orc_dense_matrix (over both orientations of every shell
pair), reproduce the reference fixture K_alk26 within the H2 tolerance, and be
linear and symmetric in D.
"""
import os

import numpy as np
import pytest

import oracle_bind
from conftest import GOLDEN, relerr


def pair_rows(tr, n):
    """(a, b, shells differ) of every row of A_H2 (pairs stored smaller shell first)."""
    sp, sb, bf = tr["sp_shell_idx"], tr["sp_bfp_sidx"], tr["shell_bf_sidx"]
    npair = tr["num_sp"]
    ra = np.empty(n, int)
    rb = np.empty(n, int)
    same = np.empty(n, bool)
    for i in range(npair):
        s0, s1 = sp[i], sp[i + npair]
        n0, n1 = bf[s0 + 1] - bf[s0], bf[s1 + 1] - bf[s1]
        for j in range(n1):
            for k in range(n0):
                r = sb[i] + j * n0 + k
                ra[r], rb[r], same[r] = bf[s0] + k, bf[s1] + j, s0 == s1
    return ra, rb, same


def dense_exchange(A, D, ra, rb, same, drop_swap=False):
    """K[m,q] = sum_{r,c} A[r,c] sum over orientations (m,n) of r, (p,q) of c of D[n,p]."""
    nbf, n = D.shape[0], A.shape[0]
    K = np.zeros((nbf, nbf))
    ori = [(ra, rb, np.ones(n, bool))] + ([] if drop_swap else [(rb, ra, ~same)])
    for m, nn, wr in ori:
        Mr = np.zeros((nbf, n))
        Mr[m[wr], np.nonzero(wr)[0]] = 1.0
        for p, q, wc in ori:
            W = A * D[np.ix_(nn, p)]
            W[~wr, :] = 0.0
            W[:, ~wc] = 0.0
            Qc = np.zeros((n, nbf))
            Qc[np.nonzero(wc)[0], q[wc]] = 1.0
            K += Mr @ W @ Qc
    return K


@pytest.fixture(scope="module")
def ethane(tmp_path_factory):
    import h2eri
    from h2eri import molecules as M
    rng = M.rng_for(200)
    atoms = M.alkane(3, rng)
    D = M.density(atoms, 13, rng)
    h = h2eri.H2ERI(1e-10, 1e-10, 1e-6)
    h.build_from_mol(text=M.mol_text(atoms))
    p = str(tmp_path_factory.mktemp("c3") / "c3.rep")
    h.save(p)
    return h, p, D


def test_oracle_exchange_equals_dense_contraction(ethane):
    h, p, D = ethane
    o = oracle_bind.Oracle(p)
    A = o.dense()
    ra, rb, same = pair_rows(h.tree(), h.n)
    Kd = dense_exchange(A, D, ra, rb, same)
    K = o.exchange(D)
    assert relerr(K, Kd) < 1e-13
    # planted mistake (CPU): dropping the swapped orientation changes K
    assert relerr(dense_exchange(A, D, ra, rb, same, drop_swap=True), Kd) > 1e-2


def test_oracle_exchange_linear_symmetric(ethane):
    h, p, D = ethane
    o = oracle_bind.Oracle(p)
    rng = np.random.default_rng(1)
    E = rng.standard_normal(D.shape)
    E = E + E.T
    K1, K2 = o.exchange(D), o.exchange(E)
    assert relerr(o.exchange(2.0 * D - 3.0 * E), 2.0 * K1 - 3.0 * K2) < 1e-13
    # A_H2 is symmetric; same-shell rows (ab| and (ba| are compressed separately,
    # so K is symmetric to the H2 tolerance
    assert relerr(K1, K1.T) < 1e-5


def test_oracle_exchange_alk26_fixture(tmp_path):
    """Reference fixture tests/K_alk26_ccpvdz.bin (K of tests/D_alk26_ccpvdz.bin
    for tests/alk26_ccpvdz.mol); measured 1.37e-6 at reltol 1e-6."""
    import h2eri
    h = h2eri.H2ERI(1e-10, 1e-10, 1e-6)
    h.build_from_mol(os.path.join(GOLDEN, "alk26_ccpvdz.mol"))
    p = str(tmp_path / "alk26.rep")
    h.save(p)
    nbf = h.nbf
    D = np.fromfile(os.path.join(GOLDEN, "D_alk26_ccpvdz.bin")).reshape(nbf, nbf)
    Kref = np.fromfile(os.path.join(GOLDEN, "K_alk26_ccpvdz.bin")).reshape(nbf, nbf)
    K = oracle_bind.Oracle(p).exchange(D)
    assert relerr(K, Kref) < 5e-6


def np_dlist(D, bf):
    """The reference's density screening rule (src/H2ERI_build_exchange.c:
    774-838): block (N,P) is kept when some |D_np| > 1e-10."""
    ns = len(bf) - 1
    keep = np.zeros((ns, ns), bool)
    for n in range(ns):
        for p in range(ns):
            keep[n, p] = np.abs(D[bf[n]:bf[n + 1], bf[p]:bf[p + 1]]).max() > 1e-10
    return keep


def screened_test_density(D, bf, seed):
    """A symmetric density whose shell blocks sit just above or below the
    screening threshold: 70% of the blocks scaled to a maximum of 9e-11
    (dropped), the rest to 2e-9 (kept), plus one block exactly at 1e-10
    (dropped: the rule is strict) and one at 1.0000001e-10 (kept)."""
    rng = np.random.default_rng(seed)
    ns = len(bf) - 1
    out = np.zeros_like(D)
    special = [(0, 1, 1e-10), (1, 2, 1.0000001e-10)]
    for n in range(ns):
        for p in range(n, ns):
            blk = D[bf[n]:bf[n + 1], bf[p]:bf[p + 1]]
            mx = np.abs(blk).max()
            target = 9e-11 if rng.random() < 0.7 else 2e-9
            for a, b, v in special:
                if (n, p) == (a, b):
                    target = v
            scaled = blk * (target / mx) if mx > 0 else blk
            out[bf[n]:bf[n + 1], bf[p]:bf[p + 1]] = scaled
            out[bf[p]:bf[p + 1], bf[n]:bf[n + 1]] = scaled.T
    return out


def test_dlist_matches_restatement(ethane):
    """Oracle and library density screening lists equal the numpy
    restatement, on a damped density and on one with blocks around the
    threshold (boundary cases included)."""
    h, p, D = ethane
    o = oracle_bind.Oracle(p)
    bf = h.tree()["shell_bf_sidx"]
    for Dx in (D, screened_test_density(D, bf, 3)):
        keep = np_dlist(Dx, bf)
        sidx_ref = np.concatenate([[0], np.cumsum(keep.sum(axis=1))])
        list_ref = np.concatenate([np.nonzero(keep[n])[0] for n in range(len(bf) - 1)])
        for sidx, lst in (o.dlist(Dx), h.exchange_dlist(Dx)):
            assert np.array_equal(sidx, sidx_ref)
            assert np.array_equal(lst, list_ref)
    Dx = screened_test_density(D, bf, 3)
    keep = np_dlist(Dx, bf)
    assert not keep[0, 1] and keep[1, 2] and 0.2 < keep.mean() < 0.5


def test_oracle_exchange_screens_density(ethane):
    """K of a density with sub-threshold blocks equals the dense contraction
    with those blocks zeroed, and differs from the unscreened contraction."""
    h, p, D = ethane
    o = oracle_bind.Oracle(p)
    bf = h.tree()["shell_bf_sidx"]
    Dx = screened_test_density(D, bf, 4)
    keep = np_dlist(Dx, bf)
    Ds = Dx.copy()
    for n in range(len(bf) - 1):
        for q in range(len(bf) - 1):
            if not keep[n, q]:
                Ds[bf[n]:bf[n + 1], bf[q]:bf[q + 1]] = 0.0
    A = o.dense()
    ra, rb, same = pair_rows(h.tree(), h.n)
    K = o.exchange(Dx)
    assert relerr(K, dense_exchange(A, Ds, ra, rb, same)) < 1e-13
    assert relerr(K, dense_exchange(A, Dx, ra, rb, same)) > 1e-2
