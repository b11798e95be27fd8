"""Python host mirror of the H2P-ERI C API, backed by libh2eri_amd.so.

The method names follow the reference's C entry points (H2ERI_init,
CMS_read_mol_file, H2ERI_process_shells, H2ERI_partition, H2ERI_build_H2,
H2ERI_matvec, H2ERI_build_Coulomb, H2ERI_build_exchange).  Every compute call
goes through the C-ABI of the HIP library; there is no Python or CPU
fallback, and a library error raises H2ERIError.
"""
import ctypes as C
import os
import tempfile

import numpy as np

from . import capi

__all__ = ["H2ERI", "Group", "H2ERIError", "capi"]


class H2ERIError(RuntimeError):
    pass


def _check(lib):
    code = lib.h2eri_last_error()
    if code:
        msg = lib.h2eri_last_error_msg().decode(errors="replace")
        lib.h2eri_clear_error()
        raise H2ERIError(f"h2eri error {code}: {msg}")


def _dptr(a):
    return a.ctypes.data_as(capi.c_dbl_p)


def _iarr(p, n):
    if not p or n <= 0:
        return np.zeros(0, dtype=np.int32)
    return np.ctypeslib.as_array(p, shape=(n,)).copy()


def _darr(p, n):
    if not p or n <= 0:
        return np.zeros(0, dtype=np.float64)
    return np.ctypeslib.as_array(p, shape=(n,)).copy()


class Group:
    """Several representations whose matvecs run together in one persistent
    launch (H2ERI_group_create / _matvec / _status / _destroy)."""

    def __init__(self, handles, x_ptrs, y_ptrs):
        self.lib = capi.load()
        n = len(handles)
        hs = (capi.H2ERI_p * n)(*[h.h for h in handles])
        xs = (C.c_void_p * n)(*x_ptrs)
        ys = (C.c_void_p * n)(*y_ptrs)
        self.handles = list(handles)   # keep the handles alive
        self.g = self.lib.H2ERI_group_create(n, hs, xs, ys)
        _check(self.lib)

    def matvec(self, stream=None):
        self.lib.H2ERI_group_matvec(self.g, C.c_void_p(stream or 0))
        _check(self.lib)

    def status(self):
        """Synchronises; raises if a dependency wait exceeded its bound."""
        r = self.lib.H2ERI_group_status(self.g)
        _check(self.lib)
        return r

    def set_mode(self, mode):
        """0: one grouped persistent launch; 1: each molecule's own pass in turn."""
        self.lib.H2ERI_group_set_mode(self.g, int(mode))
        _check(self.lib)

    def autotune(self, reps=5):
        """Times both modes and keeps the faster: {mode0_ms, mode1_ms, mode}."""
        out = np.zeros(3)
        self.lib.H2ERI_group_autotune(self.g, int(reps), _dptr(out))
        _check(self.lib)
        return dict(mode0_ms=float(out[0]), mode1_ms=float(out[1]), mode=int(out[2]))

    def close(self):
        if self.g:
            self.lib.H2ERI_group_destroy(self.g)
            self.g = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class H2ERI:
    """One H2 representation of the ERI tensor (handle of struct H2ERI)."""

    def __init__(self, scr_tol=1e-10, ext_tol=1e-10, qr_tol=1e-6):
        self.lib = capi.load()
        self.h = capi.H2ERI_p()
        self.lib.H2ERI_init(C.byref(self.h), scr_tol, ext_tol, qr_tol)
        _check(self.lib)

    # -- lifecycle --------------------------------------------------------
    def close(self):
        if self.h:
            self.lib.H2ERI_destroy(self.h)
            self.h = capi.H2ERI_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def s(self):
        return self.h.contents

    # -- construction (host) -------------------------------------------------
    def read_mol(self, path):
        s = self.s
        self.lib.CMS_read_mol_file(os.fsencode(path), C.byref(C.c_int.from_buffer(s, capi.H2ERIStruct.natom.offset)),
                                   C.byref(C.c_int.from_buffer(s, capi.H2ERIStruct.nshell.offset)),
                                   C.cast(C.addressof(s) + capi.H2ERIStruct.shells.offset,
                                          C.POINTER(C.POINTER(capi.Shell))))
        _check(self.lib)
        return self

    def read_mol_text(self, text):
        with tempfile.NamedTemporaryFile("w", suffix=".mol", delete=False) as f:
            f.write(text)
            path = f.name
        try:
            self.read_mol(path)
        finally:
            os.unlink(path)
        return self

    def process_shells(self):
        self.lib.H2ERI_process_shells(self.h)
        _check(self.lib)
        return self

    def partition(self):
        self.lib.H2ERI_partition(self.h)
        _check(self.lib)
        return self

    def build_h2(self, bd_jit=0):
        self.lib.H2ERI_build_H2(self.h, bd_jit)
        _check(self.lib)
        return self

    def build_from_mol(self, path=None, text=None):
        if text is not None:
            self.read_mol_text(text)
        else:
            self.read_mol(path)
        return self.process_shells().partition().build_h2()

    def partition_synthetic(self, coords, nbfp, max_leaf_points=300):
        coords = np.ascontiguousarray(coords, dtype=np.float64)
        assert coords.ndim == 2 and coords.shape[0] == 3
        nbfp = np.ascontiguousarray(nbfp, dtype=np.int32)
        n = coords.shape[1]
        assert nbfp.shape == (n,)
        self.lib.H2ERI_partition_points_synthetic(self.h, n, _dptr(coords), nbfp.ctypes.data_as(capi.c_int_p),
                                                  max_leaf_points)
        _check(self.lib)
        return self

    def build_synthetic(self, seed=1, max_rank=40, lr_frac=0.3, diag_lr_frac=0.0):
        self.lib.H2ERI_build_H2_synthetic(self.h, seed, max_rank, lr_frac, diag_lr_frac)
        _check(self.lib)
        return self

    def save(self, path):
        self.lib.H2ERI_save_rep(self.h, os.fsencode(path))
        _check(self.lib)

    @classmethod
    def load(cls, path):
        o = cls()
        o.lib.H2ERI_load_rep(o.h, os.fsencode(path))
        _check(o.lib)
        return o

    # -- device hot path -----------------------------------------------------
    @property
    def n(self):
        return self.s.krnl_mat_size

    @property
    def nbf(self):
        return self.s.num_bf

    def set_precision(self, bits):
        self.lib.H2ERI_set_precision(self.h, bits)
        _check(self.lib)

    def set_tile_bytes(self, nbytes):
        self.lib.H2ERI_set_tile_bytes(self.h, float(nbytes))
        _check(self.lib)

    def plan_check(self, bits=64):
        """Host-only build and bounds check of the device plan (no GPU needed)."""
        r = self.lib.H2ERI_plan_check(self.h, bits)
        _check(self.lib)
        return r

    def plan_emulate(self, x, bits=64):
        """Host execution of the device plan (tests only): (y, conflicts).
        x is one vector (N,) or a batch (nv, N) with nv in {1, 2, 4, 8, 16}."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        nv = 1 if x.ndim == 1 else x.shape[0]
        y = np.zeros_like(x)
        r = self.lib.H2ERI_plan_emulate(self.h, bits, nv, _dptr(x), _dptr(y))
        _check(self.lib)
        return y, r

    def jit_emulate(self, x):
        """Host execution of the on-the-fly (BD_JIT) matvec (tests only)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        self.lib.H2ERI_jit_emulate(self.h, _dptr(x), _dptr(y))
        _check(self.lib)
        return y

    def jit_stats(self):
        """Work of one on-the-fly matvec (host count)."""
        o = np.zeros(4)
        self.lib.H2ERI_jit_stats(self.h, _dptr(o))
        _check(self.lib)
        return {"tasks": int(o[0]), "quartets": o[1], "prim_quartets": o[2], "eri_values": o[3]}

    def set_block_compression(self, on):
        """B and D blocks compressed where it pays (1, default) or stored dense
        (0), from the next build on."""
        self.lib.H2ERI_set_block_compression(self.h, int(on))
        _check(self.lib)

    def set_schedule(self, mode):
        self.lib.H2ERI_set_schedule(self.h, mode)
        _check(self.lib)

    def set_wave_tile_bytes(self, nbytes):
        """Staged executor: matrix parts up to nbytes run as wave tiles (0: none)."""
        self.lib.H2ERI_set_wave_tile_bytes(self.h, float(nbytes))
        _check(self.lib)

    def set_near_field_grid(self, n):
        """Staged executor: at most n workgroups per near-field launch (0: one per task)."""
        self.lib.H2ERI_set_near_field_grid(self.h, int(n))
        _check(self.lib)

    def set_executor(self, mode):
        """0: one launch per stage; 1: one persistent launch per pass."""
        self.lib.H2ERI_set_executor(self.h, mode)
        _check(self.lib)

    def executor(self):
        return self.lib.H2ERI_get_executor(self.h)

    def device_status(self):
        """Synchronises; raises if a persistent-executor wait exceeded its bound."""
        r = self.lib.H2ERI_device_status(self.h)
        _check(self.lib)
        return r

    def direct_matvec_rows(self, pair_idx, x):
        """Exact rows (every integral evaluated directly) of the ERI matrix
        times x for the listed shell pairs; returns (rows, y_rows) with rows
        the indices into y of the values."""
        pair_idx = np.ascontiguousarray(pair_idx, dtype=np.int32)
        x = np.ascontiguousarray(x, dtype=np.float64)
        assert x.shape == (self.n,)
        sidx = self.tree()["sp_bfp_sidx"]
        rows = np.concatenate([np.arange(sidx[i], sidx[i + 1]) for i in pair_idx]) if len(pair_idx) else np.zeros(0, int)
        y = np.empty(len(rows))
        self.lib.H2ERI_direct_matvec_rows(self.h, len(pair_idx), pair_idx.ctypes.data_as(C.POINTER(C.c_int)),
                                          _dptr(x), _dptr(y))
        _check(self.lib)
        return rows, y

    def matvec_dev_noalloc(self, x_ptr, y_ptr, stream=None):
        """H2ERI_matvec_dev without any allocation (workspace in the
        library's scratch array); the handle must be prepared by upload()."""
        self.lib.H2ERI_matvec_dev_noalloc(self.h, C.c_void_p(x_ptr), C.c_void_p(y_ptr), C.c_void_p(stream or 0))
        _check(self.lib)

    def plan_stats(self, bits=64, nv=1):
        """Work figures of the device plan (host only)."""
        out = np.zeros(8)
        self.lib.H2ERI_plan_stats(self.h, bits, nv, _dptr(out))
        _check(self.lib)
        keys = ["tiles", "reduces", "stages", "nW", "mat_bytes", "ws_written", "ws_read", "out_elems"]
        return dict(zip(keys, out.tolist()))

    def set_spin_limit(self, polls):
        """Bound of the persistent executor's dependency waits, in polls."""
        self.lib.H2ERI_set_spin_limit(self.h, int(polls))
        _check(self.lib)

    def queue_stats(self, nv=1):
        out = np.zeros(4)
        self.lib.H2ERI_queue_stats(self.h, nv, _dptr(out))
        _check(self.lib)
        return dict(tasks=int(out[0]), static=int(out[1]), workgroups=int(out[2]), est_us=out[3])

    def queue_emulate(self, x, bits=64, workgroups=1024):
        """Host execution of the persistent executor's queue (tests only):
        (y, uncovered reads + output errors)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        nv = 1 if x.ndim == 1 else x.shape[0]
        y = np.zeros_like(x)
        r = self.lib.H2ERI_queue_emulate(self.h, bits, nv, workgroups, _dptr(x), _dptr(y))
        _check(self.lib)
        return y, r

    def upload(self):
        r = self.lib.H2ERI_upload(self.h)
        _check(self.lib)
        return r

    def matvec(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        assert x.shape == (self.n,)
        y = np.empty_like(x)
        self.lib.H2ERI_matvec(self.h, _dptr(x), _dptr(y))
        _check(self.lib)
        return y

    def matvec_multi(self, X):
        X = np.ascontiguousarray(X, dtype=np.float64)
        assert X.ndim == 2 and X.shape[1] == self.n
        Y = np.empty_like(X)
        self.lib.H2ERI_matvec_multi(self.h, X.shape[0], _dptr(X), _dptr(Y))
        _check(self.lib)
        return Y

    def matvec_dev(self, x_ptr, y_ptr, stream=None):
        self.lib.H2ERI_matvec_dev(self.h, C.c_void_p(x_ptr), C.c_void_p(y_ptr), C.c_void_p(stream or 0))
        _check(self.lib)

    def matvec_multi_dev(self, nvec, x_ptr, y_ptr, stream=None):
        self.lib.H2ERI_matvec_multi_dev(self.h, nvec, C.c_void_p(x_ptr), C.c_void_p(y_ptr), C.c_void_p(stream or 0))
        _check(self.lib)

    def build_coulomb(self, D):
        D = np.ascontiguousarray(D, dtype=np.float64)
        assert D.shape == (self.nbf, self.nbf)
        J = np.empty_like(D)
        self.lib.H2ERI_build_Coulomb(self.h, _dptr(D), _dptr(J))
        _check(self.lib)
        return J

    def build_coulomb_dev(self, d_ptr, j_ptr, stream=None):
        self.lib.H2ERI_build_Coulomb_dev(self.h, C.c_void_p(d_ptr), C.c_void_p(j_ptr), C.c_void_p(stream or 0))
        _check(self.lib)

    def build_coulomb_multi(self, Ds):
        Ds = np.ascontiguousarray(Ds, dtype=np.float64)
        assert Ds.ndim == 3 and Ds.shape[1:] == (self.nbf, self.nbf)
        Js = np.empty_like(Ds)
        self.lib.H2ERI_build_Coulomb_multi(self.h, Ds.shape[0], _dptr(Ds), _dptr(Js))
        _check(self.lib)
        return Js

    def build_coulomb_multi_dev(self, ndens, d_ptr, j_ptr, stream=None):
        self.lib.H2ERI_build_Coulomb_multi_dev(self.h, ndens, C.c_void_p(d_ptr), C.c_void_p(j_ptr),
                                               C.c_void_p(stream or 0))
        _check(self.lib)

    def graph_capture(self, kind, n, in_ptr, out_ptr):
        """Capture the device matvec (kind "matvec"), Coulomb build (kind
        "coulomb") or exchange build (kind "exchange", n = 1) on fixed device
        buffers into a HIP graph."""
        k = {"matvec": 0, "coulomb": 1, "exchange": 2}[kind]
        self.lib.H2ERI_graph_capture(self.h, k, n, C.c_void_p(in_ptr), C.c_void_p(out_ptr))
        _check(self.lib)

    def graph_replay(self, stream=None):
        self.lib.H2ERI_graph_replay(self.h, C.c_void_p(stream or 0))
        _check(self.lib)

    def build_exchange(self, D):
        D = np.ascontiguousarray(D, dtype=np.float64)
        assert D.shape == (self.nbf, self.nbf)
        K = np.empty_like(D)
        self.lib.H2ERI_build_exchange(self.h, _dptr(D), _dptr(K))
        _check(self.lib)
        return K

    def build_exchange_dev(self, d_ptr, k_ptr, stream=None):
        self.lib.H2ERI_build_exchange_dev(self.h, C.c_void_p(d_ptr), C.c_void_p(k_ptr), C.c_void_p(stream or 0))
        _check(self.lib)

    def exchange_dlist(self, D):
        """The density screening list (H2ERI_exchange_dlist): (dlist_sidx, dlist)."""
        D = np.ascontiguousarray(D, dtype=np.float64)
        assert D.shape == (self.nbf, self.nbf)
        self.lib.H2ERI_exchange_dlist(self.h, _dptr(D))
        _check(self.lib)
        ns = self.s.nshell
        sidx = _iarr(self.s.dlist_sidx, ns + 1)
        return sidx, _iarr(self.s.dlist, int(sidx[-1]))

    def exchange_workbuf_init(self):
        """Builds and uploads the exchange plan now (H2ERI_exchange_workbuf_init)."""
        self.lib.H2ERI_exchange_workbuf_init(self.h)
        _check(self.lib)

    def exchange_workbuf_free(self):
        """Releases the exchange plan's device memory (H2ERI_exchange_workbuf_free)."""
        self.lib.H2ERI_exchange_workbuf_free(self.h)
        _check(self.lib)

    def set_exchange_scratch(self, nbytes):
        """Bound of the exchange build's intermediate G (bytes); the terms run in chunks that fit it."""
        self.lib.H2ERI_set_exchange_scratch(self.h, float(nbytes))
        _check(self.lib)

    def exchange_stats(self):
        out = np.zeros(4)
        self.lib.H2ERI_exchange_stats(self.h, _dptr(out))
        _check(self.lib)
        return dict(zip(["gemm_flops", "terms", "resident_bytes", "scratch_doubles"], out.tolist()))

    def exchange_plan_check(self):
        """Host-only build and bounds check of the exchange plan (no GPU needed)."""
        out = np.zeros(4)
        n = self.lib.H2ERI_exchange_plan_check(self.h, _dptr(out))
        _check(self.lib)
        return dict(terms=n, **dict(zip(["gemm_flops", "terms_f", "resident_bytes", "scratch_doubles"], out.tolist())))

    def autotune(self, nv=1, reps=10):
        """Times executors x tile sizes x near-field schedules on the device,
        keeps the best."""
        out = np.zeros(6)
        self.lib.H2ERI_autotune(self.h, nv, reps, _dptr(out))
        _check(self.lib)
        return dict(ms=float(out[0]), tile_bytes=int(out[1]), schedule=int(out[2]), executor=int(out[3]),
                    wave_tile_bytes=int(out[4]), near_field_grid=int(out[5]))

    def set_phase_timing(self, on=True):
        self.lib.H2ERI_set_phase_timing(self.h, 1 if on else 0)

    def phase_ms(self):
        out = np.zeros(7)
        self.lib.H2ERI_get_phase_ms(self.h, _dptr(out))
        return dict(zip(["up", "B", "y1", "down", "D", "gather", "total"], out.tolist()))

    def launch_stats(self):
        """Per-launch records of the last timed call (phase timing on)."""
        n = self.lib.H2ERI_get_launch_stats(self.h, 0, None)
        out = np.zeros(5 * max(n, 1))
        n = self.lib.H2ERI_get_launch_stats(self.h, n, _dptr(out))
        names = ["up", "B", "y1", "down", "D", "gather", "pass"]
        return [dict(phase=names[int(out[5 * i])], start=out[5 * i + 1], blocks=int(out[5 * i + 2]),
                     mat_bytes=out[5 * i + 3], ms=out[5 * i + 4]) for i in range(n)]

    def matvec_bytes(self):
        out = np.zeros(6)
        self.lib.H2ERI_matvec_bytes(self.h, _dptr(out))
        _check(self.lib)
        return dict(zip(["alg", "ref_accounting", "U", "B", "D", "vec"], out.tolist()))

    def print_statistic(self):
        self.lib.H2ERI_print_statistic(self.h)

    # -- read-only views of the host representation ---------------------------
    def tree(self):
        s = self.s
        nn, L, nl = s.n_node, s.max_level + 1, s.n_leaf_node
        return dict(
            n_node=nn, n_leaf_node=nl, max_level=s.max_level, min_adm_level=s.min_adm_level,
            krnl_mat_size=s.krnl_mat_size, num_sp=s.num_sp, num_bf=s.num_bf, nshell=s.nshell,
            parent=_iarr(s.parent, nn), children=_iarr(s.children, 8 * nn).reshape(nn, 8),
            n_child=_iarr(s.n_child, nn), node_level=_iarr(s.node_level, nn),
            pt_cluster=_iarr(s.pt_cluster, 2 * nn).reshape(nn, 2),
            mat_cluster=_iarr(s.mat_cluster, 2 * nn).reshape(nn, 2),
            level_n_node=_iarr(s.level_n_node, L), level_nodes=_iarr(s.level_nodes, L * nl).reshape(L, nl),
            height_n_node=_iarr(s.height_n_node, L), height_nodes=_iarr(s.height_nodes, L * nl).reshape(L, nl),
            r_adm_pairs=_iarr(s.r_adm_pairs, 2 * s.n_r_adm_pair).reshape(-1, 2),
            r_inadm_pairs=_iarr(s.r_inadm_pairs, 2 * s.n_r_inadm_pair).reshape(-1, 2),
            sp_bfp_sidx=_iarr(s.sp_bfp_sidx, s.num_sp + 1),
            shell_bf_sidx=_iarr(s.shell_bf_sidx, s.nshell + 1) if s.shell_bf_sidx else None,
            sp_shell_idx=_iarr(s.sp_shell_idx, 2 * s.num_sp) if s.sp_shell_idx else None,
        )

    @staticmethod
    def _blk(p):
        if not p:
            return None
        b = p.contents
        if b.ld >= 0:
            return ("dense", _darr(b.data, b.nrow * b.ncol).reshape(b.nrow, b.ncol))
        k = -b.ld
        d = _darr(b.data, (b.nrow + b.ncol) * k)
        return ("lowrank", d[: b.nrow * k].reshape(b.nrow, k), d[b.nrow * k:].reshape(b.ncol, k))

    def blocks(self):
        s = self.s
        U = [self._blk(s.U[i]) for i in range(s.n_node)]
        B = [self._blk(s.c_B_blks[i]) for i in range(s.n_B)]
        D = [self._blk(s.c_D_blks[i]) for i in range(s.n_D)]
        return U, B, D
