"""twochnrg_amd -- MI355X (gfx950) NRG iterative-diagonalisation hot path.

Python face of the C-ABI in include/twochnrg_amd.h (libtwochnrg_amd.so,
built in-tree by twochnrg_amd/build.py).  The HIP library is required: every
entry point raises if it is missing; there is no CPU fallback.
"""
import ctypes
import weakref
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtwochnrg_amd.so")

# every symbol declared in include/twochnrg_amd.h
EXPORTS = [
    "nrg_ctx_create", "nrg_ctx_destroy", "nrg_last_error", "nrg_device_count",
    "nrg_diag_blocks", "nrg_diag_blocks_dev", "nrg_truncate", "nrg_ecut",
    "nrg_rotate_ops", "nrg_run_onechqs", "nrg_run_twochqs", "nrg_set_timing", "nrg_get_stats", "nrg_reset_stats",
    "nrg_set_eig_impl", "nrg_set_plain", "nrg_set_red_streams", "nrg_set_two_stage", "nrg_set_graph", "nrg_graph_stats", "nrg_autotune_eig", "nrg_eig2_stages", "nrg_compact",
    "nrg_dm_create", "nrg_dm_destroy", "nrg_dm_info", "nrg_dm_timing", "nrg_dm_memory", "nrg_dm_rho", "nrg_dm_spec",
    "nrg_dm_spectrum", "nrg_run_onechqs_multi", "nrg_dm_spec_omega", "nrg_dm_norm", "nrg_dm_weights",
    "nrg_run_twochqsz", "nrg_onechqs_thermo", "nrg_onechqs_jvp", "nrg_dm_create_jvp", "nrg_dm_spec_jvp",
    "nrg_spectrum_plan", "nrg_spectrum_combine", "nrg_rotate_ops_noalloc",
    "nrg_dm_create_from", "nrg_save_arrays", "nrg_dm_create_from_files", "nrg_dm_save_rho",
]

KSTAT_NAMES = ["eig", "hn", "fbasis", "gemm", "tdb_symv"]

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_i64p = ctypes.POINTER(ctypes.c_int64)


class NrgStep(ctypes.Structure):
    _fields_ = [
        ("nshell", ctypes.c_int), ("nq", ctypes.c_int), ("nblocks", ctypes.c_int),
        ("qn", _dp), ("bbe", _ip), ("nstates", ctypes.c_int), ("en", _dp),
        ("kept", ctypes.POINTER(ctypes.c_ubyte)),
        ("emin", ctypes.c_double), ("ecut", ctypes.c_double), ("nkept", ctypes.c_int),
        ("hpacked", _dp), ("evec", _dp),
        ("thermo_T", ctypes.c_double * 4), ("suscep", ctypes.c_double * 4),
        ("entropy", ctypes.c_double * 4),
        ("T_avg", ctypes.c_double), ("nops_avg", ctypes.c_int), ("opavg", ctypes.c_double * 8),
        ("cut_nblocks", ctypes.c_int), ("cut_qn", _dp), ("cut_bbe", _ip), ("cut_bbe_bc", _ip),
    ]


STEP_CB = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.POINTER(NrgStep))


class NrgTangentStep(ctypes.Structure):
    _fields_ = [
        ("nshell", ctypes.c_int), ("ntan", ctypes.c_int), ("nstates", ctypes.c_int), ("den", _dp),
        ("dsuscep", (ctypes.c_double * 4) * 4), ("dentropy", (ctypes.c_double * 4) * 4),
        ("nops_avg", ctypes.c_int), ("dopavg", (ctypes.c_double * 8) * 4),
    ]


TAN_CB = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.POINTER(NrgStep), ctypes.POINTER(NrgTangentStep))
MULTI_CB = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(NrgStep))


class OneChQSParams(ctypes.Structure):
    _fields_ = [
        ("U", ctypes.c_double), ("Gamma", ctypes.c_double), ("ed", ctypes.c_double),
        ("Lambda", ctypes.c_double), ("nsitesmax", ctypes.c_int), ("ncut", ctypes.c_int),
        ("calcdens", ctypes.c_int), ("disc_scheme", ctypes.c_int), ("myeps", ctypes.c_double),
        ("betabar", ctypes.c_double), ("record_blocks", ctypes.c_int),
    ]


class TwoChQSParams(ctypes.Structure):
    _fields_ = [
        ("U", ctypes.c_double), ("Gamma1", ctypes.c_double), ("Gamma2", ctypes.c_double),
        ("ed", ctypes.c_double), ("Lambda", ctypes.c_double), ("nsitesmax", ctypes.c_int),
        ("ncut", ctypes.c_int), ("betabar", ctypes.c_double), ("record_blocks", ctypes.c_int),
    ]


class DmParams(ctypes.Structure):
    _fields_ = [
        ("U", ctypes.c_double), ("Gamma", ctypes.c_double), ("ed", ctypes.c_double),
        ("Lambda", ctypes.c_double), ("nsitesmax", ctypes.c_int), ("ncut", ctypes.c_int),
        ("disc_scheme", ctypes.c_int), ("myeps", ctypes.c_double), ("betabar_nrg", ctypes.c_double),
        ("Mtemp", ctypes.c_int), ("betabar", ctypes.c_double), ("twindow", ctypes.c_double),
        ("broadtemp", ctypes.c_double), ("bbroad", ctypes.c_double), ("Nw", ctypes.c_int),
        ("method", ctypes.c_int), ("update_bef_cut", ctypes.c_int),
        ("model", ctypes.c_int), ("Gamma2", ctypes.c_double), ("fdm_site_dim", ctypes.c_double),
        ("fdm_fix", ctypes.c_int),
    ]


# bits of nrg_dmnrg_params.fdm_fix (NRG_FDM_FIX_* of the C header)
FDM_FIX_RHOMAX, FDM_FIX_DN, FDM_FIX_LAST, FDM_FIX_PAIRS, FDM_FIX_ONESIDED, FDM_FIX_NORM = 1, 2, 4, 8, 16, 32
FDM_EXACT = 63


class DmShellView(ctypes.Structure):
    """nrg_dm_shell: one shell's Acut_N, Abasis_N and operator arrays (host pointers)."""
    _fields_ = [
        ("nblocks", ctypes.c_int), ("nq", ctypes.c_int), ("nstates", ctypes.c_int),
        ("qn", _dp), ("bbe", _ip), ("bbe_bc", _ip), ("bbe_ev", _ip), ("en", _dp),
        ("kept", ctypes.POINTER(ctypes.c_ubyte)), ("evec", _dp), ("evec_len", ctypes.c_int64),
        ("basis_nblocks", ctypes.c_int), ("basis_nstates", ctypes.c_int), ("basis_qn", _dp),
        ("basis_bbe", _ip), ("basis_itype", _ip), ("basis_stcf", _ip),
        ("op_nblocks", ctypes.c_int), ("op_map", _ip), ("op_mbe", _ip), ("op", _dp), ("op_len", ctypes.c_int64),
    ]


class PairDesc(ctypes.Structure):
    _fields_ = [("ni", ctypes.c_int), ("nj", ctypes.c_int), ("ki", ctypes.c_int), ("kj", ctypes.c_int),
                ("zi_off", ctypes.c_int64), ("zj_off", ctypes.c_int64),
                ("f_off", ctypes.c_int64), ("out_off", ctypes.c_int64)]


class NrgStats(ctypes.Structure):
    _fields_ = [("ms", ctypes.c_double * 5), ("launches", ctypes.c_long * 5),
                ("flops", ctypes.c_double * 5), ("bytes", ctypes.c_double * 5)]


_lib = None


class TuneEntry(ctypes.Structure):
    _fields_ = [("red_streams", ctypes.c_int), ("ts_min", ctypes.c_int), ("graph", ctypes.c_int),
                ("pad", ctypes.c_int), ("ms", ctypes.c_double)]


class NrgError(RuntimeError):
    pass


def lib():
    """Load libtwochnrg_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise NrgError("libtwochnrg_amd.so not built: run twochnrg_amd/build.py (or __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    L.nrg_last_error.restype = ctypes.c_char_p
    L.nrg_device_count.restype = ctypes.c_int
    L.nrg_ctx_create.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
    L.nrg_ctx_destroy.argtypes = [ctypes.c_void_p]
    L.nrg_ctx_destroy.restype = None
    L.nrg_diag_blocks.argtypes = [ctypes.c_void_p, ctypes.c_int, _ip, _i64p, _dp, ctypes.c_int, _dp, _dp, _ip]
    L.nrg_diag_blocks_dev.argtypes = [ctypes.c_void_p, ctypes.c_int, _ip, _i64p, ctypes.c_void_p, _i64p,
                                      ctypes.c_void_p, _ip]
    L.nrg_truncate.argtypes = [ctypes.c_int, _dp, ctypes.c_int, ctypes.c_double,
                               ctypes.POINTER(ctypes.c_uint8), _dp, _ip]
    L.nrg_ecut.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _dp]
    L.nrg_rotate_ops.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(PairDesc), _dp, ctypes.c_int64,
                                 _dp, ctypes.c_int64, _dp, ctypes.c_int64, ctypes.c_int]
    L.nrg_set_timing.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.nrg_set_eig_impl.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.nrg_set_plain.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.nrg_set_red_streams.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.nrg_set_two_stage.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.nrg_set_graph.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.nrg_autotune_eig.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_double,
                                   ctypes.c_int, ctypes.POINTER(TuneEntry), ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_int), ctypes.POINTER(TuneEntry)]
    L.nrg_graph_stats.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_long),
                                  ctypes.POINTER(ctypes.c_long)]
    L.nrg_compact.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                              ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int), ctypes.c_void_p,
                              ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]
    L.nrg_eig2_stages.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
    L.nrg_get_stats.argtypes = [ctypes.c_void_p, ctypes.POINTER(NrgStats)]
    L.nrg_reset_stats.argtypes = [ctypes.c_void_p]
    L.nrg_run_onechqs.argtypes = [ctypes.c_void_p, ctypes.POINTER(OneChQSParams), STEP_CB, ctypes.c_void_p, _dp]
    L.nrg_run_twochqs.argtypes = [ctypes.c_void_p, ctypes.POINTER(TwoChQSParams), STEP_CB, ctypes.c_void_p, _dp]
    L.nrg_run_twochqsz.argtypes = [ctypes.c_void_p, ctypes.POINTER(TwoChQSParams), STEP_CB, ctypes.c_void_p, _dp]
    L.nrg_onechqs_thermo.argtypes = [ctypes.c_void_p, ctypes.POINTER(OneChQSParams), _dp, ctypes.c_int, _ip]
    L.nrg_onechqs_jvp.argtypes = [ctypes.c_void_p, ctypes.POINTER(OneChQSParams), ctypes.POINTER(ctypes.c_ubyte),
                                  _i64p, ctypes.c_int, _dp, TAN_CB, ctypes.c_void_p]
    L.nrg_run_onechqs_multi.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(OneChQSParams), MULTI_CB,
                                        ctypes.c_void_p, _dp]
    L.nrg_dm_create.argtypes = [ctypes.c_void_p, ctypes.POINTER(DmParams), ctypes.POINTER(ctypes.c_void_p)]
    L.nrg_dm_destroy.argtypes = [ctypes.c_void_p]
    L.nrg_dm_create_jvp.argtypes = [ctypes.c_void_p, ctypes.POINTER(DmParams), ctypes.POINTER(ctypes.c_ubyte), _i64p,
                                    ctypes.c_int, _dp, ctypes.POINTER(ctypes.c_void_p)]
    L.nrg_dm_spec_jvp.argtypes = [ctypes.c_void_p, ctypes.c_int, _ip, _dp, _dp]
    L.nrg_dm_destroy.restype = None
    L.nrg_dm_info.argtypes = [ctypes.c_void_p, _ip, _ip, _dp]
    L.nrg_dm_timing.argtypes = [ctypes.c_void_p, _dp]
    L.nrg_dm_memory.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)]
    L.nrg_dm_rho.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, ctypes.c_int64, _ip, _dp]
    L.nrg_dm_spec.argtypes = [ctypes.c_void_p, ctypes.c_int, _ip, _dp, _dp]
    L.nrg_dm_spectrum.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int, _ip]
    L.nrg_dm_spec_omega.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp]
    L.nrg_dm_norm.argtypes = [ctypes.c_void_p, _dp]
    L.nrg_dm_weights.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int, _dp]
    L.nrg_spectrum_plan.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _ip,
                                    _ip, _dp, _dp, _ip]
    L.nrg_spectrum_combine.argtypes = [ctypes.c_int, _dp, _ip, _dp, ctypes.c_int, _dp, _dp]
    L.nrg_dm_create_from.argtypes = [ctypes.c_void_p, ctypes.POINTER(DmParams), ctypes.c_int,
                                     ctypes.POINTER(DmShellView), ctypes.POINTER(ctypes.c_void_p)]
    L.nrg_save_arrays.argtypes = [ctypes.c_void_p, ctypes.POINTER(DmParams), ctypes.c_char_p, ctypes.c_char_p, _ip]
    L.nrg_dm_create_from_files.argtypes = [ctypes.c_void_p, ctypes.POINTER(DmParams), ctypes.c_char_p,
                                           ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
    L.nrg_dm_save_rho.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_char_p]
    L.nrg_rotate_ops_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(PairDesc), ctypes.c_void_p,
                                         ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                         ctypes.c_int64, ctypes.c_int]
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise NrgError("nrg error %d: %s" % (rc, lib().nrg_last_error().decode()))


def _ptr(a, t=_dp):
    return a.ctypes.data_as(t)


def device_count():
    return lib().nrg_device_count()


def spectrum_plan(Lambda, nshell, windowed, Nw):
    """Evaluation points of the omega-grid drivers (nrg_spectrum_plan):
    (shell, omegabar, omega, parity) arrays in the reference's recording order."""
    n = ctypes.c_int()
    _check(lib().nrg_spectrum_plan(Lambda, nshell, int(windowed), Nw, 0, ctypes.byref(n), None, None, None, None))
    sh, par = np.zeros(n.value, np.int32), np.zeros(n.value, np.int32)
    wb, w = np.zeros(n.value), np.zeros(n.value)
    _check(lib().nrg_spectrum_plan(Lambda, nshell, int(windowed), Nw, n.value, ctypes.byref(n), _ptr(sh, _ip),
                                   _ptr(wb), _ptr(w), _ptr(par, _ip)))
    return sh, wb, w, par


def spectrum_combine(omega, parity, value, Nw):
    """The spectrum from the values at the plan's points (nrg_spectrum_combine)."""
    w = np.ascontiguousarray(omega, np.float64)
    par = np.ascontiguousarray(parity, np.int32)
    v = np.ascontiguousarray(value, np.float64)
    wo, ro = np.zeros(len(w)), np.zeros(len(w))
    _check(lib().nrg_spectrum_combine(len(w), _ptr(w), _ptr(par, _ip), _ptr(v), Nw, _ptr(wo), _ptr(ro)))
    return wo, ro


class Context:
    """One HIP stream + device workspaces (nrg_ctx)."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        _check(lib().nrg_ctx_create(device, ctypes.byref(h)))
        self.h = h
        self._dms = weakref.WeakSet()   # DM-NRG handles made on this context

    def close(self):
        if self.h:
            # the handles' buffers live in the context: release them first
            for d in list(self._dms):
                d.close()
            lib().nrg_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_eig_impl(self, impl):
        """1: one workgroup per block (Householder + QL); 2: tridiagonalisation
        + divide and conquer + blocked back-transformation (default)."""
        _check(lib().nrg_set_eig_impl(self.h, int(impl)))

    def set_plain(self, on=True):
        """One block (pair) per launch: the unbatched baseline."""
        _check(lib().nrg_set_plain(self.h, int(on)))

    def set_red_streams(self, n):
        """Streams (block groups) of the blocked Householder reduction, 1..4
        (default 3); the results do not depend on it."""
        _check(lib().nrg_set_red_streams(self.h, int(n)))

    def set_two_stage(self, min_n):
        """Blocks with n >= min_n use the two-stage reduction (dense -> band ->
        tridiagonal); 0 turns it off."""
        _check(lib().nrg_set_two_stage(self.h, int(min_n)))

    def set_graph(self, on=True):
        """hipGraph mode: each eigensolve's launch sequences captured and launched as a graph."""
        _check(lib().nrg_set_graph(self.h, int(on)))

    def graph_stats(self):
        """(host ms capturing + instantiating + launching, nodes, graphs) since the last call."""
        ms, nodes, graphs = ctypes.c_double(), ctypes.c_long(), ctypes.c_long()
        _check(lib().nrg_graph_stats(self.h, ctypes.byref(ms), ctypes.byref(nodes), ctypes.byref(graphs)))
        return ms.value, nodes.value, graphs.value

    def autotune(self, sizes, krow_frac=0.07, reps=1):
        """Time the eigensolver's launch configurations on random blocks of these
        sizes (nrg_autotune_eig), apply the choice, return (chosen, table) with
        dicts {red_streams, ts_min, graph, ms}."""
        n = np.ascontiguousarray(sizes, dtype=np.int32)
        tab = (TuneEntry * 16)()
        ne, ch = ctypes.c_int(), TuneEntry()
        _check(lib().nrg_autotune_eig(self.h, len(n), n.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                      float(krow_frac), int(reps), tab, 16, ctypes.byref(ne), ctypes.byref(ch)))
        rows = [dict(red_streams=e.red_streams, ts_min=e.ts_min, graph=e.graph, ms=e.ms) for e in tab[:ne.value]]
        return dict(red_streams=ch.red_streams, ts_min=ch.ts_min, graph=ch.graph, ms=ch.ms), rows

    def compact(self, n, z_off, keep, dZ=None, dZcut=None):
        """CutStates' eigenvector compaction on device arrays (nrg_compact):
        the first keep[b] rows of each n[b] x n[b] block at z_off[b] of dZ, in
        block order, to dZcut (device addresses as ints; dZcut None only sizes
        the result).  Returns (offset of each block in dZcut or -1, length)."""
        n = np.ascontiguousarray(n, dtype=np.int32)
        z_off = np.ascontiguousarray(z_off, dtype=np.int64)
        keep = np.ascontiguousarray(keep, dtype=np.int32)
        nb = len(n)
        off = np.zeros(max(nb, 1), dtype=np.int64)
        tot = ctypes.c_int64(0)
        _check(lib().nrg_compact(self.h, nb, n.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                 z_off.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                 keep.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                 ctypes.c_void_p(dZ) if dZ else None, ctypes.c_void_p(dZcut) if dZcut else None,
                                 off.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), ctypes.byref(tot)))
        return off[:nb], int(tot.value)

    def set_timing(self, on=True):
        _check(lib().nrg_set_timing(self.h, int(on)))

    def reset_stats(self):
        _check(lib().nrg_reset_stats(self.h))

    def stats(self):
        st = NrgStats()
        _check(lib().nrg_get_stats(self.h, ctypes.byref(st)))
        return {k: {"ms": st.ms[i], "launches": st.launches[i], "flops": st.flops[i], "bytes": st.bytes[i]}
                for i, k in enumerate(KSTAT_NAMES)}

    def eig2_stages(self, a):
        """Stage outputs of the batched solver for one matrix (diagnostics)."""
        a = np.ascontiguousarray(a, np.float64)
        n = a.shape[0]
        d, e, dd, w = (np.zeros(n) for _ in range(4))
        qv, z = np.zeros(n * n), np.zeros(n * n)
        _check(lib().nrg_eig2_stages(self.h, n, _ptr(a), _ptr(d), _ptr(e), _ptr(dd), _ptr(qv), _ptr(w), _ptr(z)))
        return dict(d=d, e=e, dd=dd, qv=qv.reshape(n, n), w=w, z=z.reshape(n, n))

    # DiagBlock replacement -------------------------------------------------
    def diag_blocks(self, blocks, packed=False):
        """blocks: list of (n x n) symmetric arrays, or of packed upper vectors
        with packed=True.  Returns (list of W, list of Z with row i = vector i, info)."""
        ns = []
        for b in blocks:
            b = np.asarray(b)
            if packed:
                n = int(round((np.sqrt(8 * b.size + 1) - 1) / 2))
            else:
                n = b.shape[0]
            ns.append(n)
        n_arr = np.array(ns, dtype=np.int32)
        flat = np.concatenate([np.ascontiguousarray(b, dtype=np.float64).ravel() for b in blocks]) \
            if blocks else np.zeros(1)
        lens = [np.asarray(b).size for b in blocks]
        h_off = np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.int64)
        W = np.zeros(max(1, int(n_arr.sum())))
        Z = np.zeros(max(1, int((n_arr.astype(np.int64) ** 2).sum())))
        info = np.zeros(max(1, len(ns)), dtype=np.int32)
        _check(lib().nrg_diag_blocks(self.h, len(ns), _ptr(n_arr, _ip), _ptr(h_off, _i64p), _ptr(flat),
                                     int(packed), _ptr(W), _ptr(Z), _ptr(info, _ip)))
        outW, outZ = [], []
        wo = zo = 0
        for n in ns:
            outW.append(W[wo:wo + n].copy())
            outZ.append(Z[zo:zo + n * n].reshape(n, n).copy())
            wo += n
            zo += n * n
        return outW, outZ, info[: len(ns)].copy()

    def rotate(self, problems):
        """problems: list of (Zi (ki x ni), F (ni x nj), Zj (kj x nj)).
        Returns list of Zi F Zj^T (FP64 on the device)."""
        zs, fs, descs = [], [], []
        zoff = foff = ooff = 0
        for zi, f, zj in problems:
            zi = np.ascontiguousarray(zi, np.float64)
            zj = np.ascontiguousarray(zj, np.float64)
            f = np.ascontiguousarray(f, np.float64)
            ki, ni = zi.shape
            kj, nj = zj.shape
            d = PairDesc(ni, nj, ki, kj, zoff, zoff + zi.size, foff, ooff)
            zs += [zi.ravel(), zj.ravel()]
            fs.append(f.ravel())
            zoff += zi.size + zj.size
            foff += f.size
            ooff += ki * kj
            descs.append(d)
        Z = np.concatenate(zs) if zs else np.zeros(1)
        F = np.concatenate(fs) if fs else np.zeros(1)
        out = np.zeros(max(1, ooff))
        arr = (PairDesc * max(1, len(descs)))(*descs)
        _check(lib().nrg_rotate_ops(self.h, len(descs), arr, _ptr(Z), Z.size, _ptr(F), F.size, _ptr(out),
                                    out.size, 0))
        res = []
        o = 0
        for d in descs:
            res.append(out[o:o + d.ki * d.kj].reshape(d.ki, d.kj).copy())
            o += d.ki * d.kj
        return res

    # NRG_main loop ---------------------------------------------------------
    def run_onechqs(self, U=0.5, Gamma=0.0282691, ed=-0.25, Lambda=2.5, nsitesmax=4, ncut=1000,
                    calcdens=0, disc_scheme=0, myeps=1e-18, betabar=0.727, record_blocks=False,
                    keep_steps=True, on_step=None):
        steps = []

        def cb(_u, st):
            if on_step is not None:
                on_step(st)
            if keep_steps:
                steps.append(step_to_dict(st))

        p = OneChQSParams(U, Gamma, ed, Lambda, nsitesmax, ncut, calcdens, disc_scheme, myeps, betabar,
                          int(record_blocks))
        fn = STEP_CB(cb)
        ms = ctypes.c_double(0.0)
        _check(lib().nrg_run_onechqs(self.h, ctypes.byref(p), fn, None, ctypes.byref(ms)))
        return steps, ms.value


    def onechqs_jvp(self, U=0.5, Gamma=0.02, ed=-0.25, Lambda=2.5, nsitesmax=4, ncut=1000, calcdens=2,
                    disc_scheme=0, myeps=1e-18, betabar=0.727, kept=None, dpar=None):
        """Derivative pass of the fixed-truncation loop (nrg_onechqs_jvp).  kept =
        (flags, offsets) as from kept_arrays() of a plain run, or None; dpar rows
        are tangent directions (dU, dGamma, ded).  Returns (steps, tangents)."""
        steps, tans = [], []

        def cb(_u, st, dt):
            steps.append(step_to_dict(st))
            tans.append(tangent_to_dict(dt))

        p = OneChQSParams(U, Gamma, ed, Lambda, nsitesmax, ncut, calcdens, disc_scheme, myeps, betabar, 0)
        dp_ = np.ascontiguousarray(np.zeros((0, 3)) if dpar is None else np.asarray(dpar, np.float64).reshape(-1, 3))
        if kept is not None:
            kf = np.ascontiguousarray(kept[0], dtype=np.uint8)
            ko = np.ascontiguousarray(kept[1], dtype=np.int64)
            if len(ko) < nsitesmax + 1:
                raise ValueError("kept offsets need nsitesmax+1 entries")
            kfp, kop = kf.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)), _ptr(ko, _i64p)
        else:
            kfp = kop = None
        fn = TAN_CB(cb)
        _check(lib().nrg_onechqs_jvp(self.h, ctypes.byref(p), kfp, kop, dp_.shape[0],
                                     _ptr(dp_) if dp_.size else None, fn, None))
        return steps, tans

    def onechqs_thermo(self, U=0.5, Gamma=0.02, ed=-0.25, Lambda=2.5, nsitesmax=40, ncut=400, disc_scheme=0,
                       myeps=1e-18, betabar=0.727):
        """Impurity thermodynamics with the chain subtraction (nrg_onechqs_thermo):
        rows (4 per step) of T, Tchi_tot, Tchi_chain, Tchi_imp, S_tot, S_chain, S_imp."""
        p = OneChQSParams(U, Gamma, ed, Lambda, nsitesmax, ncut, 2, disc_scheme, myeps, betabar, 0)
        n = ctypes.c_int()
        _check(lib().nrg_onechqs_thermo(self.h, ctypes.byref(p), None, 0, ctypes.byref(n)))
        out = np.zeros((n.value, 7))
        _check(lib().nrg_onechqs_thermo(self.h, ctypes.byref(p), _ptr(out), n.value, ctypes.byref(n)))
        return out

    def run_onechqs_multi(self, instances, nsitesmax, calcdens=0, disc_scheme=0, myeps=1e-18, betabar=0.727,
                          keep_steps=True, on_step=None):
        """Lockstep run of many OneChQS instances (nrg_run_onechqs_multi).
        instances: list of dicts with U, Gamma, ed, Lambda, ncut."""
        n = len(instances)
        arr = (OneChQSParams * n)()
        for i, d in enumerate(instances):
            arr[i] = OneChQSParams(d["U"], d["Gamma"], d["ed"], d.get("Lambda", 2.5), nsitesmax, d["ncut"],
                                   calcdens, disc_scheme, myeps, betabar, 0)
        steps = [[] for _ in range(n)]

        def cb(_u, inst, st):
            if on_step is not None:
                on_step(inst, st)
            if keep_steps:
                steps[inst].append(step_to_dict(st))

        fn = MULTI_CB(cb)
        ms = ctypes.c_double(0.0)
        _check(lib().nrg_run_onechqs_multi(self.h, n, arr, fn, None, ctypes.byref(ms)))
        return steps, ms.value

    def dm(self, U=0.5, Gamma=0.02, ed=-0.25, Lambda=2.5, nsitesmax=40, ncut=400, Mtemp=38, betabar=1.0,
           disc_scheme=0, twindow=0.0, broadtemp=0.0, bbroad=0.0, Nw=0, myeps=0.0, betabar_nrg=0.0,
           method=0, update_bef_cut=0, kept=None, dpar=None, model=0, Gamma2=0.0,
           fdm_site_dim=0.0, fdm_fix=0):
        """DM-NRG handle (nrg_dm_create): forward chain + rho sweep on the device.
        method 0: DM-NRG, 1: CFS (DM_NRG -C), 2: FDM (DM_NRG -F); CFS/FDM want
        update_bef_cut=1 (the chain keeps the discarded states per shell).
        kept = (flags, offsets) from kept_arrays of a calcdens-1 run freezes every
        step's Kept set; dpar rows are tangent directions (dU, dGamma, ded)
        (nrg_dm_create_jvp; DM-NRG method only).  model=1 runs the TwoChQSz
        two-channel chain (Gamma, Gamma2 the two hybridisations; DM-NRG, CFS or
        FDM over shells that keep every state) and the spectral function of d_up.  fdm_site_dim: the FDM environment count
        per later site (0: 4 for model 0 as the reference, 16 for model 1).
        fdm_fix: FDM_FIX_* bits switching the reference's FDM quirks off
        (FDM_EXACT: the full density matrix, whose CalcNorm is 1)."""
        return DMNRG(self, DmParams(U, Gamma, ed, Lambda, nsitesmax, ncut, disc_scheme, myeps, betabar_nrg,
                                    Mtemp, betabar, twindow, broadtemp, bbroad, Nw, method, update_bef_cut,
                                    model, Gamma2, fdm_site_dim, fdm_fix),
                     kept=kept, dpar=dpar)

    @staticmethod
    def _dm_params(U=0.5, Gamma=0.02, ed=-0.25, Lambda=2.5, nsitesmax=40, ncut=400, Mtemp=38, betabar=1.0,
                   disc_scheme=0, twindow=0.0, broadtemp=0.0, bbroad=0.0, Nw=0, myeps=0.0, betabar_nrg=0.0,
                   method=0, update_bef_cut=0, model=0, Gamma2=0.0, fdm_site_dim=0.0, fdm_fix=0):
        return DmParams(U, Gamma, ed, Lambda, nsitesmax, ncut, disc_scheme, myeps, betabar_nrg, Mtemp, betabar,
                        twindow, broadtemp, bbroad, Nw, method, update_bef_cut, model, Gamma2, fdm_site_dim, fdm_fix)

    def save_arrays(self, directory, ident, **kw):
        """nrg_save_arrays: run the chain of kw and write SaveArrays' Abasis_/Acut_/Mat0_
        .bin files of every shell into directory; returns the number of shells."""
        n = ctypes.c_int()
        p = self._dm_params(**kw)
        _check(lib().nrg_save_arrays(self.h, ctypes.byref(p), str(directory).encode(), ident.encode(),
                                     ctypes.byref(n)))
        return n.value

    def dm_from_files(self, directory, ident, nshells, **kw):
        """nrg_dm_create_from_files: DM_NRG over the .bin files of shells 0 .. nshells-1."""
        p = self._dm_params(**kw)
        h = ctypes.c_void_p()
        _check(lib().nrg_dm_create_from_files(self.h, ctypes.byref(p), str(directory).encode(), ident.encode(),
                                              nshells, ctypes.byref(h)))
        return DMNRG(self, p, handle=h)

    def dm_from_shells(self, shells, **kw):
        """nrg_dm_create_from: shells = list of dicts with the nrg_dm_shell arrays
        (qn, bbe, bbe_bc, bbe_ev, en, kept, evec, basis_qn, basis_bbe, basis_itype,
        basis_stcf, op_map, op_mbe, op)."""
        keep, views = [], []
        for sh in shells:
            a = {k: np.ascontiguousarray(v, dtype=(np.float64 if k in ("qn", "en", "evec", "basis_qn", "op")
                                                   else np.uint8 if k == "kept" else np.int32))
                 for k, v in sh.items() if v is not None}
            keep.append(a)
            v = DmShellView()
            v.nblocks, v.nq, v.nstates = len(a["qn"]) // 2, 2, len(a["en"])
            v.qn, v.en, v.evec = _ptr(a["qn"]), _ptr(a["en"]), _ptr(a["evec"])
            v.bbe, v.bbe_bc, v.bbe_ev = _ptr(a["bbe"], _ip), _ptr(a["bbe_bc"], _ip), _ptr(a["bbe_ev"], _ip)
            v.kept = a["kept"].ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)) if "kept" in a else None
            v.evec_len = len(a["evec"])
            v.basis_nblocks, v.basis_nstates = len(a["basis_qn"]) // 2, len(a["basis_stcf"])
            v.basis_qn, v.basis_bbe = _ptr(a["basis_qn"]), _ptr(a["basis_bbe"], _ip)
            v.basis_itype, v.basis_stcf = _ptr(a["basis_itype"], _ip), _ptr(a["basis_stcf"], _ip)
            v.op_nblocks = len(a["op_map"]) // 2
            v.op_map, v.op_mbe, v.op = _ptr(a["op_map"], _ip), _ptr(a["op_mbe"], _ip), _ptr(a["op"])
            v.op_len = len(a["op"])
            views.append(v)
        arr = (DmShellView * len(views))(*views)
        p = self._dm_params(**kw)
        h = ctypes.c_void_p()
        _check(lib().nrg_dm_create_from(self.h, ctypes.byref(p), len(views), arr, ctypes.byref(h)))
        return DMNRG(self, p, handle=h)

    def run_twochqs(self, U=0.4, Gamma1=0.03, Gamma2=0.03, ed=-0.2, Lambda=3.0, nsitesmax=3, ncut=3000,
                    betabar=0.727, record_blocks=False, keep_steps=True, on_step=None, qsz=False):
        """TwoChQS (SU(2), multiplets) or, with qsz=True, TwoChQSz ((Q, Sz) blocks, states)."""
        steps = []

        def cb(_u, st):
            if on_step is not None:
                on_step(st)
            if keep_steps:
                steps.append(step_to_dict(st))

        p = TwoChQSParams(U, Gamma1, Gamma2, ed, Lambda, nsitesmax, ncut, betabar, int(record_blocks))
        fn = STEP_CB(cb)
        ms = ctypes.c_double(0.0)
        run = lib().nrg_run_twochqsz if qsz else lib().nrg_run_twochqs
        _check(run(self.h, ctypes.byref(p), fn, None, ctypes.byref(ms)))
        return steps, ms.value

    def run_twochqsz(self, U=0.5, Gamma1=0.03, Gamma2=0.01, ed=-0.25, Lambda=2.5, nsitesmax=3, ncut=4000,
                     betabar=0.727, record_blocks=False, keep_steps=True, on_step=None):
        return self.run_twochqs(U, Gamma1, Gamma2, ed, Lambda, nsitesmax, ncut, betabar, record_blocks, keep_steps,
                                on_step, qsz=True)


def truncate(en, ncut, myeps=1e-18):
    """SetDegen + Ecut + SetKept (host part of the library; no device needed)."""
    e = np.ascontiguousarray(en, np.float64).copy()
    kept = np.zeros(len(e), np.uint8)
    ec = ctypes.c_double()
    nk = ctypes.c_int()
    _check(lib().nrg_truncate(len(e), _ptr(e), ncut, myeps, kept.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                              ctypes.byref(ec), ctypes.byref(nk)))
    return e, kept.astype(bool), ec.value, nk.value


def ecut(en, ncut):
    e = np.ascontiguousarray(en, np.float64)
    ec = ctypes.c_double()
    _check(lib().nrg_ecut(len(e), _ptr(e), ncut, ctypes.byref(ec)))
    return ec.value


def step_to_dict(st):
    s = st.contents
    nb, nq, ns = s.nblocks, s.nq, s.nstates
    d = {
        "N": s.nshell, "nq": nq, "nblocks": nb,
        "qn": np.ctypeslib.as_array(s.qn, (nb * nq,)).reshape(nb, nq).copy(),
        "bbe": np.ctypeslib.as_array(s.bbe, (2 * nb,)).reshape(nb, 2).copy(),
        "en": np.ctypeslib.as_array(s.en, (ns,)).copy(),
        "kept": np.ctypeslib.as_array(s.kept, (ns,)).astype(bool).copy(),
        "emin": s.emin, "ecut": s.ecut, "nkept": s.nkept,
        "thermo_T": np.array(s.thermo_T[:]), "suscep": np.array(s.suscep[:]),
        "entropy": np.array(s.entropy[:]), "T_avg": s.T_avg,
        "opavg": np.array(s.opavg[: s.nops_avg]), "cut_nblocks": s.cut_nblocks,
    }
    cb = s.cut_nblocks
    d["cut_qn"] = np.ctypeslib.as_array(s.cut_qn, (cb * nq,)).reshape(cb, nq).copy() if cb else np.zeros((0, nq))
    d["cut_bbe"] = np.ctypeslib.as_array(s.cut_bbe, (2 * cb,)).reshape(cb, 2).copy() if cb else np.zeros((0, 2), int)
    d["cut_bbe_bc"] = (np.ctypeslib.as_array(s.cut_bbe_bc, (2 * cb,)).reshape(cb, 2).copy()
                       if cb else np.zeros((0, 2), int))
    sizes = d["bbe"][:, 1] - d["bbe"][:, 0] + 1
    if bool(s.hpacked):
        d["hpacked"] = np.ctypeslib.as_array(s.hpacked, (int(np.sum(sizes * (sizes + 1) // 2)),)).copy()
    if bool(s.evec):
        d["evec"] = np.ctypeslib.as_array(s.evec, (int(np.sum(sizes * sizes)),)).copy()
    return d


def tangent_to_dict(dt):
    d = dt.contents
    nt, ns = d.ntan, d.nstates
    return {"N": d.nshell,
            "den": (np.ctypeslib.as_array(d.den, (nt * ns,)).reshape(nt, ns).copy() if nt else np.zeros((0, ns))),
            "dsuscep": np.array([[d.dsuscep[t][i] for i in range(4)] for t in range(nt)]).reshape(nt, 4),
            "dentropy": np.array([[d.dentropy[t][i] for i in range(4)] for t in range(nt)]).reshape(nt, 4),
            "dopavg": np.array([[d.dopavg[t][i] for i in range(d.nops_avg)] for t in range(nt)]).reshape(
                nt, d.nops_avg)}


def kept_arrays(steps):
    """Concatenated Kept flags of the records N >= 0 and their offsets: the
    frozen-truncation input of onechqs_jvp."""
    ks = [s["kept"].astype(np.uint8) for s in steps if s["N"] >= 0]
    off = np.zeros(len(ks) + 1, dtype=np.int64)
    off[1:] = np.cumsum([len(k) for k in ks])
    return (np.concatenate(ks) if ks else np.zeros(0, np.uint8)), off


class DMNRG:
    """Handle of a device DM-NRG run (include/twochnrg_amd.h nrg_dm_*)."""

    def __init__(self, ctx, params, kept=None, dpar=None, handle=None):
        self.ctx = ctx
        self.h = None
        h = ctypes.c_void_p()
        if handle is not None:
            self.ntan = 0
            h = handle
        elif kept is None and dpar is None:
            self.ntan = 0
            _check(lib().nrg_dm_create(ctx.h, ctypes.byref(params), ctypes.byref(h)))
        else:
            dp_ = np.ascontiguousarray(np.zeros((0, 3)) if dpar is None else np.asarray(dpar, np.float64).reshape(-1, 3))
            self.ntan = dp_.shape[0]
            if kept is not None:
                kf = np.ascontiguousarray(kept[0], dtype=np.uint8)
                ko = np.ascontiguousarray(kept[1], dtype=np.int64)
                kfp, kop = kf.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)), _ptr(ko, _i64p)
            else:
                kfp = kop = None
            _check(lib().nrg_dm_create_jvp(ctx.h, ctypes.byref(params), kfp, kop, self.ntan,
                                           _ptr(dp_) if dp_.size else None, ctypes.byref(h)))
        self.h = h
        ctx._dms.add(self)
        a, b, t = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        _check(lib().nrg_dm_info(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(t)))
        self.nshell_max, self.spec_nshell_max, self.temp = a.value, b.value, t.value

    def close(self):
        if self.h:
            lib().nrg_dm_destroy(self.h)
            self.h = None
            self.ctx._dms.discard(self)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def save_rho(self, directory, ident):
        """nrg_dm_save_rho: rhoDM_<ident>_N<N>.bin of every shell."""
        _check(lib().nrg_dm_save_rho(self.h, str(directory).encode(), ident.encode()))

    def timing(self):
        ms = np.zeros(4)
        _check(lib().nrg_dm_timing(self.h, _ptr(ms)))
        return dict(forward=ms[0], sweep=ms[1], poles=ms[2], broaden=ms[3])

    def memory(self):
        """Device bytes held by the context: when the sweep starts (the peak), and now."""
        a, b = ctypes.c_uint64(), ctypes.c_uint64()
        _check(lib().nrg_dm_memory(self.h, ctypes.byref(a), ctypes.byref(b)))
        return dict(before_sweep=a.value, now=b.value)

    def rho(self, N, sizes):
        """rho_N blocks (list of k x k) for the given block sizes, and its trace."""
        tot = int(sum(k * k for k in sizes))
        out = np.zeros(max(tot, 1))
        nb, tr = ctypes.c_int(), ctypes.c_double()
        _check(lib().nrg_dm_rho(self.h, N, _ptr(out), out.size, ctypes.byref(nb), ctypes.byref(tr)))
        blocks, o = [], 0
        for k in sizes:
            blocks.append(out[o:o + k * k].reshape(k, k).copy())
            o += k * k
        return blocks, tr.value

    def spec(self, shells, omegabar):
        sh = np.ascontiguousarray(shells, np.int32)
        w = np.ascontiguousarray(omegabar, np.float64)
        out = np.zeros(len(sh))
        _check(lib().nrg_dm_spec(self.h, len(sh), _ptr(sh, _ip), _ptr(w), _ptr(out)))
        return out

    def spec_jvp(self, shells, omegabar):
        """d/dtheta of spec(): array (len(shells), ntan)."""
        sh = np.ascontiguousarray(shells, np.int32)
        w = np.ascontiguousarray(omegabar, np.float64)
        out = np.zeros(max(len(sh) * self.ntan, 1))
        _check(lib().nrg_dm_spec_jvp(self.h, len(sh), _ptr(sh, _ip), _ptr(w), _ptr(out)))
        return out[:len(sh) * self.ntan].reshape(len(sh), self.ntan)

    def spec_omega(self, omega):
        """CalcSpecDM_NRG(omega): the shell-window sum (CFS/FDM)."""
        w = np.ascontiguousarray(np.atleast_1d(omega), np.float64)
        out = np.zeros(len(w))
        _check(lib().nrg_dm_spec_omega(self.h, len(w), _ptr(w), _ptr(out)))
        return out

    def norm(self):
        """CalcNorm (CFS/FDM)."""
        v = ctypes.c_double()
        _check(lib().nrg_dm_norm(self.h, ctypes.byref(v)))
        return v.value

    def weights(self):
        """FDM w_N (N = 0 .. nshell_max) and Z_tot."""
        wn = np.zeros(self.nshell_max + 1)
        z = ctypes.c_double()
        _check(lib().nrg_dm_weights(self.h, _ptr(wn), len(wn), ctypes.byref(z)))
        return wn, z.value

    def spectrum(self):
        n = ctypes.c_int()
        _check(lib().nrg_dm_spectrum(self.h, None, None, 0, ctypes.byref(n)))
        w, r = np.zeros(n.value), np.zeros(n.value)
        _check(lib().nrg_dm_spectrum(self.h, _ptr(w), _ptr(r), n.value, ctypes.byref(n)))
        return w, r
