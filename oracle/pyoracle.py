"""ctypes binding of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module.  The product package never does.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


class OrcStep(ctypes.Structure):
    _fields_ = [
        ("nshell", ctypes.c_int), ("nq", ctypes.c_int), ("nblocks", ctypes.c_int),
        ("qn", ctypes.POINTER(ctypes.c_double)), ("bbe", ctypes.POINTER(ctypes.c_int)),
        ("nstates", ctypes.c_int), ("en", ctypes.POINTER(ctypes.c_double)),
        ("kept", ctypes.POINTER(ctypes.c_ubyte)),
        ("emin", ctypes.c_double), ("ecut", ctypes.c_double), ("nkept", ctypes.c_int),
        ("hpacked", ctypes.POINTER(ctypes.c_double)), ("evec", ctypes.POINTER(ctypes.c_double)),
        ("thermo_T", ctypes.c_double * 4), ("suscep", ctypes.c_double * 4),
        ("entropy", ctypes.c_double * 4),
        ("T_avg", ctypes.c_double), ("nops_avg", ctypes.c_int), ("opavg", ctypes.c_double * 8),
        ("cut_nblocks", ctypes.c_int), ("cut_qn", ctypes.POINTER(ctypes.c_double)),
        ("cut_bbe", ctypes.POINTER(ctypes.c_int)), ("cut_bbe_bc", ctypes.POINTER(ctypes.c_int)),
    ]


STEP_CB = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.POINTER(OrcStep))


class OneChQSParams(ctypes.Structure):
    _fields_ = [
        ("U", ctypes.c_double), ("Gamma", ctypes.c_double), ("ed", ctypes.c_double),
        ("Lambda", ctypes.c_double), ("nsitesmax", ctypes.c_int), ("ncut", ctypes.c_int),
        ("calcdens", ctypes.c_int), ("disc_scheme", ctypes.c_int), ("myeps", ctypes.c_double),
        ("betabar", ctypes.c_double), ("record_blocks", ctypes.c_int),
    ]


class OrcTanStep(ctypes.Structure):
    _fields_ = [
        ("nshell", ctypes.c_int), ("ntan", ctypes.c_int), ("nstates", ctypes.c_int),
        ("den", ctypes.POINTER(ctypes.c_double)),
        ("dsuscep", (ctypes.c_double * 4) * 4), ("dentropy", (ctypes.c_double * 4) * 4),
        ("nops_avg", ctypes.c_int), ("dopavg", (ctypes.c_double * 8) * 4),
    ]


TAN_CB = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.POINTER(OrcStep), ctypes.POINTER(OrcTanStep))


class TwoChQSParams(ctypes.Structure):
    _fields_ = [
        ("U", ctypes.c_double), ("Gamma1", ctypes.c_double), ("Gamma2", ctypes.c_double),
        ("ed", ctypes.c_double), ("Lambda", ctypes.c_double), ("nsitesmax", ctypes.c_int),
        ("ncut", ctypes.c_int), ("betabar", ctypes.c_double), ("record_blocks", ctypes.c_int),
    ]


_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", HERE])
    lib = ctypes.CDLL(LIB_PATH)
    lib.orc_run_onechqs.argtypes = [ctypes.POINTER(OneChQSParams), STEP_CB, ctypes.c_void_p]
    lib.orc_run_onechqs.restype = ctypes.c_int
    lib.orc_run_twochqs.argtypes = [ctypes.POINTER(TwoChQSParams), STEP_CB, ctypes.c_void_p]
    lib.orc_run_twochqs.restype = ctypes.c_int
    lib.orc_run_twochqsz.argtypes = [ctypes.POINTER(TwoChQSParams), STEP_CB, ctypes.c_void_p]
    lib.orc_run_twochqsz.restype = ctypes.c_int
    lib.orc_run_onechqs_jvp.argtypes = [ctypes.POINTER(OneChQSParams), ctypes.POINTER(ctypes.c_ubyte),
                                        ctypes.POINTER(ctypes.c_longlong), ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_double), TAN_CB, ctypes.c_void_p]
    lib.orc_run_onechqs_jvp.restype = ctypes.c_int
    lib.orc_cgordan.argtypes = [ctypes.c_double] * 6
    lib.orc_cgordan.restype = ctypes.c_double
    dp = ctypes.POINTER(ctypes.c_double)
    lib.orc_eigh_packed.argtypes = [ctypes.c_int, dp, dp, dp]
    lib.orc_eigh_full.argtypes = [ctypes.c_int, dp, dp, dp]
    lib.orc_setdegen.argtypes = [ctypes.c_int, dp, ctypes.c_double]
    lib.orc_ecut.argtypes = [ctypes.c_int, dp, ctypes.c_int]
    lib.orc_ecut.restype = ctypes.c_double
    lib.orc_rotate.argtypes = [ctypes.c_int] * 4 + [dp, dp, dp, dp]
    lib.orc_chin.argtypes = [ctypes.c_int, ctypes.c_double]
    lib.orc_chin.restype = ctypes.c_double
    _lib = lib
    return lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def step_to_dict(st):
    """Copy one step record (shared layout of oracle and product) into numpy."""
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
        "opavg": np.array(s.opavg[: s.nops_avg]),
        "cut_nblocks": s.cut_nblocks,
    }
    cb = s.cut_nblocks
    d["cut_qn"] = np.ctypeslib.as_array(s.cut_qn, (cb * nq,)).reshape(cb, nq).copy() if cb else np.zeros((0, nq))
    d["cut_bbe"] = np.ctypeslib.as_array(s.cut_bbe, (2 * cb,)).reshape(cb, 2).copy() if cb else np.zeros((0, 2), int)
    d["cut_bbe_bc"] = (np.ctypeslib.as_array(s.cut_bbe_bc, (2 * cb,)).reshape(cb, 2).copy()
                       if cb else np.zeros((0, 2), int))
    sizes = d["bbe"][:, 1] - d["bbe"][:, 0] + 1
    if bool(s.hpacked):
        tot = int(np.sum(sizes * (sizes + 1) // 2))
        d["hpacked"] = np.ctypeslib.as_array(s.hpacked, (tot,)).copy()
    if bool(s.evec):
        tot = int(np.sum(sizes * sizes))
        d["evec"] = np.ctypeslib.as_array(s.evec, (tot,)).copy()
    return d


def run_onechqs(U=0.5, Gamma=0.0282691, ed=-0.25, Lambda=2.5, nsitesmax=4, ncut=1000,
                calcdens=0, disc_scheme=0, myeps=1e-18, betabar=0.727, record_blocks=False):
    lib = load()
    steps = []

    def cb(_user, st):
        steps.append(step_to_dict(st))

    p = OneChQSParams(U, Gamma, ed, Lambda, nsitesmax, ncut, calcdens, disc_scheme, myeps,
                      betabar, int(record_blocks))
    fn = STEP_CB(cb)
    rc = lib.orc_run_onechqs(ctypes.byref(p), fn, None)
    if rc != 0:
        raise RuntimeError("oracle run failed: %d" % rc)
    return steps


def tan_step_to_dict(dt):
    """Copy one tangent record into numpy (ntan x ... arrays)."""
    d = dt.contents
    nt, ns = d.ntan, d.nstates
    return {"N": d.nshell,
            "den": (np.ctypeslib.as_array(d.den, (nt * ns,)).reshape(nt, ns).copy() if nt else np.zeros((0, ns))),
            "dsuscep": np.array([[d.dsuscep[t][i] for i in range(4)] for t in range(nt)]).reshape(nt, 4),
            "dentropy": np.array([[d.dentropy[t][i] for i in range(4)] for t in range(nt)]).reshape(nt, 4),
            "dopavg": np.array([[d.dopavg[t][i] for i in range(d.nops_avg)] for t in range(nt)]).reshape(
                nt, d.nops_avg)}


def kept_arrays(steps):
    """Concatenated Kept flags of steps N >= 0 and their offsets (the frozen-truncation input)."""
    ks = [s["kept"].astype(np.uint8) for s in steps if s["N"] >= 0]
    off = np.zeros(len(ks) + 1, dtype=np.int64)
    off[1:] = np.cumsum([len(k) for k in ks])
    return (np.concatenate(ks) if ks else np.zeros(0, np.uint8)), off


def run_onechqs_jvp(U=0.5, Gamma=0.02, ed=-0.25, Lambda=2.5, nsitesmax=4, ncut=1000, calcdens=2,
                    disc_scheme=0, myeps=1e-18, betabar=0.727, kept=None, dpar=None):
    """The NRG_main loop with frozen Kept sets (kept = (flags, offsets) from
    kept_arrays, or None) and tangents along the rows of dpar (dU, dGamma, ded).
    Returns (steps, tangents)."""
    lib = load()
    steps, tans = [], []

    def cb(_user, st, dt):
        steps.append(step_to_dict(st))
        tans.append(tan_step_to_dict(dt))

    p = OneChQSParams(U, Gamma, ed, Lambda, nsitesmax, ncut, calcdens, disc_scheme, myeps, betabar, 0)
    dp_ = np.ascontiguousarray(np.zeros((0, 3)) if dpar is None else np.asarray(dpar, np.float64).reshape(-1, 3))
    if kept is not None:
        kf = np.ascontiguousarray(kept[0], dtype=np.uint8)
        ko = np.ascontiguousarray(kept[1], dtype=np.int64)
        kfp = kf.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte))
        kop = ko.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong))
    else:
        kfp = kop = None
    fn = TAN_CB(cb)
    rc = lib.orc_run_onechqs_jvp(ctypes.byref(p), kfp, kop, dp_.shape[0], _dp(dp_) if dp_.size else None, fn, None)
    if rc != 0:
        raise RuntimeError("oracle jvp run failed: %d" % rc)
    return steps, tans


def run_twochqs(U=0.4, Gamma1=0.03, Gamma2=0.03, ed=-0.2, Lambda=3.0, nsitesmax=3, ncut=3000,
                betabar=0.727, record_blocks=False):
    lib = load()
    steps = []

    def cb(_user, st):
        steps.append(step_to_dict(st))

    p = TwoChQSParams(U, Gamma1, Gamma2, ed, Lambda, nsitesmax, ncut, betabar, int(record_blocks))
    fn = STEP_CB(cb)
    rc = lib.orc_run_twochqs(ctypes.byref(p), fn, None)
    if rc != 0:
        raise RuntimeError("oracle run failed: %d" % rc)
    return steps


def onechqs_thermo(U=0.5, Gamma=0.02, ed=-0.25, Lambda=2.5, nsitesmax=40, ncut=400, disc_scheme=0,
                   myeps=0.0, betabar=0.727):
    """Impurity thermodynamics of CalcThermo with the chain subtraction
    (src/NRGcodehandlerClass.cpp:627-652, src/NRGthermoClass.cpp:10-151): the
    chain run (U = ed = Gamma = 0) minus the free orbital's 1/8 and 2 ln 2
    (src/CodeHandler_ModelSwitch.cpp:217-220), then the impurity run; rows
    (4 per step N >= 0) of T, Tchi_tot, Tchi_chain, Tchi_imp, S_tot, S_chain, S_imp."""
    kw = dict(Lambda=Lambda, nsitesmax=nsitesmax, ncut=ncut, calcdens=2, disc_scheme=disc_scheme, myeps=myeps,
              betabar=betabar)
    ch = [s for s in run_onechqs(U=0.0, Gamma=0.0, ed=0.0, **kw) if s["N"] >= 0]
    im = [s for s in run_onechqs(U=U, Gamma=Gamma, ed=ed, **kw) if s["N"] >= 0]
    rows = []
    for a, b in zip(im, ch):
        for nc in range(4):
            cchi = b["suscep"][nc] - 1.0 / 8.0
            cS = b["entropy"][nc] - 2.0 * np.log(2.0)
            rows.append([a["thermo_T"][nc], a["suscep"][nc], cchi, a["suscep"][nc] - cchi,
                         a["entropy"][nc], cS, a["entropy"][nc] - cS])
    return np.array(rows)


def run_twochqsz(U=0.5, Gamma1=0.03, Gamma2=0.01, ed=-0.25, Lambda=2.5, nsitesmax=3, ncut=4000,
                 betabar=0.727, record_blocks=False):
    """TwoChQSz standalone driver: (Q, Sz) blocks, parameters as run_twochqs."""
    lib = load()
    steps = []

    def cb(_user, st):
        steps.append(step_to_dict(st))

    p = TwoChQSParams(U, Gamma1, Gamma2, ed, Lambda, nsitesmax, ncut, betabar, int(record_blocks))
    fn = STEP_CB(cb)
    rc = lib.orc_run_twochqsz(ctypes.byref(p), fn, None)
    if rc != 0:
        raise RuntimeError("oracle run failed: %d" % rc)
    return steps


def cgordan(J1, M1, J2, M2, J, M):
    return load().orc_cgordan(J1, M1, J2, M2, J, M)


def eigh_packed(n, ap):
    ap = np.ascontiguousarray(ap, dtype=np.float64)
    w = np.zeros(n)
    z = np.zeros(n * n)
    load().orc_eigh_packed(n, _dp(ap), _dp(w), _dp(z))
    return w, z.reshape(n, n)


def eigh_full(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    n = a.shape[0]
    w = np.zeros(n)
    z = np.zeros(n * n)
    load().orc_eigh_full(n, _dp(a), _dp(w), _dp(z))
    return w, z.reshape(n, n)


def setdegen(en, myeps):
    e = np.ascontiguousarray(en, dtype=np.float64).copy()
    load().orc_setdegen(len(e), _dp(e), myeps)
    return e


def ecut(en, ncut):
    e = np.ascontiguousarray(en, dtype=np.float64)
    return load().orc_ecut(len(e), _dp(e), ncut)


def rotate(zi, f, zj):
    zi = np.ascontiguousarray(zi, dtype=np.float64)
    zj = np.ascontiguousarray(zj, dtype=np.float64)
    f = np.ascontiguousarray(f, dtype=np.float64)
    ki, ni = zi.shape
    kj, nj = zj.shape
    out = np.zeros((ki, kj))
    load().orc_rotate(ni, nj, ki, kj, _dp(zi), _dp(f), _dp(zj), _dp(out))
    return out


def chin(n, Lambda):
    return load().orc_chin(n, Lambda)


# ---------------------------------------------------------------- DM-NRG
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


# bits of fdm_fix (nrg_oracle.h ORC_FDM_FIX_*): each switches one quirk of the reference's FDM off
FDM_FIX_RHOMAX, FDM_FIX_DN, FDM_FIX_LAST, FDM_FIX_PAIRS, FDM_FIX_ONESIDED, FDM_FIX_NORM = 1, 2, 4, 8, 16, 32
FDM_EXACT = 63


_PD = ctypes.POINTER(ctypes.c_double)
_PI = ctypes.POINTER(ctypes.c_int)


class DmShellView(ctypes.Structure):
    _fields_ = [
        ("nblocks", ctypes.c_int), ("nq", ctypes.c_int), ("nstates", ctypes.c_int),
        ("qn", _PD), ("bbe", _PI), ("en", _PD), ("bbe_bc", _PI), ("ev", _PD), ("bbe_ev", _PI),
        ("rho", _PD), ("rho_mbe", _PI), ("trace", ctypes.c_double),
        ("op_nblocks", ctypes.c_int), ("op_map", _PI), ("op_mbe", _PI), ("op", _PD),
        ("basis_nstates", ctypes.c_int), ("basis_nblocks", ctypes.c_int),
        ("basis_stcf", _PI), ("basis_itype", _PI), ("basis_qn", _PD), ("basis_bbe", _PI),
        ("kept", _PI), ("wn", ctypes.c_double),
    ]


def _arr(p, n, dt=np.float64):
    if n <= 0 or not p:
        return np.zeros(0, dt)
    return np.ctypeslib.as_array(p, shape=(n,)).astype(dt).copy()


class DmRun:
    """DM-NRG restatement on a OneChQS chain: see orc_dm_run.
    method 0: DM-NRG, 1: CFS, 2: FDM; update_bef_cut: the NRG run's UpdateBefCut."""

    def __init__(self, U, Gamma, ed, Lambda, nsitesmax, ncut, Mtemp, betabar, disc_scheme=0,
                 twindow=0.0, broadtemp=0.0, bbroad=0.0, Nw=0, myeps=0.0, betabar_nrg=0.0,
                 method=0, update_bef_cut=0, kept=None, dpar=None, model=0, Gamma2=0.0, fdm_site_dim=0.0, fdm_fix=0):
        """kept = (flags, offsets) from kept_arrays of an NRG run freezes every
        step's Kept set; dpar rows are tangent directions (dU, dGamma, ded)
        carried through the chain and the sweep (orc_dm_run_jvp)."""
        L = load()
        L.orc_dm_run_jvp.restype = ctypes.c_void_p
        L.orc_dm_run_jvp.argtypes = [ctypes.POINTER(DmParams), ctypes.POINTER(ctypes.c_ubyte),
                                     ctypes.POINTER(ctypes.c_longlong), ctypes.c_int, _PD]
        L.orc_dm_spec_shell_jvp.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, _PD]
        L.orc_dm_run.restype = ctypes.c_void_p
        L.orc_dm_run.argtypes = [ctypes.POINTER(DmParams)]
        L.orc_dm_free.argtypes = [ctypes.c_void_p]
        L.orc_dm_info.argtypes = [ctypes.c_void_p, _PI, _PI, _PD]
        L.orc_dm_shell.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(DmShellView)]
        L.orc_dm_spec_shell.restype = ctypes.c_double
        L.orc_dm_spec_shell.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double]
        L.orc_dm_spectrum.argtypes = [ctypes.c_void_p, ctypes.POINTER(_PD), ctypes.POINTER(_PD)]
        for f in (L.orc_dm_spec_omega, L.orc_dm_norm, L.orc_dm_ztot):
            f.restype = ctypes.c_double
        L.orc_dm_spec_omega.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.orc_dm_norm.argtypes = [ctypes.c_void_p]
        L.orc_dm_ztot.argtypes = [ctypes.c_void_p]
        p = DmParams(U, Gamma, ed, Lambda, nsitesmax, ncut, disc_scheme, myeps, betabar_nrg, Mtemp, betabar,
                     twindow, broadtemp, bbroad, Nw, method, update_bef_cut, model, Gamma2, fdm_site_dim, fdm_fix)
        self.method = method
        self._L = L
        dp_ = np.ascontiguousarray(np.zeros((0, 3)) if dpar is None else np.asarray(dpar, np.float64).reshape(-1, 3))
        self.ntan = dp_.shape[0]
        if kept is not None:
            self._kf = np.ascontiguousarray(kept[0], dtype=np.uint8)
            self._ko = np.ascontiguousarray(kept[1], dtype=np.int64)
            kfp = self._kf.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte))
            kop = self._ko.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong))
        else:
            kfp = kop = None
        self.h = L.orc_dm_run_jvp(ctypes.byref(p), kfp, kop, self.ntan, _dp(dp_) if dp_.size else None)
        if not self.h:
            raise RuntimeError("orc_dm_run_jvp refused the arguments")
        a, b, t = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        self.nsh = L.orc_dm_info(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(t))
        self.nshell_max, self.spec_nshell_max, self.temp = a.value, b.value, t.value
        self.ztot = L.orc_dm_ztot(self.h)

    def close(self):
        if self.h:
            self._L.orc_dm_free(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def shell(self, N):
        v = DmShellView()
        if self._L.orc_dm_shell(self.h, N, ctypes.byref(v)) != 0:
            raise IndexError(N)
        nb, nq = v.nblocks, v.nq
        bbe = _arr(v.bbe, 2 * nb, np.int64).reshape(nb, 2)
        bbe_bc = _arr(v.bbe_bc, 2 * nb, np.int64).reshape(nb, 2)
        rmbe = _arr(v.rho_mbe, 2 * nb, np.int64).reshape(nb, 2)
        rho_all = _arr(v.rho, int(rmbe[-1, 1]) + 1 if nb else 0)
        rho = []
        for b in range(nb):
            n = int(bbe[b, 1] - bbe[b, 0] + 1)
            m = np.zeros((n, n))
            m[np.triu_indices(n)] = rho_all[rmbe[b, 0]:rmbe[b, 1] + 1]
            rho.append(m + np.triu(m, 1).T)
        ev_bbe = _arr(v.bbe_ev, 2 * nb, np.int64).reshape(nb, 2)
        ev_all = _arr(v.ev, int(ev_bbe[-1, 1]) + 1 if nb else 0)
        Z = []
        for b in range(nb):
            k = int(bbe[b, 1] - bbe[b, 0] + 1)
            n = int(bbe_bc[b, 1] - bbe_bc[b, 0] + 1)
            Z.append(ev_all[ev_bbe[b, 0]:ev_bbe[b, 0] + k * n].reshape(k, n))
        nop = v.op_nblocks
        omap = _arr(v.op_map, 2 * nop, np.int64).reshape(nop, 2)
        ombe = _arr(v.op_mbe, 2 * nop, np.int64).reshape(nop, 2)
        op_all = _arr(v.op, int(ombe[-1, 1]) + 1 if nop else 0)
        ops = {}
        for i in range(nop):
            b1, b2 = int(omap[i, 0]), int(omap[i, 1])
            n1 = int(bbe[b1, 1] - bbe[b1, 0] + 1)
            n2 = int(bbe[b2, 1] - bbe[b2, 0] + 1)
            ops[(b1, b2)] = op_all[ombe[i, 0]:ombe[i, 1] + 1].reshape(n1, n2)
        bnb = v.basis_nblocks
        return dict(N=N, qn=_arr(v.qn, nb * nq).reshape(nb, nq), bbe=bbe, bbe_bc=bbe_bc,
                    en=_arr(v.en, v.nstates), rho=rho, trace=v.trace, Z=Z, op=ops,
                    basis_stcf=_arr(v.basis_stcf, v.basis_nstates, np.int64),
                    basis_itype=_arr(v.basis_itype, v.basis_nstates, np.int64),
                    basis_qn=_arr(v.basis_qn, bnb * nq).reshape(bnb, nq),
                    basis_bbe=_arr(v.basis_bbe, 2 * bnb, np.int64).reshape(bnb, 2),
                    kept=_arr(v.kept, v.nstates, np.int64) if v.kept else np.ones(v.nstates, np.int64),
                    wn=v.wn)

    def spec_shell(self, N, w):
        return self._L.orc_dm_spec_shell(self.h, N, w)

    def spec_omega(self, w):
        return self._L.orc_dm_spec_omega(self.h, w)

    def spec_shell_jvp(self, N, w):
        out = np.zeros(max(self.ntan, 1))
        if self._L.orc_dm_spec_shell_jvp(self.h, N, w, _dp(out)) != 0:
            raise RuntimeError("orc_dm_spec_shell_jvp failed")
        return out[:self.ntan]

    def norm(self):
        return self._L.orc_dm_norm(self.h)

    def spectrum(self):
        po, pr = _PD(), _PD()
        n = self._L.orc_dm_spectrum(self.h, ctypes.byref(po), ctypes.byref(pr))
        return _arr(po, n), _arr(pr, n)


def grid_plan(Lambda, nshell, windowed, Nw):
    """orc_grid_plan: (shell, omegabar, omega, parity) of the omega-grid drivers."""
    L = load()
    L.orc_grid_plan.restype = ctypes.c_int
    L.orc_grid_plan.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                ctypes.POINTER(ctypes.c_int), _PD, _PD, ctypes.POINTER(ctypes.c_int)]
    n = L.orc_grid_plan(Lambda, nshell, int(windowed), Nw, None, None, None, None)
    sh, par = np.zeros(n, np.int32), np.zeros(n, np.int32)
    wb, w = np.zeros(n), np.zeros(n)
    L.orc_grid_plan(Lambda, nshell, int(windowed), Nw, sh.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _dp(wb),
                    _dp(w), par.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    return sh, wb, w, par


def grid_combine(omega, parity, value, Nw):
    """orc_grid_combine: the spectrum from the values at the plan's points."""
    L = load()
    L.orc_grid_combine.restype = ctypes.c_int
    L.orc_grid_combine.argtypes = [ctypes.c_int, _PD, ctypes.POINTER(ctypes.c_int), _PD, ctypes.c_int, _PD, _PD]
    w = np.ascontiguousarray(omega, np.float64)
    par = np.ascontiguousarray(parity, np.int32)
    v = np.ascontiguousarray(value, np.float64)
    wo, ro = np.zeros(len(w)), np.zeros(len(w))
    L.orc_grid_combine(len(w), _dp(w), par.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _dp(v), Nw, _dp(wo),
                       _dp(ro))
    return wo, ro
