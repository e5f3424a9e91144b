"""Brute-force exact diagonalisation of impurity + Wilson-chain Hamiltonians in
the full Fock space (Jordan-Wigner), block-diagonal in (particle number, 2Sz).
Used to pin the CPU oracle independently of the reference's golden log.

Modes are spin-orbitals; hop(i, j, t) adds t * sum_sigma (c+_i c_j + h.c.)
for spatial orbitals i, j.
"""
import itertools

import numpy as np


class FockED:
    def __init__(self, norb):
        self.norb = norb
        self.nmode = 2 * norb  # mode 2*o + s, s=0 up, s=1 down
        self.hops = []         # (o1, o2, t)
        self.onsite = []       # (o, eps)   eps * n_o
        self.hubbard = []      # (o, U)     U * n_o,up n_o,dn
        self.const = 0.0

    def hop(self, o1, o2, t):
        self.hops.append((o1, o2, t))

    def eps(self, o, e):
        self.onsite.append((o, e))

    def U(self, o, u):
        self.hubbard.append((o, u))

    def spectrum(self, scale=1.0):
        """All many-body eigenvalues (times scale), sorted."""
        nm = self.nmode
        out = []
        states = np.arange(1 << nm, dtype=np.int64)
        popc = np.array([bin(s).count("1") for s in range(1 << nm)])
        up_mask = sum(1 << (2 * o) for o in range(self.norb))
        nup = np.array([bin(s & up_mask).count("1") for s in range(1 << nm)])
        for ne in range(nm + 1):
            for nu in range(0, min(ne, self.norb) + 1):
                sel = states[(popc == ne) & (nup == nu)]
                if len(sel) == 0:
                    continue
                idx = {int(s): i for i, s in enumerate(sel)}
                H = np.zeros((len(sel), len(sel)))
                for i, s in enumerate(sel):
                    s = int(s)
                    d = self.const
                    for o, e in self.onsite:
                        d += e * (((s >> (2 * o)) & 1) + ((s >> (2 * o + 1)) & 1))
                    for o, u in self.hubbard:
                        d += u * (((s >> (2 * o)) & 1) * ((s >> (2 * o + 1)) & 1))
                    H[i, i] += d
                    for o1, o2, t in self.hops:
                        for sp in (0, 1):
                            for a, b in ((o1, o2), (o2, o1)):
                                ma, mb = 2 * a + sp, 2 * b + sp
                                # c+_ma c_mb |s>
                                if not (s >> mb) & 1:
                                    continue
                                s1 = s ^ (1 << mb)
                                sign = (-1) ** bin(s1 & ((1 << mb) - 1)).count("1")
                                if (s1 >> ma) & 1:
                                    continue
                                sign *= (-1) ** bin(s1 & ((1 << ma) - 1)).count("1")
                                s2 = s1 | (1 << ma)
                                H[idx[s2], i] += sign * t
                out.append(np.linalg.eigvalsh(H))
        return np.sort(np.concatenate(out)) * scale


def onechqs_chain_spectrum(U, Gamma, ed, Lam, N, a_lambda=True):
    """H_N of NRG_main's OneChQS Anderson recursion, rescaled, minus its minimum:
    H_{-1} = (ed n_d + U n_up n_dn + U/2)/(Lam*HLF),
    H_N = Lam^{(N+1)/2} H_{-1} + sum_{n=0}^{N} Lam^{(N-n)/2} t_n (f+_{n-1} f_n + h.c.),
    f_{-1} = d, t_0 = chi_{-1}, t_n = chi_Wilson(n-1)."""
    HLF = 0.5 * (1 + 1 / Lam)
    chi = np.sqrt(2 * Gamma / np.pi) / (np.sqrt(Lam) * HLF)
    if a_lambda:
        chi *= np.sqrt(0.5 * np.log(Lam) * (Lam + 1) / (Lam - 1))

    def chin(n):
        return (1 - Lam ** -(n + 1.0)) / np.sqrt((1 - Lam ** -(2.0 * n + 1)) * (1 - Lam ** -(2.0 * n + 3)))

    ed_ = FockED(N + 2)  # orbital 0 = d, 1.. = sites 0..N
    s = Lam ** ((N + 1) / 2.0) / (Lam * HLF)
    ed_.eps(0, ed * s)
    ed_.U(0, U * s)
    ed_.const = 0.5 * U * s
    for n in range(0, N + 1):
        t = chi if n == 0 else chin(n - 1)
        ed_.hop(n, n + 1, t * Lam ** ((N - n) / 2.0))
    e = ed_.spectrum()
    return e - e.min()


def twochqs_chain_spectrum(U, G1, G2, ed, Lam, N):
    """H_N of the legacy TwoChQS Anderson recursion, rescaled, minus its minimum:
    H_0 = (ed n_d + U n_up n_dn + U/2)/(HLF sqrt(Lam)) + sum_ch v_ch (f+_{0,ch} d + h.c.),
    v_ch = sqrt(Gamma_ch (2/pi)/HLF^2/Lam); H_N = Lam^{N/2} H_0 +
    sum_{n=1}^{N} Lam^{(N-n)/2} chi_n sum_ch (f+_{n-1,ch} f_{n,ch} + h.c.)."""
    HLF = 0.5 * (1 + 1 / Lam)
    v = [np.sqrt(g * (2 / np.pi) / HLF ** 2 / Lam) for g in (G1, G2)]

    def chi(n):
        return (1 - Lam ** -float(n)) / np.sqrt((1 - Lam ** -(2.0 * n - 1)) * (1 - Lam ** -(2.0 * n + 1)))

    norb = 1 + 2 * (N + 1)
    E = FockED(norb)

    def orb(n, ch):
        return 1 + 2 * n + ch

    s = Lam ** (N / 2.0) / (HLF * np.sqrt(Lam))
    E.eps(0, ed * s)
    E.U(0, U * s)
    E.const = 0.5 * U * s
    for ch in range(2):
        E.hop(0, orb(0, ch), v[ch] * Lam ** (N / 2.0))
        for n in range(1, N + 1):
            E.hop(orb(n - 1, ch), orb(n, ch), chi(n) * Lam ** ((N - n) / 2.0))
    e = E.spectrum()
    return e - e.min()


def expand_multiplets(en, qn, bbe, s_col=1):
    """NRG multiplet energies -> state energies (each repeated 2S+1 times)."""
    out = []
    for b in range(len(bbe)):
        S = qn[b, s_col]
        for i in range(bbe[b, 0], bbe[b, 1] + 1):
            out += [en[i]] * int(round(2 * S + 1))
    return np.sort(np.array(out))


# ---------------------------------------------------------------- dense Fock-space operators
def fock_annihilate(nmode, mode):
    """Dense c_mode on the full 2^nmode Fock space (Jordan-Wigner sign: occupied
    modes below `mode`)."""
    dim = 1 << nmode
    C = np.zeros((dim, dim))
    for s in range(dim):
        if (s >> mode) & 1:
            sign = (-1) ** bin(s & ((1 << mode) - 1)).count("1")
            C[s ^ (1 << mode), s] = sign
    return C


def fock_hamiltonian(E):
    """Dense H of a FockED on the full Fock space."""
    nm = E.nmode
    dim = 1 << nm
    C = [fock_annihilate(nm, m) for m in range(nm)]
    H = np.eye(dim) * E.const
    for o, e in E.onsite:
        for sp in (0, 1):
            m = 2 * o + sp
            H += e * C[m].T @ C[m]
    for o, u in E.hubbard:
        H += u * (C[2 * o].T @ C[2 * o]) @ (C[2 * o + 1].T @ C[2 * o + 1])
    for o1, o2, t in E.hops:
        for sp in (0, 1):
            a, b = C[2 * o1 + sp], C[2 * o2 + sp]
            H += t * (a.T @ b + b.T @ a)
    return H, C


def onechqs_chain_fock(U, Gamma, ed, Lam, N, a_lambda=True):
    """Dense rescaled H_N (minus its minimum) of the OneChQS recursion on the full
    Fock space of d + sites 0..N, with the annihilators of every mode
    (mode 2*orbital + spin, orbital 0 = d)."""
    HLF = 0.5 * (1 + 1 / Lam)
    chi = np.sqrt(2 * Gamma / np.pi) / (np.sqrt(Lam) * HLF)
    if a_lambda:
        chi *= np.sqrt(0.5 * np.log(Lam) * (Lam + 1) / (Lam - 1))

    def chin(n):
        return (1 - Lam ** -(n + 1.0)) / np.sqrt((1 - Lam ** -(2.0 * n + 1)) * (1 - Lam ** -(2.0 * n + 3)))

    E = FockED(N + 2)
    s = Lam ** ((N + 1) / 2.0) / (Lam * HLF)
    E.eps(0, ed * s)
    E.U(0, U * s)
    E.const = 0.5 * U * s
    for n in range(0, N + 1):
        t = chi if n == 0 else chin(n - 1)
        E.hop(n, n + 1, t * Lam ** ((N - n) / 2.0))
    H, C = fock_hamiltonian(E)
    w = np.linalg.eigvalsh(H)
    return H - w.min() * np.eye(len(H)), C


def partial_trace_last_orbital(rho, nmode):
    """Trace out the two highest modes (the last orbital): with Jordan-Wigner
    strings counting lower modes only, operators of the remaining modes act as
    O (x) 1 on (low bits, high bits)."""
    dl = 1 << (nmode - 2)
    out = np.zeros((dl, dl))
    for h in range(4):
        idx = np.arange(dl) + (h << (nmode - 2))
        out += rho[np.ix_(idx, idx)]
    return out
