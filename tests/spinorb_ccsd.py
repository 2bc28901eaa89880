"""Independent checkers for the oracle (test-only).

* ``spinorb_ccsd``: spin-orbital CCSD with the textbook equations of
  J. F. Stanton, J. Gauss, J. D. Watts, R. J. Bartlett, J. Chem. Phys. 94,
  4334 (1991), written for a general (non-canonical) Fock matrix.  It shares no
  code or intermediates with the spin-adapted closed-shell equations the
  oracle restates, so agreement pins the oracle's residual.
* ``fci_two_electron``: exact diagonalisation for a closed-shell two-electron
  system (CCSD is exact there).
* ``random_hamiltonian``: a random real Hamiltonian with 8-fold ERI symmetry.
"""
from __future__ import annotations

import numpy as np


def random_hamiltonian(no, nv, seed=0, eri_scale=0.05, fock_offdiag=0.02, f_ov=0.0):
    """Random spatial Hamiltonian: Fock f (n x n, symmetric) and chemist ERIs
    (pq|rs) with 8-fold permutational symmetry.  Occupied orbitals first."""
    rng = np.random.default_rng(seed)
    n = no + nv
    eps = np.concatenate([np.sort(rng.uniform(-1.5, -0.4, no)),
                          np.sort(rng.uniform(0.3, 2.5, nv))])
    f = np.diag(eps)
    x = rng.normal(0, fock_offdiag, (n, n))
    x = 0.5 * (x + x.T)
    f = f + x - np.diag(np.diag(x))
    f[:no, no:] = 0.0
    f[no:, :no] = 0.0
    if f_ov:
        y = rng.normal(0, f_ov, (no, nv))
        f[:no, no:] = y
        f[no:, :no] = y.T
    g = rng.normal(0, eri_scale, (n, n, n, n))
    # symmetrise over the 8 real-orbital permutations of (pq|rs)
    g = (g + g.transpose(1, 0, 2, 3) + g.transpose(0, 1, 3, 2) + g.transpose(1, 0, 3, 2)
         + g.transpose(2, 3, 0, 1) + g.transpose(3, 2, 0, 1) + g.transpose(2, 3, 1, 0)
         + g.transpose(3, 2, 1, 0)) / 8.0
    return f, g


def mpqc_blocks(f, g, no):
    """Slice a spatial Hamiltonian into the reference's ``cc::Integrals``
    index orders (physicist <pq|rs> = (pr|qs))."""
    o = slice(0, no)
    v = slice(no, f.shape[0])
    phys = g.transpose(0, 2, 1, 3)  # <pq|rs>
    return dict(
        Fia=f[o, v].copy(), Fij=f[o, o].copy(), Fab=f[v, v].copy(),
        Gijab=phys[o, o, v, v].copy(), Gijkl=phys[o, o, o, o].copy(),
        Giajb=phys[o, v, o, v].copy(), Giabc=phys[o, v, v, v].copy(),
        Gijka=phys[o, o, o, v].copy(), Gabcd=phys[v, v, v, v].copy())


def _spin_blocks(f, g, no):
    n = f.shape[0]
    ns = 2 * n
    spin = np.arange(ns) % 2
    orb = np.arange(ns) // 2
    fs = f[orb][:, orb] * (spin[:, None] == spin[None, :])
    phys = g.transpose(0, 2, 1, 3)  # <PQ|RS>
    P = phys[np.ix_(orb, orb, orb, orb)]
    s = spin
    d_pr = s[:, None, None, None] == s[None, None, :, None]
    d_qs = s[None, :, None, None] == s[None, None, None, :]
    d_ps = s[:, None, None, None] == s[None, None, None, :]
    d_qr = s[None, :, None, None] == s[None, None, :, None]
    asym = P * (d_pr & d_qs) - P.transpose(0, 1, 3, 2) * (d_ps & d_qr)
    # occupied spin orbitals first: spatial occupied are 0..no-1 -> spin 0..2no-1
    return fs, asym, 2 * no


def spinorb_ccsd(f, g, no, conv=1e-12, max_iter=500):
    """Returns (E_corr, t1[a,i] spin-orbital, t2[a,b,i,j], nocc_so)."""
    fs, G, O = _spin_blocks(f, g, no)
    ns = fs.shape[0]
    o = slice(0, O)
    v = slice(O, ns)
    V = ns - O
    fo = fs[o, o]
    fv = fs[v, v]
    fov = fs[o, v]
    Dia = np.diag(fo)[:, None] - np.diag(fv)[None, :]
    Dijab = (np.diag(fo)[:, None, None, None] + np.diag(fo)[None, :, None, None]
             - np.diag(fv)[None, None, :, None] - np.diag(fv)[None, None, None, :])
    oooo = G[o, o, o, o]
    ooov = G[o, o, o, v]
    oovv = G[o, o, v, v]
    ovvo = G[o, v, v, o]
    ovvv = G[o, v, v, v]
    vvvv = G[v, v, v, v]
    vvvo = G[v, v, v, o]
    ovoo = G[o, v, o, o]
    ovov = G[o, v, o, v]
    t1 = np.zeros((O, V))  # t[i,a]
    t2 = np.zeros((O, O, V, V))  # t[i,j,a,b]
    e = np.einsum
    fo_off = fo - np.diag(np.diag(fo))
    fv_off = fv - np.diag(np.diag(fv))

    def energy(t1, t2):
        return (e("ia,ia->", fov, t1) + 0.25 * e("ijab,ijab->", oovv, t2)
                + 0.5 * e("ijab,ia,jb->", oovv, t1, t1))

    hist_t, hist_r = [], []
    Eold = 0.0
    for it in range(max_iter):
        tau_t = t2 + 0.5 * (e("ia,jb->ijab", t1, t1) - e("ib,ja->ijab", t1, t1))
        tau = t2 + e("ia,jb->ijab", t1, t1) - e("ib,ja->ijab", t1, t1)
        Fae = (fv_off - 0.5 * e("me,ma->ae", fov, t1) + e("mf,mafe->ae", t1, ovvv)
               - 0.5 * e("mnaf,mnef->ae", tau_t, oovv))
        Fmi = (fo_off + 0.5 * e("ie,me->mi", t1, fov) + e("ne,mnie->mi", t1, ooov)
               + 0.5 * e("inef,mnef->mi", tau_t, oovv))
        Fme = fov + e("nf,mnef->me", t1, oovv)
        Wmnij = (oooo + e("je,mnie->mnij", t1, ooov) - e("ie,mnje->mnij", t1, ooov)
                 + 0.25 * e("ijef,mnef->mnij", tau, oovv))
        Wabef = (vvvv - e("mb,amef->abef", t1, G[v, o, v, v]) + e("ma,bmef->abef", t1, G[v, o, v, v])
                 + 0.25 * e("mnab,mnef->abef", tau, oovv))
        Wmbej = (ovvo + e("jf,mbef->mbej", t1, ovvv) - e("nb,mnej->mbej", t1, G[o, o, v, o])
                 - e("jnfb,mnef->mbej", 0.5 * t2 + e("jf,nb->jnfb", t1, t1), oovv))
        # T1
        r1 = (fov + e("ie,ae->ia", t1, Fae) - e("ma,mi->ia", t1, Fmi)
              + e("imae,me->ia", t2, Fme) - e("nf,naif->ia", t1, ovov)
              - 0.5 * e("imef,maef->ia", t2, ovvv) - 0.5 * e("mnae,nmei->ia", t2, G[o, o, v, o]))
        # T2
        Ftmp_be = Fae - 0.5 * e("mb,me->be", t1, Fme)
        Ftmp_mj = Fmi + 0.5 * e("je,me->mj", t1, Fme)
        x = e("ijae,be->ijab", t2, Ftmp_be)
        r2 = oovv + x - x.transpose(0, 1, 3, 2)
        x = e("imab,mj->ijab", t2, Ftmp_mj)
        r2 -= x - x.transpose(1, 0, 2, 3)
        r2 += 0.5 * e("mnab,mnij->ijab", tau, Wmnij)
        r2 += 0.5 * e("ijef,abef->ijab", tau, Wabef)
        x = e("imae,mbej->ijab", t2, Wmbej) - e("ie,ma,mbej->ijab", t1, t1, ovvo)
        r2 += x - x.transpose(1, 0, 2, 3) - x.transpose(0, 1, 3, 2) + x.transpose(1, 0, 3, 2)
        x = e("ie,abej->ijab", t1, vvvo)
        r2 += x - x.transpose(1, 0, 2, 3)
        x = e("ma,mbij->ijab", t1, ovoo)
        r2 -= x - x.transpose(0, 1, 3, 2)
        # residual of the full equation: rhs - D t
        res1 = r1 - Dia * t1
        res2 = r2 - Dijab * t2
        t1n = r1 / Dia
        t2n = r2 / Dijab
        Enew = energy(t1n, t2n)
        rn = np.sqrt(np.sum(res1 ** 2) + np.sum(res2 ** 2))
        # simple DIIS
        hist_t.append(np.concatenate([t1n.ravel(), t2n.ravel()]))
        hist_r.append(np.concatenate([(t1n - t1).ravel(), (t2n - t2).ravel()]))
        if len(hist_t) > 8:
            hist_t.pop(0)
            hist_r.pop(0)
        if len(hist_t) >= 2:
            m = len(hist_t)
            B = -np.ones((m + 1, m + 1))
            B[m, m] = 0
            for a_ in range(m):
                for b_ in range(m):
                    B[a_, b_] = hist_r[a_] @ hist_r[b_]
            rhs = np.zeros(m + 1)
            rhs[m] = -1
            c = np.linalg.lstsq(B, rhs, rcond=None)[0][:m]
            vec = sum(ci * ti for ci, ti in zip(c, hist_t))
            t1n = vec[: t1.size].reshape(t1.shape)
            t2n = vec[t1.size:].reshape(t2.shape)
        t1, t2 = t1n, t2n
        if abs(Enew - Eold) < conv and rn < conv:
            break
        Eold = Enew
    else:
        raise RuntimeError("spin-orbital CCSD did not converge")
    Ecorr = energy(t1, t2)
    # transpose to the (a,i) / (a,b,i,j) orders
    return Ecorr, t1.T.copy(), t2.transpose(2, 3, 0, 1).copy(), O


def spin_to_spatial(t1_so, t2_so, no, nv):
    """Closed-shell spatial amplitudes from spin-orbital ones:
    t1[a,i] = t(a alpha, i alpha), t2[a,b,i,j] = t(a alpha, b beta, i alpha, j beta)."""
    t1 = t1_so[0::2, 0::2][:nv, :no]
    t2 = t2_so[0::2, 1::2, 0::2, 1::2]
    return t1.copy(), t2.copy()


def fci_two_electron(f, g):
    """Exact singlet ground state for 2 electrons; orbital 0 is the occupied
    one.  Returns E_corr = E_FCI - E_ref."""
    n = f.shape[0]
    # core Hamiltonian from the Fock matrix of the closed-shell determinant
    h = f - (2.0 * g[:, :, 0, 0] - g[:, 0, 0, :])
    e_ref = 2.0 * h[0, 0] + g[0, 0, 0, 0]
    I = np.eye(n)
    H = (np.einsum("pr,qs->pqrs", h, I) + np.einsum("pr,qs->pqrs", I, h)
         + g.transpose(0, 2, 1, 3))  # (pr|qs) as <pq|rs>
    H = H.reshape(n * n, n * n)
    # project onto symmetric (singlet spatial) functions
    S = np.zeros((n * n, n * (n + 1) // 2))
    col = 0
    for p in range(n):
        for q in range(p, n):
            if p == q:
                S[p * n + q, col] = 1.0
            else:
                S[p * n + q, col] = S[q * n + p, col] = np.sqrt(0.5)
            col += 1
    Hs = S.T @ H @ S
    w = np.linalg.eigvalsh(Hs)
    return float(w[0] - e_ref)


def spinorb_triples(f, g, no, t1_ai, t2_abij):
    """Spin-orbital (T) of the textbook equations (K. Raghavachari, G. W. Trucks,
    J. A. Pople, M. Head-Gordon, Chem. Phys. Lett. 157, 479 (1989); written as in
    T. D. Crawford, H. F. Schaefer, Rev. Comp. Chem. 14, 33 (2000)):
      D t(c) = P(i/jk) P(a/bc) [t_jk^ae <ei||bc> - t_im^bc <ma||jk>]
      D t(d) = P(i/jk) P(a/bc) t_i^a <jk||bc>
      E(T) = 1/36 sum t(c) D (t(c) + t(d)),  P(i/jk) f = f(ijk) - f(jik) - f(kji)
    from spin-orbital amplitudes t1[a,i], t2[a,b,i,j] (as spinorb_ccsd returns)."""
    fs, G, O = _spin_blocks(f, g, no)
    ns = fs.shape[0]
    o = slice(0, O)
    v = slice(O, ns)
    e = np.einsum
    t1 = t1_ai.T
    t2 = t2_abij.transpose(2, 3, 0, 1)
    eo, ev = np.diag(fs)[o], np.diag(fs)[v]
    D = (eo[:, None, None, None, None, None] + eo[None, :, None, None, None, None]
         + eo[None, None, :, None, None, None] - ev[None, None, None, :, None, None]
         - ev[None, None, None, None, :, None] - ev[None, None, None, None, None, :])
    X = e("ia,jkbc->ijkabc", t1, G[o, o, v, v])
    Y = e("jkae,eibc->ijkabc", t2, G[v, o, v, v]) - e("imbc,majk->ijkabc", t2, G[o, v, o, o])

    def P(Z):
        Z = Z - Z.transpose(1, 0, 2, 3, 4, 5) - Z.transpose(2, 1, 0, 3, 4, 5)
        return Z - Z.transpose(0, 1, 2, 4, 3, 5) - Z.transpose(0, 1, 2, 5, 4, 3)
    tc = P(Y) / D
    td = P(X) / D
    return float(np.sum(tc * D * (tc + td))) / 36.0
