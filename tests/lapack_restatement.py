"""Second restatement of calc_bdw_lanc / calc_bdw_blanc / calc_bdw_cholesky
that calls the very LAPACK routines the reference calls (through SciPy's
LAPACK wrappers): dgeqp3 + dorgqr (brown_m.f90:359, 373, 411, 426), dsbevd
on the band packed exactly as brown_m.f90:442-467, dstev for the s = 1
tridiagonal (the reference calls dstevd: the same eigenproblem, another
driver), and dpotrf('U') (brown_m.f90:539).

Test infrastructure: it pins the C oracle's own QR/eigen code against the
library routines at the reference's call boundary.
"""
import numpy as np
import scipy.linalg.lapack as lp


def lanczos_s1(D, xi, lanc_mxitr, tol, reorth=False):
    n = D.shape[0]
    dimk_min = 2
    dimk_max = max(dimk_min, lanc_mxitr) + 1
    V = np.zeros((n, dimk_max + 1))
    hd = np.zeros(dimk_max + 1)
    hsd = np.zeros(dimk_max + 1)
    normz = np.sqrt(np.sum(xi * xi))
    V[:, 0] = xi / normz
    old = None
    bdw = None
    f = 0
    for j in range(1, dimk_max + 1):
        f += 1
        w = D @ V[:, j - 1]
        if j > 1:
            w = w - hsd[j - 1] * V[:, j - 2]
        hd[j] = w @ V[:, j - 1]
        if j < dimk_max:
            w = w - hd[j] * V[:, j - 1]
            if reorth:
                for _ in range(2):
                    w = w - V[:, :j] @ (V[:, :j].T @ w)
            hsd[j] = np.sqrt(np.sum(w * w))
            V[:, j] = w / hsd[j]
        if j >= dimk_min:
            d, e = hd[1:j + 1].copy(), hsd[1:j].copy()
            lam, P, info = lp.dstev(d, e, compute_v=1)
            assert info == 0
            t1 = np.sqrt(lam) * P[0, :]
            t2 = P @ t1
            bdw = normz * (V[:, :j] @ t2)
            if j > dimk_min:
                err = np.sqrt(np.sum((bdw - old) ** 2)) / np.sqrt(np.sum(old ** 2))
                if err <= tol:
                    return bdw, True, f
            old = bdw.copy()
    return bdw, False, f


def _qr_piv(W):
    qr, jpvt, tau, work, info = lp.dgeqp3(np.asfortranarray(W))
    assert info == 0
    s = W.shape[1]
    Rp = np.zeros((s, s))
    for jc in range(s):
        Rp[: jc + 1, jpvt[jc] - 1] = qr[: jc + 1, jc]
    q, work, info = lp.dorgqr(qr, tau)
    assert info == 0
    return q[:, :s], Rp


def block_lanczos(D, xi, lanc_mxitr, tol, reorth=False):
    n, s = xi.shape
    dimk_min = 2
    dimk_max = max(dimk_min, lanc_mxitr) + 1
    V = np.zeros((n, s * dimk_max))
    hbd = np.zeros((s, s * dimk_max))
    hbsd = np.zeros((s, s * dimk_max))
    q, matr = _qr_piv(xi)
    V[:, :s] = q
    old = None
    bdw = None
    f = 0
    for j in range(1, dimk_max + 1):
        f += 1
        c0 = (j - 1) * s
        W = D @ V[:, c0:c0 + s]
        if j > 1:
            W = W - V[:, c0 - s:c0] @ hbsd[:, c0 - s:c0].T
        hbd[:, c0:c0 + s] = V[:, c0:c0 + s].T @ W
        if j < dimk_max:
            A = hbd[:, c0:c0 + s]
            As = np.tril(A) + np.tril(A, -1).T
            W = W - V[:, c0:c0 + s] @ As
            if reorth:
                for _ in range(2):
                    W = W - V[:, :c0 + s] @ (V[:, :c0 + s].T @ W)
            q, B = _qr_piv(W)
            hbsd[:, c0:c0 + s] = B
            V[:, c0 + s:c0 + 2 * s] = q
        if j >= dimk_min:
            js = j * s
            hb = np.zeros((2 * s, js))
            for jb in range(j):
                for jc in range(s):
                    jg = jb * s + jc
                    for ir in range(jc, s):
                        ig = jb * s + ir
                        hb[ig - jg, jg] = hbd[ir, jg]
            for jb in range(j - 1):
                for jc in range(s):
                    jg = jb * s + jc
                    for ir in range(s):
                        ig = (jb + 1) * s + ir
                        hb[ig - jg, jg] = hbsd[ir, jg]
            lam, P, info = lp.dsbevd(np.asfortranarray(hb), compute_v=1, lower=1)
            assert info == 0
            t = (P[:s, :].T @ matr) * np.sqrt(lam)[:, None]
            bdw = V[:, :js] @ (P @ t)
            if j > dimk_min:
                err = np.max(np.sum(np.abs(bdw - old), 0)) / np.max(np.sum(np.abs(old), 0))
                if err <= tol:
                    return bdw, True, j
            old = bdw.copy()
    return bdw, False, f


def cholesky_u_xi(D, xi):
    u, info = lp.dpotrf(np.asfortranarray(D), lower=0)
    assert info == 0
    U = np.triu(u)
    return U @ xi
