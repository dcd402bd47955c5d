"""FP64 numpy/scipy restatement of the reference HODLR path (oracle; test use only).

Every function names the reference file:line it follows.  The structure is
kept literal where it matters for parity (ACA pivot and stop rules, tree
orientation, landmark choice, dense M = blkdiag(Q_V, Q_U) in lrsymfact, the
order of the Woodbury cascade), so that ranks, factors, solves, logdets and
stochastic gradients can be compared entry-for-entry with the HIP library.

Index convention: the reference's 1-based inclusive [a, b, c, d] block
(rows a:b = second child, cols c:d = first child, indices.jl:40-51) is stored
here 0-based half-open as (a0, b1, c0, d1) with rows a0:b1, cols c0:d1.
"""
import math

import numpy as np
import scipy.linalg as sla

from .kernels import kernel_matrix


# ---------------------------------------------------------------------------
# Levels and tree indices  (src/HODLR/src/structstypes.jl:30-37, indices.jl)
# ---------------------------------------------------------------------------

class FixedLevel:
    """min(lv, floor(log2 n)) dyadic splits (structstypes.jl:30-32)."""

    def __init__(self, lv):
        self.lv = int(lv)


class LogLevel:
    """floor(log2 n) - lv dyadic splits (structstypes.jl:35-37)."""

    def __init__(self, lv):
        self.lv = int(lv)


def floor_log2(n):
    return int(n).bit_length() - 1


def leafsizes(N, lvl):
    """indices.jl:5-24: number of levels and leaf sizes (remainder to the
    first leaves)."""
    fl = floor_log2(N)
    if isinstance(lvl, FixedLevel):
        n = min(max(0, lvl.lv), fl)
    else:
        if lvl.lv < 0:
            raise ValueError("Please provide a positive value for the level offset.")
        n = min(max(0, fl - lvl.lv), fl)
    nleaf = 2 ** n
    base = N // nleaf
    szs = [base] * nleaf
    rem = N - base * nleaf
    for j in range(rem):
        szs[j] += 1
    return n, szs


def leafindices(szs):
    """indices.jl:29-37 -> list of (a0, b1, c0, d1) with a0 == c0, b1 == d1."""
    out, s = [], 0
    for m in szs:
        out.append((s, s + m, s, s + m))
        s += m
    return out


def nextlevel(leafinds, lvl):
    """indices.jl:40-51: level-lvl off-diagonal blocks = rows of the second
    child x cols of the first child."""
    skp, half = 2 ** lvl, 2 ** (lvl - 1)
    out = []
    for j in range(0, len(leafinds) - 1, skp):
        out.append((leafinds[j + half][0], leafinds[j + skp - 1][1],
                    leafinds[j][2], leafinds[j + half - 1][3]))
    return out


def hodlr_indices(N, lvl):
    """indices.jl:55-65."""
    lv, szs = leafsizes(N, lvl)
    lf = leafindices(szs)
    nonlf = [nextlevel(lf, j) for j in range(1, lv + 1)]
    return lv, lf, nonlf


# ---------------------------------------------------------------------------
# Matrix sources: lazy KernelMatrix blocks and dense matrices
# ---------------------------------------------------------------------------

class KernelMatrix:
    """Lazy kernel matrix (src/structstypes.jl:2-23): K[i, j] = k(x1_i, x2_j)."""

    def __init__(self, x1, x2, kernel):
        self.x1 = np.ascontiguousarray(x1, dtype=np.float64)
        self.x2 = np.ascontiguousarray(x2, dtype=np.float64)
        self.kernel = kernel

    @property
    def shape(self):
        return (self.x1.shape[0], self.x2.shape[0])

    def row(self, i):
        return kernel_matrix(self.kernel, self.x1[i:i + 1], self.x2)[0]

    def col(self, j):
        return kernel_matrix(self.kernel, self.x1, self.x2[j:j + 1])[:, 0]

    def full(self):
        """src/baseoverloads.jl:20-29."""
        return kernel_matrix(self.kernel, self.x1, self.x2)

    def sub(self, a0, b1, c0, d1):
        """submatrix (src/utils.jl:27-33)."""
        return KernelMatrix(self.x1[a0:b1], self.x2[c0:d1], self.kernel)


class DenseSource:
    def __init__(self, M):
        self.M = np.asarray(M, dtype=np.float64)

    @property
    def shape(self):
        return self.M.shape

    def row(self, i):
        return self.M[i, :].copy()

    def col(self, j):
        return self.M[:, j].copy()

    def full(self):
        return self.M.copy()


# ---------------------------------------------------------------------------
# ACA  (src/factorizations.jl:9-84, src/utils.jl:2-15)
# ---------------------------------------------------------------------------

def restrictedmaxabs(x, noinds):
    """src/utils.jl:2-15: first index of the strict maximum |x_j| over
    j not in noinds, with its signed value; (0.0, first free index) if all
    candidates are zero.  0-based."""
    n = len(x)
    ind = 0
    while ind in noinds and ind < n - 1:
        ind += 1
    a = np.abs(x)
    if noinds:
        a = a.copy()
        a[list(noinds)] = -1.0
    a[:ind] = -1.0
    j = int(np.argmax(a))          # argmax returns the first maximum
    if a[j] > 0.0:
        return float(x[j]), j
    return 0.0, ind


class ACAError(RuntimeError):
    pass


def aca(M, rtol, maxrank=0, trace=None):
    """Partial-pivot adaptive cross approximation, src/factorizations.jl:9-84.

    M is a matrix source (row(i), col(j), shape).  Returns (U, V) with
    M ~= U V^T, U of size nro x r and V of size ncl x r.  If `trace` is a
    list, the pivots (row, col) are appended to it.
    """
    nro, ncl = M.shape
    maxsz = min(nro, ncl) if maxrank == 0 else min(maxrank, min(nro, ncl))
    U, V = [], []
    strt = 1                                    # 1-based, as :22-27
    tmprow = M.row(strt - 1)
    while np.max(np.abs(tmprow)) == 0.0 and strt < maxsz:
        strt += 1
        tmprow = M.row(strt - 1)
    if strt == maxsz:
        raise ACAError("Incompatible maxrank or matrix dimensions.")
    ucol = set()
    urow = {strt - 1}
    mv, mi = restrictedmaxabs(tmprow, ucol)     # :33
    tmprow = tmprow / mv
    tmpcol = M.col(mi)
    U.append(tmpcol)
    V.append(tmprow)
    if trace is not None:
        trace.append((strt - 1, mi))
    znorm2 = np.dot(tmprow, tmprow) * np.dot(tmpcol, tmpcol)
    ucol.add(mi)
    frnk = 1
    while (np.linalg.norm(tmprow) * np.linalg.norm(tmpcol) > rtol * math.sqrt(znorm2)
           and frnk < maxsz):                  # :43
        mv, ri = restrictedmaxabs(tmpcol, urow)
        urow.add(ri)
        tmprow = M.row(ri)
        for j in range(frnk):                   # :48-50
            tmprow = tmprow - U[j][ri] * V[j]
        mv, ci = restrictedmaxabs(tmprow, ucol)
        if mv == 0.0:                           # :55
            break
        ucol.add(ci)
        tmprow = tmprow / mv
        tmpcol = M.col(ci)
        for j in range(frnk):                   # :59-61
            tmpcol = tmpcol - V[j][ci] * U[j]
        for j in range(frnk):                   # :63-65
            znorm2 += 2.0 * abs(np.dot(U[j], tmpcol)) * abs(np.dot(V[j], tmprow))
        znorm2 += np.dot(tmprow, tmprow) * np.dot(tmpcol, tmpcol)
        frnk += 1
        V.append(tmprow)
        U.append(tmpcol)
        if trace is not None:
            trace.append((ri, ci))
    return np.stack(U, axis=1), np.stack(V, axis=1)


# ---------------------------------------------------------------------------
# Nystrom  (constructor.jl:24-33, src/factorizations.jl:86-94)
# ---------------------------------------------------------------------------

def nystrom_landmarks(n, r):
    """Int64.(round.(LinRange(1, n, r))) (constructor.jl:31), 0-based.

    LinRange evaluates (1 - t) * 1 + t * n with t = j / (r - 1); Julia's
    round is half-to-even, as numpy's rint."""
    j = np.arange(r, dtype=np.float64)
    t = j / (r - 1)
    v = (1.0 - t) * 1.0 + t * float(n)
    return np.rint(v).astype(np.int64) - 1


class NystromKernel:
    """src/structstypes.jl:25-28: landmarks + Cholesky of K(P, P)."""

    def __init__(self, lndmk, kernel):
        self.lndmk = np.ascontiguousarray(lndmk, dtype=np.float64)
        self.KPP = kernel_matrix(kernel, self.lndmk, self.lndmk)
        self.F = sla.cho_factor(self.KPP, lower=False)


def nystrom_uvt(Kb, nyker):
    """src/factorizations.jl:86-94: U = K(rows, P), V = (F \\ K(P, cols))^T."""
    U = kernel_matrix(Kb.kernel, Kb.x1, nyker.lndmk)
    K2 = kernel_matrix(Kb.kernel, nyker.lndmk, Kb.x2)
    V = sla.cho_solve(nyker.F, K2).T.copy()
    return U, V


# ---------------------------------------------------------------------------
# KernelHODLR and its factor  (structstypes.jl:40-63, constructor.jl:6-58)
# ---------------------------------------------------------------------------

class LowRankW:
    """W = I + M X M^T (structstypes.jl:40-43)."""

    def __init__(self, M, X):
        self.M = M
        self.X = X

    @property
    def size(self):
        return self.M.shape[0]

    def solterm(self, v):
        """lrx_solterm (HODLR/src/utils.jl:37-41)."""
        luf = self.X + np.eye(self.X.shape[0])
        Xv = self.X @ v
        return Xv - self.X @ sla.solve_triangular(luf, Xv, lower=True)

    def solterm_t(self, v):
        """lrx_solterm_t (HODLR/src/utils.jl:43-47)."""
        luf = self.X.T + np.eye(self.X.shape[0])
        Xv = self.X.T @ v
        return Xv - self.X.T @ sla.solve_triangular(luf, Xv, lower=False)

    def ldiv(self, t):
        """ldiv!(W, target) (HODLR/src/baseoverloads.jl:60-63), returns new."""
        return t - self.M @ self.solterm(self.M.T @ t)

    def ldiv_t(self, t):
        """ldiv!(W', target) (HODLR/src/baseoverloads.jl:65-69)."""
        return t - self.M @ self.solterm_t(self.M.T @ t)

    def mul(self, t):
        """mul!(target, W, src) (HODLR/src/baseoverloads.jl:36-45)."""
        return t + self.M @ (self.X @ (self.M.T @ t))

    def mul_t(self, t):
        """mul!(target, W', src) (HODLR/src/baseoverloads.jl:47-58)."""
        return t + self.M @ (self.X.T @ (self.M.T @ t))

    def det(self):
        """det(I + M^T M X) (HODLR/src/baseoverloads.jl:26-28)."""
        return np.linalg.det(np.eye(self.X.shape[0]) + (self.M.T @ self.M) @ self.X)

    def full(self):
        return np.eye(self.M.shape[0]) + (self.M @ self.X) @ self.M.T


def lrsymfact(U12, U21):
    """HODLR/src/utils.jl:29-35: QR of both panels, X = chol([I R;R' I]).L - I,
    M = blkdiag(Q12, Q21) stored dense."""
    Q1, R1 = np.linalg.qr(U12)
    Q2, R2 = np.linalg.qr(U21)
    R = R1 @ R2.T
    k1, k2 = R1.shape[0], R2.shape[0]
    C = np.block([[np.eye(k1), R], [R.T, np.eye(k2)]])
    Lc = np.linalg.cholesky(C)
    X = Lc - np.eye(k1 + k2)
    M = np.zeros((Q1.shape[0] + Q2.shape[0], Q1.shape[1] + Q2.shape[1]))
    M[:Q1.shape[0], :Q1.shape[1]] = Q1
    M[Q1.shape[0]:, Q1.shape[1]:] = Q2
    return LowRankW(M, X)


def _bd_apply(blocks, t, op):
    """BDiagonal ldiv!/mul! loop (BlockDiagonal.jl:51-70): apply op(block_j)
    to consecutive row ranges of t."""
    out = np.empty_like(t)
    s = 0
    for B in blocks:
        m = B.shape[0] if isinstance(B, np.ndarray) else B.size
        out[s:s + m] = op(B, t[s:s + m])
        s += m
    assert s == t.shape[0]
    return out


def _tri_ldiv(L, t):
    return sla.solve_triangular(L, t, lower=True)


def _tri_ldiv_t(L, t):
    return sla.solve_triangular(L, t, lower=True, trans='T')


class FactorHODLR:
    """Symmetric factor W = leafW * W_1 * ... * W_L (structstypes.jl:46-49)."""

    def __init__(self, leafW, nonleafW):
        self.leafW = leafW
        self.nonleafW = nonleafW

    def mul(self, src):
        """W x (HODLR/src/baseoverloads.jl:95-109)."""
        tmp = np.array(src, dtype=np.float64, copy=True)
        for j in range(len(self.nonleafW) - 1, -1, -1):
            tmp = _bd_apply(self.nonleafW[j], tmp, lambda W, t: W.mul(t))
        return _bd_apply(self.leafW, tmp, lambda L, t: L @ t)

    def mul_t(self, src):
        """W^T x (HODLR/src/baseoverloads.jl:111-124)."""
        tmp = _bd_apply(self.leafW, np.asarray(src, dtype=np.float64), lambda L, t: L.T @ t)
        for j in range(len(self.nonleafW)):
            tmp = _bd_apply(self.nonleafW[j], tmp, lambda W, t: W.mul_t(t))
        return tmp

    def ldiv(self, src):
        """W \\ x (HODLR/src/baseoverloads.jl:126-137)."""
        t = _bd_apply(self.leafW, np.asarray(src, dtype=np.float64), _tri_ldiv)
        for j in range(len(self.nonleafW)):
            t = _bd_apply(self.nonleafW[j], t, lambda W, v: W.ldiv(v))
        return t

    def ldiv_t(self, src):
        """W^T \\ x (HODLR/src/baseoverloads.jl:139-150)."""
        t = np.array(src, dtype=np.float64, copy=True)
        for j in range(len(self.nonleafW) - 1, -1, -1):
            t = _bd_apply(self.nonleafW[j], t, lambda W, v: W.ldiv_t(v))
        return _bd_apply(self.leafW, t, _tri_ldiv_t)

    def full(self):
        """HODLR/src/baseoverloads.jl:329-336 (testing only)."""
        out = sla.block_diag(*self.leafW)
        for lev in self.nonleafW:
            out = out @ sla.block_diag(*[W.full() for W in lev])
        return out


class NotFactorizedError(RuntimeError):
    pass


class KernelHODLR:
    """HODLR matrix (structstypes.jl:52-63), built as constructor.jl:6-58."""

    def __init__(self, K, ep, maxrank, lvl, nystrom=False, plel=False, sorted=False,
                 aca_trace=None):
        if not (K.x1.shape == K.x2.shape and np.array_equal(K.x1, K.x2)):
            raise ValueError("This function builds symmetric matrices.")   # :16-19
        n = K.shape[0]
        level, leafinds, nonleafinds = hodlr_indices(n, lvl)
        self.ep, self.lvl, self.mrnk = float(ep), level, int(maxrank)
        self.leafindices, self.nonleafindices = leafinds, nonleafinds
        self.nys = bool(nystrom)
        self.kernel = K.kernel
        self.x = K.x1
        if nystrom:                                                        # :24-33
            if maxrank < 2:
                raise ValueError("Please supply a fixed off-diagonal rank that is >=2.")
            if maxrank >= min(b1 - 1 - a0 for (a0, b1, c0, d1) in leafinds):
                raise ValueError("Your nystrom rank is too big. Reduce the HODLR level or nystrom rank.")
            self.nyind = nystrom_landmarks(n, maxrank)
            nyker = NystromKernel(K.x1[self.nyind], K.kernel)
        self.L = [K.sub(*li).full() for li in leafinds]                    # :39
        self.U, self.V = [], []
        self.ranks = []
        for j in range(level):                                             # :42-53
            Us, Vs, rk = [], [], []
            for bi, x in enumerate(nonleafinds[j]):
                if nystrom:
                    u, v = nystrom_uvt(K.sub(*x), nyker)
                else:
                    tr = None
                    if aca_trace is not None:
                        tr = []
                        aca_trace.append(((j, bi), tr))
                    u, v = aca(K.sub(*x), ep, maxrank, trace=tr)
                Us.append(u)
                Vs.append(v)
                rk.append(u.shape[1])
            self.U.append(Us)
            self.V.append(Vs)
            self.ranks.append(rk)
        self.W = None

    @property
    def n(self):
        return sum(L.shape[0] for L in self.L)

    def full(self):
        """HODLR/src/baseoverloads.jl:313-326 (requires unfactorized U/V)."""
        if self.U is None:
            raise NotFactorizedError("full(HK) needs the unfactorized U, V.")
        n = self.n
        out = np.empty((n, n))
        for j, (a0, b1, c0, d1) in enumerate(self.leafindices):
            out[a0:b1, c0:d1] = self.L[j]
        for lev in range(len(self.U)):
            for j, (a0, b1, c0, d1) in enumerate(self.nonleafindices[lev]):
                B = self.U[lev][j] @ self.V[lev][j].T
                out[a0:b1, c0:d1] = B
                out[c0:d1, a0:b1] = B.T
        return out

    def matvec(self, src):
        """mul!(target, K::KernelHODLR, src) (HODLR/src/baseoverloads.jl:152-179)."""
        src = np.asarray(src, dtype=np.float64)
        if self.W is None:
            tgt = np.zeros_like(src)
            for j, (a0, b1, c0, d1) in enumerate(self.leafindices):
                tgt[c0:d1] = self.L[j] @ src[c0:d1]
            for j in range(len(self.U)):
                for k, (a0, b1, c0, d1) in enumerate(self.nonleafindices[j]):
                    tgt[c0:d1] += self.V[j][k] @ (self.U[j][k].T @ src[a0:b1])
                    tgt[a0:b1] += self.U[j][k] @ (self.V[j][k].T @ src[c0:d1])
            return tgt
        return self.W.mul(self.W.mul_t(src))

    def solve(self, src):
        """ldiv!(target, K::KernelHODLR, src) (HODLR/src/baseoverloads.jl:181-192);
        a matrix right-hand side is solved column by column (:257-263)."""
        if self.W is None:
            raise NotFactorizedError("No solves without factorization.")
        return self.W.ldiv_t(self.W.ldiv(src))

    def logdet(self):
        """HODLR/src/baseoverloads.jl:194-209: leaf Cholesky logdets (leaves
        re-factorized, :200) + 2 log|det(I + M^T M X)| per block (:204)."""
        if self.W is None:
            raise NotFactorizedError("No logdet without factorization.")
        out = 0.0
        for L in self.L:
            c = np.linalg.cholesky(L)
            out += 2.0 * np.sum(np.log(np.diag(c)))
        for lev in self.W.nonleafW:
            for W in lev:
                out += 2.0 * math.log(abs(W.det()))
        return out

    @property
    def is_factorized(self):
        return self.W is not None


def _invapply(Wvec, Uvec, Vvec, ldiv):
    """HODLR/src/utils.jl:54-65: V_j gets the first stepsz factors of its
    slot, U_j the next stepsz."""
    stepsz = len(Wvec) // (2 * len(Uvec))
    base = 0
    for k in range(len(Uvec)):
        Vvec[k] = _bd_apply(Wvec[base:base + stepsz], Vvec[k], ldiv)
        Uvec[k] = _bd_apply(Wvec[base + stepsz:base + 2 * stepsz], Uvec[k], ldiv)
        base += 2 * stepsz


def symmetricfactorize(HK):
    """symmetricfactorize! (HODLR/src/symfact.jl:6-50)."""
    LW = [np.linalg.cholesky(L) for L in HK.L]                     # :17
    for lv in range(len(HK.U)):                                    # :21
        _invapply(LW, HK.U[lv], HK.V[lv], _tri_ldiv)
    nonleafW = []
    for lv in range(len(HK.U)):                                    # :28-38
        Ws = [lrsymfact(V, U) for V, U in zip(HK.V[lv], HK.U[lv])]
        nonleafW.append(Ws)
        for lev in range(lv + 1, len(HK.U)):
            _invapply(Ws, HK.U[lev], HK.V[lev], lambda W, t: W.ldiv(t))
    HK.W = FactorHODLR(LW, nonleafW)                                # :41
    HK.U = None                                                    # :45-46
    HK.V = None


def negloglik(HK, y):
    """maxlikfunctions.jl:2-9: 0.5 logdet + 0.5 y' K^-1 y."""
    if not HK.is_factorized:
        raise NotFactorizedError("The HODLR matrix needs to be factorized for this.")
    return 0.5 * HK.logdet() + 0.5 * float(np.dot(y, HK.solve(y)))


# ---------------------------------------------------------------------------
# Derivative HODLR and the stochastic gradient / Fisher / Hessian
# (constructor.jl:66-121, HODLR/src/utils.jl:67-177,
#  stochasticgradient.jl:2-45, stochastichessian.jl:3-147)
# ---------------------------------------------------------------------------

class DerivativeBlock:
    """HODLR/src/utils.jl:67-74."""

    def __init__(self, Kb, dkernel, lndmk):
        self.K1p = kernel_matrix(Kb.kernel, Kb.x1, lndmk)
        self.Kp2 = kernel_matrix(Kb.kernel, lndmk, Kb.x2)
        self.K1pd = kernel_matrix(dkernel, Kb.x1, lndmk)
        self.Kp2d = kernel_matrix(dkernel, lndmk, Kb.x2)


def _cs(S, b):
    return sla.cho_solve(S, b)


def dblock_mul(B, src, S, Sj):
    """HODLR/src/utils.jl:76-82."""
    out = B.K1pd @ _cs(S, B.Kp2 @ src)
    out = out - B.K1p @ _cs(S, Sj @ _cs(S, B.Kp2 @ src))
    out = out + B.K1p @ _cs(S, B.Kp2d @ src)
    return out


def dblock_mul_t(B, src, S, Sj):
    """HODLR/src/utils.jl:84-90."""
    out = B.Kp2.T @ _cs(S, B.K1pd.T @ src)
    out = out - B.Kp2.T @ _cs(S, Sj @ _cs(S, B.K1p.T @ src))
    out = out + B.Kp2d.T @ _cs(S, B.K1p.T @ src)
    return out


def dblock_full(B, S, Sj):
    """HODLR/src/utils.jl:92-98."""
    out = B.K1pd @ _cs(S, B.Kp2)
    out = out - B.K1p @ _cs(S, Sj @ _cs(S, B.Kp2))
    out = out + B.K1p @ _cs(S, B.Kp2d)
    return out


class DerivativeHODLR:
    """Exact derivative of a Nystrom HODLR matrix (constructor.jl:66-95)."""

    def __init__(self, K, dkernel, HK):
        if not HK.nys:
            raise ValueError("This is only valid for Nystrom-block matrices.")
        n = K.shape[0]
        self.lndmk = K.x1[nystrom_landmarks(n, HK.mrnk)]
        self.S = sla.cho_factor(kernel_matrix(K.kernel, self.lndmk, self.lndmk), lower=False)
        self.Sj = kernel_matrix(dkernel, self.lndmk, self.lndmk)
        self.dkernel = dkernel
        self.leafindices = HK.leafindices
        self.nonleafindices = HK.nonleafindices
        dK = KernelMatrix(K.x1, K.x2, dkernel)
        self.L = [dK.sub(*li).full() for li in HK.leafindices]
        self.B = [[DerivativeBlock(K.sub(*x), dkernel, self.lndmk) for x in lev]
                  for lev in HK.nonleafindices]
        self.n = n

    def matvec(self, src):
        """HODLR/src/baseoverloads.jl:216-234."""
        src = np.asarray(src, dtype=np.float64)
        tgt = np.zeros_like(src)
        for j, (a0, b1, c0, d1) in enumerate(self.leafindices):
            tgt[c0:d1] += self.L[j] @ src[c0:d1]
        for j in range(len(self.B)):
            for k, (a0, b1, c0, d1) in enumerate(self.nonleafindices[j]):
                tgt[c0:d1] += dblock_mul_t(self.B[j][k], src[a0:b1], self.S, self.Sj)
                tgt[a0:b1] += dblock_mul(self.B[j][k], src[c0:d1], self.S, self.Sj)
        return tgt

    def full(self):
        """HODLR/src/baseoverloads.jl:338-351."""
        out = np.zeros((self.n, self.n))
        for j, (a0, b1, c0, d1) in enumerate(self.leafindices):
            out[a0:b1, c0:d1] = self.L[j]
        for lev in range(len(self.B)):
            for j, (a0, b1, c0, d1) in enumerate(self.nonleafindices[lev]):
                blk = dblock_full(self.B[lev][j], self.S, self.Sj)
                out[a0:b1, c0:d1] = blk
                out[c0:d1, a0:b1] = blk.T
        return out


def hodlr_trace_apply(HK, HKj, v):
    """stochasticgradient.jl:2-10: t = W^-T v; dot(t, dK t)."""
    t = HK.W.ldiv_t(v)
    return float(np.dot(t, HKj.matvec(t)))


def hodlr_grad_term(K, HK, dkernel, y, vecs):
    """stochasticgradient.jl:14-26."""
    HKj = DerivativeHODLR(K, dkernel, HK)
    prvec = HK.solve(HKj.matvec(HK.solve(y)))
    trtm = sum(hodlr_trace_apply(HK, HKj, v) for v in vecs) / len(vecs)
    return 0.5 * (trtm - float(np.dot(y, prvec)))


def stoch_gradient(K, HK, y, dkernels, vecs, nparams=None):
    """stochasticgradient.jl:30-45."""
    if not HK.is_factorized:
        raise NotFactorizedError("The matrix needs to be factorized for this to work.")
    if nparams is not None and len(dkernels) != nparams:
        raise ValueError("You didn't supply the right number of gradient funs.")
    return np.array([hodlr_grad_term(K, HK, dk, y, vecs) for dk in dkernels])


def profile_negloglik(HK, y):
    """maxlikfunctions.jl:11-19: 0.5 (logdet K + n log(y' K^-1 y))."""
    if not HK.is_factorized:
        raise NotFactorizedError("The HODLR matrix needs to be factorized for this.")
    return 0.5 * (HK.logdet() + len(y) * math.log(float(np.dot(y, HK.solve(y)))))


def hodlr_p_grad_term(K, HK, dkernel, y, vecs):
    """stochasticgradient.jl:48-64: 0.5 (trace - n y'K^-1 dK K^-1 y / y'K^-1 y)."""
    HKj = DerivativeHODLR(K, dkernel, HK)
    prvec = HK.solve(HKj.matvec(HK.solve(y)))
    trtm = sum(hodlr_trace_apply(HK, HKj, v) for v in vecs) / len(vecs)
    return 0.5 * (trtm - len(y) * float(np.dot(y, prvec)) / float(np.dot(y, HK.solve(y))))


def stoch_profile_gradient(K, HK, y, dkernels, vecs):
    """stochasticgradient.jl:68-86."""
    if not HK.is_factorized:
        raise NotFactorizedError("The matrix needs to be factorized for this to work.")
    return np.array([hodlr_p_grad_term(K, HK, dk, y, vecs) for dk in dkernels])


def givesaa(nvec, n, rng):
    """Rademacher SAA vectors (HODLR/src/utils.jl:15-27), from a numpy RNG."""
    return [rng.choice(np.array([-1.0, 1.0]), size=n) for _ in range(nvec)]


def _tr1_sym_diag(HK, DKj, v):
    """stochastichessian.jl:58-66."""
    t1 = HK.W.ldiv_t(v)
    t2 = HK.W.ldiv(DKj.matvec(t1))
    return float(np.dot(t2, t2))


def _tr1_sym_offdiag(HK, DKj, DKk, v):
    """stochastichessian.jl:68-77."""
    t1 = HK.W.ldiv_t(v)
    t2 = HK.W.ldiv(DKj.matvec(t1) + DKk.matvec(t1))
    return float(np.dot(t2, t2))


def stoch_fisher(K, HK, DKs, vecs):
    """stochastichessian.jl:91-115."""
    if not HK.is_factorized:
        raise NotFactorizedError("The matrix needs to be factorized.")
    p = len(DKs)
    out = np.zeros((p, p))
    for j in range(p):
        out[j, j] = 0.5 * sum(_tr1_sym_diag(HK, DKs[j], v) for v in vecs) / len(vecs)
    for j in range(p):
        for k in range(j + 1, p):
            tmp1 = 0.5 * sum(_tr1_sym_offdiag(HK, DKs[j], DKs[k], v) for v in vecs) / len(vecs)
            out[j, k] = 0.5 * (tmp1 - out[j, j] - out[k, k])
    return np.triu(out) + np.triu(out, 1).T


class SecondDerivativeBlock:
    """HODLR/src/utils.jl:100-105."""

    def __init__(self, Kb, djk, lndmk):
        self.K1pjk = kernel_matrix(djk, Kb.x1, lndmk)
        self.Kp2jk = kernel_matrix(djk, lndmk, Kb.x2)


def sdblock_mul(Bj, Bk, Bjk, src, Sp, Spj, Spk, Spjk):
    """HODLR/src/utils.jl:110-132."""
    c = lambda b: _cs(Sp, b)
    out = Bjk.K1pjk @ c(Bj.Kp2 @ src)
    out = out - Bj.K1pd @ c(Spk @ c(Bj.Kp2 @ src))
    out = out + Bj.K1pd @ c(Bk.Kp2d @ src)
    out = out + Bk.K1pd @ c(Spj @ c(Bj.Kp2 @ src))
    out = out - Bk.K1p @ c(Spk @ c(Spj @ c(Bk.Kp2 @ src)))
    out = out + Bk.K1p @ c(Spjk @ c(Bk.Kp2 @ src))
    out = out - Bk.K1p @ c(Spj @ c(Spk @ c(Bk.Kp2 @ src)))
    out = out + Bk.K1p @ c(Spj @ c(Bj.Kp2d @ src))
    out = out + Bk.K1pd @ c(Bk.Kp2 @ src)
    out = out - Bk.K1p @ c(Spk @ c(Bk.Kp2 @ src))
    out = out + Bk.K1p @ c(Bjk.Kp2jk @ src)
    return out


def sdblock_mul_t(Bj, Bk, Bjk, src, Sp, Spj, Spk, Spjk):
    """HODLR/src/utils.jl:134-155."""
    c = lambda b: _cs(Sp, b)
    out = Bj.Kp2.T @ c(Bjk.K1pjk.T @ src)
    out = out - Bj.Kp2.T @ c(Spk @ c(Bj.K1pd.T @ src))
    out = out + Bk.Kp2d.T @ c(Bk.K1pd.T @ src)
    out = out + Bj.Kp2.T @ c(Spj @ c(Bk.K1pd.T @ src))
    out = out - Bj.Kp2.T @ c(Spj @ c(Spk @ c(Bk.K1p.T @ src)))
    out = out + Bj.Kp2.T @ c(Spjk @ c(Bk.K1p.T @ src))
    out = out - Bj.Kp2.T @ c(Spk @ c(Spj @ c(Bk.K1p.T @ src)))
    out = out + Bk.Kp2d.T @ c(Spj @ c(Bk.K1p.T @ src))
    out = out + Bj.Kp2d.T @ c(Bk.K1pd.T @ src)
    out = out - Bj.Kp2d.T @ c(Spk @ c(Bk.K1p.T @ src))
    out = out + Bjk.Kp2jk.T @ c(Bk.K1p.T @ src)
    return out


def sdblock_full_exact(Bj, Bk, Bjk, Sp, Spj, Spk, Spjk):
    """The exact second parameter derivative of a Nystrom block
    K1 S^-1 K2 (K1 = K(rows,P), S = K(P,P), K2 = K(P,cols)):
    the eleven terms of d_k d_j (K1 S^-1 K2).  The reference's
    SDBlock_mul (HODLR/src/utils.jl:110-132) flips the signs of the terms
    from d_k(-K1 S^-1 S_j S^-1 K2) and uses K2 instead of K2_j in two
    terms; this version is pinned by finite differences in the tests."""
    c = lambda b: _cs(Sp, b)
    K1, K2 = Bk.K1p, Bj.Kp2
    K1j, K1k, K2j, K2k = Bj.K1pd, Bk.K1pd, Bj.Kp2d, Bk.Kp2d
    out = Bjk.K1pjk @ c(K2)
    out = out - K1j @ c(Spk @ c(K2))
    out = out + K1j @ c(K2k)
    out = out - K1k @ c(Spj @ c(K2))
    out = out + K1 @ c(Spk @ c(Spj @ c(K2)))
    out = out - K1 @ c(Spjk @ c(K2))
    out = out + K1 @ c(Spj @ c(Spk @ c(K2)))
    out = out - K1 @ c(Spj @ c(K2k))
    out = out + K1k @ c(K2j)
    out = out - K1 @ c(Spk @ c(K2j))
    out = out + K1 @ c(Bjk.Kp2jk)
    return out


class SecondDerivative:
    """Leaves + blocks + S_jk of an exact second derivative
    (constructor.jl:100-121, stochastichessian.jl:38-41)."""

    def __init__(self, K, djk, HK):
        n = K.shape[0]
        lndmk = K.x1[nystrom_landmarks(n, HK.mrnk)]
        d2K = KernelMatrix(K.x1, K.x2, djk)
        self.L = [d2K.sub(*li).full() for li in HK.leafindices]
        self.B = [[SecondDerivativeBlock(d2K.sub(*x), djk, lndmk) for x in lev]
                  for lev in HK.nonleafindices]
        self.Sjk = kernel_matrix(djk, lndmk, lndmk)


def deriv2mul(DKj, DKk, D2, src, exact=False):
    """HODLR/src/utils.jl:157-177 (exact=False, the reference's block terms);
    exact=True uses sdblock_full_exact for the off-diagonal blocks."""
    src = np.asarray(src, dtype=np.float64)
    tgt = np.zeros_like(src)
    for j, (a0, b1, c0, d1) in enumerate(DKj.leafindices):
        tgt[c0:d1] += D2.L[j] @ src[c0:d1]
    for j in range(len(DKj.B)):
        for k, (a0, b1, c0, d1) in enumerate(DKj.nonleafindices[j]):
            args = (DKj.B[j][k], DKk.B[j][k], D2.B[j][k])
            if exact:
                blk = sdblock_full_exact(*args, DKj.S, DKj.Sj, DKk.Sj, D2.Sjk)
                tgt[c0:d1] += blk.T @ src[a0:b1]
                tgt[a0:b1] += blk @ src[c0:d1]
            else:
                tgt[c0:d1] += sdblock_mul_t(*args, src[a0:b1], DKj.S, DKj.Sj, DKk.Sj, D2.Sjk)
                tgt[a0:b1] += sdblock_mul(*args, src[c0:d1], DKj.S, DKj.Sj, DKk.Sj, D2.Sjk)
    return tgt


def stoch_hessian(K, HK, y, d1kernels, d2kernels, vecs, exact=False):
    """stochastichessian.jl:117-147 (non-profile).  d2kernels[j][k-j] is the
    (j, k) second-derivative kernel, or None for an identically zero one
    (the ZeroFunction shortcut, :135-143)."""
    if not HK.is_factorized:
        raise NotFactorizedError("The matrix needs to be factorized.")
    p = len(d1kernels)
    HKsd = HK.solve(y)
    DKs = [DerivativeHODLR(K, dk, HK) for dk in d1kernels]
    out = -stoch_fisher(K, HK, DKs, vecs)
    for j in range(p):
        for k in range(j, p):
            djk = d2kernels[j][k - j]
            if djk is not None:
                D2 = SecondDerivative(K, djk, HK)
                o = -HK.solve(DKs[k].matvec(HK.solve(DKs[j].matvec(HKsd))))   # :9-12
                o = o - HK.solve(DKs[j].matvec(HK.solve(DKs[k].matvec(HKsd))))
                o = o + HK.solve(deriv2mul(DKs[j], DKs[k], D2, HKsd, exact=exact))
                sterm = float(np.dot(y, o))
                t2 = 0.0
                for v in vecs:                                                  # :79-89
                    t1 = HK.W.ldiv_t(v)
                    t2 += float(np.dot(t1, deriv2mul(DKs[j], DKs[k], D2, t1, exact=exact)))
                t2 /= len(vecs)
                out[j, k] += 0.5 * t2 - 0.5 * sterm
            else:
                o = -HK.solve(DKs[k].matvec(HK.solve(DKs[j].matvec(HKsd))))   # :50-56
                o = o - HK.solve(DKs[j].matvec(HK.solve(DKs[k].matvec(HKsd))))
                out[j, k] -= 0.5 * float(np.dot(y, o))
    return np.triu(out) + np.triu(out, 1).T


# ---------------------------------------------------------------------------
# RKernelHODLR: recursive HODLR for (possibly non-symmetric) matvecs
# (constructor.jl:130-161, structstypes.jl:22-27, baseoverloads.jl:285-305)
# ---------------------------------------------------------------------------

class UVt:
    """UVt (structstypes.jl): block = U V^T; U*(V'v) (baseoverloads.jl:289)."""

    def __init__(self, U, V):
        self.U, self.V = U, V

    def mul(self, v):
        return self.U @ (self.V.T @ v)


class RKernelHODLR:
    """constructor.jl:130-161.  A11/A22 are dense matrices at the last level,
    else RKernelHODLR; A12 = ACA(K12), A21 = ACA(K21).  Rows and columns
    split at floor(len/2) (the first half is the smaller one)."""

    def __init__(self, A11, A22, A12, A21):
        self.A11, self.A22, self.A12, self.A21 = A11, A22, A12, A21

    @property
    def shape(self):
        s11 = self.A11.shape
        s22 = self.A22.shape
        return (s11[0] + s22[0], s11[1] + s22[1])

    def mul(self, v):
        """baseoverloads.jl:299-303."""
        n11 = self.A11.shape[1]
        v1, v2 = v[:n11], v[n11:]
        m11 = self.A11 @ v1 if isinstance(self.A11, np.ndarray) else self.A11.mul(v1)
        m22 = self.A22 @ v2 if isinstance(self.A22, np.ndarray) else self.A22.mul(v2)
        return np.concatenate([m11 + self.A12.mul(v2), self.A21.mul(v1) + m22])

    def ranks(self):
        """[(rank K12, rank K21) per node] per depth, root first, nodes left to
        right (the order the library reports)."""
        out = []
        level = [self]
        while level and isinstance(level[0], RKernelHODLR):
            out.append([(b.A12.U.shape[1], b.A21.U.shape[1]) for b in level])
            nxt = []
            for b in level:
                if isinstance(b.A11, RKernelHODLR):
                    nxt += [b.A11, b.A22]
            level = nxt
        return out


def rkernelhodlr(K, tol, maxrank=0, lvl=None):
    """RKernelHODLR(K, tol, maxrank, lvl) (constructor.jl:130-161): returns a
    dense matrix when the level is 0 (:141)."""
    if lvl is None:
        lvl = LogLevel(7)
    if isinstance(lvl, LogLevel):   # :133-136
        lv = int(math.floor(math.log2(min(K.shape)))) - lvl.lv
        return rkernelhodlr(K, tol, maxrank, FixedLevel(lv))
    if lvl.lv == 0:                 # :141
        return K.full()
    len1, len2 = K.shape
    mpt1, mpt2 = len1 // 2, len2 // 2
    K11 = K.sub(0, mpt1, 0, mpt2)
    K22 = K.sub(mpt1, len1, mpt2, len2)
    K12 = K.sub(0, mpt1, mpt2, len2)
    K21 = K.sub(mpt1, len1, 0, mpt2)
    A12 = UVt(*aca(K12, tol, maxrank))
    A21 = UVt(*aca(K21, tol, maxrank))
    if lvl.lv == 1:
        return RKernelHODLR(K11.full(), K22.full(), A12, A21)
    A11 = rkernelhodlr(K11, tol, maxrank, FixedLevel(lvl.lv - 1))
    A22 = rkernelhodlr(K22, tol, maxrank, FixedLevel(lvl.lv - 1))
    return RKernelHODLR(A11, A22, A12, A21)
