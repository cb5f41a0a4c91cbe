"""Independent restatements used to pin the CPU oracle (dense algebra,
brute force, textbook routines).  Nothing here calls the oracle."""
from __future__ import annotations

import numpy as np


def ring_columns(S_all, Y_all, head1, col, m):
    """S, Y (n x col) in ring order from the 1-based head slot."""
    slots = [(head1 - 1 + j) % m for j in range(col)]
    return S_all[:, slots], Y_all[:, slots]


def middle_matrix(S, Y, theta):
    """M = [[-D, L'], [L, theta S'S]]^{-1} (Byrd, Nocedal, Schnabel 1994)."""
    SY = S.T @ Y
    D = np.diag(np.diag(SY))
    L = np.tril(SY, -1)
    K = np.block([[-D, L.T], [L, theta * (S.T @ S)]])
    return np.linalg.inv(K)


def dense_B(S, Y, theta, n):
    """Compact L-BFGS matrix B = theta I - W M W', W = [Y, theta S]."""
    if S.shape[1] == 0:
        return theta * np.eye(n)
    W = np.hstack([Y, theta * S])
    return theta * np.eye(n) - W @ middle_matrix(S, Y, theta) @ W.T


def cauchy_bruteforce(x, g, l, u, nbd, iwhere_in, B):
    """Generalized Cauchy point from its definition: the first local
    minimiser of m(z) = g'z + z'Bz/2 along z(t) = P(x - t g) - x, each
    segment's derivative and curvature formed densely from scratch.
    Returns (xcp, iwhere, passed) where passed is the number of breakpoints
    crossed."""
    n = len(x)
    d = np.zeros(n)
    iw = iwhere_in.copy()
    tb = np.full(n, np.inf)
    for i in range(n):
        neg = -g[i]
        if iw[i] not in (3, -1):
            tl = x[i] - l[i] if nbd[i] <= 2 else 0.0
            tu = u[i] - x[i] if nbd[i] >= 2 else 0.0
            xlower = nbd[i] <= 2 and tl <= 0
            xupper = nbd[i] >= 2 and tu <= 0
            iw[i] = 0
            if xlower:
                if neg <= 0:
                    iw[i] = 1
            elif xupper:
                if neg >= 0:
                    iw[i] = 2
            elif abs(neg) <= 0:
                iw[i] = -3
        if iw[i] in (0, -1):
            d[i] = neg
            if nbd[i] in (1, 2) and neg < 0:
                tb[i] = (x[i] - l[i]) / (-neg)
            elif nbd[i] >= 2 and neg > 0:
                tb[i] = (u[i] - x[i]) / neg
    order = [i for i in np.argsort(tb, kind="stable") if np.isfinite(tb[i])]
    t_prev = 0.0
    z = np.zeros(n)
    dcur = d.copy()
    passed = 0
    tstar = 0.0
    for k in range(len(order) + 1):
        f1 = g @ dcur + z @ (B @ dcur)
        f2 = dcur @ (B @ dcur)
        t_next = tb[order[k]] if k < len(order) else np.inf
        if not np.any(dcur):
            tstar = t_prev
            break
        dtm = -f1 / f2
        if dtm < t_next - t_prev:
            tstar = t_prev + max(dtm, 0.0)
            break
        # cross breakpoint k
        i = order[k]
        z = z + (t_next - t_prev) * dcur
        z[i] = (u[i] if dcur[i] > 0 else l[i]) - x[i]
        iw[i] = 2 if dcur[i] > 0 else 1
        dcur[i] = 0.0
        t_prev = t_next
        passed += 1
    xcp = x + z + (tstar - t_prev) * dcur
    # passed variables sit exactly on their bound
    crossed = (d != 0) & (dcur == 0)
    xcp[crossed] = np.where(d[crossed] > 0, u[crossed], l[crossed])
    return xcp, iw, passed


def subspace_dense(x, xcp, g, l, u, nbd, free, B):
    """L-BFGS-B 2.1 subspace step: Newton step on the free variables of the
    model at the Cauchy point, then backtracked into the box with the
    limiting variable pinned (Byrd et al. 1995, section 5.1)."""
    F = np.where(free)[0]
    r = g + B @ (xcp - x)
    BFF = B[np.ix_(F, F)]
    dF = -np.linalg.solve(BFF, r[F])
    alpha, ibd = 1.0, -1
    for j, k in enumerate(F):
        dk = dF[j]
        cand = None
        if nbd[k] != 0:
            if dk < 0 and nbd[k] <= 2:
                t2 = l[k] - xcp[k]
                cand = 0.0 if t2 >= 0 else (t2 / dk if dk * alpha < t2 else None)
            elif dk > 0 and nbd[k] >= 2:
                t2 = u[k] - xcp[k]
                cand = 0.0 if t2 <= 0 else (t2 / dk if dk * alpha > t2 else None)
        if cand is not None and cand < alpha:
            alpha, ibd = cand, j
    z = xcp.copy()
    if alpha < 1 and ibd >= 0:
        k = F[ibd]
        z[k] = u[k] if dF[ibd] > 0 else l[k]
        dF[ibd] = 0.0
    z[F] = z[F] + alpha * dF
    return z, alpha


def projected_gauss_seidel_penta(p, tol=1e-15, max_sweeps=200000):
    """Unique minimiser of the strictly convex pentadiagonal box QP by
    projected Gauss-Seidel over the three colour classes i mod 3."""
    n = p.n
    x = np.zeros(n)
    a0, a1, a2, b = p.a0, p.a1, p.a2, p.b
    for _ in range(max_sweeps):
        xo = x.copy()
        for c in range(3):
            idx = np.arange(c, n, 3)
            s = np.zeros(len(idx))
            for off, coef in ((1, a1), (2, a2)):
                up = idx + off
                ok = up < n
                s[ok] += coef[idx[ok]] * x[up[ok]]
                dn = idx - off
                ok = dn >= 0
                s[ok] += coef[dn[ok]] * x[dn[ok]]
            x[idx] = np.clip((b[idx] - s) / a0[idx], p.l[idx], p.u[idx])
        if np.max(np.abs(x - xo)) < tol:
            break
    return x


def projected_gauss_seidel_grid(p, tol=1e-15, max_sweeps=200000):
    """Same for the 5-point grid quadratic (red-black ordering)."""
    nx, ny = p.nx, p.ny
    v = np.zeros((ny, nx))
    b = p.b.reshape(ny, nx) * p.h * p.h
    lo = p.l.reshape(ny, nx)
    hi = p.u.reshape(ny, nx)
    ii, jj = np.meshgrid(np.arange(nx), np.arange(ny))
    for _ in range(max_sweeps):
        vo = v.copy()
        for colr in (0, 1):
            pad = np.zeros((ny + 2, nx + 2))
            pad[1:-1, 1:-1] = v
            nb = pad[1:-1, :-2] + pad[1:-1, 2:] + pad[:-2, 1:-1] + pad[2:, 1:-1]
            upd = np.clip((b + nb) / 4.0, lo, hi)
            mask = ((ii + jj) % 2) == colr
            v = np.where(mask, upd, v)
        if np.max(np.abs(v - vo)) < tol:
            break
    return v.reshape(-1)
