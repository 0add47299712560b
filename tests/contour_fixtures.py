"""Closed-form contour functions used by the reference's integral tests.

Restates the inputs of tests/set_up_g.h:13-243 (set_up_G_herm and set_up_G:
2x2 functions of t, t' and tau with explicit "cc" partners), so that the
golden scalars of tests/integrals-test.cpp can be reproduced.  tau runs over
the Gauss-Lobatto-Legendre nodes, tau_m = beta (x_m + 1)/2.
"""
import numpy as np

from nedyson_amd.green import GreenFunc


def _taus(x, beta):
    return beta * (np.asarray(x) + 1) / 2.0


def set_up_G_herm(x, Nt, Ntau, beta, dt):
    """tests/set_up_g.h:13-109.  Returns (G, S)."""
    G = GreenFunc(Nt, Ntau, 2, -1)
    S = GreenFunc(Nt, Ntau, 2, -1)
    i = 1j
    taus = _taus(x, beta)
    for t in range(Nt + 1):
        for tp in range(t + 1):
            r = S.retptr(t, tp).reshape(-1)
            r[0] = i * (dt ** 4 * t * t * tp * tp)
            r[1] = i * (dt * dt * t * tp)
            r[2] = i * (dt * dt * t * tp)
            r[3] = i * (dt ** 6 * t ** 3 * tp ** 3)
    for tp in range(Nt + 1):
        for t in range(tp + 1):
            g = G.lesptr(t, tp).reshape(-1)
            g[0] = i * np.exp(i * (t - tp) * dt)
            g[1] = (t - tp) * dt
            g[2] = (t - tp) * dt
            g[3] = i * np.exp(3 * i * (t - tp) * dt)
            s = S.lesptr(t, tp).reshape(-1)
            s[0] = i * (dt * dt * t * tp)
            s[1] = i * np.exp(2 * i * (dt * (t - tp)))
            s[2] = i * np.exp(2 * i * (dt * (t - tp)))
            s[3] = dt * (t - tp)
    for t in range(Nt + 1):
        for tp in range(t + 1):
            g = G.retptr(t, tp).reshape(-1)
            g[0] = -i * np.cos(3 * dt * (t - tp))
            g[1] = -np.sin(2 * dt * (t - tp))
            g[2] = -np.sin(2 * dt * (t - tp))
            g[3] = -i * np.cos(dt * (t - tp))
    for t in range(Nt + 1):
        for m in range(Ntau + 1):
            tau = taus[m]
            g = G.tv[t, m].reshape(-1)
            g[0] = tau * tau
            g[1] = i * (dt * t)
            g[2] = np.exp(i * tau)
            g[3] = dt * dt * t * t - i * tau
            s = S.tv[t, m].reshape(-1)
            s[0] = dt * t * tau
            s[1] = dt * dt * t * t * tau
            s[2] = np.sin(dt * t)
            s[3] = i * (dt * t - tau)
    for m in range(Ntau + 1):
        tau = taus[m]
        G.mat[m].reshape(-1)[:] = [tau, i * tau, tau * tau, np.exp(i * tau)]
        S.mat[m].reshape(-1)[:] = [tau * i, tau, np.sin(tau * tau), np.exp(i * tau)]
    return G, S


def set_up_G(x, Nt, Ntau, beta, dt):
    """tests/set_up_g.h:115-243.  Returns (G, Gcc, S, Scc)."""
    G, Gcc, S, Scc = (GreenFunc(Nt, Ntau, 2, -1) for _ in range(4))
    i = 1j
    Ssig = Gsig = -1
    taus = _taus(x, beta)
    cj = np.conj
    for t in range(Nt + 1):
        for tp in range(t + 1):
            s = S.retptr(t, tp).reshape(-1)
            s[0] = dt ** 3 * t * tp * tp
            s[1] = i * (dt * t) - dt * tp
            s[2] = np.sin(dt * t)
            s[3] = np.cos(dt * tp)
            c = Scc.retptr(t, tp).reshape(-1)
            c[0] = -cj(dt ** 3 * tp * t * t)
            c[2] = -cj(i * (dt * tp) - dt * t)
            c[1] = -cj(np.sin(dt * tp))
            c[3] = -cj(np.cos(dt * t))
    for tp in range(Nt + 1):
        for t in range(tp + 1):
            s = S.lesptr(t, tp).reshape(-1)
            s[0] = dt ** 3 * t * t * tp
            s[1] = dt * t + 2. * dt * tp
            s[2] = np.exp(i * (dt * t))
            s[3] = np.exp(i * (t - tp) * dt)
            c = Scc.lesptr(t, tp).reshape(-1)
            c[0] = -cj(dt ** 3 * tp * tp * t)
            c[2] = -cj(dt * tp + 2. * dt * t)
            c[1] = -cj(np.exp(i * (dt * tp)))
            c[3] = -cj(np.exp(i * (tp - t) * dt))
    for t in range(Nt + 1):
        for m in range(Ntau + 1):
            tau = taus[m]
            s = S.tv[t, m].reshape(-1)
            s[0] = dt * t * tau
            s[1] = np.cos(dt * t * tau)
            s[2] = np.exp(i * tau)
            s[3] = dt * t - tau
            c = Scc.tv[t, m].reshape(-1)
            c[0] = -Ssig * cj(i - dt * t * (beta - tau))
            c[2] = -Ssig * cj(np.sin(2 * (beta - tau)))
            c[1] = -Ssig * cj(3 * (beta - tau) + dt * t)
            c[3] = -Ssig * cj(dt * t * i)
    for m in range(Ntau + 1):
        tau = taus[m]
        S.mat[m].reshape(-1)[:] = [1, tau, 2, i * tau * tau]
    for t in range(Nt + 1):
        for tp in range(t + 1):
            g = G.retptr(t, tp).reshape(-1)
            g[0] = dt * (t + tp)
            g[1] = dt * (t + 2 * tp)
            g[2] = dt * (t + 3 * tp)
            g[3] = i * np.cos(dt * tp)
            c = Gcc.retptr(t, tp).reshape(-1)
            c[0] = -cj(dt * (tp + t))
            c[2] = -cj(dt * (tp + 2 * t))
            c[1] = -cj(dt * (tp + 3 * t))
            c[3] = -cj(i * np.cos(dt * t))
    for tp in range(Nt + 1):
        for t in range(tp + 1):
            g = G.lesptr(t, tp).reshape(-1)
            g[0] = t * dt * i
            g[1] = tp * dt * i
            g[2] = dt * (t - tp)
            g[3] = 4
            c = Gcc.lesptr(t, tp).reshape(-1)
            c[0] = -cj(tp * dt * i)
            c[2] = -cj(t * dt * i)
            c[1] = -cj(dt * (tp - t))
            c[3] = -cj(4)
    for t in range(Nt + 1):
        for m in range(Ntau + 1):
            tau = taus[m]
            g = G.tv[t, m].reshape(-1)
            g[0] = dt * t
            g[1] = i * 2.
            g[2] = tau
            g[3] = tau * tau
            c = Gcc.tv[t, m].reshape(-1)
            c[0] = -Gsig * cj(dt * t * (beta - tau))
            c[2] = -Gsig * cj(i * (beta - tau))
            c[1] = -Gsig * cj(i - (beta - tau))
            c[3] = -Gsig * cj(dt * t + (beta - tau))
    for m in range(Ntau + 1):
        tau = taus[m]
        G.mat[m].reshape(-1)[:] = [np.exp(i * tau), np.exp(i * tau), np.cos(tau), i]
    return G, Gcc, S, Scc
