"""Differentiable FP64 restatement of one SD-ADER step with frozen blending
coefficients, in PyTorch on the CPU.

TEST INFRASTRUCTURE ONLY.  This module is the autograd checker of the HIP
adjoint (``spd_adj_backward``): only ``tests/`` may import it.  The product
package ``spd_amd`` never imports, links or calls anything under ``oracle/``.

It restates ``SDADER_Simulator.perform_update`` (``src/sdader_simulator.py:
361-378``) exactly as the NumPy oracle (``oracle/spd_oracle.py``) does -- the
same functions, the same reference lines and quirks (Q1, Q7, Q8, Q14, Q15,
Q19) -- but without in-place updates, so that ``torch.autograd`` can run
through it.  The trouble detection is replaced by given face coefficients
(``affected_faces`` per substep and dimension, ``src/trouble_detection.py:
80-87``) and dt is a given number: the map differentiated is the step with
flags and dt frozen.  Geometry, operators and options are read from an
``OracleSimulator``.  The minmod limiter is written without its division
(same values; the reference's ``(R/L) * L`` has an undefined derivative at
L = 0).
"""
from __future__ import annotations

import numpy as np
import torch

from oracle.spd_oracle import _sl

F64 = torch.float64


def _t(x):
    return torch.as_tensor(np.asarray(x, dtype=np.float64), dtype=F64)


def cs2(P, rho, gamma, min_c2):
    """``src/hydro.py:3-25``."""
    c2 = gamma * P / rho
    return torch.where(c2 > min_c2, c2, c2.new_tensor(min_c2))


def prims(U, vels, P, gamma):
    """``src/hydro.py:48-86``."""
    W = list(U.unbind(0))
    K = U[0] * 0
    for v in vels:
        W[v] = U[v] / U[0]
        K = K + W[v] ** 2
    K = K * (0.5 * U[0])
    W[P] = (gamma - 1) * (U[P] - K)
    return torch.stack(W)


def cons(W, vels, P, gamma):
    """``src/hydro.py:88-125``."""
    U = list(W.unbind(0))
    K = W[0] * 0
    for v in vels:
        U[v] = W[v] * W[0]
        K = K + W[v] ** 2
    K = K * (0.5 * W[0])
    U[P] = W[P] / (gamma - 1) + K
    return torch.stack(U)


def euler_flux(W, vels, P, gamma):
    """``src/hydro.py:127-170`` (kinetic sum over ``vels[::-1]``, Q14)."""
    F = list(W.unbind(0))
    K = W[0] * 0
    v1 = vels[0]
    m = None
    for v in vels[::-1]:
        m = W[0] * W[v]
        K = K + m * W[v]
        F[v] = m * W[v1]
    E = W[P] / (gamma - 1) + 0.5 * K
    F[0] = m
    F[v1] = m * W[v1] + W[P]
    F[P] = W[v1] * (E + W[P])
    return torch.stack(F)


def riemann_llf(WL, WR, UL, UR, vels, P, gamma, min_c2):
    """``src/riemann_solver.py:31-65``."""
    v1 = vels[0]
    FL = euler_flux(WL, vels, P, gamma)
    FR = euler_flux(WR, vels, P, gamma)
    cL = torch.sqrt(cs2(WL[P], WL[0], gamma, min_c2)) + torch.abs(WL[v1])
    cR = torch.sqrt(cs2(WR[P], WR[0], gamma, min_c2)) + torch.abs(WR[v1])
    cmax = torch.where(cL > cR, cL, cR)[None]
    return 0.5 * (FR + FL) - 0.5 * cmax * (UR - UL)


def riemann_hllc(WL, WR, UL, UR, vels, P, gamma, min_c2, low_mach=False):
    """``src/riemann_solver.py:67-212`` (Q8; lhllc with ``low_mach``)."""
    v1 = vels[0]
    cL = torch.sqrt(cs2(WL[P], WL[0], gamma, min_c2)) + torch.abs(WL[v1])
    cR = torch.sqrt(cs2(WR[P], WR[0], gamma, min_c2)) + torch.abs(WR[v1])
    cmax = torch.where(cL > cR, cL, cR)
    vL, vR = WL[v1], WR[v1]
    sL = torch.where(vL < vR, vL, vR) - cmax
    sR = torch.where(vL > vR, vL, vR) + cmax
    rcL = WL[0] * (vL - sL)
    rcR = WR[0] * (sR - vR)
    vs = (rcR * vR + rcL * vL + (WL[P] - WR[P])) / (rcR + rcL)
    jump = rcL * rcR * (vL - vR)
    if low_mach:
        mach = torch.minimum(torch.ones_like(cmax), torch.maximum(torch.abs(vL), torch.abs(vR)) / cmax)
        jump = mach * (2 - mach) * rcL * rcR * (vL - vR)
    Ps = (rcR * WL[P] + rcL * WR[P] + jump) / (rcR + rcL)
    rsL = WL[0] * (sL - vL) / (sL - vs)
    esL = ((sL - vL) * UL[P] - WL[P] * vL + Ps * vs) / (sL - vs)
    rsR = WR[0] * (sR - vR) / (sR - vs)
    esR = ((sR - vR) * UR[P] - WR[P] * vR + Ps * vs) / (sR - vs)
    r = torch.where(sL > 0, WL[0], torch.where(vs > 0, rsL, torch.where(sR > 0, rsR, WR[0])))
    v = torch.where(sL > 0, vL, torch.where(vs > 0, vs, torch.where(sR > 0, vs, vR)))
    p = torch.where(sL > 0, WL[P], torch.where(vs > 0, Ps, torch.where(sR > 0, Ps, WR[P])))
    e = torch.where(sL > 0, UL[P], torch.where(vs > 0, esL, torch.where(sR > 0, esR, UR[P])))
    F = list(WL.unbind(0))
    F[0] = r * v
    F[v1] = F[0] * v + p
    F[P] = v * (e + p)
    for vt in vels[1:]:
        F[vt] = F[0] * torch.where(vs > 0, WL[vt], WR[vt])
    return torch.stack(F)


def riemann(name, ML, MR, vels, P, gamma, min_c2, prims_in):
    """``src/riemann_solver.py:10-28``."""
    if prims_in:
        WL, WR = ML, MR
        UL, UR = cons(WL, vels, P, gamma), cons(WR, vels, P, gamma)
    else:
        UL, UR = ML, MR
        WL, WR = prims(UL, vels, P, gamma), prims(UR, vels, P, gamma)
    if name == "llf":
        return riemann_llf(WL, WR, UL, UR, vels, P, gamma, min_c2)
    return riemann_hllc(WL, WR, UL, UR, vels, P, gamma, min_c2, low_mach=(name == "lhllc"))


def minmod(dL, dR):
    """``src/muscl.py:12-35`` (Q7) without the division."""
    pick = torch.where(torch.abs(dR) < torch.abs(dL), dR, dL)
    return torch.where(dL * dR > 0, pick, dL * 0)


def moncen(dL, dR, xL, xR, xM):
    """``src/muscl.py:37-59``."""
    dC = (xL * dL + xR * dR) / (xL + xR)
    slope = torch.minimum(torch.abs(2 * dL * xL / xM), torch.abs(2 * dR * xR / xM))
    slope = torch.sign(dC) * torch.minimum(slope, torch.abs(dC))
    return torch.where(dL * dR >= 0, slope, slope * 0)


def _contract(A, M, axis):
    """``src/transforms.py:3-25``."""
    out = torch.einsum("fs,...s->...f", M, torch.movedim(A, axis, -1))
    return torch.movedim(out, -1, axis)


def _flip(X, k):
    """Negate variable k (axis 0)."""
    return torch.cat([X[:k], -X[k:k + 1], X[k + 1:]])


class FrozenStep:
    """One step of ``o`` (an ``OracleSimulator``) as a differentiable map of
    ``U_sp`` with dt and the face coefficients given."""

    def __init__(self, o):
        self.o = o
        ops = o.ops
        self.S2F, self.D2S = _t(ops.sp_to_fp), _t(ops.dfp_to_sp)
        self.S2C, self.C2S = _t(ops.sp_to_cv), _t(ops.cv_to_sp)
        self.INV, self.W_TP = _t(ops.invader), ops.w_tp

    # -- layouts -------------------------------------------------------------
    def to_fv(self, M):
        o = self.o
        nd = o.ndim
        lead = M.dim() - 2 * nd
        E = M.shape[lead:lead + nd]
        L = M.shape[lead + nd:]
        perm = list(range(lead))
        for k in range(nd):
            perm += [lead + k, lead + nd + k]
        shape = list(M.shape[:lead]) + [E[k] * L[k] for k in range(nd)]
        return M.permute(*perm).reshape(shape)

    def to_sd(self, M):
        o = self.o
        nd = o.ndim
        lead = M.dim() - nd
        shape = list(M.shape[:lead])
        for k in range(nd):
            shape += [o.N[nd - 1 - k], o.n]
        perm = list(range(lead)) + [lead + 2 * k for k in range(nd)] + \
            [lead + 2 * k + 1 for k in range(nd)]
        return M.reshape(shape).permute(*perm)

    def _full(self, M, A):
        nd = self.o.ndim
        if nd == 1:
            return torch.einsum("il,uxl->uxi", A, M)
        if nd == 2:
            return torch.einsum("jm,il,uyxml->uyxji", A, A, M)
        return torch.einsum("kn,jm,il,uzyxnml->uzyxkji", A, A, A, M)

    def _bshape(self, a, arr, lead=1):
        o = self.o
        shape = [1] * (lead + o.ndim)
        shape[lead + o.ndim - 1 - a] = arr.size
        return _t(arr).reshape(shape)

    # -- predictor (src/sdader_simulator.py:148-240) -----------------------------
    def solve_faces(self, M):
        o = self.o
        nd, lead = o.ndim, 2
        Ffp = []
        for a in range(nd):
            vels = [int(v) for v in np.roll(o.vels, -a)]
            la = o._local_axis(a, lead)
            ea = o._elem_axis(a, lead)
            Mfp = _contract(M, self.S2F, la)
            F = euler_flux(prims(Mfp, vels, o.P, o.gamma), vels, o.P, o.gamma)
            left = Mfp.select(la, 0)
            right = Mfp.select(la, o.nf - 1)
            Na = o.N[a]
            bl, br = o.BC[a]
            if bl == "periodic":
                gL = right.narrow(ea, Na - 1, 1)
            else:
                gL = left.narrow(ea, 0, 1)
                if bl == "reflective":
                    gL = _flip(gL, a + 1)
            if br == "periodic":
                gR = left.narrow(ea, 0, 1)
            else:
                gR = right.narrow(ea, Na - 1, 1)
                if br == "reflective":
                    gR = _flip(gR, a + 1)
            ML = torch.cat([gL, right], dim=ea)
            MR = torch.cat([left, gR], dim=ea)
            Fr = riemann(o.rs_sd, ML, MR, vels, o.P, o.gamma, o.min_c2, prims_in=False)
            lo = Fr.narrow(ea, 0, Na)
            hi = Fr.narrow(ea, 1, Na)
            Fm = torch.movedim(F, la, -1)
            Fm = torch.cat([lo.unsqueeze(-1), Fm[..., 1:-1], hi.unsqueeze(-1)], dim=-1)
            Ffp.append(torch.movedim(Fm, -1, la))
        return Ffp

    def ader_dudt(self, Ffp):
        o = self.o
        d = None
        for a in range(o.ndim):
            term = _contract(Ffp[a], self.D2S, o._local_axis(a, 2)) / o.h[a]
            d = term if d is None else d + term
        return d

    def ader_predictor(self, U_sp, dt):
        o = self.o
        Uad = U_sp[:, None].expand((o.V, o.T) + tuple(U_sp.shape[1:]))
        Ffp = None
        for it in range(o.m + 1):
            Ffp = self.solve_faces(Uad)
            if it < o.m:
                dudt = self.ader_dudt(Ffp)
                Uad = U_sp[:, None] - torch.einsum("np,up...->un...", self.INV, dudt) * dt
        return Ffp

    # -- FV corrector (src/sdader_simulator.py:288-349, src/fv_simulator.py) -----
    def _face_dx(self, a):
        o = self.o
        g = o.ngh
        f = o.faces[a]
        return f[g + 1:-g] - f[g:-(g + 1)]

    def fv_apply(self, U_cv, Ff, dt):
        o = self.o
        nt = U_cv.dim()
        dU = U_cv * 0
        for a in range(o.ndim):
            ax = nt - 1 - a
            dx = self._bshape(a, self._face_dx(a))
            dU = dU + (Ff[a][_sl(nt, ax, slice(1, None))] - Ff[a][_sl(nt, ax, slice(None, -1))]) / dx
        return U_cv - dU * dt

    def ho_faces(self, Favg, i):
        o = self.o
        nd = o.ndim
        out = []
        for a in range(nd):
            F = Favg[a][:, i]
            la = 1 + 2 * nd - 1 - a
            ea = 1 + nd - 1 - a
            body = F.narrow(la, 0, o.n)
            fv = self.to_fv(body)
            last = F.narrow(ea, o.N[a] - 1, 1).narrow(la, o.nf - 1, 1)
            out.append(torch.cat([fv, self.to_fv(last)], dim=1 + nd - 1 - a))
        return out

    def ghost_fill(self, W):
        """All ghost layers of an active-region FV array, x first: periodic
        wrap, or the non-mirrored copy of the first / last layers (reflective:
        normal velocity negated) (``src/fv_simulator.py:224-273``)."""
        o = self.o
        g = o.ngh
        X = W
        nt = X.dim()
        for a in range(o.ndim):
            ax = nt - 1 - a
            n = X.shape[ax]
            first = X.narrow(ax, 0, g)
            last = X.narrow(ax, n - g, g)
            kl, kr = o.BC[a]
            lo = last if kl == "periodic" else first
            hi = first if kr == "periodic" else last
            if kl == "reflective":
                lo = _flip(lo, a + 1)
            if kr == "reflective":
                hi = _flip(hi, a + 1)
            X = torch.cat([lo, X, hi], dim=ax)
        return X

    def muscl_hancock(self, W_cv, dt):
        """``src/fv_simulator.py:167-175``, ``src/muscl.py:61-75,129-221`` (Q1, Q7, Q19)."""
        o = self.o
        nd, g = o.ndim, o.ngh
        Mfv = self.ghost_fill(W_cv)
        nt = Mfv.dim()
        S, grads = {}, {}
        for a in range(nd):
            ax = nt - 1 - a
            hcv = self._bshape(a, o.h_cv[a])
            hfp = self._bshape(a, o.h_fp[a])
            dM = (Mfv[_sl(nt, ax, slice(1, None))] - Mfv[_sl(nt, ax, slice(None, -1))]) / hcv
            L = dM[_sl(nt, ax, slice(None, -1))]
            R = dM[_sl(nt, ax, slice(1, None))]
            if o.slope_limiter == "moncen":
                dMh = moncen(L, R, hcv[_sl(nt, ax, slice(None, -1))],
                             hcv[_sl(nt, ax, slice(1, None))], hfp[_sl(nt, ax, slice(1, -1))])
            else:
                dMh = minmod(L, R)
            S[a] = 0.5 * dMh * hfp[_sl(nt, ax, slice(1, -1))]
            crop = [slice(None)] * nt
            for b in range(nd):
                if b != a:
                    crop[nt - 1 - b] = slice(1, -1)
            grads[a] = dMh[tuple(crop)]
        inner = (slice(None),) + (slice(1, -1),) * nd
        W = Mfv[inner]
        P = o.P
        vels = [int(v) for v in o.vels]
        dtW = [W[0] * 0 for _ in range(o.V)]
        for a in range(nd):
            v = vels[a]
            dW = grads[a]
            dtW[0] = dtW[0] - (W[v] * dW[0] + W[0] * dW[v])
            dtW[P] = dtW[P] - (W[v] * dW[P] + o.gamma * W[P] * dW[v])
            dtW[v] = dtW[v] - (W[v] * dW[v] + dW[P] / W[0])
            for v2 in vels[1:]:
                dtW[v2] = dtW[v2] - W[v] * dW[v2]
        dtW[P] = dtW[P] * 0
        dtW = torch.stack(dtW)
        Mfv = Mfv + torch.nn.functional.pad(0.5 * dtW * dt, [1, 1] * nd)
        F = []
        for a in range(nd):
            ax = nt - 1 - a
            sR = [slice(None)] + [slice(g, -g)] * nd
            sL = list(sR)
            tR = list(sR)
            tL = list(sR)
            sR[ax] = slice(g, -1)
            tR[ax] = slice(1, None)
            sL[ax] = slice(1, -g)
            tL[ax] = slice(None, -1)
            MR = Mfv[tuple(sR)] - S[a][tuple(tR)]
            ML = Mfv[tuple(sL)] + S[a][tuple(tL)]
            vr = [int(v) for v in np.roll(o.vels, -a)]
            F.append(riemann(o.rs_fv, ML, MR, vr, P, o.gamma, o.min_c2, prims_in=True))
        return F

    def __call__(self, U_sp, dt, aff=None):
        """U_sp (element layout, float64 tensor) -> U_sp after one step.  ``aff``:
        per substep, per dimension, the face coefficients (shape (1, faces...))
        of the FV fallback (ignored without it)."""
        o = self.o
        nd = o.ndim
        Ffp = self.ader_predictor(U_sp, dt)
        if o.update == "SD":
            dudt = self.ader_dudt(Ffp)
            return U_sp - torch.einsum("t,ut...->u...", _t(self.W_TP), dudt) * dt
        U_cv = self.to_fv(self._full(U_sp, self.S2C))
        Favg = []
        for a in range(nd):
            F = Ffp[a]
            for b in range(nd):
                if b != a:
                    F = _contract(F, self.S2C, o._local_axis(b, 2))
            Favg.append(F)
        vels = [int(v) for v in o.vels]
        for i in range(o.T):
            dti = dt * o.ops.w_tp[i]
            Fho = self.ho_faces(Favg, i)
            if o.FB:
                W_cv = prims(U_cv, vels, o.P, o.gamma)
                Ffb = self.muscl_hancock(W_cv, dti)
                Fb = []
                for a in range(nd):
                    th = 1.0 if o.godunov else aff[i][a]
                    Fb.append(th * Ffb[a] + (1 - th) * Fho[a])
                U_cv = self.fv_apply(U_cv, Fb, dti)
            else:
                U_cv = self.fv_apply(U_cv, Fho, dti)
        return self._full(self.to_sd(U_cv), self.C2S)


def face_coefficients(o, theta):
    """Face coefficients max(theta_L, theta_R) of every dimension from theta on
    the ghosted grid (``src/trouble_detection.py:80-87``); theta: array of
    shape ((Ns+4) per dim, reversed)."""
    nd, g = o.ndim, o.ngh
    theta = _t(theta)
    aff = []
    for a in range(nd):
        ax = nd - 1 - a
        s1 = [slice(g, -g)] * nd
        s2 = [slice(g, -g)] * nd
        s1[ax] = slice(g - 1, -g)
        s2[ax] = slice(g, -(g - 1))
        aff.append(torch.maximum(theta[tuple(s1)], theta[tuple(s2)])[None])
    return aff
