"""FP64 PyTorch model of the PT tables, differentiated by autograd: the reference for the
hand-written adjoint (test infrastructure only).

The model is the reference's gridded algorithm with its parameters as torch tensors:
  * geometry -- which grid node each line samples in each record -- is read from the
    oracle's single-line deposits (oracle/marfa_oracle.c restates src/LineGridCalc.f90 and
    is pinned bit for bit against the compiled reference); X = nu* - origin - node position;
  * spectroscopy as src/Spectroscopy.f90:38-145 (mo_spec_f64), line shapes as
    src/Shapes.f90 / src/Spectroscopy.f90:194-261 / src/ChiFactors.f90 (mo_voigt_f64),
    cascade as src/Grids.f90:80-139 -- each checked against the oracle's values in
    tests/test_adjoint.py before its gradients are used.
The branch decisions (regions, limiter, grid selection) are held fixed by autograd, which
is the derivative the adjoint computes.
"""
from __future__ import annotations

import numpy as np
import torch

import pyoracle as po

NT, NGRID = 20481, 30690
NNODE = NGRID + NT


def _f32(bits):
    return float(np.frombuffer(np.uint32(bits).tobytes(), np.float32)[0])


PI, SQLN2, C2 = _f32(0x40490FDA), _f32(0x3F55224D), _f32(0x3FB829CC)
DOPC, SQRTPI, SQRT14, LOG2 = _f32(0x34C04328), _f32(0x3FE2DFC4), _f32(0x3F97739F), _f32(0x3F317218)
BOLSI, PATM, TREF = float(np.float32(1.3806503e-23)), 101325.0, 296.0
ASYM_U = [1.0, 1.5, 2.0, 2.5, 3.0, 3.5, 4.0, 4.5, 5.0]
ASYM_W = [float(np.float32(w)) for w in (-0.688, 0.2667, 0.6338, 0.4405, 0.2529, 0.1601, 0.1131, 0.0853, 0.068)]


def node_positions() -> np.ndarray:
    """Position (cm-1 from the subinterval origin) of every node of the grid layout."""
    pos = np.zeros(NNODE)
    for k in range(10):
        n, base, h = 10 << k, 30 * ((1 << k) - 1), 2.0 ** -k
        i = np.arange(n)
        pos[base + i] = i * h
        pos[base + n + i] = i * h + h / 2
        pos[base + 2 * n + i] = (i + 1) * h
    pos[NGRID:] = np.arange(NT) * 2.0 ** -11
    return pos


class Geometry:
    """Sample pairs (node of record r, line) of every level, from the oracle, with the
    oracle's FP64 sample values for checking the torch shapes."""

    def __init__(self, job, tips):
        oj = po.OracleJob(job, tips)
        self.origin = oj.origin.copy()
        self.nrec, self.nlev = oj.nrec, oj.nlev
        pos = node_positions()
        self.levels = []
        for lev in range(self.nlev):
            node, line, val = [], [], []
            for r in range(self.nrec):
                for i in range(int(oj.lo[r]), int(oj.hi[r])):
                    g = oj.line_grids(lev, r, i)
                    nz = np.nonzero(g)[0]
                    node.append(r * NNODE + nz)
                    line.append(np.full(nz.size, i))
                    val.append(g[nz])
            node = np.concatenate(node)
            self.levels.append({"node": torch.as_tensor(node), "line": torch.as_tensor(np.concatenate(line)),
                                "rec": torch.as_tensor(node // NNODE),
                                "pos": torch.as_tensor(pos[node % NNODE]),
                                "val": np.concatenate(val)})


def _f(v):
    """the FP32 value of a REAL literal of the reference"""
    return float(np.float32(v))


def _chi(ax, chi_id, mol, T):
    one = torch.ones_like(ax)
    f = _f
    if mol != 2 or chi_id == 0:
        return one
    if chi_id == 1:
        return torch.where(ax > 3, torch.where(ax <= 150, f(1.084) * torch.exp(-f(0.027) * ax),
                                               torch.where(ax <= 300, f(0.208) * torch.exp(-f(0.016) * ax),
                                                           f(0.025) * torch.exp(-f(0.009) * ax))), one)
    if chi_id == 2:
        return torch.where(ax >= 3, torch.where(ax < 10, f(1.35) * torch.exp(-ax / 10.0),
                                                f(0.614) * torch.exp(-ax / 47.0)), one)
    if chi_id == 3:  # Perrin: temperature-dependent exponents (src/ChiFactors.f90:82-122)
        T = torch.as_tensor(T, dtype=ax.dtype)
        B1 = f(0.0888) - f(0.16) * torch.exp(-f(0.0041) * T)
        B2 = f(0.0526) * torch.exp(-f(0.00152) * T)
        B3 = f(0.0232)
        return torch.where(ax > 3, torch.where(ax < 30, torch.exp(-B1 * (ax - 3.0)),
                                               torch.where(ax < 120, torch.exp(-B1 * 27.0 - B2 * (ax - 30.0)),
                                                           torch.exp(-B1 * 27.0 - B2 * 90.0 - B3 * (ax - 120.0)))), one)
    raise NotImplementedError("chi factor")


def _kuntz(region, y, x2):
    f = _f
    y2 = y * y
    if region == 1:
        num = (f(0.2820948) + f(0.5641896) * y2) * y + f(0.5641896) * y * x2
        den = 0.25 + y2 + y2 * y2 + (y2 + y2 - 1) * x2 + x2 * x2
        return num / den
    if region == 2:
        a3 = y * (((f(0.56419) * y2 + f(3.10304)) * y2 + f(4.65456)) * y2 + f(1.05786))
        b3 = y * ((f(1.69257) * y2 + f(0.56419)) * y2 + f(2.962))
        c3 = y * (f(1.69257) * y2 - f(2.53885))
        d3 = y * f(0.56419)
        a4 = (((y2 + 6) * y2 + 10.5) * y2 + 4.5) * y2 + 0.5625
        b4 = ((4 * y2 + 6) * y2 + 9) * y2 - 4.5
        c4 = 10.5 + 6 * (y2 - 1) * y2
        d4 = 4 * y2 - 6
        return (((d3 * x2 + c3) * x2 + b3) * x2 + a3) / ((((x2 + d4) * x2 + c4) * x2 + b4) * x2 + a4)

    def poly(cs):
        p = torch.zeros_like(y) + cs[0]
        for c in cs[1:]:
            p = p * y + c
        return p
    a5 = poly([f(c) for c in (0.564224, 7.55895, 49.5213, 204.510, 581.746, 1174.8, 1678.33, 1629.76, 973.778, 272.102)])
    b5 = poly([f(c) for c in (2.25689, 22.6778, 100.705, 247.198, 336.364, 220.843, -2.34403, -60.5644)])
    c5 = poly([f(c) for c in (3.38534, 22.6798, 52.8454, 42.5683, 18.546, 4.58029)])
    d5 = poly([f(c) for c in (2.25689, 7.56186, 1.66203, -0.128922)])
    e5 = f(0.971457e-3) + f(0.564224) * y
    a6 = poly([1.0] + [f(c) for c in (13.3988, 88.2674, 369.199, 1074.41, 2256.98, 3447.63, 3764.97, 2802.87, 1280.83, 272.102)])
    b6 = poly([5.0] + [f(c) for c in (53.5952, 266.299, 793.427, 1549.68, 2037.31, 1758.34, 902.306, 211.678)])
    c6 = poly([10.0] + [f(c) for c in (80.3928, 269.292, 479.258, 497.302, 308.186, 78.866)])
    d6 = poly([10.0] + [f(c) for c in (53.5952, 92.7586, 55.0293, 22.0353)])
    e6 = 5 * y + f(13.3988)
    e6 = e6 * y + f(1.49645)
    num = (((e5 * x2 + d5) * x2 + c5) * x2 + b5) * x2 + a5
    den = ((((x2 + e6) * x2 + d6) * x2 + c6) * x2 + b6) * x2 + a6
    return num / den


def shape(X, gD, gL, S, shape_id=0, chi_id=0, mol=2, T=296.0):
    """Profile x S(T) at signed offset X (FP64), branch by branch on the selected samples."""
    out = torch.zeros_like(X)
    ax = X.abs()

    def lor(i):
        return S[i] * gL[i] / PI / (X[i] ** 2 + gL[i] ** 2) * _chi(ax[i], chi_id, mol, T)

    if shape_id == 1:
        return lor(torch.ones_like(X, dtype=torch.bool))
    VX = (SQLN2 * X / gD).abs()
    VY = SQLN2 * gL / gD
    r0 = VX >= 15
    r1 = ~r0 & (VX + VY >= 15)
    r2 = ~r0 & ~r1 & (VX + VY >= 5.5)
    r3 = ~r0 & ~r1 & ~r2 & ((VX <= 1) | (VY >= float(np.float32(0.02))))
    r4 = ~r0 & ~r1 & ~r2 & ~r3
    out = out.index_put((r0,), lor(r0))
    for reg, m in ((1, r1), (2, r2), (3, r3)):
        out = out.index_put((m,), _kuntz(reg, VY[m], VX[m] ** 2) * S[m] / SQRTPI / (gD[m] / SQLN2))
    a1 = r4 & (VX > 5)
    out = out.index_put((a1,), lor(a1) * (1 + 1.5 / VX[a1] ** 2))
    dop = r4 & ~a1

    def doppler(i):
        return SQLN2 / (SQRTPI * gD[i]) * torch.exp(-((X[i] / gD[i]) ** 2) * LOG2) * S[i]
    a2 = dop & (VX > SQRT14)
    vx = VX[a2]
    idx = (vx.detach() / 0.5 - float(np.float32(1.00001))).to(torch.int64)
    U = torch.tensor(ASYM_U, dtype=X.dtype)
    W = torch.tensor(ASYM_W, dtype=X.dtype)
    F = 2 * (W[idx - 1] * (U[idx] - vx) + W[idx] * (vx - U[idx - 1]))
    out = out.index_put((a2,), doppler(a2) + S[a2] * gL[a2] / PI / X[a2] ** 2 * (1 + F))
    d0 = dop & ~a2
    out = out.index_put((d0,), doppler(d0))
    return out


def cascade(g):
    """src/Grids.f90:80-139 on a batch of records: g (nrec, NNODE) -> (nrec, NT)."""
    nr = g.shape[0]
    P, M, L = g[:, 0:10], g[:, 10:20], g[:, 20:30]

    def a12(P, M, L):
        A1 = 0.375 * P + 0.75 * M - 0.125 * L
        A1 = torch.where((A1 < P) | (A1 < M), 0.5 * (P + M), A1)
        A2 = 0.375 * L + 0.75 * M - 0.125 * P
        A2 = torch.where((A2 < L) | (A2 < M), 0.5 * (L + M), A2)
        return A1, A2

    for k in range(9):
        n1, b1 = 20 << k, 30 * ((1 << (k + 1)) - 1)
        A1, A2 = a12(P, M, L)
        inter = lambda x, y: torch.stack([x, y], -1).reshape(nr, n1)
        P, M, L = (g[:, b1:b1 + n1] + inter(P, M), g[:, b1 + n1:b1 + 2 * n1] + inter(A1, A2),
                   g[:, b1 + 2 * n1:b1 + 3 * n1] + inter(M, L))
    A1, A2 = a12(P, M, L)
    rk = g[:, NGRID:]
    add = torch.stack([A1, M, A2, L], -1).reshape(nr, NT - 1)
    return torch.cat([rk[:, :1], rk[:, 1:] + add], 1)


class Model:
    """records(S, gair, nT, delta, T, P) for a fixed job; all other inputs held fixed."""

    def __init__(self, job, tips):
        self.job, self.geo = job, Geometry(job, tips)
        L = job.lines
        t = lambda a: torch.as_tensor(np.asarray(a, np.float64))
        self.nu, self.gself, self.E = t(L.nu), t(L.gself), t(L.elow)
        iso = np.asarray(L.molIso) // 100
        self.iso = iso
        self.M = t([po.wiso(int(i)) for i in iso])
        self.tips = np.asarray(tips, np.float64)
        self.rho = t(job.atm.rho)

    def params(self):
        L, A = self.job.lines, self.job.atm
        t = lambda a: torch.as_tensor(np.asarray(a, np.float32).astype(np.float64)).requires_grad_(True)
        return [t(L.S), t(L.gair), t(L.nT), t(L.delta), t(A.T), t(A.P)]

    def level_spec(self, S, gair, nT, delta, T, P, lev):
        Tl, Pl = T[lev], P[lev]
        ps = self.rho[lev] * 10.0 * BOLSI * Tl / PATM
        shift = delta * Pl
        shift = shift + (shift.detach().float().double() - shift.detach())  # the FP32 product's value
        nus = self.nu + shift
        gD = DOPC * nus * torch.sqrt(Tl / self.M)
        gL = (TREF / Tl) ** nT * (gair * (Pl - ps) + self.gself * ps)
        ti = int((float(Tl.detach()) - 20.0) / 2.0 + 1.0)
        wu = (Tl - (ti * 2.0 + 18.0)) / 2.0
        rows = torch.as_tensor(self.tips[self.iso - 1])
        Q = (1 - wu) * rows[:, ti - 1] + wu * rows[:, ti]
        tf = rows[:, 138] / Q
        bf = torch.exp(-(C2 * self.E / Tl)) / torch.exp(-(C2 * self.E / TREF))
        ef = (1 - torch.exp(-(C2 * self.nu / Tl))) / (1 - torch.exp(-(C2 * self.nu / TREF)))
        return nus, gD, gL, S * tf * bf * ef

    def level_samples(self, params, lev):
        S, gair, nT, delta, T, P = params
        nus, gD, gL, ST = self.level_spec(S, gair, nT, delta, T, P, lev)
        d = self.geo.levels[lev]
        li = d["line"]
        org = torch.as_tensor(self.geo.origin)[d["rec"]]
        X = (nus[li] - org) - d["pos"]
        return shape(X, gD[li], gL[li], ST[li], self.job.shape_id, self.job.chi_id, self.job.mol_code, T[lev])

    def records(self, params):
        out = []
        for lev in range(self.geo.nlev):
            d = self.geo.levels[lev]
            f = self.level_samples(params, lev)
            g = torch.zeros(self.geo.nrec * NNODE, dtype=torch.float64).index_add(0, d["node"], f)
            rec = cascade(g.reshape(self.geo.nrec, NNODE))
            if self.job.target == 1:
                rec = rec * self.rho[lev]
            out.append(rec)
        return torch.stack(out)


def loss(records, target):
    m = (records > 0) & (target > 0)
    lr = torch.log(torch.where(m, records, torch.ones_like(records)))
    lt = torch.log(torch.where(m, target, torch.ones_like(target)))
    return (torch.where(m, lr - lt, torch.zeros_like(lr)) ** 2).sum()
