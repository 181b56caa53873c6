"""The graph-network preconditioner restated with numpy (small graphs).

TEST INFRASTRUCTURE ONLY: the checker of nlmech_amd's network kernels.

Same parameter vector layout, features, layers, decoders, loss and backward
pass as the library.  With bf16=True every tensor the library stores in BF16
(features, layer states, messages, mean messages, the gradients fed to the
matrix products, the weights) is rounded to BF16 (round to nearest even) at
the same point; sums are evaluated in float64.  With bf16=False the whole
computation is float64 without rounding, which makes the gradient checkable
against finite differences of the loss.
"""
from __future__ import annotations

import numpy as np

L, H, FN, FE = 4, 64, 8, 8
OFF_WN, OFF_BN, OFF_WE, OFF_BE = 0, 512, 576, 1088
OFF_LAYER, LAYER = 1152, 20608
L_WA, L_WB, L_WC, L_BM, L_WD, L_WF, L_BU = 0, 4096, 8192, 12288, 12352, 16448, 20544
OFF_WOE = OFF_LAYER + L * LAYER
OFF_BOE = OFF_WOE + 256
OFF_WON = OFF_BOE + 4
OFF_BON = OFF_WON + 256
NPARAMS = OFF_BON + 4


def bf16(a):
    """Round float32 values to BF16 (nearest even), returned as float64."""
    f = np.ascontiguousarray(np.asarray(a, np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    u = np.where(nan, 0x7FC00000, u).astype(np.uint32)
    return u.view(np.float32).astype(np.float64)


class Graph:
    """Directed bonds (i, j) in reference CSR order with tangent blocks."""

    def __init__(self, x, horizon, row_ptr, col, blocks, diag, fixity):
        self.n = x.shape[0]
        self.x = np.asarray(x, np.float64)[:, :2]
        self.horizon = horizon
        self.rows = np.repeat(np.arange(self.n), np.diff(row_ptr))
        self.cols = np.asarray(col, np.int64)
        self.deg = np.diff(row_ptr).astype(np.float64)
        self.blocks = np.asarray(blocks, np.float64).reshape(-1, 4)
        self.diag = np.asarray(diag, np.float64).reshape(-1, 4)
        self.fix = np.asarray(fixity, np.int64)
        key = {(int(i), int(j)): k for k, (i, j) in
               enumerate(zip(self.rows, self.cols))}
        self.rev = np.array([key[(int(j), int(i))] for i, j in
                             zip(self.rows, self.cols)], np.int64)

    def mask(self, i, j):
        fi0 = (self.fix[i] & 1) == 0
        fi1 = (self.fix[i] & 2) == 0
        fj0 = (self.fix[j] & 1) == 0
        fj1 = (self.fix[j] & 2) == 0
        return np.stack([fi0 & fj0, fi0 & fj1, fi1 & fj0, fi1 & fj1],
                        1).astype(np.float64)


class Net:
    def __init__(self, w, bf16_mode=True):
        self.w = np.asarray(w, np.float64)
        self.rb = bf16 if bf16_mode else (lambda a: np.asarray(a, np.float64))
        self.f32 = (lambda a: np.asarray(a, np.float32).astype(np.float64)) \
            if bf16_mode else (lambda a: np.asarray(a, np.float64))

    def W(self, off, k, n=64):
        return self.w[off:off + k * n].reshape(k, n)

    def Wr(self, off, k, n=64):
        return self.rb(self.W(off, k, n))

    def layer(self, l):
        return OFF_LAYER + l * LAYER

    # -- forward -----------------------------------------------------------
    def forward(self, g: Graph):
        rb, f32 = self.rb, self.f32
        s = 0.5 * (g.diag[:, 0] + g.diag[:, 3])
        s = np.where(s > 0, s, 1.0)
        isq = f32(1.0 / np.sqrt(s))
        fx = np.stack([(g.fix & 1) != 0, (g.fix & 2) != 0], 1).astype(float)
        xn = np.concatenate([g.diag / s[:, None], fx, g.deg[:, None] / 32.0,
                             np.ones((g.n, 1))], 1)
        xn = rb(f32(xn))
        wgt = 1.0 / np.sqrt(s[g.rows] * s[g.cols])
        dx = (g.x[g.cols] - g.x[g.rows]) / g.horizon
        xe = np.concatenate([g.blocks * wgt[:, None], dx,
                             np.sqrt((dx * dx).sum(1))[:, None],
                             np.ones((len(g.rows), 1))], 1)
        xe = rb(f32(xe))
        c = {"xn": xn, "xe": xe, "isq": isq, "h": [], "g": [], "agg": []}
        h = rb(np.maximum(xn @ self.Wr(OFF_WN, FN) + self.w[OFF_BN:OFF_BN + 64],
                          0))
        ge = rb(np.maximum(xe @ self.Wr(OFF_WE, FE) +
                           self.w[OFF_BE:OFF_BE + 64], 0))
        c["h"].append(h)
        c["g"].append(ge)
        cnt = np.maximum(g.deg, 1.0)
        for l in range(L):
            o = self.layer(l)
            P = f32(h @ self.Wr(o + L_WA, 64))
            Q = f32(h @ self.Wr(o + L_WB, 64))
            m = ge @ self.Wr(o + L_WC, 64) + self.w[o + L_BM:o + L_BM + 64] \
                + P[g.rows] + Q[g.cols]
            m = rb(np.maximum(m, 0))
            agg = np.zeros((g.n, 64))
            np.add.at(agg, g.rows, m)
            agg = rb(f32(agg) * f32(1.0 / cnt)[:, None])
            hu = np.concatenate([h, agg], 1) @ np.concatenate(
                [self.Wr(o + L_WD, 64), self.Wr(o + L_WF, 64)], 0)
            h = rb(np.maximum(hu + self.w[o + L_BU:o + L_BU + 64], 0))
            ge = m
            c["g"].append(ge)
            c["agg"].append(agg)
            c["h"].append(h)
        Woe = self.W(OFF_WOE, 64, 4)
        Won = self.W(OFF_WON, 64, 4)
        me = ge @ Woe + self.w[OFF_BOE:OFF_BOE + 4]
        mn = h @ Won + self.w[OFF_BON:OFF_BON + 4] + np.array([1., 0, 0, 1])
        sc_e = isq[g.rows] * isq[g.cols]
        Me = me * g.mask(g.rows, g.cols) * sc_e[:, None]
        Mn = mn * g.mask(np.arange(g.n), np.arange(g.n)) * (isq * isq)[:, None]
        return Me, Mn, c

    # -- loss ----------------------------------------------------------------
    @staticmethod
    def apply(g: Graph, Me, Mn, r):
        """z = M r for r [K, n, 2]."""
        z = np.einsum("nab,knb->kna", Mn.reshape(-1, 2, 2), r)
        contrib = np.einsum("eab,keb->kea", Me.reshape(-1, 2, 2), r[:, g.cols])
        for k in range(r.shape[0]):
            np.add.at(z[k], g.rows, contrib[k])
        return z

    def loss_grad(self, g: Graph, r, e):
        """r, e: [K, n, 2]; returns (loss, grad[NPARAMS])."""
        rb, f32 = self.rb, self.f32
        K = r.shape[0]
        Me, Mn, c = self.forward(g)
        z = self.apply(g, Me, Mn, r)
        ck = np.array([1.0 / (K * (ek ** 2).sum()) for ek in e])
        wv = z - e
        loss = float(sum(ck[k] * (wv[k] ** 2).sum() for k in range(K)))
        isq = c["isq"]
        dMe = 2 * np.einsum("k,kea,keb->eab", ck, wv[:, g.rows],
                            r[:, g.cols]).reshape(-1, 4)
        dMe = f32(dMe * g.mask(g.rows, g.cols) *
                  (isq[g.rows] * isq[g.cols])[:, None])
        ar = np.arange(g.n)
        dMn = 2 * np.einsum("k,kna,knb->nab", ck, wv, r).reshape(-1, 4)
        dMn = f32(dMn * g.mask(ar, ar) * (isq * isq)[:, None])
        grad = np.zeros(NPARAMS)
        Woe = self.W(OFF_WOE, 64, 4)
        Won = self.W(OFF_WON, 64, 4)
        gL, hL = c["g"][L], c["h"][L]
        grad[OFF_WOE:OFF_WOE + 256] = (gL.T @ dMe).reshape(-1)
        grad[OFF_BOE:OFF_BOE + 4] = dMe.sum(0)
        grad[OFF_WON:OFF_WON + 256] = (hL.T @ dMn).reshape(-1)
        grad[OFF_BON:OFF_BON + 4] = dMn.sum(0)
        dg = rb(f32(dMe @ Woe.T))
        dh = f32(dMn @ Won.T)
        cnt = np.maximum(g.deg, 1.0)
        for l in range(L - 1, -1, -1):
            o = self.layer(l)
            h_in, h_out = c["h"][l], c["h"][l + 1]
            g_in, m = c["g"][l], c["g"][l + 1]
            agg = c["agg"][l]
            dpre_u = rb(f32(dh * (h_out > 0)))
            grad[o + L_WD:o + L_WD + 4096] += (h_in.T @ dpre_u).reshape(-1)
            grad[o + L_WF:o + L_WF + 4096] += (agg.T @ dpre_u).reshape(-1)
            grad[o + L_BU:o + L_BU + 64] += dpre_u.sum(0)
            dh_new = f32(dpre_u @ self.Wr(o + L_WD, 64).T)
            dagg = f32(dpre_u @ self.Wr(o + L_WF, 64).T)
            dm = f32(dg + f32(dagg[g.rows] / cnt[g.rows, None]))
            dpre_m = rb(dm * (m > 0))
            grad[o + L_WC:o + L_WC + 4096] += (g_in.T @ dpre_m).reshape(-1)
            grad[o + L_BM:o + L_BM + 64] += dpre_m.sum(0)
            dg_prev = rb(f32(dpre_m @ self.Wr(o + L_WC, 64).T))
            dP = np.zeros((g.n, 64))
            np.add.at(dP, g.rows, dpre_m)
            dQ = np.zeros((g.n, 64))
            np.add.at(dQ, g.cols, dpre_m)
            dP16, dQ16 = rb(f32(dP)), rb(f32(dQ))
            grad[o + L_WA:o + L_WA + 4096] += (h_in.T @ dP16).reshape(-1)
            grad[o + L_WB:o + L_WB + 4096] += (h_in.T @ dQ16).reshape(-1)
            dh = f32(dh_new + dP16 @ self.Wr(o + L_WA, 64).T +
                     dQ16 @ self.Wr(o + L_WB, 64).T)
            dg = dg_prev
        dpre_n = rb(f32(dh * (c["h"][0] > 0)))
        grad[OFF_WN:OFF_WN + 512] += (c["xn"].T @ dpre_n).reshape(-1)
        grad[OFF_BN:OFF_BN + 64] += dpre_n.sum(0)
        dpre_e = rb(dg * (c["g"][0] > 0))
        grad[OFF_WE:OFF_WE + 512] += (c["xe"].T @ dpre_e).reshape(-1)
        grad[OFF_BE:OFF_BE + 64] += dpre_e.sum(0)
        return loss, grad
