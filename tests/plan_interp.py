"""Execute a scheduled jaqsi_amd program on the CPU (test infrastructure).

Mirrors what the tile-pass kernels do with a program: tile gathers, per
substep register layouts, gate records (targets on register bits, controls
and diagonal factors anywhere), the reversed adjoint sweep, the cotangent
slot map and the chain rule to angles.  Used to check the native planner
against the oracle without a GPU; the GPU tests then check the kernels
against the same oracle.
"""

from __future__ import annotations

import math

import numpy as np

HDR = 16
GW, BW, MW = 16, 12, 8
AK_NOP, AK_MC1, AK_MC2, AK_MCX, AK_MCSWAP, AK_DIAG, AK_DENSEK, AK_WIDE = range(8)
SIGMA = [np.array([[0, 1], [1, 0]], dtype=complex), np.array([[0, -1j], [1j, 0]]), np.diag([1.0, -1.0]).astype(complex)]
PF_WIDE = 4
PAULI = [
    np.eye(2, dtype=complex),
    np.array([[0, 1], [1, 0]], dtype=complex),
    np.array([[0, -1j], [1j, 0]], dtype=complex),
    np.array([[1, 0], [0, -1]], dtype=complex),
]


def gen_matrix(mt, theta, xm, kb, deriv=False):
    kind, flags, nc, word, param, moff = mt[:6]
    dim = 1 << kb
    if kind == 2:
        m = xm[moff : moff + dim * dim].reshape(dim, dim)
        return np.zeros_like(m) if deriv else m.copy()
    t = kb - nc
    s = dim - (1 << t)
    out = np.zeros((dim, dim), dtype=complex) if deriv else np.eye(dim, dtype=complex)
    if kind == 1:
        blk = np.diag([0] * ((1 << t) - 1) + [1j * np.exp(1j * theta)]) if deriv else np.diag([1] * ((1 << t) - 1) + [np.exp(1j * theta)])
    else:
        P = np.ones((1, 1), dtype=complex)
        for j in range(t):
            P = np.kron(P, PAULI[(word >> (2 * (t - 1 - j))) & 3])
        I = np.eye(1 << t)
        if deriv:
            blk = -0.5 * math.sin(theta / 2) * I - 0.5j * math.cos(theta / 2) * P
        else:
            blk = math.cos(theta / 2) * I - 1j * math.sin(theta / 2) * P
    out[s:, s:] = blk
    return out


def member_full(mt, theta, xm, deriv=False):
    flags, kb = mt[1], mt[6]
    if flags & 2:
        g = gen_matrix(mt, theta, xm, kb)
        if deriv:
            dg = gen_matrix(mt, theta, xm, kb, True)
            v = np.kron(dg, np.conj(g)) + np.kron(g, np.conj(dg))
        else:
            v = np.kron(g, np.conj(g))
    else:
        v = gen_matrix(mt, theta, xm, kb, deriv)
    return np.conj(v) if flags & 1 else v


class Program:
    def __init__(self, plan, quat=False, backward_table=4):
        # quat: the cotangent of MC1 gates flagged in word 15 is kept as
        # t_k = Im Tr(bar^T U sigma_k) (t_x + i t_y in the bar00 slot, t_z in
        # the bar10 slot's real part), and the chain turns it into
        # bar00 = -be (i t_x + t_y) + i t_z conj(al), bar10 = al (i t_x + t_y)
        # + i t_z conj(be) for U = [[al, .], [be, .]], as the plan-specialised
        # backward and the chain kernel do
        self.quat = quat
        self.w = plan.program()
        # backward program: 4 the prebuilt kernels', 5 the plan-specialised kernels'
        self.wb = plan.table(backward_table)
        self.bt = plan.table(0).reshape(-1, BW)
        self.ml = plan.table(1)
        self.mt = plan.table(2).reshape(-1, MW)
        sm = plan.table(3)
        self.sm = sm.reshape(-1, 5) if sm.size >= 5 else np.zeros((0, 5), dtype=np.int32)
        w = self.w
        self.P, self.T, self.RB, self.n = int(w[1]), int(w[2]), int(w[3]), int(w[4])
        self.nbars_reals = int(w[6])

    def members(self, b):
        rec = self.bt[b]
        return [int(m) for m in self.ml[rec[3] : rec[4]]]

    def theta(self, m, angles):
        p = int(self.mt[m][4])
        return angles[p] if p >= 0 else 0.0

    def eff(self, m, angles, xm, b, deriv=False):
        rec = self.bt[b]
        D, offs, diagf = rec[2], rec[9], rec[8]
        full = member_full(self.mt[m], self.theta(m, angles), xm, deriv)
        e = full[offs : offs + D, offs : offs + D]
        return np.diag(e).copy() if diagf else e

    def block_matrix(self, b, angles, xm):
        rec = self.bt[b]
        D, diagf = rec[2], rec[8]
        M = np.ones(D, dtype=complex) if diagf else np.eye(D, dtype=complex)
        for m in self.members(b):
            E = self.eff(m, angles, xm, b)
            M = E * M if diagf else E @ M
        return M

    # ------------------------------------------------------------------
    def passes(self, backward=False):
        w, T = (self.wb if backward else self.w), self.T
        RB = int(w[3])
        for p in range(self.P):
            rec = int(w[HDR + p])
            nsub, flags, nslots, sbase = (int(x) for x in w[rec : rec + 4])
            tbits = [int(x) for x in w[rec + 4 : rec + 4 + T]]
            soff = [int(x) for x in w[rec + 4 + T : rec + 4 + T + nsub]]
            subs = []
            for o in soff:
                ng = int(w[o])
                reg = [int(x) for x in w[o + 1 : o + 1 + RB]]
                thr = [int(x) for x in w[o + 1 + RB : o + 1 + T]]
                gates = [w[o + 1 + T + q * GW : o + 1 + T + (q + 1) * GW].astype(np.int64) for q in range(ng)]
                subs.append((reg, thr, gates))
            yield p, flags, nslots, sbase, tbits, subs

    def tiles(self, tbits):
        n = self.n
        real = [g for g in tbits if g < n]
        rest = [g for g in range(n) if g not in real]
        for t in range(1 << len(rest)):
            base = 0
            for j, g in enumerate(rest):
                if (t >> j) & 1:
                    base |= 1 << g
            yield t, base

    def gidx(self, tbits, base):
        T = self.T
        x = np.arange(1 << T, dtype=np.int64)
        g = np.full(x.shape, base, dtype=np.int64)
        for j, gb in enumerate(tbits):
            g |= ((x >> j) & 1) << gb
        return g

    # ------------------------------------------------------------------
    def _gate(self, x, g, reg, base, pool, mode, tbits):
        """Apply gate record g to tile vector x (tile-local indexing)."""
        kind = int(g[0])
        if kind == AK_NOP:
            return x
        T = self.T
        t = np.arange(1 << T, dtype=np.int64)
        creg, cthr = int(g[4]), int(g[5])
        cout = (int(g[7]) << 32) | (int(g[6]) & 0xFFFFFFFF)
        cmask = cthr
        for b in range(len(reg)):
            if (creg >> b) & 1:
                cmask |= 1 << reg[b]
        ok = ((base & cout) == cout) and True
        act = ((t & cmask) == cmask) & ok
        b = int(g[1])
        if kind == AK_DIAG:
            k = int(g[11])
            d = pool[b][1] if (mode == 1 and pool[b][1] is not None) else pool[b][0]
            if mode == 1 and pool[b][1] is None:
                d = np.conj(d)
            gl = self.gidx(tbits, base)
            idx = np.zeros_like(t)
            for j in range(k):
                loc = (int(g[12 + j // 4]) >> (8 * (j % 4))) & 0xFF
                typ, pos = loc >> 6, loc & 63
                if typ == 0:
                    bit = (t >> reg[pos]) & 1
                elif typ == 1:
                    bit = (t >> pos) & 1
                else:
                    bit = (gl >> pos) & 1
                idx |= bit << (k - 1 - j)
            return x * d[idx]
        if kind in (AK_MC1, AK_MC2, AK_DENSEK):
            M = pool[b][0]
            if mode == 1:
                M = pool[b][1] if pool[b][1] is not None else np.conj(M).T
            elif mode == 2:
                M = M.T
            if kind == AK_MC1:
                tb = [reg[int(g[2])]]
            elif kind == AK_MC2:
                tb = [reg[int(g[2])], reg[int(g[3])]]
            else:
                k = int(g[2])
                tb = [reg[k - 1 - j] for j in range(k)]
        elif kind == AK_MCX:
            M = np.array([[0, 1], [1, 0]], dtype=complex)
            tb = [reg[int(g[2])]]
        else:
            M = np.eye(4, dtype=complex)[[0, 2, 1, 3]]
            tb = [reg[int(g[2])], reg[int(g[3])]]
        k = len(tb)
        tmask = sum(1 << q for q in tb)
        sel = ((t & tmask) == 0) & act
        i0 = t[sel]
        idx = []
        for q in range(1 << k):
            o = i0.copy()
            for j in range(k):
                if (q >> (k - 1 - j)) & 1:
                    o |= 1 << tb[j]
            idx.append(o)
        v = np.stack([x[o] for o in idx])
        y = M @ v
        x = x.copy()
        for q in range(1 << k):
            x[idx[q]] = y[q]
        return x

    def _wide_index(self, g):
        """[rest, c] amplitude indices of a wide block: rest over the
        indices with every wire bit zero, c over the 2^k wire values (wire j
        = bit k-1-j of c)."""
        n, k = self.n, int(g[11])
        gb = [(int(g[12 + j // 4]) >> (8 * (j % 4))) & 0xFF for j in range(k)]
        i = np.arange(1 << n, dtype=np.int64)
        mask = sum(1 << b for b in gb)
        rest = i[(i & mask) == 0]
        cbits = np.zeros(1 << k, dtype=np.int64)
        for c in range(1 << k):
            for j in range(k):
                if (c >> (k - 1 - j)) & 1:
                    cbits[c] |= 1 << gb[j]
        return rest[:, None] | cbits[None, :]

    def _wide_forward(self, psi, g, pool):
        idx = self._wide_index(g)
        M = pool[int(g[1])][0]
        X = psi[idx]
        out = psi.copy()
        out[idx] = X * M[None, :] if int(g[2]) else X @ M.T
        return out

    def _wide_backward(self, psi_in, lam, g, pool, partial, p):
        idx = self._wide_index(g)
        M = pool[int(g[1])][0]
        X, L = psi_in[idx], lam[idx]
        if int(g[10]) >= 0:
            vals = (L * X).sum(0) if int(g[2]) else (L.T @ X).reshape(-1)
            flat = np.stack([vals.real, vals.imag], axis=-1).reshape(-1)
            key = (p, int(g[10]))
            partial[key] = partial.get(key, 0) + flat
        out = lam.copy()
        out[idx] = L * M[None, :] if int(g[2]) else L @ M
        return out

    def _pool(self, angles, xm):
        pool = {}
        for b in range(len(self.bt)):
            M = self.block_matrix(b, angles, xm)
            inv = None
            if self.bt[b][6] >= 0:
                inv = 1.0 / M if self.bt[b][8] else np.linalg.inv(M)
            pool[b] = (M, inv)
        return pool

    def forward(self, angles, xm=None, psi0=None):
        n = self.n
        psi = np.zeros(1 << n, dtype=complex)
        if psi0 is None:
            psi[0] = 1
        else:
            psi[:] = psi0
        pool = self._pool(angles, xm)
        for p, flags, nslots, sbase, tbits, subs in self.passes():
            if flags & PF_WIDE:
                psi = self._wide_forward(psi, subs[0][2][0], pool)
                continue
            out = psi.copy()
            for t, base in self.tiles(tbits):
                gl = self.gidx(tbits, base)
                valid = gl < (1 << n)
                x = np.zeros(gl.shape, dtype=complex)
                x[valid] = psi[gl[valid]]
                for reg, thr, gates in subs:
                    for g in gates:
                        x = self._gate(x, g, reg, base, pool, 0, tbits)
                out[gl[valid]] = x[valid]
            psi = out
        return psi

    def adjoint(self, angles, lam_seed_fn, xm=None, psi0=None):
        """Returns (final psi, bars per block, angle grads, lambda0)."""
        n = self.n
        pool = self._pool(angles, xm)
        # forward with checkpoints
        psi = np.zeros(1 << n, dtype=complex)
        if psi0 is None:
            psi[0] = 1
        else:
            psi[:] = psi0
        ck = []
        psi_init = psi.copy()
        for p, flags, nslots, sbase, tbits, subs in self.passes():
            if flags & PF_WIDE:
                psi = self._wide_forward(psi, subs[0][2][0], pool)
                ck.append(psi)
                continue
            out = psi.copy()
            for t, base in self.tiles(tbits):
                gl = self.gidx(tbits, base)
                valid = gl < (1 << n)
                x = np.zeros(gl.shape, dtype=complex)
                x[valid] = psi[gl[valid]]
                for reg, thr, gates in subs:
                    for g in gates:
                        x = self._gate(x, g, reg, base, pool, 0, tbits)
                out[gl[valid]] = x[valid]
            psi = out
            ck.append(psi)
        final = psi
        lam = lam_seed_fn(final)
        partial = {}
        for p, flags, nslots, sbase, tbits, subs in reversed(list(self.passes(backward=True))):
            if flags & PF_WIDE:  # reads the pass's checkpointed input
                lam = self._wide_backward(ck[p - 1] if p > 0 else psi_init, lam, subs[0][2][0], pool, partial, p)
                continue
            src = ck[p]
            newlam = lam.copy()
            for t, base in self.tiles(tbits):
                gl = self.gidx(tbits, base)
                valid = gl < (1 << n)
                x = np.zeros(gl.shape, dtype=complex)
                l = np.zeros(gl.shape, dtype=complex)
                x[valid] = src[gl[valid]]
                l[valid] = lam[gl[valid]]
                for reg, thr, gates in reversed(subs):
                    for g in reversed(gates):
                        x = self._gate(x, g, reg, base, pool, 1, tbits)
                        if g[10] >= 0:
                            self._grad(x, l, g, reg, base, tbits, partial, p, t, pool)
                        l = self._gate(l, g, reg, base, pool, 2, tbits)
                newlam[gl[valid]] = l[valid]
            lam = newlam
        bars = np.zeros(self.nbars_reals, dtype=float)
        for base_, nsl, local, bar, size in self.sm:
            pass
        # partial[(p, slot)] accumulated over tiles; map to bars by slot map
        for p, flags, nslots, sbase, tbits, subs in self.passes(backward=True):
            for reg, thr, gates in subs:
                for g in gates:
                    if g[10] < 0:
                        continue
                    vals = partial[(p, int(g[10]))]
                    ent = [e for e in self.sm if e[0] == sbase and e[2] == g[10]]
                    assert len(ent) == 1
                    bars[ent[0][3] : ent[0][3] + ent[0][4]] = vals
        grads = self.chain(angles, xm, bars)
        return final, bars, grads, lam

    def _grad(self, psi_in, lam_out, g, reg, base, tbits, partial, p, tile, pool=None):
        kind = int(g[0])
        T = self.T
        t = np.arange(1 << T, dtype=np.int64)
        creg, cthr = int(g[4]), int(g[5])
        cout = (int(g[7]) << 32) | (int(g[6]) & 0xFFFFFFFF)
        cmask = cthr
        for b in range(len(reg)):
            if (creg >> b) & 1:
                cmask |= 1 << reg[b]
        act = ((t & cmask) == cmask) & ((base & cout) == cout)
        if kind == AK_DIAG:
            k = int(g[11])
            gl = self.gidx(tbits, base)
            idx = np.zeros_like(t)
            for j in range(k):
                loc = (int(g[12 + j // 4]) >> (8 * (j % 4))) & 0xFF
                typ, pos = loc >> 6, loc & 63
                bit = ((t >> reg[pos]) & 1) if typ == 0 else ((t >> pos) & 1) if typ == 1 else ((gl >> pos) & 1)
                idx |= bit << (k - 1 - j)
            vals = np.zeros(1 << k, dtype=complex)
            np.add.at(vals, idx, lam_out * psi_in)
        else:
            tb = [reg[int(g[2])]] if kind == AK_MC1 else [reg[int(g[2])], reg[int(g[3])]]
            k = len(tb)
            tmask = sum(1 << q for q in tb)
            sel = ((t & tmask) == 0) & act
            i0 = t[sel]
            idx = []
            for q in range(1 << k):
                o = i0.copy()
                for j in range(k):
                    if (q >> (k - 1 - j)) & 1:
                        o |= 1 << tb[j]
                idx.append(o)
            vals = np.zeros((1 << k, 1 << k), dtype=complex)
            for o in range(1 << k):
                for c in range(1 << k):
                    vals[o, c] = np.sum(lam_out[idx[o]] * psi_in[idx[c]])
            vals = vals.reshape(-1)
            if self.quat and kind == AK_MC1 and int(g[15]):
                U = pool[int(g[1])][0]
                tk = [np.imag(np.trace(vals.reshape(2, 2).T @ U @ s)) for s in SIGMA]
                vals = np.array([tk[0] + 1j * tk[1], 0, tk[2], 0])
        key = (p, int(g[10]))
        flat = np.stack([vals.real, vals.imag], axis=-1).reshape(-1)
        partial[key] = partial.get(key, 0) + flat

    def chain(self, angles, xm, bars):
        grads = np.zeros(len(angles))
        barc = bars[0::2] + 1j * bars[1::2]
        for b in range(len(self.bt)):
            rec = self.bt[b]
            if rec[7] < 0:
                continue
            D, diagf = rec[2], rec[8]
            ne = D if diagf else D * D
            bar = barc[rec[7] : rec[7] + ne]
            if not diagf:
                bar = bar.reshape(D, D)
            if self.quat and rec[10]:
                U = self.block_matrix(b, angles, xm)
                al, be = U[0, 0], U[1, 0]
                tx, ty, tz = bar[0, 0].real, bar[0, 0].imag, bar[1, 0].real
                w = 1j * tx + ty
                bar = np.array([[-be * w + 1j * tz * np.conj(al), 0], [al * w + 1j * tz * np.conj(be), 0]])
            mem = self.members(b)
            Es = [self.eff(m, angles, xm, b) for m in mem]
            for j, m in enumerate(mem):
                if self.mt[m][4] < 0:
                    continue
                L = np.ones(D, dtype=complex) if diagf else np.eye(D, dtype=complex)
                R = np.ones(D, dtype=complex) if diagf else np.eye(D, dtype=complex)
                for E in Es[j + 1 :]:
                    L = E * L if diagf else E @ L
                for E in Es[:j]:
                    R = E * R if diagf else E @ R
                dE = self.eff(m, angles, xm, b, True)
                W = L * bar * R if diagf else L.T @ bar @ R.T
                grads[self.mt[m][4]] += np.real(np.sum(W * dE))
        return grads
