"""Second, independent FP64 transcription of the ZSC step, in NumPy, that keeps
the reference's own structure: z ghost planes, x/y neighbour tables, push
streaming and the three pack/unpack halo exchanges of one MPI rank talking to
itself on a periodic communicator.

Structure followed (carlosrosales/lbm_bench_gpu):
  neighbour tables        src/init.c:62-82
  ghost planes, z offset  src/vgrid.c:38-52 (one rank: zl=1, zu=zmax, NZ=zmax+2)
  step order              src/main.cu:498-589
  update_*                src/update_phi.cu:40-44, update_rho.cu:44-51,
                          update_velocity.cu:45-61
  collisions              src/collision_f.cu:44-116, collision_g.cu:49-196
  push streaming          src/stream_f.cu:40-61, stream_g.cu:44-78
  halo                    src/pack_mpi_phi.cu:41-42 / unpack_mpi_phi.cu:41-42,
                          src/pack_mpi_f.cu:41-42 / unpack_mpi_f.cu:42-43,
                          src/pack_mpi.cu:42-59 / unpack_mpi.cu:42-59
It exists only to cross-check the oracle's periodic-wrap formulation.
"""
from __future__ import annotations

import numpy as np

W0, W1, W2 = 1.0 / 3.0, 1.0 / 18.0, 1.0 / 36.0
INV6, INV12 = 1.0 / 6.0, 1.0 / 12.0


class GhostBox:
    def __init__(self, nx, ny, zmax, coef):
        self.nx, self.ny, self.zmax = nx, ny, zmax
        self.NZ = zmax + 2
        self.c = coef
        i = np.arange(nx)
        j = np.arange(ny)
        self.ie = np.where(i == nx - 1, 0, i + 1)
        self.iw = np.where(i == 0, nx - 1, i - 1)
        self.jn = np.where(j == ny - 1, 0, j + 1)
        self.js = np.where(j == 0, ny - 1, j - 1)
        shape = (self.NZ, ny, nx)
        self.fA = np.zeros((7,) + shape)
        self.fB = np.zeros((7,) + shape)
        self.gA = np.zeros((19,) + shape)
        self.gB = np.zeros((19,) + shape)
        self.phi = np.zeros(shape)
        self.rho = np.zeros(shape)
        self.ux = np.zeros(shape)
        self.uy = np.zeros(shape)
        self.uz = np.zeros(shape)

    # A lattice of real planes <-> pop-major (q, zmax, ny, nx)
    def set_real(self, f, g):
        self.fA[:, 1:-1] = f
        self.gA[:, 1:-1] = g

    def real(self):
        return self.fA[:, 1:-1].copy(), self.gA[:, 1:-1].copy()

    def _p(self, phi, k):
        """19 neighbour planes of phi around real plane k, via nb tables."""
        ie, iw, jn, js = self.ie, self.iw, self.jn, self.js
        kt, kb = k + 1, k - 1
        P = lambda kk, jj, ii: phi[kk][np.ix_(jj, ii)]
        J = np.arange(self.ny)
        I = np.arange(self.nx)
        return {
            "00": P(k, J, I),
            "e": P(k, J, ie), "w": P(k, J, iw), "n": P(k, jn, I), "s": P(k, js, I),
            "t": P(kt, J, I), "b": P(kb, J, I),
            "en": P(k, jn, ie), "ws": P(k, js, iw), "es": P(k, js, ie), "wn": P(k, jn, iw),
            "et": P(kt, J, ie), "wb": P(kb, J, iw), "eb": P(kb, J, ie), "wt": P(kt, J, iw),
            "nt": P(kt, jn, I), "sb": P(kb, js, I), "nb": P(kb, jn, I), "st": P(kt, js, I),
        }

    def _force(self, p, grav_on):
        c = self.c
        phin = p["00"]
        lap = (p["en"] + p["ws"] + p["es"] + p["wn"] + p["et"] + p["wb"] + p["eb"] + p["wt"]
               + p["nt"] + p["sb"] + p["nb"] + p["st"]
               + 2.0 * (p["e"] + p["w"] + p["n"] + p["s"] + p["t"] + p["b"] - 12.0 * phin)) * INV6
        mu = c.alpha4 * phin * (phin * phin - c.phi_star_sq) - c.kappa * lap
        fx = mu * (2.0 * (p["e"] - p["w"]) + p["en"] - p["ws"] + p["es"] - p["wn"]
                   + p["et"] - p["wb"] + p["eb"] - p["wt"]) * INV12
        fy = mu * (2.0 * (p["n"] - p["s"]) + p["en"] - p["ws"] + p["wn"] - p["es"]
                   + p["nt"] - p["sb"] + p["nb"] - p["st"]) * INV12
        fz = mu * (2.0 * (p["t"] - p["b"]) + p["et"] - p["wb"] + p["wt"] - p["eb"]
                   + p["nt"] - p["sb"] + p["st"] - p["nb"]) * INV12
        g = c.grav if grav_on else 0.0
        fz = np.where(phin > 0.0, fz + g, fz)
        return phin, mu, fx, fy, fz

    def step(self, grav_on=True):
        c = self.c
        fA, fB, gA, gB = self.fA, self.fB, self.gA, self.gB
        R = slice(1, self.NZ - 1)
        # update_phi / update_rho / update_velocity on real planes
        self.phi[R] = fA[0, R] + fA[1, R] + fA[2, R] + fA[3, R] + fA[4, R] + fA[5, R] + fA[6, R]
        s = gA[0, R]
        for q in range(1, 19):
            s = s + gA[q, R]
        self.rho[R] = s
        G = gA[:, R]
        inv_rho = 1.0 / self.rho[R]
        self.ux[R] = (G[1] - G[2] + G[7] - G[8] + G[9] - G[10] + G[11] - G[12] + G[13] - G[14]) * inv_rho
        self.uy[R] = (G[3] - G[4] + G[7] - G[8] - G[9] + G[10] + G[15] - G[16] + G[17] - G[18]) * inv_rho
        self.uz[R] = (G[5] - G[6] + G[11] - G[12] - G[13] + G[14] + G[15] - G[16] - G[17] + G[18]) * inv_rho
        # phi halo: pack top/bottom real planes, self-exchange, unpack to ghosts
        top_snd, bot_snd = self.phi[self.NZ - 2].copy(), self.phi[1].copy()
        self.phi[0], self.phi[self.NZ - 1] = top_snd, bot_snd
        # collisions, A -> B, plane by plane
        for k in range(1, self.NZ - 1):
            p = self._p(self.phi, k)
            phin, mu, fx, fy, fz = self._force(p, grav_on)
            rhon = self.rho[k]
            inv_rho = 1.0 / rhon
            uxn = self.ux[k] + 0.5 * fx * inv_rho
            uyn = self.uy[k] + 0.5 * fy * inv_rho
            uzn = self.uz[k] + 0.5 * fz * inv_rho
            af = 0.5 * c.gamma * c.inv_tau_phi * mu
            cf = c.inv_tau_phi * c.inv_eta2 * phin
            t = c.inv_tau_phi_one
            fB[0, k] = t * fA[0, k] - 6.0 * af + c.inv_tau_phi * phin
            fB[1, k] = t * fA[1, k] + af + cf * uxn
            fB[2, k] = t * fA[2, k] + af - cf * uxn
            fB[3, k] = t * fA[3, k] + af + cf * uyn
            fB[4, k] = t * fA[4, k] + af - cf * uyn
            fB[5, k] = t * fA[5, k] + af + cf * uzn
            fB[6, k] = t * fA[6, k] + af - cf * uzn
            # collision_g recomputes u' with a division by rho
            ux = self.ux[k] + 0.5 * fx / rhon
            uy = self.uy[k] + 0.5 * fy / rhon
            uz = self.uz[k] + 0.5 * fz / rhon
            K1, K2, KC1, KC2, tg = c.k1, c.k2, c.kc1, c.kc2, c.inv_tau_rho_one
            temp = K1 * (3.0 * phin * mu + rhon * (1.0 - 1.5 * (ux * ux + uy * uy + uz * uz)))
            uf = ux * fx + uy * fy + uz * fz
            ux, uy, uz = ux * 3.0, uy * 3.0, uz * 3.0
            gB[0, k] = (tg * gA[0, k] + c.k0 * (rhon - 6.0 * phin * mu
                        - rhon * (ux * ux + uy * uy + uz * uz) / 6.0) - c.kc0 * uf)
            e1 = temp + K1 * rhon * ux * ux * 0.5 + KC1 * (ux * fx - uf)
            e2 = ux * (K1 * rhon + KC1 * fx)
            gB[1, k] = tg * gA[1, k] + e1 + e2
            gB[2, k] = tg * gA[2, k] + e1 - e2
            e1 = temp + K1 * rhon * uy * 0.5 * uy + KC1 * (uy * fy - uf)
            e2 = uy * (K1 * rhon + KC1 * fy)
            gB[3, k] = tg * gA[3, k] + e1 + e2
            gB[4, k] = tg * gA[4, k] + e1 - e2
            e1 = temp + K1 * rhon * uz * 0.5 * uz + KC1 * (uz * fz - uf)
            e2 = uz * (K1 * rhon + KC1 * fz)
            gB[5, k] = tg * gA[5, k] + e1 + e2
            gB[6, k] = tg * gA[6, k] + e1 - e2
            temp = K2 * (3.0 * phin * mu + rhon * (1.0 - (ux * ux + uy * uy + uz * uz) / 6.0))
            for (qp, qm, sv, Sv) in ((7, 8, ux + uy, fx + fy), (9, 10, ux - uy, fx - fy),
                                     (11, 12, ux + uz, fx + fz), (13, 14, ux - uz, fx - fz),
                                     (15, 16, uy + uz, fy + fz), (17, 18, uy - uz, fy - fz)):
                e1 = temp + K2 * rhon * sv * 0.5 * sv + KC2 * (sv * Sv - uf)
                e2 = K2 * rhon * sv + KC2 * Sv
                gB[qp, k] = tg * gA[qp, k] + e1 + e2
                gB[qm, k] = tg * gA[qm, k] + e1 - e2
        # f halo before smoothing: outward f_B,6 / f_B,5 into the ghosts
        fB[6, 0] = fB[6, self.NZ - 2]
        fB[5, self.NZ - 1] = fB[5, 1]
        # push streaming from every real plane
        J, I = np.meshgrid(np.arange(self.ny), np.arange(self.nx), indexing="ij")
        IE, IW, JN, JS = self.ie[I], self.iw[I], self.jn[J], self.js[J]
        eta, eta2 = c.eta, c.eta2
        for k in range(1, self.NZ - 1):
            kt, kb = k + 1, k - 1
            fA[0, k] = fB[0, k]
            fA[1, k][J, IE] = eta * fB[1, k][J, I] + eta2 * fB[1, k][J, IE]
            fA[2, k][J, IW] = eta * fB[2, k][J, I] + eta2 * fB[2, k][J, IW]
            fA[3, k][JN, I] = eta * fB[3, k][J, I] + eta2 * fB[3, k][JN, I]
            fA[4, k][JS, I] = eta * fB[4, k][J, I] + eta2 * fB[4, k][JS, I]
            fA[5, kt] = eta * fB[5, k] + eta2 * fB[5, kt]
            fA[6, kb] = eta * fB[6, k] + eta2 * fB[6, kb]
            gA[0, k] = gB[0, k]
            for q, (jj, ii, kk) in {1: (J, IE, k), 2: (J, IW, k), 3: (JN, I, k), 4: (JS, I, k),
                                    5: (J, I, kt), 6: (J, I, kb), 7: (JN, IE, k), 8: (JS, IW, k),
                                    9: (JS, IE, k), 10: (JN, IW, k), 11: (J, IE, kt),
                                    12: (J, IW, kb), 13: (J, IE, kb), 14: (J, IW, kt),
                                    15: (JN, I, kt), 16: (JS, I, kb), 17: (JN, I, kb),
                                    18: (JS, I, kt)}.items():
                gA[q, kk][jj, ii] = gB[q, k][J, I]
        # population halo: outward A pops in the ghosts re-enter at the far face
        NZ = self.NZ
        for q in (5,):
            fA[q, 1] = fA[q, NZ - 1]
        for q in (6,):
            fA[q, NZ - 2] = fA[q, 0]
        for q in (5, 11, 14, 15, 18):
            gA[q, 1] = gA[q, NZ - 1]
        for q in (6, 12, 13, 16, 17):
            gA[q, NZ - 2] = gA[q, 0]
