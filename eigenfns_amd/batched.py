"""Batched, continuously admitted deflated block LOBPCG for many small problems.

The serving path: many (structure, Bloch k-point, band count)
solves of one grid size advance in lockstep, one LOBPCG iteration per engine
step, through shared launches — multi-operator Θ / preconditioner
(efn_theta_multi: each slot's vectors carry their own frame tables and ε),
multi-problem Gram and combine (efn_gram_multi / efn_combine_multi), strided
row reductions over every slot's rows at once — and new requests take a slot
as soon as one frees (continuous batching).  Each slot runs the reference's
algorithm (eigenfns/solver.py:136-382): guard warm start, re-deflation every
iteration, SVQB (1e-5 drop), fixed-shape P whitening, Rayleigh–Ritz with the
dead-row penalty, lock the lowest m − guard.  A slot without P keeps zero P
rows, which the dead-row penalty removes from the Rayleigh–Ritz — the
reference's [X, W] basis.

Batch invariance: every per-slot quantity is computed from that slot's data
only, with reduction orders that depend only on the grid size and block width
(Gram splits depend on the row length only, Θ is per vector, random rows come
from the request's own counter-based stream), so a request's eigenvalues are
bitwise the same alone and in any mix.

Storage per slot (fixed, HBM-resident): [X; W; P] and [ΘX; ΘW; ΘP] (3m rows
each), R, two m-row temporaries, the carry and the locked rows (zero-padded to
the engine capacity so deflation runs at one shape).
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np
import torch

from .linalg import (COMBINE_IN_PLACE_MAX, DEAD_SPREAD, ORTHO_METHODS, RR_DEAD_PENALTY, SVQB_DROP,
                     cholqr_coeffs_batched, combine_multi,
                     combine_multi_split, copy_multi, dead_spread, eigh_batched_device, eigh_stack_host,
                     gram_multi, one_blas_thread, rowreduce_strided, row_axpby_norm_strided,
                     rowstats_strided, MAX_PROBLEMS)
from .operator import MaxwellOperator, precondition_multi, theta_multi
from .solver import STAGNATION_WINDOW, StagnationLock


@dataclass
class Request:
    """One eigenproblem: lowest `nev` bands of `op`."""

    op: MaxwellOperator
    nev: int
    seed: int = 0
    rid: int = 0
    meta: dict = field(default_factory=dict)
    # filled by the engine
    vals: np.ndarray | None = None
    res: np.ndarray | None = None
    vecs: torch.Tensor | None = None
    iterations: int = 0
    theta_applications: int = 0
    converged: bool = True       # every locked block met the residual tolerance
    stagnated_blocks: int = 0    # blocks locked at the FP32 residual floor above tol
    t_arrive: float = 0.0
    t_admit: float = 0.0
    t_done: float = 0.0


class _Slot:
    def __init__(self, req: Request, m: int, guard: int):
        self.req = req
        self.op = req.op
        self.nev = req.nev
        self.m = m
        self.guard = guard
        self.locked_vals = np.empty((0,))
        self.locked_res = np.empty((0,))
        self.n_locked = 0
        self.nc = 0                  # carry rows waiting in the carry buffer
        self.block_idx = 0
        self.phase = "init"
        self.it = 0
        self.n_lock = 0
        self.stag = None


class BatchedLOBPCG:
    """Lockstep LOBPCG over up to `nslots` (≤ 32) problems of grid size G.

    Parameters follow lobpcg_blocks (m, guard, tol, maxit); `max_nev` sizes the
    per-slot locked storage.  `keep_vectors` copies each finished request's
    locked vectors out of the engine (otherwise only eigenvalues and residuals
    are returned)."""

    def __init__(self, grid_size: int, m: int = 24, guard: int = 8, nslots: int = 8,
                 max_nev: int = 64, tol: float = 1e-4, maxit: int = 400,
                 keep_vectors: bool = False, normalized_rq: bool = True,
                 verify_dead_rows: bool = False, small_dense: str = "host",
                 stagnation: int = STAGNATION_WINDOW, ortho: str = "cholqr"):
        if not 0 < guard < m:
            raise ValueError("need 0 < guard < m")
        if not 1 <= nslots <= MAX_PROBLEMS:
            raise ValueError(f"nslots must be in [1, {MAX_PROBLEMS}]")
        self.G = int(grid_size)
        self.m = int(m)
        self.guard = int(guard)
        self.nslots = int(nslots)
        self.tol = float(tol)
        self.maxit = int(maxit)
        self.stagnation = int(stagnation)  # lobpcg_blocks' FP32-floor lock, per slot
        self.keep_vectors = keep_vectors
        self.normalized_rq = normalized_rq
        self.verify_dead_rows = verify_dead_rows  # also measure the row norms and compare (tests)
        # the SVQB / whitening / Rayleigh–Ritz eigensolves: "host" (LAPACK over a
        # thread pool, masks and normalisation in NumPy; the default), "device"
        # (efn_heev_batched and the masks on the device: no host round trip) or
        # "hybrid" (SVQB and whitening on the device, Rayleigh–Ritz on the
        # host).  The device Jacobi takes 7.7 ms per 72 × 72 Rayleigh–Ritz
        # matrix (its eigenvectors live in global memory: A and V do not both
        # fit in LDS at n = 72) against 0.4 ms for one host LAPACK call, and the
        # engines' syncs are already overlapped by the second engine per grid.
        if small_dense not in ("device", "host", "hybrid"):
            raise ValueError(f"small_dense must be 'device', 'host' or 'hybrid', got {small_dense!r}")
        self.small_dense = small_dense
        # the SVQB steps and the P whitening by pivoted Cholesky-QR on the
        # device ("cholqr", the default: no host round trip for them in any
        # small_dense mode) or by eigendecomposition ("svqb", the reference's)
        if ortho not in ORTHO_METHODS:
            raise ValueError(f"ortho must be one of {ORTHO_METHODS}, got {ortho!r}")
        self.ortho = ortho
        per = self.m - self.guard
        self.cap = int(np.ceil(max_nev / per)) * per + self.m
        self.max_nev = int(max_nev)
        G = self.G
        shape = (2, G, G, G)
        self.n = 2 * G ** 3
        # workspace owner for the multi launches (ε = 1; only its workspace is used)
        self.ws = MaxwellOperator(np.ones((G, G, G), np.float32), 1.0)
        dev = self.ws._dev
        c64 = torch.complex64
        m3 = 3 * self.m
        self.S = torch.zeros((nslots * m3,) + shape, dtype=c64, device=dev)
        self.HS = torch.zeros_like(self.S)
        self.R = torch.zeros((nslots * self.m,) + shape, dtype=c64, device=dev)
        # per slot [T1_s; T2_s] (2m rows) of scratch, needed only when a combine
        # cannot run in place (2m > COMBINE_IN_PLACE_MAX)
        self.T12 = (torch.zeros((2 * nslots * self.m,) + shape, dtype=c64, device=dev)
                    if 2 * self.m > COMBINE_IN_PLACE_MAX else None)
        self.CARRY = torch.zeros_like(self.R)
        self.L = torch.zeros((nslots * self.cap,) + shape, dtype=c64, device=dev)
        self.slots: list[_Slot | None] = [None] * nslots
        self.steps = 0
        self.timing = {}  # host seconds per phase (includes waits at the phase's syncs)

    def _tick(self, name, t0):
        t1 = time.perf_counter()
        self.timing[name] = self.timing.get(name, 0.0) + (t1 - t0)
        return t1

    # ---- bookkeeping --------------------------------------------------------
    def free_slots(self) -> int:
        return sum(s is None for s in self.slots)

    def active(self) -> int:
        return self.nslots - self.free_slots()

    def bytes_per_slot(self) -> int:
        return (2 * 3 * self.m + 4 * self.m + self.cap) * self.n * 8

    def admit(self, req: Request) -> int:
        if req.op.grid_size != self.G:
            raise ValueError("request grid size does not match the engine")
        if req.nev > self.max_nev:
            raise ValueError(f"nev {req.nev} exceeds the engine capacity {self.max_nev}")
        for s, sl in enumerate(self.slots):
            if sl is None:
                self.slots[s] = _Slot(req, self.m, self.guard)
                self.L[s * self.cap:(s + 1) * self.cap].zero_()
                req.t_admit = time.perf_counter()
                return s
        raise RuntimeError("no free slot")

    # ---- row views ----------------------------------------------------------
    def _X(self, s):
        return self.S[s * 3 * self.m:s * 3 * self.m + self.m]

    def _W(self, s):
        return self.S[s * 3 * self.m + self.m:s * 3 * self.m + 2 * self.m]

    def _P(self, s):
        return self.S[s * 3 * self.m + 2 * self.m:(s + 1) * 3 * self.m]

    def _XW(self, s):
        return self.S[s * 3 * self.m:s * 3 * self.m + 2 * self.m]

    def _Sb(self, s):
        return self.S[s * 3 * self.m:(s + 1) * 3 * self.m]

    def _HSb(self, s):
        return self.HS[s * 3 * self.m:(s + 1) * 3 * self.m]

    def _rows(self, buf, s, width):
        return buf[s * width:(s + 1) * width]

    def _t1(self, s):
        return self.T12[2 * s * self.m:(2 * s + 1) * self.m]

    def _t2(self, s):
        return self.T12[(2 * s + 1) * self.m:(2 * s + 2) * self.m]

    def _combine_back(self, idx, C, blocks):
        """block_j ← C_jᵀ·block_j for slots idx: in place up to
        COMBINE_IN_PLACE_MAX rows, else through T1 and a copy."""
        if self.m <= COMBINE_IN_PLACE_MAX:
            combine_multi(C, blocks, blocks)
        else:
            Ts = [self._t1(s) for s in idx]
            combine_multi(C, blocks, Ts)
            copy_multi(Ts, blocks)

    def _Lk(self, s, k):
        return self.L[s * self.cap:s * self.cap + k]

    # ---- batched building blocks ---------------------------------------------
    def _deflate(self, idx, target):
        """target(s) ← target(s) − L_s(L_sᴴ target(s)) for slots idx (zero-padded locked rows)."""
        nl = max(self.slots[s].n_locked for s in idx)
        if nl == 0:
            return
        Ls = [self._Lk(s, nl) for s in idx]
        Ts = [target(s) for s in idx]
        C = gram_multi(Ls, Ts)
        combine_multi(-C, Ls, Ts, Xs=Ts, alpha=1.0)

    @staticmethod
    def _whiten(w, V):
        """V·w^{-1/2} with directions below SVQB_DROP·max(w) (or all, if max(w) = 0)
        zeroed (solver.py:119-124, :265-272), and the keep mask."""
        wmax = w.max(dim=1, keepdim=True).values
        keep = (w > SVQB_DROP * torch.clamp(wmax, min=1e-300)) & (wmax > 0)
        scale = torch.where(keep, 1.0 / torch.sqrt(torch.where(keep, w, torch.ones_like(w))),
                            torch.zeros_like(w))
        # select rather than scale by zero: a dropped direction's vector may be non-finite
        return torch.where(keep[:, None, :], V * scale[:, None, :], torch.zeros_like(V)), keep

    def _whiten_small(self, G):
        """(V·w^{-1/2} with dropped directions zeroed, keep mask) of a Gram stack:
        on the host in NumPy when the eigensolve is there (one copy of the
        coefficients back; masks as NumPy), on the device otherwise.  With
        Cholesky-QR the coefficients and mask (rows below the rank) come from
        the device kernel in every mode."""
        if self.ortho == "cholqr":
            C, rank = cholqr_coeffs_batched(G)
            return C, torch.arange(G.shape[1], device=G.device)[None, :] < rank[:, None]
        if self.small_dense == "host":
            w, V = eigh_stack_host(G.cpu().numpy())
            wmax = w.max(axis=1, keepdims=True)
            keep = (w > SVQB_DROP * np.maximum(wmax, 1e-300)) & (wmax > 0)
            scale = np.where(keep, 1.0 / np.sqrt(np.where(keep, w, 1.0)), 0.0)
            return torch.as_tensor(V * scale[:, None, :], device=G.device), keep
        return self._whiten(*eigh_batched_device(G))

    def _svqb(self, idx, block, passes=2):
        """SVQB of block(s) for slots idx, two sweeps, written back in place.
        Returns the last sweep's keep mask (len(idx), rows): a dropped direction
        is an exactly zero row of the result."""
        keep = None
        for _ in range(passes):
            Bs = [block(s) for s in idx]
            Cw, keep = self._whiten_small(gram_multi(Bs, Bs, hermitian=True))
            self._combine_back(idx, Cw, Bs)
        return keep

    def _check_dead(self, idx, dead):
        """Compare the dead-row bookkeeping with measured row norms (tests)."""
        m = self.m
        rn2 = rowreduce_strided(0, self.S, None, 3 * m, 3 * m, 3 * m, self.nslots * 3 * m,
                                self.n).real.cpu().numpy().reshape(self.nslots, 3 * m)
        for j, s in enumerate(idx):
            if not np.array_equal(dead[j], np.sqrt(rn2[s]) < 0.5):
                raise AssertionError(f"slot {s}: dead-row bookkeeping disagrees with row norms")

    def _rr_host(self, idx, A, keep_w, keepp):
        """Rayleigh–Ritz coefficients [c/‖c‖ | c'/‖c'‖] with the eigensolve on the
        host: penalty, eigensolve and normalisation in NumPy, one copy back.
        Dead rows (solver.py:296-297, norm < 0.5) without another pass over S:
        X norms from this step's residual statistics; W and P rows are unit or
        exactly zero after SVQB / whitening (their keep masks)."""
        m = self.m
        A = A.cpu().numpy()
        keep_w = keep_w.cpu().numpy() if torch.is_tensor(keep_w) else keep_w
        keepp = keepp.cpu().numpy() if torch.is_tensor(keepp) else keepp
        dead = np.stack([~np.concatenate([self._xn[s] >= 0.5, keep_w[j], keepp[j]])
                         for j, s in enumerate(idx)])
        if self.verify_dead_rows:
            self._check_dead(idx, dead)
        for j in range(len(idx)):   # symmetrised, dead rows penalised (solver.py:294-297)
            A[j] = (A[j] + A[j].conj().T) / 2
            penalty = max(RR_DEAD_PENALTY, 100.0 * float(np.abs(np.diag(A[j])).max()))
            A[j][np.diag_indices(3 * m)] += np.where(dead[j], penalty * (1.0 + DEAD_SPREAD * np.arange(3 * m)),
                                                     0.0)
        _, C = eigh_stack_host(A, lowest=m)
        Cp = C.copy()
        Cp[:, :m] = 0
        nx = np.maximum(np.linalg.norm(C, axis=1), 1e-20)
        npn = np.maximum(np.linalg.norm(Cp, axis=1), 1e-20)
        return torch.as_tensor(np.concatenate([C / nx[:, None, :], Cp / npn[:, None, :]], axis=2),
                               device=self.S.device)

    def _rr_device(self, idx, A, keep_w, keepp):
        """_rr_host with the penalty, eigensolve and normalisation on the device."""
        m = self.m
        xn = torch.stack([self._xn_dev[s] for s in idx])   # views: no host-to-device copy
        dead = ~torch.cat([xn >= 0.5, keep_w, keepp], dim=1)
        if self.verify_dead_rows:
            self._check_dead(idx, dead.cpu().numpy())
        A = (A + A.conj().transpose(1, 2)) / 2
        penalty = torch.clamp(100.0 * torch.diagonal(A, dim1=1, dim2=2).abs().max(dim=1).values,
                              min=RR_DEAD_PENALTY)
        A = A + torch.diag_embed(torch.where(dead, penalty[:, None] * dead_spread(3 * m, A.device)[None, :],
                                             0.0).to(A.dtype))
        _, C = eigh_batched_device(A, lowest=m)
        Cp = C.clone()
        Cp[:, :m] = 0
        nx = torch.clamp(torch.linalg.vector_norm(C, dim=1), min=1e-20)
        npn = torch.clamp(torch.linalg.vector_norm(Cp, dim=1), min=1e-20)
        return torch.cat([C / nx[:, None, :], Cp / npn[:, None, :]], dim=2)

    def _theta(self, idx, src, row0_src, dst, row0_dst, stride_src, stride_dst, nrow):
        for c in range(0, len(idx), MAX_PROBLEMS):
            part = idx[c:c + MAX_PROBLEMS]
            theta_multi(self.ws, [self.slots[s].op for s in part], src,
                        [s * stride_src + row0_src for s in part], dst,
                        [s * stride_dst + row0_dst for s in part], nrow)
        for s in idx:
            self.slots[s].req.theta_applications += nrow

    # ---- phases -------------------------------------------------------------
    def _init_blocks(self, idx):
        m = self.m
        for s in idx:
            sl = self.slots[s]
            X = self._X(s)
            if sl.nc:
                X[:sl.nc].copy_(self.CARRY[s * m:s * m + sl.nc])
            if sl.nc < m:
                Z = sl.op.random_block(m, sl.req.seed, kind="normal", row0=sl.block_idx * m)
                Z = Z[:m - sl.nc]
                nrm = torch.sqrt(torch.sum(torch.abs(Z.reshape(Z.shape[0], -1)) ** 2, dim=1))
                X[sl.nc:] = Z / torch.clamp(nrm, min=1e-30)[:, None, None, None, None]
            sl.block_idx += 1
            self._P(s).zero_()
        self._deflate(idx, self._X)
        self._svqb(idx, self._X)
        self._theta(idx, self.S, 0, self.HS, 0, 3 * m, 3 * m, m)
        for s in idx:
            sl = self.slots[s]
            sl.phase = "iter"
            sl.it = 0
            sl.stag = StagnationLock(self.tol, self.stagnation)
            sl.n_lock = min(m - self.guard, sl.nev - sl.locked_vals.size)

    def _iterate(self, idx, lam):
        m = self.m
        t = time.perf_counter()
        # W = K⁻¹R, deflated, projected out of X twice, SVQB (solver.py:251-257)
        for c in range(0, len(idx), MAX_PROBLEMS):
            part = idx[c:c + MAX_PROBLEMS]
            precondition_multi(self.ws, [self.slots[s].op for s in part], self.R,
                               [s * m for s in part], self.S, [s * 3 * m + m for s in part], m)
        self._deflate(idx, self._W)
        for _ in range(2):
            Xs = [self._X(s) for s in idx]
            Ws = [self._W(s) for s in idx]
            combine_multi(-gram_multi(Xs, Ws), Xs, Ws, Xs=Ws, alpha=1.0)
        t = self._tick("precond+deflate+proj", t)
        keep_w = self._svqb(idx, self._W)
        t = self._tick("svqb_W", t)
        # P projected out of [X, W] twice, fixed-shape whitening (solver.py:258-274)
        for _ in range(2):
            XWs = [self._XW(s) for s in idx]
            Ps = [self._P(s) for s in idx]
            combine_multi(-gram_multi(XWs, Ps), XWs, Ps, Xs=Ps, alpha=1.0)
        Ps = [self._P(s) for s in idx]
        Cp, keepp = self._whiten_small(gram_multi(Ps, Ps, hermitian=True))
        self._combine_back(idx, Cp, Ps)
        t = self._tick("P", t)
        # Θ(W), Θ(P) fresh; Rayleigh–Ritz on [X; W; P] (solver.py:275-302)
        self._theta(idx, self.S, m, self.HS, m, 3 * m, 3 * m, 2 * m)
        t = self._tick("theta_WP", t)
        Sbs = [self._Sb(s) for s in idx]
        A = gram_multi(Sbs, [self._HSb(s) for s in idx], hermitian=True)
        # X ← S·c_i/‖S·c_i‖ and P ← S·c'_i/‖S·c'_i‖ (c' = c with its X rows
        # zeroed; solver.py:298-307).  [X; W; P] is orthonormal by construction
        # (SVQB, projections, whitening) and dead rows get exactly zero weight,
        # so ‖S·c‖ = ‖c‖ to FP32 rounding: the norms of the coefficient vectors
        # are folded into them, and one split combine writes both updates
        # straight into the slot's X and P rows.
        if self.small_dense == "device":
            C2 = self._rr_device(idx, A, keep_w, keepp)
        else:
            C2 = self._rr_host(idx, A, keep_w, keepp)
        t = self._tick("RR_gram+eigh", t)
        Xs = [self._X(s) for s in idx]
        Ps = [self._P(s) for s in idx]
        if 2 * m <= COMBINE_IN_PLACE_MAX:
            combine_multi_split(C2, Sbs, Xs, Ps, m)
        else:
            T1s = [self._t1(s) for s in idx]
            T2s = [self._t2(s) for s in idx]
            combine_multi_split(C2, Sbs, T1s, T2s, m)
            copy_multi(T1s, Xs)
            copy_multi(T2s, Ps)
        self._deflate(idx, self._X)
        t = self._tick("update_XP", t)
        self._theta(idx, self.S, 0, self.HS, 0, 3 * m, 3 * m, m)
        t = self._tick("theta_X", t)
        for s in idx:
            self.slots[s].it += 1
            self.slots[s].req.iterations += 1

    def _finish_block(self, s, lam, rn, order):
        """Lock the lowest n_lock rows, keep the rest as carry (solver.py:316-345)."""
        sl = self.slots[s]
        m = self.m
        if not np.isfinite(lam[order[:sl.n_lock]]).all():
            raise RuntimeError(f"request {sl.req.rid}: rank collapse in block {sl.block_idx}")
        X = self._X(s)
        lock_idx = torch.as_tensor(order[:sl.n_lock], device=X.device)
        carry_idx = torch.as_tensor(order[sl.n_lock:], device=X.device)
        base = s * self.cap
        self.L[base + sl.n_locked:base + sl.n_locked + sl.n_lock] = X[lock_idx]
        sl.nc = int(carry_idx.numel())
        self.CARRY[s * m:s * m + sl.nc] = X[carry_idx]
        sl.n_locked += sl.n_lock
        sl.locked_vals = np.concatenate([sl.locked_vals, lam[order[:sl.n_lock]]])
        sl.locked_res = np.concatenate([sl.locked_res, rn[order[:sl.n_lock]]])
        if sl.locked_vals.size >= sl.nev:
            req = sl.req
            o = np.argsort(sl.locked_vals, kind="stable")
            req.vals = sl.locked_vals[o][:sl.nev]
            req.res = sl.locked_res[o][:sl.nev]
            if self.keep_vectors:
                req.vecs = self.L[base:base + sl.n_locked][torch.as_tensor(o, device=X.device)][
                    :sl.nev].clone()
            req.t_done = time.perf_counter()
            self.slots[s] = None
            return req
        sl.phase = "init"
        return None

    # ---- one engine step = one LOBPCG iteration for every active slot -------
    def step(self) -> list:
        finished = []
        act = [s for s, sl in enumerate(self.slots) if sl is not None]
        if not act:
            return finished
        self.steps += 1
        m = self.m
        t = time.perf_counter()
        init = [s for s in act if self.slots[s].phase == "init"]
        if init:
            self._init_blocks(init)
            t = self._tick("init", t)
        # Rayleigh quotients, residuals, dead rows for the X rows of every slot
        M = self.nslots * m
        # one read of X and ΘX for ⟨x,Θx⟩, |x|², |Θx|²; R = ΘX − λX with |R|² in the same pass
        st = rowstats_strided(self.S, self.HS, m, 3 * m, 3 * m, M, self.n)
        lam, xn2, hx2 = st[:, 0], st[:, 2], st[:, 3]
        if self.normalized_rq:
            lam = lam / torch.clamp(xn2, min=1e-30)
        rr2 = row_axpby_norm_strided(self.HS, 3 * m, torch.ones_like(lam), self.S, 3 * m, -lam, m,
                                     M, self.n, self.R, m)
        self._xn_dev = torch.sqrt(xn2).view(self.nslots, m)   # device masks (small_dense="device")
        stats = torch.stack([lam, xn2, rr2, hx2]).cpu().numpy()
        t = self._tick("residuals+sync", t)
        lam_h = stats[0].reshape(self.nslots, m)
        xn = np.sqrt(stats[1]).reshape(self.nslots, m)
        self._xn = xn
        rn_h = (np.sqrt(stats[2]) / np.maximum(np.sqrt(stats[3]), 1e-30)).reshape(self.nslots, m)
        cont = []
        for s in act:
            sl = self.slots[s]
            dead = xn[s] < 0.5
            l = np.where(dead, np.inf, lam_h[s])
            r = np.where(dead, np.inf, rn_h[s])
            order = np.argsort(l)
            worst = float(r[order[:sl.n_lock]].max())
            if worst < self.tol or sl.it >= self.maxit or sl.stag.stalled(worst):
                if not worst < self.tol:
                    sl.req.converged = False
                    sl.req.stagnated_blocks += sl.it < self.maxit
                req = self._finish_block(s, l, r, order)
                if req is not None:
                    finished.append(req)
            else:
                cont.append(s)
        t = self._tick("lock", t)
        if cont:
            self._iterate(cont, lam)
        return finished


def solve_batched(requests, grid_size, m=24, guard=8, nslots=8, max_nev=None, tol=1e-4,
                  maxit=400, keep_vectors=False, verify_dead_rows=False, small_dense="host",
                  ortho="cholqr"):
    """Solve a list of same-grid requests through one engine (admission as slots free)."""
    max_nev = max_nev or max(r.nev for r in requests)
    eng = BatchedLOBPCG(grid_size, m, guard, nslots, max_nev, tol, maxit, keep_vectors,
                        verify_dead_rows=verify_dead_rows, small_dense=small_dense, ortho=ortho)
    pending = list(requests)
    done = []
    with one_blas_thread():
        while pending or eng.active():
            while pending and eng.free_slots():
                eng.admit(pending.pop(0))
            done.extend(eng.step())
    return done
