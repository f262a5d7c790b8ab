"""Batched, continuously admitted deflated block LOBPCG for many small problems.

The serving path: many (structure, Bloch k-point, band count)
solves of one grid size advance in lockstep, one LOBPCG iteration per engine
step, through shared launches — multi-operator Θ / preconditioner
(efn_theta_multi: each slot's vectors carry their own frame tables and ε),
multi-problem Gram and combine (efn_gram_multi / efn_combine_multi), row
reductions over the occupied slots' rows at once — and new requests take a slot
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

Block widths may differ per request (``Request.m`` / ``Request.guard``, at
most the engine's width): every launch is ragged — slot s runs its own m_s
rows through the same Θ / preconditioner / Gram / combine launches
(efn_theta_ragged, efn_gram_ragged, efn_combine_ragged) — so problems of
different widths (the k-points of a band-structure path) share one engine.

Storage per slot (fixed, HBM-resident, sized by the engine's width M):
[X; W; P] and [ΘX; ΘW; ΘP] (3M rows each; a slot of width m uses the leading
3m), R, two M-row temporaries, the carry and the locked rows (zero-padded to
the engine capacity so deflation runs at one shape).
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np
import torch

from .linalg import (COMBINE_IN_PLACE_MAX, DEAD_SPREAD, ORTHO_METHODS, QL_MAX_N, RR_DEAD_PENALTY,
                     SVQB_DROP, cholqr_coeffs_batched, combine_ragged, dead_spread, eigh_batched_device,
                     eigh_ragged_device,
                     eigh_list_host, gram_ragged, one_blas_thread, rowreduce_strided,
                     row_axpby_norm_sel, rowstats_sel, MAX_PROBLEMS)
from .operator import MaxwellOperator, precondition_ragged, theta_ragged
from .solver import STAGNATION_WINDOW, StagnationLock


@dataclass
class Request:
    """One eigenproblem: lowest `nev` bands of `op`."""

    op: MaxwellOperator
    nev: int
    seed: int = 0
    rid: int = 0
    meta: dict = field(default_factory=dict)
    m: int | None = None       # block width (None: the engine's)
    guard: int | None = None   # guard rows (None: the engine's)
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


def rr_dead_rows(widths, K: int, xn: torch.Tensor, keep_w: torch.Tensor,
                 keepp: torch.Tensor) -> torch.Tensor:
    """Dead rows (len(widths), K) of padded Rayleigh–Ritz stacks
    (solver.py:296-297): row r of problem j is X row r (r < w_j), W row r − w_j
    (r < 2w_j) or P row r − 2w_j (r < 3w_j), dead when the X row's norm is
    below 0.5 or the W / P direction was dropped by the whitening; rows from
    3w_j on are padding, not dead.  xn (X row norms), keep_w and keepp are
    (nprob, ≥ max width) — a ragged launch pads them to its widest problem
    only — and every gather index is clamped to the operand's own width."""
    wmax = max(widths)
    if K < 3 * wmax:
        raise ValueError(f"rr_dead_rows: K = {K} below three times the widest problem ({wmax})")
    for name, t in (("xn", xn), ("keep_w", keep_w), ("keepp", keepp)):
        if t.dim() != 2 or t.shape[0] != len(widths) or t.shape[1] < wmax:
            raise ValueError(f"rr_dead_rows: {name} {tuple(t.shape)} narrower than the widest "
                             f"problem ({len(widths)} × {wmax})")
    dev = xn.device
    r = torch.arange(K, device=dev)[None, :]
    wt = torch.as_tensor(widths, device=dev)[:, None]

    def take(mask, i):
        return torch.gather(mask, 1, torch.clamp(i, 0, mask.shape[1] - 1).expand(len(widths), K))

    alive = torch.where(r < wt, take(xn, r) >= 0.5,
                        torch.where(r < 2 * wt, take(keep_w, r - wt), take(keepp, r - 2 * wt)))
    return ~alive & (r < 3 * wt)


class _Slot:
    def __init__(self, req: Request, m: int, guard: int):
        # m: this request's block width (≤ the engine's), guard its guard rows
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
    """Lockstep LOBPCG over up to `nslots` (≤ 64) problems of grid size G.

    Parameters follow lobpcg_blocks (m, guard, tol, maxit); `m` is also the
    widest block a request may bring (Request.m), `max_nev` sizes the
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
        # (the library's QL kernel in one ragged launch up to 96 × 96,
        # rocSOLVER per distinct size above, and the masks on the device: no
        # host round trip) or "hybrid" (SVQB and whitening on the device,
        # Rayleigh–Ritz on the host).  A device call takes about 2.3 ms at
        # 72 × 72 whatever the batch, against 0.4 ms for one host LAPACK call
        # (many run at once in the pool), so the host stays the default.
        if small_dense not in ("device", "host", "hybrid"):
            raise ValueError(f"small_dense must be 'device', 'host' or 'hybrid', got {small_dense!r}")
        self.small_dense = small_dense
        # the SVQB steps and the P whitening by pivoted Cholesky-QR on the
        # device ("cholqr", the default: no host round trip for them in any
        # small_dense mode) or by eigendecomposition ("svqb", the reference's)
        if ortho not in ORTHO_METHODS:
            raise ValueError(f"ortho must be one of {ORTHO_METHODS}, got {ortho!r}")
        self.ortho = ortho
        # locked rows: at most nev plus one block's overshoot
        self.cap = int(max_nev) + self.m
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
        self._sel_key = None  # occupied (slot, width) pairs of the cached row selection

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
        m = int(req.m) if req.m is not None else self.m
        guard = int(req.guard) if req.guard is not None else self.guard
        if not 0 < m <= self.m:
            raise ValueError(f"request width {m} must be in [1, {self.m}] (the engine's width)")
        if not 0 < guard < m:
            raise ValueError("need 0 < guard < width")
        for s, sl in enumerate(self.slots):
            if sl is None:
                sl = self.slots[s] = _Slot(req, m, guard)
                # the slot's row views, fixed while it is occupied
                b, M = s * 3 * self.m, self.m
                sl.X, sl.W, sl.P = self.S[b:b + m], self.S[b + m:b + 2 * m], self.S[b + 2 * m:b + 3 * m]
                sl.XW, sl.Sb, sl.HSb = self.S[b:b + 2 * m], self.S[b:b + 3 * m], self.HS[b:b + 3 * m]
                if self.T12 is not None:
                    sl.t1 = self.T12[2 * s * M:2 * s * M + m]
                    sl.t2 = self.T12[(2 * s + 1) * M:(2 * s + 1) * M + m]
                self.L[s * self.cap:(s + 1) * self.cap].zero_()
                req.t_admit = time.perf_counter()
                return s
        raise RuntimeError("no free slot")

    # ---- row views (slot s of width w: X, W, P are rows [0, w), [w, 2w),
    # [2w, 3w) of its 3M-row region) ------------------------------------------
    def _w(self, s):
        return self.slots[s].m

    def _X(self, s):
        return self.slots[s].X

    def _W(self, s):
        return self.slots[s].W

    def _P(self, s):
        return self.slots[s].P

    def _XW(self, s):
        return self.slots[s].XW

    def _Sb(self, s):
        return self.slots[s].Sb

    def _HSb(self, s):
        return self.slots[s].HSb

    def _t1(self, s):
        return self.slots[s].t1

    def _t2(self, s):
        return self.slots[s].t2

    def _combine_back(self, idx, C, blocks):
        """block_j ← C_jᵀ·block_j for slots idx: in place up to
        COMBINE_IN_PLACE_MAX rows, else through T1 and a copy."""
        ws = [self._w(s) for s in idx]
        if self.m <= COMBINE_IN_PLACE_MAX:
            combine_ragged(C, blocks, ws, ws, blocks)
        else:
            Ts = [self._t1(s) for s in idx]
            combine_ragged(C, blocks, ws, ws, Ts)
            for T, B in zip(Ts, blocks):
                B.copy_(T)

    def _Lk(self, s, k):
        return self.L[s * self.cap:s * self.cap + k]

    # ---- batched building blocks ---------------------------------------------
    def _project_out(self, idx, basis, target, rows_b, rows_t):
        """target(s) ← target(s) − basis(s)(basis(s)ᴴ target(s)), one Gram and
        one combine launch over slots idx (ragged row counts)."""
        Bs = [basis(s) for s in idx]
        Ts = [target(s) for s in idx]
        C = gram_ragged(Bs, Ts)
        combine_ragged(-C, Bs, [rows_b(s) for s in idx], [rows_t(s) for s in idx], Ts, Xs=Ts, alpha=1.0)

    def _deflate(self, idx, target):
        """target(s) ← target(s) − L_s(L_sᴴ target(s)) for slots idx (zero-padded locked rows)."""
        nl = max(self.slots[s].n_locked for s in idx)
        if nl == 0:
            return
        self._project_out(idx, lambda s: self._Lk(s, nl), target, lambda s: nl, self._w)

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

    def _eigh_device_ragged(self, A, widths, lowest=None):
        """(w, V) of each problem's leading widths[j] block of the padded device
        stack A: one ragged launch of the library's QL kernel for the problems
        up to 96 × 96, one rocSOLVER call per width above that; padded back:
        V's leading widths[j] rows and min(lowest, width) columns, w padded
        with +inf."""
        nprob, K = int(A.shape[0]), int(A.shape[1])
        cols = K if lowest is None else lowest
        small = [j for j, wj in enumerate(widths) if 0 < wj <= QL_MAX_N]
        if len(small) == nprob and K <= QL_MAX_N and nprob <= MAX_PROBLEMS:
            return eigh_ragged_device(A, widths, cols)
        w = torch.full((nprob, cols), float("inf"), dtype=torch.float64, device=A.device)
        V = torch.zeros((nprob, K, cols), dtype=torch.complex128, device=A.device)
        if small:
            kq = max(widths[j] for j in small)
            jt = torch.as_tensor(small, device=A.device)
            lo = min(cols, kq)
            wq, Vq = eigh_ragged_device(A[jt][:, :kq, :kq].contiguous(), [widths[j] for j in small], lo)
            w[jt, :lo] = wq
            V[jt, :kq, :lo] = Vq
        for k in sorted(set(wj for wj in widths if wj > QL_MAX_N)):
            js = [j for j, wj in enumerate(widths) if wj == k]
            jt = torch.as_tensor(js, device=A.device)
            lo = None if lowest is None else min(lowest, k)
            wk, Vk = eigh_batched_device(A[jt][:, :k, :k].contiguous(), lowest=lo, method="dc")
            w[jt, :wk.shape[1]] = wk
            V[jt, :k, :Vk.shape[2]] = Vk
        return w, V

    def _whiten_small(self, G, widths):
        """(V·w^{-1/2} with dropped directions zeroed, keep mask) of a padded
        Gram stack (problem j of width widths[j]): on the host in NumPy when
        the eigensolve is there (one copy of the coefficients back; masks as
        NumPy), on the device otherwise.  With Cholesky-QR the coefficients and
        mask (rows below the rank) come from the device kernel in every mode:
        the padding is zero rows, which the pivoted Cholesky never selects."""
        if self.ortho == "cholqr":
            C, rank = cholqr_coeffs_batched(G)
            return C, torch.arange(G.shape[1], device=G.device)[None, :] < rank[:, None]
        K = int(G.shape[1])
        if self.small_dense == "host":
            Gh = G.cpu().numpy()
            res = eigh_list_host([Gh[j, :k, :k] for j, k in enumerate(widths)])
            Cw = np.zeros((len(widths), K, K), dtype=np.complex128)
            keep = np.zeros((len(widths), K), dtype=bool)
            for j, (k, (w, V)) in enumerate(zip(widths, res)):
                wmax = w.max()
                kp = (w > SVQB_DROP * max(wmax, 1e-300)) & (wmax > 0)
                Cw[j, :k, :k] = V * np.where(kp, 1.0 / np.sqrt(np.where(kp, w, 1.0)), 0.0)[None, :]
                keep[j, :k] = kp
            return torch.as_tensor(Cw, device=G.device), keep
        w, V = self._eigh_device_ragged(G, widths)
        w = torch.where(torch.isinf(w), torch.zeros_like(w), w)
        return self._whiten(w, V)

    def _svqb(self, idx, block, passes=2):
        """SVQB of block(s) for slots idx, two sweeps, written back in place.
        Returns the last sweep's keep mask (len(idx), M): a dropped direction
        is an exactly zero row of the result."""
        keep = None
        widths = [self._w(s) for s in idx]
        for _ in range(passes):
            Bs = [block(s) for s in idx]
            Cw, keep = self._whiten_small(gram_ragged(Bs, Bs, hermitian=True), widths)
            self._combine_back(idx, Cw, Bs)
        return keep

    def _check_dead(self, idx, dead):
        """Compare the dead-row bookkeeping with measured row norms (tests)."""
        M = self.m
        rn2 = rowreduce_strided(0, self.S, None, 3 * M, 3 * M, 3 * M, self.nslots * 3 * M,
                                self.n).real.cpu().numpy().reshape(self.nslots, 3 * M)
        for j, s in enumerate(idx):
            w = self._w(s)
            if not np.array_equal(dead[j], np.sqrt(rn2[s, :3 * w]) < 0.5):
                raise AssertionError(f"slot {s}: dead-row bookkeeping disagrees with row norms")

    def _pack_rr(self, idx, Cs):
        """Padded (nslots, 3M, 2M) [c/‖c‖ | c'/‖c'‖] of the per-slot lowest
        Ritz coefficient blocks Cs[j] (3w × w), c' = c with its X rows zeroed."""
        M = self.m
        out = np.zeros((len(idx), 3 * M, 2 * M), dtype=np.complex128)
        for j, (s, C) in enumerate(zip(idx, Cs)):
            w = self._w(s)
            Cp = C.copy()
            Cp[:w] = 0
            out[j, :3 * w, :w] = C / np.maximum(np.linalg.norm(C, axis=0), 1e-20)[None, :]
            out[j, :3 * w, w:2 * w] = Cp / np.maximum(np.linalg.norm(Cp, axis=0), 1e-20)[None, :]
        return out

    def _rr_host(self, idx, A, keep_w, keepp):
        """Rayleigh–Ritz coefficients [c/‖c‖ | c'/‖c'‖] with the eigensolve on the
        host: penalty, eigensolve and normalisation in NumPy, one copy back.
        Dead rows (solver.py:296-297, norm < 0.5) without another pass over S:
        X norms from this step's residual statistics; W and P rows are unit or
        exactly zero after SVQB / whitening (their keep masks)."""
        A = A.cpu().numpy()
        keep_w = keep_w.cpu().numpy() if torch.is_tensor(keep_w) else keep_w
        keepp = keepp.cpu().numpy() if torch.is_tensor(keepp) else keepp
        mats, dead = [], []
        for j, s in enumerate(idx):
            w = self._w(s)
            d = ~np.concatenate([self._xn[s, :w] >= 0.5, keep_w[j, :w], keepp[j, :w]])
            dead.append(d)
            Aj = A[j, :3 * w, :3 * w]
            Aj = (Aj + Aj.conj().T) / 2           # symmetrised, dead rows penalised (solver.py:294-297)
            penalty = max(RR_DEAD_PENALTY, 100.0 * float(np.abs(np.diag(Aj)).max()))
            Aj[np.diag_indices(3 * w)] += np.where(d, penalty * (1.0 + DEAD_SPREAD * np.arange(3 * w)), 0.0)
            mats.append(Aj)
        if self.verify_dead_rows:
            self._check_dead(idx, dead)
        res = eigh_list_host(mats, lowest=[self._w(s) for s in idx])
        return torch.as_tensor(self._pack_rr(idx, [V for _, V in res]), device=self.S.device)

    def _rr_device(self, idx, A, keep_w, keepp):
        """_rr_host with the penalty, eigensolve and normalisation on the device."""
        M = self.m
        widths = [self._w(s) for s in idx]
        # the ragged Gram pads to three times the widest slot in idx, which is
        # below 3M once the engine's widest requests have left
        K = int(A.shape[1])
        xn = torch.stack([self._xn_dev[s] for s in idx])   # views: no host-to-device copy
        dead = rr_dead_rows(widths, K, xn, keep_w, keepp)
        if self.verify_dead_rows:
            dh = dead.cpu().numpy()
            self._check_dead(idx, [dh[j, :3 * w] for j, w in enumerate(widths)])
        A = (A + A.conj().transpose(1, 2)) / 2
        penalty = torch.clamp(100.0 * torch.diagonal(A, dim1=1, dim2=2).abs().max(dim=1).values,
                              min=RR_DEAD_PENALTY)
        A = A + torch.diag_embed(torch.where(dead, penalty[:, None] * dead_spread(K, A.device)[None, :],
                                             0.0).to(A.dtype))
        _, C = self._eigh_device_ragged(A, [3 * w for w in widths], lowest=M)
        C2 = torch.zeros((len(idx), K, 2 * M), dtype=torch.complex128, device=A.device)
        for j, w in enumerate(widths):
            c = C[j, :3 * w, :w]
            cp = c.clone()
            cp[:w] = 0
            C2[j, :3 * w, :w] = c / torch.clamp(torch.linalg.vector_norm(c, dim=0), min=1e-20)[None, :]
            C2[j, :3 * w, w:2 * w] = cp / torch.clamp(torch.linalg.vector_norm(cp, dim=0), min=1e-20)[None, :]
        return C2

    def _theta(self, idx, part):
        """Θ of the X rows (part "X") or the W and P rows ("WP") of slots idx
        into the same rows of HS: one ragged launch set per 64 slots."""
        M = self.m
        for c in range(0, len(idx), MAX_PROBLEMS):
            sub = idx[c:c + MAX_PROBLEMS]
            off = [s * 3 * M + (0 if part == "X" else self._w(s)) for s in sub]
            rows = [self._w(s) * (1 if part == "X" else 2) for s in sub]
            theta_ragged(self.ws, [self.slots[s].op for s in sub], self.S, off, self.HS, off, rows)
            for s, n in zip(sub, rows):
                self.slots[s].req.theta_applications += n

    # ---- phases -------------------------------------------------------------
    def _init_blocks(self, idx):
        M = self.m
        for s in idx:
            sl = self.slots[s]
            m = sl.m
            X = self._X(s)
            if sl.nc:
                X[:sl.nc].copy_(self.CARRY[s * M:s * M + sl.nc])
            if sl.nc < m:
                Z = sl.op.random_block(m, sl.req.seed, kind="normal", row0=sl.block_idx * m)
                Z = Z[:m - sl.nc]
                nrm = torch.sqrt(torch.sum(torch.abs(Z.reshape(Z.shape[0], -1)) ** 2, dim=1))
                X[sl.nc:] = Z / torch.clamp(nrm, min=1e-30)[:, None, None, None, None]
            sl.block_idx += 1
            self._P(s).zero_()
        self._deflate(idx, self._X)
        self._svqb(idx, self._X)
        self._theta(idx, "X")
        for s in idx:
            sl = self.slots[s]
            sl.phase = "iter"
            sl.it = 0
            sl.stag = StagnationLock(self.tol, self.stagnation)
            sl.n_lock = min(sl.m - sl.guard, sl.nev - sl.locked_vals.size)

    def _iterate(self, idx, lam):
        M = self.m
        w = self._w
        t = time.perf_counter()
        # W = K⁻¹R, deflated, projected out of X twice, SVQB (solver.py:251-257)
        for c in range(0, len(idx), MAX_PROBLEMS):
            part = idx[c:c + MAX_PROBLEMS]
            precondition_ragged(self.ws, [self.slots[s].op for s in part], self.R,
                                [s * M for s in part], self.S, [s * 3 * M + w(s) for s in part],
                                [w(s) for s in part])
        self._deflate(idx, self._W)
        for _ in range(2):
            self._project_out(idx, self._X, self._W, w, w)
        t = self._tick("precond+deflate+proj", t)
        keep_w = self._svqb(idx, self._W)
        t = self._tick("svqb_W", t)
        # P projected out of [X, W] twice, fixed-shape whitening (solver.py:258-274)
        for _ in range(2):
            self._project_out(idx, self._XW, self._P, lambda s: 2 * w(s), w)
        Ps = [self._P(s) for s in idx]
        Cp, keepp = self._whiten_small(gram_ragged(Ps, Ps, hermitian=True), [w(s) for s in idx])
        self._combine_back(idx, Cp, Ps)
        t = self._tick("P", t)
        # Θ(W), Θ(P) fresh; Rayleigh–Ritz on [X; W; P] (solver.py:275-302)
        self._theta(idx, "WP")
        t = self._tick("theta_WP", t)
        Sbs = [self._Sb(s) for s in idx]
        A = gram_ragged(Sbs, [self._HSb(s) for s in idx], hermitian=True)
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
        m_in = [3 * w(s) for s in idx]
        m_out = [2 * w(s) for s in idx]
        split = [w(s) for s in idx]
        if 2 * M <= COMBINE_IN_PLACE_MAX:
            combine_ragged(C2, Sbs, m_in, m_out, Xs, Ps, split)
        else:
            T1s = [self._t1(s) for s in idx]
            T2s = [self._t2(s) for s in idx]
            combine_ragged(C2, Sbs, m_in, m_out, T1s, T2s, split)
            for T1, T2, X, P in zip(T1s, T2s, Xs, Ps):
                X.copy_(T1)
                P.copy_(T2)
        self._deflate(idx, self._X)
        t = self._tick("update_XP", t)
        self._theta(idx, "X")
        t = self._tick("theta_X", t)
        for s in idx:
            self.slots[s].it += 1
            self.slots[s].req.iterations += 1

    def _finish_block(self, s, lam, rn, order):
        """Lock the lowest n_lock rows, keep the rest as carry (solver.py:316-345)."""
        sl = self.slots[s]
        M = self.m
        if not np.isfinite(lam[order[:sl.n_lock]]).all():
            raise RuntimeError(f"request {sl.req.rid}: rank collapse in block {sl.block_idx}")
        X = self._X(s)
        lock_idx = torch.as_tensor(order[:sl.n_lock], device=X.device)
        carry_idx = torch.as_tensor(order[sl.n_lock:], device=X.device)
        base = s * self.cap
        self.L[base + sl.n_locked:base + sl.n_locked + sl.n_lock] = X[lock_idx]
        sl.nc = int(carry_idx.numel())
        self.CARRY[s * M:s * M + sl.nc] = X[carry_idx]
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

    def _row_selection(self, act):
        """(device selection slot << 16 | r, host slot and row indices, device
        linear indices slot·M + r) of the X rows of the occupied slots act,
        rebuilt only when the occupied slots or their widths change."""
        key = tuple((s, self.slots[s].m) for s in act)
        if key != self._sel_key:
            slots = np.repeat([s for s, _ in key], [w for _, w in key])
            rows = np.concatenate([np.arange(w) for _, w in key])
            dev = self.S.device
            self._sel = (torch.as_tensor(((slots << 16) | rows).astype(np.int32), device=dev),
                         slots, rows, torch.as_tensor(slots * self.m + rows, device=dev))
            self._sel_key = key
        return self._sel

    # ---- one engine step = one LOBPCG iteration for every active slot -------
    def step(self) -> list:
        finished = []
        act = [s for s, sl in enumerate(self.slots) if sl is not None]
        if not act:
            return finished
        self.steps += 1
        M = self.m
        t = time.perf_counter()
        init = [s for s in act if self.slots[s].phase == "init"]
        if init:
            self._init_blocks(init)
            t = self._tick("init", t)
        # Rayleigh quotients, residuals, dead rows of the X rows of the occupied
        # slots only (each slot's own width: no padding rows, no free slots)
        sel, sel_slot, sel_row, sel_lin = self._row_selection(act)
        # one read of X and ΘX for ⟨x,Θx⟩, |x|², |Θx|²; R = ΘX − λX with |R|² in the same pass
        st = rowstats_sel(self.S, self.HS, sel, M, self.nslots, 3 * M, 3 * M, self.n)
        lam, xn2, hx2 = st[:, 0], st[:, 2], st[:, 3]
        if self.normalized_rq:
            lam = lam / torch.clamp(xn2, min=1e-30)
        rr2 = row_axpby_norm_sel(self.HS, 3 * M, torch.ones_like(lam), self.S, 3 * M, -lam, sel, M,
                                 self.nslots, self.n, self.R, M)
        # device masks (small_dense="device"): X row norms in the (slot, row) layout
        xnd = torch.zeros(self.nslots * M, dtype=torch.float64, device=self.S.device)
        xnd[sel_lin] = torch.sqrt(xn2)
        self._xn_dev = xnd.view(self.nslots, M)
        stats = torch.stack([lam, xn2, rr2, hx2]).cpu().numpy()
        t = self._tick("residuals+sync", t)
        lam_h = np.zeros((self.nslots, M))
        xn = np.zeros((self.nslots, M))
        rn_h = np.zeros((self.nslots, M))
        lam_h[sel_slot, sel_row] = stats[0]
        xn[sel_slot, sel_row] = np.sqrt(stats[1])
        rn_h[sel_slot, sel_row] = np.sqrt(stats[2]) / np.maximum(np.sqrt(stats[3]), 1e-30)
        self._xn = xn
        cont = []
        for s in act:
            sl = self.slots[s]
            w = sl.m
            dead = xn[s, :w] < 0.5
            l = np.where(dead, np.inf, lam_h[s, :w])
            r = np.where(dead, np.inf, rn_h[s, :w])
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
