"""The Maxwell operator Θ = ∇×(ε⁻¹∇×) in the transverse-H plane-wave basis, on MI355X.

Drop-in for eigenfns.operator (reference eigenfns/operator.py): same names,
arguments and attributes, with torch complex64 device tensors where the
reference uses jax arrays.  Θ and the preconditioner run as five fused HIP
passes (libeigenfns_hip.so: efn_theta / efn_precondition); the frame t̂₁, t̂₂
and |k+G| are rebuilt inside the kernels, so the `basis` tensors exist only
for callers that read them.

Vectors are (m, 2, G, G, G) complex64, C order, z fastest, frequencies in
np.fft.fftfreq order.  Θ has physical scale (the 1/G³ of the inverse FFT is
folded into ε⁻¹).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from ._device import as_block, device, ptr, stream_ptr


@dataclass(frozen=True)
class TransverseBasis:
    """Per-(k+G) transverse frame and |k+G| for one k-point (operator.py:29-41)."""

    grid_size: int
    box_size: float
    k_frac: tuple
    kn: torch.Tensor   # (G,G,G) float32 — |k+G|
    t: torch.Tensor    # (2,3,G,G,G) float32 — t̂₁, t̂₂ (zero where |k+G| = 0)

    @property
    def n_dof(self) -> int:
        return 2 * self.grid_size ** 3


def make_basis(grid_size: int, box_size: float, k_frac=(0.0, 0.0, 0.0),
               rdtype=torch.float32, dev=None) -> TransverseBasis:
    """Frame and |k+G| built in FP64 on the host and cast to `rdtype` (operator.py:44-79).

    Reference axis ẑ, switched to x̂ where |K_z| > 0.9·|K|; t̂₁ = K×ref/|·|,
    t̂₂ = K×t̂₁/|·|, both zeroed where K = 0.
    """
    G, L = int(grid_size), float(box_size)
    freq = np.fft.fftfreq(G, d=L / G)
    k0 = 2 * np.pi * np.asarray(k_frac, np.float64) / L
    g = 2 * np.pi * freq
    K = np.stack(np.meshgrid(g + k0[0], g + k0[1], g + k0[2], indexing="ij"))
    kn = np.sqrt((K ** 2).sum(0))
    use_x = np.abs(K[2]) > 0.9 * np.maximum(kn, 1e-300)
    # K × ẑ = (Ky, −Kx, 0); K × x̂ = (0, Kz, −Ky)
    t1 = np.where(use_x[None], np.stack([np.zeros_like(K[0]), K[2], -K[1]]),
                  np.stack([K[1], -K[0], np.zeros_like(K[0])]))
    n1 = np.linalg.norm(t1, axis=0)
    zero = n1 == 0
    t1 = t1 / np.where(zero, 1.0, n1)
    t2 = np.cross(K, t1, axis=0)
    n2 = np.linalg.norm(t2, axis=0)
    t2 = t2 / np.where(n2 == 0, 1.0, n2)
    t1[:, zero] = 0.0
    t2[:, zero] = 0.0
    dev = device() if dev is None else dev
    return TransverseBasis(G, L, tuple(float(x) for x in k_frac),
                           torch.as_tensor(kn, dtype=rdtype, device=dev),
                           torch.as_tensor(np.stack([t1, t2]), dtype=rdtype, device=dev))


class MaxwellOperator:
    """Matrix-free Θ for a fixed ε(r) grid and k-point, batched over vectors.

    operator.py:130-184.  `dtype` must be complex64 (the GPU path computes in
    complex FP32; FP64 lives in the small dense algebra and the oracle).
    `max_chunk` bounds the workspace (vectors per launch set; 0 = the tuning
    table's choice for this grid, else the library default; −1 = no workspace,
    for operators that only lend their tables to multi-operator launches).
    `passes` picks Θ's pass structure for this operator (0 = the tuning table's
    choice, else the library's; 3 or 5 to force one).
    """

    def __init__(self, eps, box_size: float, k_frac=(0.0, 0.0, 0.0), dtype=torch.complex64,
                 max_chunk: int = 0, passes: int = 0, tuned: bool = True):
        eps_np = eps.detach().cpu().numpy() if isinstance(eps, torch.Tensor) else np.asarray(eps)
        if eps_np.ndim != 3 or len(set(eps_np.shape)) != 1:
            raise ValueError(f"eps must be a cube, got {eps_np.shape}")
        if dtype not in (torch.complex64, np.complex64, "complex64"):
            raise ValueError("the HIP operator computes in complex64 only")
        self.dtype = torch.complex64
        self.grid_size = int(eps_np.shape[0])
        self.box_size = float(box_size)
        self.k_frac = tuple(float(x) for x in k_frac)
        lib = _lib.load()
        if not lib.efn_grid_supported(self.grid_size):
            raise ValueError(f"grid size {self.grid_size} has no compiled kernels")
        self._dev = device()
        if tuned and max_chunk >= 0 and (max_chunk == 0 or passes == 0):
            from .autotune import choice
            entry = choice(self.grid_size)
            if entry:
                max_chunk = max_chunk or int(entry["chunk"])
                passes = passes or int(entry["passes"])
        eps32 = np.ascontiguousarray(eps_np, dtype=np.float32)
        handle = C.c_void_p()
        with torch.cuda.device(self._dev):
            _lib.check(lib.efn_ctx_create(self.grid_size, self.box_size, *self.k_frac,
                                          eps32.ctypes.data, int(max_chunk), C.byref(handle)),
                       "efn_ctx_create")
        self._ctx = handle
        if passes:
            rc = int(lib.efn_ctx_set_passes(self._ctx, int(passes)))
            if rc < 0:
                lib.efn_ctx_destroy(self._ctx)
                self._ctx = None
                _lib.check(rc, "passes")
        kf, emi, ch, wsb = C.c_double(), C.c_double(), C.c_int(), C.c_size_t()
        _lib.check(lib.efn_ctx_query(self._ctx, C.byref(kf), C.byref(emi), C.byref(ch),
                                     C.byref(wsb)))
        self.kn_floor = float(kf.value)
        self.eps_mean_inv = float(np.mean(1.0 / eps32))
        self.chunk = int(ch.value)
        self.workspace_bytes = int(wsb.value)
        self._passes_req = int(passes)
        self._eps32 = eps32
        self._basis = None
        self._inv_eps = None
        self._eps_grid = None
        self._mask = None

    @property
    def passes(self) -> int:
        """The pass structure (3 or 5) this operator's launches use now."""
        return int(_lib.load().efn_ctx_set_passes(self._ctx, -1))

    def __del__(self):
        ctx = getattr(self, "_ctx", None)
        if ctx is not None and ctx.value:
            try:
                torch.cuda.synchronize(self._dev)
                _lib.load().efn_ctx_destroy(ctx)
            except Exception:
                pass
            self._ctx = None

    # ---- lazily materialised attributes of the reference operator ---------
    @property
    def basis(self) -> TransverseBasis:
        if self._basis is None:
            self._basis = make_basis(self.grid_size, self.box_size, self.k_frac, dev=self._dev)
        return self._basis

    @property
    def inv_eps(self) -> torch.Tensor:
        if self._inv_eps is None:
            self._inv_eps = torch.as_tensor(1.0 / self._eps32, device=self._dev)
        return self._inv_eps

    @property
    def eps_grid(self) -> torch.Tensor:
        if self._eps_grid is None:
            self._eps_grid = torch.as_tensor(self._eps32, device=self._dev)
        return self._eps_grid

    @property
    def mask(self) -> torch.Tensor:
        """(G,G,G) float32, 0 at the k+G = 0 slot and 1 elsewhere (solver.py:162)."""
        if self._mask is None:
            kn = torch.empty((self.grid_size,) * 3, dtype=torch.float32, device=self._dev)
            _lib.check(_lib.load().efn_ctx_basis(self._ctx, kn.data_ptr(), None, stream_ptr()))
            self._mask = (kn > 0).to(torch.float32)
        return self._mask

    def kernel_basis(self):
        """(kn, t) exactly as the kernels rebuild them (FP32 on the device)."""
        G = self.grid_size
        kn = torch.empty((G, G, G), dtype=torch.float32, device=self._dev)
        t = torch.empty((2, 3, G, G, G), dtype=torch.float32, device=self._dev)
        _lib.check(_lib.load().efn_ctx_basis(self._ctx, kn.data_ptr(), t.data_ptr(), stream_ptr()))
        return kn, t

    # ---- operator applications --------------------------------------------
    def _check_block(self, X) -> torch.Tensor:
        X = as_block(X)
        G = self.grid_size
        if X.dim() != 5 or tuple(X.shape[1:]) != (2, G, G, G):
            raise ValueError(f"expected a block (m, 2, {G}, {G}, {G}), got {tuple(X.shape)}")
        return X

    def theta_affine(self, X, alpha: float = 1.0, beta: float = 0.0, out=None) -> torch.Tensor:
        """alpha·Θ·X + beta·X in one fused pass set."""
        X = self._check_block(X)
        if out is None:
            out = torch.empty_like(X)
        _lib.check(_lib.load().efn_theta(self._ctx, X.data_ptr(), out.data_ptr(), X.shape[0],
                                         float(alpha), float(beta), stream_ptr()), "theta")
        return out

    def theta(self, Hs) -> torch.Tensor:
        """Θ·Hs for a block (m, 2, G, G, G) of spectral transverse fields (operator.py:153-155)."""
        return self.theta_affine(Hs, 1.0, 0.0)

    def theta_into(self, X, out) -> torch.Tensor:
        """Θ·X written into the preallocated block `out` (contiguous, same shape)."""
        return self.theta_affine(X, 1.0, 0.0, out=self._check_out(out, X))

    def _check_out(self, out, X):
        if not out.is_contiguous() or out.shape != X.shape or out.dtype != torch.complex64:
            raise ValueError("out must be a contiguous complex64 block shaped like the input")
        return out

    def precondition(self, Rs, target: float = 0.0) -> torch.Tensor:
        """MPB transverse-projection preconditioner (operator.py:157-161; `target` unused)."""
        X = self._check_block(Rs)
        return self.precondition_into(X, torch.empty_like(X))

    def precondition_into(self, R, out) -> torch.Tensor:
        X = self._check_block(R)
        self._check_out(out, X)
        _lib.check(_lib.load().efn_precondition(self._ctx, X.data_ptr(), out.data_ptr(),
                                                X.shape[0], stream_ptr()), "precondition")
        return out

    def precondition_simple(self, Rs, target: float = 0.0) -> torch.Tensor:
        """Diagonal kinetic preconditioner (operator.py:163-166)."""
        X = self._check_block(Rs)
        scale = float(np.float32(max(target / self.eps_mean_inv, 1e-6) * self.eps_mean_inv + 1e-6))
        out = torch.empty_like(X)
        _lib.check(_lib.load().efn_precondition_simple(self._ctx, X.data_ptr(), out.data_ptr(),
                                                       X.shape[0], scale, stream_ptr()))
        return out

    # ---- fused recurrence steps (fused into Θ's last pass; no reference symbol) --
    def cheb_step(self, T1, T0, Y, alpha: float, ck: float) -> None:
        """T0 ← 2(alpha·Θ·T1 − T1) − T0; Y ← Y + ck·T0, in place (interior.py:347-349)."""
        _lib.check(_lib.load().efn_cheb_step(self._ctx, T1.data_ptr(), T0.data_ptr(), Y.data_ptr(),
                                             float(alpha), float(ck), T1.shape[0], stream_ptr()),
                   "cheb_step")

    def kpm_step(self, T1, T0, mom: torch.Tensor, k: int, alpha: float) -> None:
        """T0 ← 2(alpha·Θ·T1 − T1) − T0 and moments 2k, 2k+1 (exp_kpm_dos.py:145-152)."""
        _lib.check(_lib.load().efn_kpm_step(self._ctx, T1.data_ptr(), T0.data_ptr(), mom.data_ptr(),
                                            mom.shape[1], int(k), float(alpha), T1.shape[0],
                                            stream_ptr()), "kpm_step")

    def kpm_count_step(self, T1, T0, Z, est: torch.Tensor, ck: float, alpha: float) -> None:
        """T0 ← 2(alpha·Θ·T1 − T1) − T0; est += ck·Re⟨Z, T0⟩ (chebyshev.py:260-263)."""
        _lib.check(_lib.load().efn_kpm_count_step(self._ctx, T1.data_ptr(), T0.data_ptr(),
                                                  Z.data_ptr(), est.data_ptr(), float(ck),
                                                  float(alpha), T1.shape[0], stream_ptr()),
                   "kpm_count_step")

    def random_block(self, m: int, seed: int, kind: str = "normal", row0: int = 0) -> torch.Tensor:
        """Counter-based Γ-masked random block (normal complex or Rademacher)."""
        G = self.grid_size
        out = torch.empty((m, 2, G, G, G), dtype=torch.complex64, device=self._dev)
        k = {"normal": 0, "rademacher": 1}[kind]
        _lib.check(_lib.load().efn_random_block(self._ctx, k, int(seed) & (2 ** 64 - 1), int(row0),
                                                int(m), out.data_ptr(), stream_ptr()))
        return out

    # ---- conversions (operator.py:169-184) ----------------------------------
    def to_cartesian_spectral(self, Hs) -> torch.Tensor:
        """(m,2,G,G,G) transverse → (m,3,G,G,G) Cartesian spectral H(k+G)."""
        X = self._check_block(Hs)
        t = self.basis.t
        return X[:, 0:1] * t[0][None] + X[:, 1:2] * t[1][None]

    def _realspace(self, Hs, mode: int, eigvals, out_shape, dtype):
        X = self._check_block(Hs)
        m = X.shape[0]
        out = torch.empty((m,) + out_shape, dtype=dtype, device=self._dev)
        lam = None
        if eigvals is not None:
            lam = torch.as_tensor(np.asarray(eigvals, dtype=np.float32).reshape(-1),
                                  device=self._dev)
            if lam.numel() != m:
                raise ValueError(f"{m} vectors but {lam.numel()} eigenvalues")
        _lib.check(_lib.load().efn_realspace(self._ctx, mode, X.data_ptr(),
                                             None if lam is None else lam.data_ptr(),
                                             out.data_ptr(), m, stream_ptr()), "realspace")
        return out

    def h_realspace(self, Hs) -> torch.Tensor:
        """Real-space Cartesian H(r) = ifftn(a·t̂₁ + b·t̂₂) for a block (operator.py:174-176),
        (m,3,G,G,G) complex64."""
        G = self.grid_size
        return self._realspace(Hs, 0, None, (3, G, G, G), torch.complex64)

    def e_realspace(self, Hs, eigvals) -> torch.Tensor:
        """Real-space E(r) = ε⁻¹·ifftn(∇×H)/λ^{1/2} for a block (operator.py:178-184),
        (m,3,G,G,G) complex64."""
        G = self.grid_size
        return self._realspace(Hs, 1, eigvals, (3, G, G, G), torch.complex64)

    def energy_density(self, Hs, eigvals) -> torch.Tensor:
        """ε·|E(r)|² per voxel for a block, (m,G,G,G) float32 — the per-band
        energy density of run_modes.py:116-124 and run_interior.py:252-262,
        fused with the field transform (E is never stored)."""
        G = self.grid_size
        return self._realspace(Hs, 2, eigvals, (G, G, G), torch.float32)


def theta_engine(engine: int | None = None) -> int:
    """Process-wide pass structure of Θ, the preconditioner and the fused steps:
    0 = the library's per-grid choice, 3 = three passes (xin | yzmid | xout,
    grids 128 and 192), 5 = five passes.  Returns the previous setting (None:
    query only)."""
    lib = _lib.load()
    if engine is None:
        return int(lib.efn_theta_engine(-1))
    prev = int(lib.efn_theta_engine(int(engine)))
    if prev < 0:
        _lib.check(prev, "theta_engine")
    return prev


def theta_passes(grid_size: int) -> int:
    """The pass structure (3 or 5) Θ uses now at this grid size."""
    rc = int(_lib.load().efn_theta_passes(int(grid_size)))
    if rc < 0:
        _lib.check(rc, "theta_passes")
    return rc


MAX_SLOTS = 64   # operators per multi / ragged launch


def _multi(kind, ws_op, ops, in_base, in_off, out_base, out_off, nrow, alpha=1.0, beta=0.0):
    import ctypes as C
    n = len(ops)
    if n > MAX_SLOTS:
        raise ValueError(f"at most {MAX_SLOTS} operators per multi launch")
    for op in ops:
        if op.grid_size != ws_op.grid_size:
            raise ValueError("multi launch: all operators must share the grid size")
    ctxs = (C.c_void_p * n)(*[op._ctx.value for op in ops])
    io = (C.c_int64 * n)(*[int(x) for x in in_off])
    oo = (C.c_int64 * n)(*[int(x) for x in out_off])
    lib = _lib.load()
    if kind == "theta":
        rc = lib.efn_theta_multi(ws_op._ctx, n, ctxs, in_base.data_ptr(), io, out_base.data_ptr(),
                                 oo, int(nrow), float(alpha), float(beta), stream_ptr())
    else:
        rc = lib.efn_precondition_multi(ws_op._ctx, n, ctxs, in_base.data_ptr(), io,
                                        out_base.data_ptr(), oo, int(nrow), stream_ptr())
    _lib.check(rc, f"{kind}_multi")


def theta_multi(ws_op, ops, in_base, in_off, out_base, out_off, nrow, alpha=1.0, beta=0.0):
    """One launch set applying Θ of several operators (same G) to row ranges of two buffers.

    Slot i: rows in_off[i] .. in_off[i]+nrow−1 of `in_base` (a contiguous
    (rows, 2, G, G, G) buffer) → rows out_off[i] .. of `out_base`; `ws_op`'s
    workspace is used for the passes."""
    _multi("theta", ws_op, ops, in_base, in_off, out_base, out_off, nrow, alpha, beta)


def precondition_multi(ws_op, ops, in_base, in_off, out_base, out_off, nrow):
    """Multi-operator transverse-projection preconditioner (see theta_multi)."""
    _multi("precond", ws_op, ops, in_base, in_off, out_base, out_off, nrow)


def _ragged(kind, ws_op, ops, in_base, in_off, out_base, out_off, rows, alpha=1.0, beta=0.0):
    import ctypes as C
    n = len(ops)
    if n > MAX_SLOTS:
        raise ValueError(f"at most {MAX_SLOTS} operators per ragged launch")
    if not (len(in_off) == len(out_off) == len(rows) == n):
        raise ValueError("ragged launch: one offset pair and row count per operator")
    for op in ops:
        if op.grid_size != ws_op.grid_size:
            raise ValueError("ragged launch: all operators must share the grid size")
    ctxs = (C.c_void_p * n)(*[op._ctx.value for op in ops])
    io = (C.c_int64 * n)(*[int(x) for x in in_off])
    oo = (C.c_int64 * n)(*[int(x) for x in out_off])
    rw = (C.c_int * n)(*[int(x) for x in rows])
    lib = _lib.load()
    if kind == "theta":
        rc = lib.efn_theta_ragged(ws_op._ctx, n, ctxs, in_base.data_ptr(), io, out_base.data_ptr(),
                                  oo, rw, float(alpha), float(beta), stream_ptr())
    else:
        rc = lib.efn_precondition_ragged(ws_op._ctx, n, ctxs, in_base.data_ptr(), io,
                                         out_base.data_ptr(), oo, rw, stream_ptr())
    _lib.check(rc, f"{kind}_ragged")


def theta_ragged(ws_op, ops, in_base, in_off, out_base, out_off, rows, alpha=1.0, beta=0.0):
    """theta_multi with a row count per operator: slot i maps rows in_off[i] ..
    in_off[i]+rows[i]−1 of `in_base` to rows out_off[i] .. of `out_base`, so
    one launch set covers blocks of different widths."""
    _ragged("theta", ws_op, ops, in_base, in_off, out_base, out_off, rows, alpha, beta)


def precondition_ragged(ws_op, ops, in_base, in_off, out_base, out_off, rows):
    """Ragged multi-operator preconditioner (see theta_ragged)."""
    _ragged("precond", ws_op, ops, in_base, in_off, out_base, out_off, rows)
