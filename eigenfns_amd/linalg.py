"""Block linear algebra on the device: Gram, combine, deflation, SVQB, row reductions.

Kernels: libeigenfns_hip.so (efn_gram / efn_combine / efn_rowreduce /
efn_row_axpby).  Small dense FP64 eigenproblems (SVQB Gram matrices,
Rayleigh–Ritz, ≤ 3m × 3m) run in host LAPACK by default (as the reference
does, solver.py:119-121); EFN_SMALL_DENSE=device routes them to rocSOLVER's
batched divide and conquer on the device (efn_heevd_batched), which removes
their host round trips.

Reference: eigenfns/solver.py:27-126 and the row reductions of
solver.py:228-238 / interior.py:75-82.
"""
from __future__ import annotations

import contextlib
import os
import threading

import numpy as np
import torch

from . import _lib
from ._device import SCRATCH, as_block, device, stream_ptr

SVQB_DROP = 1e-5        # solver.py:27 — relative Gram-eigenvalue drop threshold
RR_DEAD_PENALTY = 1e4   # solver.py:28
# Dead rows of a Rayleigh–Ritz matrix get penalty·(1 + DEAD_SPREAD·row) on the
# diagonal instead of one common value: the dead directions only need to sit
# far above the live spectrum, and an exactly degenerate, exactly decoupled
# cluster made rocSOLVER's batched divide and conquer return non-finite vectors.
DEAD_SPREAD = 1e-6


def dead_spread(n: int, device=None, dtype=torch.float64) -> torch.Tensor:
    """1 + DEAD_SPREAD·j for j < n (multiplies the dead-row penalties)."""
    return 1.0 + DEAD_SPREAD * torch.arange(n, device=device, dtype=dtype)
DEFLATE_CHUNK = 128     # solver.py:60


def _flat_n(X: torch.Tensor) -> int:
    return int(X.numel()) // int(X.shape[0]) if X.shape[0] else int(np.prod(X.shape[1:]))


def _small_dense_on_host() -> bool:
    return os.environ.get("EFN_SMALL_DENSE", "host").lower() not in ("device", "gpu")


GRAM_ENGINES = {"valu_f64": 0, "mfma_f64": 1, "mfma_f32": 2}


def gram_engine(name: str | None = None) -> str:
    """Select the process-wide Gram engine by name; returns the previous one.

    "mfma_f64" (default): matrix cores on the exact FP64 images of the FP32
    inputs; "valu_f64": vector ALU, same precision; "mfma_f32": FP32 matrix
    cores with FP32 partial sums of ≤64 products added into FP64."""
    inv = {v: k for k, v in GRAM_ENGINES.items()}
    code = -1 if name is None else GRAM_ENGINES[name]
    rc = _lib.load().efn_gram_engine(code)
    if rc < 0:
        _lib.check(rc, "gram_engine")
    return inv[rc]


def combine_engine(name: str | None = None) -> str:
    """Select the process-wide combine kernel ("mfma": FP32 matrix cores, the
    default for more than 8 output rows; "valu": vector ALU); returns the
    previous one."""
    names = {"valu": 0, "mfma": 1}
    inv = {v: k for k, v in names.items()}
    rc = _lib.load().efn_combine_engine(-1 if name is None else names[name])
    if rc < 0:
        _lib.check(rc, "combine_engine")
    return inv[rc]


def gram(A: torch.Tensor, B: torch.Tensor) -> torch.Tensor:
    """⟨A_i, B_j⟩ as an (m_a, m_b) complex128 device tensor (solver.py:35-43).

    Products of the FP32 inputs are exact in FP64 and summed in FP64."""
    A = as_block(A)
    B = as_block(B)
    ma, mb = int(A.shape[0]), int(B.shape[0])
    n = _flat_n(A)
    if mb and _flat_n(B) != n:
        raise ValueError("gram: row lengths differ")
    out = torch.empty((ma, mb), dtype=torch.complex128, device=A.device)
    if ma == 0 or mb == 0:
        return out
    lib = _lib.load()
    need = lib.efn_gram_work_bytes(ma, mb, n)
    work = SCRATCH.get(need, "gram")
    _lib.check(lib.efn_gram(A.data_ptr(), ma, B.data_ptr(), mb, n, out.data_ptr(),
                            work.data_ptr(), work.numel(), stream_ptr()), "gram")
    return out


def gram_herm(A: torch.Tensor, B: torch.Tensor | None = None) -> torch.Tensor:
    """AᴴB (B = A if None) for a product known to be Hermitian: the upper block
    triangle of tiles is computed and mirrored (solver.py:119, :275-294)."""
    A = as_block(A)
    B = A if B is None else as_block(B)
    if int(A.shape[0]) != int(B.shape[0]):
        raise ValueError("gram_herm: blocks must have the same number of rows")
    if int(A.shape[0]) == 0:
        return torch.empty((0, 0), dtype=torch.complex128, device=A.device)
    return gram_multi([A.contiguous()], [B.contiguous()], hermitian=True)[0]


# combine may write into its own input block (out in S) up to this many output
# rows: one row tile, each column read in full before it is written
COMBINE_IN_PLACE_MAX = 64


def combine_into(C: torch.Tensor, S: torch.Tensor, out: torch.Tensor, X=None,
                 alpha: float = 0.0) -> torch.Tensor:
    """out_i = alpha·X_i + Σ_j C[j, i]·S_j with S one contiguous block (m_in rows)."""
    C = C.to(device=S.device, dtype=torch.complex64).contiguous()
    m_in, m_out = int(C.shape[0]), int(C.shape[1])
    if int(S.shape[0]) != m_in:
        raise ValueError(f"combine: C has {m_in} rows but the block has {int(S.shape[0])}")
    n = _flat_n(S) if m_in else _flat_n(out)
    _lib.check(_lib.load().efn_combine(C.data_ptr(), m_in, m_out, S.data_ptr(), n,
                                       None if X is None else X.data_ptr(), float(alpha),
                                       out.data_ptr(), stream_ptr()), "combine")
    return out


def combine(C, *blocks) -> torch.Tensor:
    """Linear combination out_i = Σ_j C[j, i]·S_j over the concatenated blocks (solver.py:49-57)."""
    blocks = [as_block(b) for b in blocks]
    C = torch.as_tensor(C).to(device=blocks[0].device)
    m_out = int(C.shape[1])
    shape = (m_out,) + tuple(blocks[0].shape[1:])
    out = torch.empty(shape, dtype=torch.complex64, device=blocks[0].device)
    off = 0
    for i, Bk in enumerate(blocks):
        r = int(Bk.shape[0])
        combine_into(C[off:off + r], Bk, out, X=out if i else None, alpha=1.0 if i else 0.0)
        off += r
    return out


def rownorm2(X: torch.Tensor) -> torch.Tensor:
    """Σ|x|² per row in FP64 (device float64 tensor)."""
    return _rowreduce(0, X, None).real


def rowdot(A: torch.Tensor, B: torch.Tensor) -> torch.Tensor:
    """⟨a_i, b_i⟩ per row in FP64 (device complex128 tensor)."""
    return _rowreduce(1, A, B)


def _rowreduce(op, A, B):
    A = as_block(A)
    m, n = int(A.shape[0]), _flat_n(A)
    out = torch.empty(m, dtype=torch.complex128, device=A.device)
    if m == 0:
        return out
    lib = _lib.load()
    need = lib.efn_rowreduce_work_bytes(m, n)
    work = SCRATCH.get(need, "rowred")
    _lib.check(lib.efn_rowreduce(op, A.data_ptr(), None if B is None else as_block(B).data_ptr(),
                                 m, n, out.data_ptr(), work.data_ptr(), work.numel(),
                                 stream_ptr()), "rowreduce")
    return out


def rownorm(X) -> torch.Tensor:
    return torch.sqrt(rownorm2(X))


def row_axpby(X, a, Y=None, b=None, out=None) -> torch.Tensor:
    """out_i = a_i·X_i + b_i·Y_i with per-row coefficients (FP64 device tensors)."""
    X = as_block(X)
    m, n = int(X.shape[0]), _flat_n(X)
    a = torch.as_tensor(a, dtype=torch.float64, device=X.device).reshape(-1).contiguous()
    if Y is not None:
        b = torch.as_tensor(b, dtype=torch.float64, device=X.device).reshape(-1).contiguous()
    if out is None:
        out = torch.empty_like(X)
    _lib.check(_lib.load().efn_row_axpby(X.data_ptr(), a.data_ptr(),
                                         None if Y is None else Y.data_ptr(),
                                         None if Y is None else b.data_ptr(), m, n,
                                         out.data_ptr(), stream_ptr()), "row_axpby")
    return out


def axpby(X, a: float, Y=None, b: float = 0.0, out=None) -> torch.Tensor:
    """out = a·X + b·Y elementwise (complex64)."""
    if out is None:
        out = torch.empty_like(X)
    _lib.check(_lib.load().efn_axpby(X.data_ptr(), float(a), None if Y is None else Y.data_ptr(),
                                     float(b), X.numel(), out.data_ptr(), stream_ptr()), "axpby")
    return out


def normalize_rows(X, floor: float = 1e-30, out=None) -> torch.Tensor:
    """X_i / max(‖X_i‖, floor) (solver.py:204-205, :306-307)."""
    s = 1.0 / torch.clamp(rownorm(X), min=floor)
    return row_axpby(X, s, out=out)


# ---- small dense FP64 ------------------------------------------------------
def eigh(A: torch.Tensor):
    """Hermitian eigendecomposition in FP64 (ascending): host LAPACK by default,
    rocSOLVER on the device with EFN_SMALL_DENSE=device (no host round trip)."""
    A = (A + A.conj().T) / 2
    if _small_dense_on_host():
        w, V = np.linalg.eigh(A.detach().cpu().numpy())
        return (torch.as_tensor(w, device=A.device), torch.as_tensor(V, device=A.device))
    w, V = eigh_batched_device(A[None])
    return w[0], V[0]


def svqb_coeffs(Gm: torch.Tensor, drop: float = SVQB_DROP) -> torch.Tensor:
    """V·w^{-1/2} with directions below drop·max(w) zeroed (solver.py:119-124)."""
    w, V = eigh(Gm)
    wmax = torch.clamp(w.max(), min=1e-300)
    keep = w > drop * wmax
    scale = torch.where(keep, 1.0 / torch.sqrt(torch.where(keep, w, torch.ones_like(w))),
                        torch.zeros_like(w))
    return V * scale[None, :]


def deflate_chunk_rows(grid_size: int, target_bytes: float = 0.75e9) -> int:
    """Rows per deflation chunk (solver.py:85-92).  The device-resident locked
    set makes chunking unnecessary; kept for API compatibility."""
    return max(16, int(target_bytes // (2 * grid_size ** 3 * 8)))


def deflate(X, locked, chunk_rows: int = DEFLATE_CHUNK, n_locked: int | None = None):
    """X ← X − L(LᴴX) against the locked rows (solver.py:95-110), one pass.

    The locked set is device-resident (no host streaming); `chunk_rows` only
    bounds the Gram size per call."""
    if locked is None:
        return X
    L = locked if n_locked is None else locked[:n_locked]
    if isinstance(L, np.ndarray):
        L = as_block(L)
    if L.shape[0] == 0:
        return X
    X = as_block(X)
    out = X
    for s in range(0, int(L.shape[0]), chunk_rows):
        Lc = L[s:s + chunk_rows]
        Cg = gram(Lc, out)
        res = torch.empty_like(X) if out is X else out
        combine_into(-Cg, Lc, res, X=out, alpha=1.0)
        out = res
    return out


ORTHO_METHODS = ("svqb", "cholqr")


def cholqr_coeffs_batched(Gs: torch.Tensor, drop: float = SVQB_DROP):
    """(C, rank) for a (nprob, k, k) stack of Hermitian PSD Gram matrices, on
    the device (efn_cholqr_batched, FP64): C = Π·L⁻ᴴ of a pivoted Cholesky
    stopped at pivots ≤ drop·max(diag G), zero columns past the rank, so
    Cᴴ·G·C = diag(I_rank, 0); rank stays on the device (no host sync)."""
    Gs = Gs.to(torch.complex128).contiguous()
    nprob, k = int(Gs.shape[0]), int(Gs.shape[1])
    C = torch.empty_like(Gs)
    rank = torch.empty(nprob, dtype=torch.int32, device=Gs.device)
    _lib.check(_lib.load().efn_cholqr_batched(nprob, k, Gs.data_ptr(), C.data_ptr(), float(drop),
                                              rank.data_ptr(), stream_ptr()), "cholqr_batched")
    return C, rank


def cholqr_coeffs(Gm: torch.Tensor, drop: float = SVQB_DROP) -> torch.Tensor:
    """Cholesky-QR coefficients of one Gram matrix: the SVQB step's C
    (Cᴴ·G·C = I on the kept rows, dropped rows zero) from a pivoted FP64
    Cholesky on the device instead of an eigendecomposition."""
    Gm = Gm.to(torch.complex128)
    return cholqr_coeffs_batched(((Gm + Gm.conj().T) / 2)[None], drop)[0][0]


def ortho_coeffs(Gm: torch.Tensor, drop: float = SVQB_DROP, method: str = "svqb") -> torch.Tensor:
    """Orthonormalising coefficients of a Gram matrix by `method` ("svqb":
    solver.py:119-124; "cholqr": pivoted Cholesky)."""
    if method == "svqb":
        return svqb_coeffs(Gm, drop)
    if method == "cholqr":
        return cholqr_coeffs(Gm, drop)
    raise ValueError(f"ortho method must be one of {ORTHO_METHODS}, got {method!r}")


def ortho_block(X, passes: int = 2, drop: float = SVQB_DROP, method: str = "svqb"):
    """SVQB orthonormalisation, FP64 Gram algebra, `passes` sweeps (solver.py:113-126);
    method="cholqr" uses Cholesky-QR coefficients instead (same passes and drop).

    Rank-dropped directions come back as zero rows."""
    X = as_block(X)
    for _ in range(passes):
        Cw = ortho_coeffs(gram_herm(X), drop, method)
        Y = torch.empty_like(X)
        combine_into(Cw, X, Y)
        X = Y
    return X


def apply_chunked(fn, X, rows: int):
    """Apply a per-vector operator in sub-chunks of `rows` vectors (solver.py:74-82)."""
    if X.shape[0] <= rows:
        return fn(X)
    return torch.cat([fn(X[s:s + rows]) for s in range(0, X.shape[0], rows)], dim=0)


# ---- multi-problem forms (one launch for up to 64 independent problems) ------
MAX_PROBLEMS = 64


def _ptrs(ts):
    import ctypes as C
    return (C.c_void_p * len(ts))(*[None if t is None else t.data_ptr() for t in ts])


def gram_multi(As, Bs, hermitian: bool = False) -> torch.Tensor:
    """(nprob, m_a, m_b) complex128: A_pᴴ·B_p for each pair (FP64-accumulated).

    hermitian=True (square products known to be Hermitian, e.g. AᴴA or the
    Rayleigh–Ritz AᴴΘA that is symmetrised afterwards): only the upper block
    triangle of tiles is computed and mirrored."""
    nprob = len(As)
    if nprob == 0:
        raise ValueError("gram_multi: no problems")
    if nprob > MAX_PROBLEMS:
        raise ValueError(f"gram_multi: at most {MAX_PROBLEMS} problems per launch")
    ma, mb = int(As[0].shape[0]), int(Bs[0].shape[0])
    n = _flat_n(As[0])
    for A, B in zip(As, Bs):
        if int(A.shape[0]) != ma or int(B.shape[0]) != mb or not A.is_contiguous() \
                or not B.is_contiguous():
            raise ValueError("gram_multi: blocks must be contiguous and uniform in shape")
    out = torch.empty((nprob, ma, mb), dtype=torch.complex128, device=As[0].device)
    lib = _lib.load()
    need = lib.efn_gram_multi_work_bytes(nprob, ma, mb, n)
    work = SCRATCH.get(need, "gram")
    if hermitian:
        if ma != mb:
            raise ValueError("gram_multi: a Hermitian product needs square blocks")
        _lib.check(lib.efn_gram_multi_herm(nprob, _ptrs(As), _ptrs(Bs), ma, n, out.data_ptr(),
                                           work.data_ptr(), work.numel(), stream_ptr()), "gram_multi")
    else:
        _lib.check(lib.efn_gram_multi(nprob, _ptrs(As), _ptrs(Bs), ma, mb, n, out.data_ptr(),
                                      work.data_ptr(), work.numel(), stream_ptr()), "gram_multi")
    return out


def combine_multi(C: torch.Tensor, Ss, outs, Xs=None, alpha: float = 0.0) -> None:
    """out_p = alpha·X_p + C_pᵀ·S_p for each problem (C: (nprob, m_in, m_out))."""
    nprob = len(Ss)
    if nprob > MAX_PROBLEMS:
        raise ValueError(f"combine_multi: at most {MAX_PROBLEMS} problems per launch")
    C = C.to(device=outs[0].device, dtype=torch.complex64).contiguous()
    m_in, m_out = int(C.shape[1]), int(C.shape[2])
    n = _flat_n(outs[0])
    _lib.check(_lib.load().efn_combine_multi(nprob, C.data_ptr(), _ptrs(Ss), m_in, m_out, n,
                                             None if Xs is None else _ptrs(Xs), float(alpha),
                                             _ptrs(outs), stream_ptr()), "combine_multi")


def combine_multi_split(C: torch.Tensor, Ss, outs, outs2, split: int) -> None:
    """Rows i < split of C_pᵀ·S_p into outs[p], rows i ≥ split into outs2[p]
    (C: (nprob, m_in, m_out)); both may lie inside S_p when m_out ≤
    COMBINE_IN_PLACE_MAX."""
    nprob = len(Ss)
    if nprob > MAX_PROBLEMS:
        raise ValueError(f"combine_multi_split: at most {MAX_PROBLEMS} problems per launch")
    C = C.to(device=Ss[0].device, dtype=torch.complex64).contiguous()
    m_in, m_out = int(C.shape[1]), int(C.shape[2])
    n = _flat_n(Ss[0])
    _lib.check(_lib.load().efn_combine_multi_split(nprob, C.data_ptr(), _ptrs(Ss), m_in, m_out, n,
                                                   _ptrs(outs), _ptrs(outs2), int(split),
                                                   stream_ptr()), "combine_multi_split")


def gram_ragged(As, Bs, hermitian: bool = False) -> torch.Tensor:
    """(nprob, max m_a, max m_b) complex128: A_pᴴ·B_p per pair, each problem
    with its own row counts, in one launch; zero outside each problem's own
    m_a × m_b (problems may be empty).  hermitian: square Hermitian products."""
    nprob = len(As)
    if nprob == 0 or nprob != len(Bs):
        raise ValueError("gram_ragged: one A and one B per problem")
    if nprob > MAX_PROBLEMS:
        raise ValueError(f"gram_ragged: at most {MAX_PROBLEMS} problems per launch")
    import ctypes as C
    ma = [int(A.shape[0]) for A in As]
    mb = [int(B.shape[0]) for B in Bs]
    ref = next((A for A in As if A.shape[0]), As[0])
    n = _flat_n(ref)
    row = ref.shape[1:]
    for A, B in zip(As, Bs):
        if (A.shape[0] and (A.shape[1:] != row or not A.is_contiguous())) or \
                (B.shape[0] and (B.shape[1:] != row or not B.is_contiguous())):
            raise ValueError("gram_ragged: blocks must be contiguous with one row length")
    mam, mbm = max(ma), max(mb)
    out = torch.empty((nprob, mam, mbm), dtype=torch.complex128, device=ref.device)
    if mam == 0 or mbm == 0:
        return out
    lib = _lib.load()
    work = SCRATCH.get(lib.efn_gram_multi_work_bytes(nprob, mam, mbm, n), "gram")
    _lib.check(lib.efn_gram_ragged(nprob, _ptrs([A if A.shape[0] else None for A in As]),
                                   _ptrs([B if B.shape[0] else None for B in Bs]),
                                   (C.c_int * nprob)(*ma), (C.c_int * nprob)(*mb), n, int(bool(hermitian)),
                                   out.data_ptr(), work.data_ptr(), work.numel(), stream_ptr()),
               "gram_ragged")
    return out


def combine_ragged(C: torch.Tensor, Ss, m_in, m_out, outs, outs2=None, split=None, Xs=None,
                   alpha: float = 0.0) -> None:
    """out_p (rows < split[p]) and out2_p (the rest) of alpha·X_p + C_pᵀ·S_p with
    problem p's own m_in[p] rows of S and m_out[p] outputs; C is (nprob,
    max m_in, max m_out), each problem reading its leading block."""
    import ctypes as Ct
    nprob = len(Ss)
    if nprob > MAX_PROBLEMS:
        raise ValueError(f"combine_ragged: at most {MAX_PROBLEMS} problems per launch")
    C = C.to(device=outs[0].device, dtype=torch.complex64).contiguous()
    if C.shape[1] < max(m_in) or C.shape[2] < max(m_out):
        raise ValueError("combine_ragged: C smaller than the largest problem")
    if C.shape[1] != max(m_in) or C.shape[2] != max(m_out):
        C = C[:, :max(m_in), :max(m_out)].contiguous()
    split = list(m_out) if split is None else list(split)
    ref = next((S for S in Ss if S is not None and S.shape[0]), outs[0])
    n = _flat_n(ref)
    iv = lambda xs: (Ct.c_int * nprob)(*[int(x) for x in xs])
    _lib.check(_lib.load().efn_combine_ragged(nprob, C.data_ptr(), _ptrs(Ss), iv(m_in), iv(m_out), n,
                                              None if Xs is None else _ptrs(Xs), float(alpha), _ptrs(outs),
                                              None if outs2 is None else _ptrs(outs2), iv(split),
                                              stream_ptr()), "combine_ragged")


def copy_multi(srcs, dsts) -> None:
    """dst_p ← src_p (src None: zero fill) for equally sized contiguous blocks."""
    nprob = len(dsts)
    count = int(dsts[0].numel())
    _lib.check(_lib.load().efn_copy_multi(nprob, None if srcs is None else _ptrs(srcs),
                                          _ptrs(dsts), count, stream_ptr()), "copy_multi")


def rowreduce_strided(op: int, A, B, nrow: int, stride_a: int, stride_b: int, m: int, n: int):
    """Row reductions over logical rows i → physical (i // nrow)·stride + i % nrow."""
    out = torch.empty(m, dtype=torch.complex128, device=A.device)
    lib = _lib.load()
    need = lib.efn_rowreduce_work_bytes(m, n)
    work = SCRATCH.get(need, "rowred")
    _lib.check(lib.efn_rowreduce_strided(op, A.data_ptr(), stride_a,
                                         None if B is None else B.data_ptr(), stride_b, nrow, m, n,
                                         out.data_ptr(), work.data_ptr(), work.numel(),
                                         stream_ptr()), "rowreduce_strided")
    return out


def row_axpby_strided(X, stride_x, a, Y, stride_y, b, nrow, m, n, out, stride_out):
    """out_i = a_i·X_i + b_i·Y_i over strided logical rows (see rowreduce_strided)."""
    a = torch.as_tensor(a, dtype=torch.float64, device=X.device).reshape(-1).contiguous()
    if Y is not None:
        b = torch.as_tensor(b, dtype=torch.float64, device=X.device).reshape(-1).contiguous()
    _lib.check(_lib.load().efn_row_axpby_strided(X.data_ptr(), stride_x, a.data_ptr(),
                                                 None if Y is None else Y.data_ptr(), stride_y,
                                                 None if Y is None else b.data_ptr(), nrow, m, n,
                                                 out.data_ptr(), stride_out, stream_ptr()),
               "row_axpby_strided")
    return out


def rowstats_strided(A, B, nrow: int, stride_a: int, stride_b: int, m: int, n: int):
    """(m, 4) FP64 per logical row: Re⟨A_i,B_i⟩, Im⟨A_i,B_i⟩, |A_i|², |B_i|² in one
    read of each block (Rayleigh quotient and residual inputs, solver.py:228-238)."""
    out = torch.empty((m, 4), dtype=torch.float64, device=A.device)
    lib = _lib.load()
    work = SCRATCH.get(lib.efn_rowstats_work_bytes(m, n), "rowstats")
    _lib.check(lib.efn_rowstats_strided(A.data_ptr(), stride_a, B.data_ptr(), stride_b, nrow, m, n,
                                        out.data_ptr(), work.data_ptr(), work.numel(),
                                        stream_ptr()), "rowstats_strided")
    return out


def _check_sel(sel, nrow: int, nslot: int, n: int, blocks) -> int:
    """Host checks of a row selection (slot << 16 | r entries, device int32)
    against the blocks it addresses: each (block, stride) must hold nslot
    slots of stride rows of length n with nrow ≤ stride.  Returns len(sel)."""
    if not (torch.is_tensor(sel) and sel.dtype == torch.int32 and sel.dim() == 1 and sel.is_contiguous()):
        raise ValueError("row selection must be a contiguous 1-D int32 tensor")
    if not 0 < nrow < 65536 or not 0 < nslot < 32768:
        raise ValueError("row selection needs 0 < nrow < 65536 and 0 < nslot < 32768")
    for T, stride in blocks:
        if T is None:
            continue
        if stride < nrow or _flat_n(T) != n or T.shape[0] < (nslot - 1) * stride + nrow:
            raise ValueError(f"block {tuple(T.shape)} does not hold {nslot} slots of stride {stride} "
                             f"(nrow {nrow}, row length {n})")
        if sel.device != T.device:
            raise ValueError("row selection and blocks on different devices")
    return int(sel.numel())


def rowstats_sel(A, B, sel, nrow: int, nslot: int, stride_a: int, stride_b: int, n: int):
    """rowstats_strided over selected rows: sel[i] = slot << 16 | r names row r
    of slot `slot` (physical row slot·stride + r of A and of B); (len(sel), 4)."""
    m = _check_sel(sel, nrow, nslot, n, [(A, stride_a), (B, stride_b)])
    out = torch.empty((m, 4), dtype=torch.float64, device=A.device)
    if m == 0:
        return out
    lib = _lib.load()
    work = SCRATCH.get(lib.efn_rowstats_work_bytes(m, n), "rowstats")
    _lib.check(lib.efn_rowstats_sel(A.data_ptr(), stride_a, B.data_ptr(), stride_b, sel.data_ptr(), nrow,
                                    nslot, m, n, out.data_ptr(), work.data_ptr(), work.numel(),
                                    stream_ptr()), "rowstats_sel")
    return out


def row_axpby_norm_sel(X, stride_x, a, Y, stride_y, b, sel, nrow: int, nslot: int, n: int, out,
                       stride_out):
    """row_axpby_norm_strided over selected rows (see rowstats_sel); a, b per
    selected row; returns |out_i|² (FP64, len(sel))."""
    m = _check_sel(sel, nrow, nslot, n, [(X, stride_x), (Y, stride_y), (out, stride_out)])
    a = torch.as_tensor(a, dtype=torch.float64, device=X.device).reshape(-1).contiguous()
    if Y is not None:
        b = torch.as_tensor(b, dtype=torch.float64, device=X.device).reshape(-1).contiguous()
    if a.numel() != m or (Y is not None and b.numel() != m):
        raise ValueError("row_axpby_norm_sel: one coefficient per selected row")
    nrm = torch.empty(m, dtype=torch.complex128, device=X.device)
    if m == 0:
        return nrm.real
    lib = _lib.load()
    work = SCRATCH.get(lib.efn_rowreduce_work_bytes(m, n), "rowred")
    _lib.check(lib.efn_row_axpby_norm_sel(X.data_ptr(), stride_x, a.data_ptr(),
                                          None if Y is None else Y.data_ptr(), stride_y,
                                          None if Y is None else b.data_ptr(), sel.data_ptr(), nrow, nslot,
                                          m, n, out.data_ptr(), stride_out, nrm.data_ptr(),
                                          work.data_ptr(), work.numel(), stream_ptr()),
               "row_axpby_norm_sel")
    return nrm.real


def row_axpby_norm_strided(X, stride_x, a, Y, stride_y, b, nrow, m, n, out, stride_out):
    """out_i = a_i·X_i + b_i·Y_i over strided rows; returns |out_i|² (FP64, m) from
    the same pass."""
    a = torch.as_tensor(a, dtype=torch.float64, device=X.device).reshape(-1).contiguous()
    if Y is not None:
        b = torch.as_tensor(b, dtype=torch.float64, device=X.device).reshape(-1).contiguous()
    nrm = torch.empty(m, dtype=torch.complex128, device=X.device)
    lib = _lib.load()
    work = SCRATCH.get(lib.efn_rowreduce_work_bytes(m, n), "rowred")
    _lib.check(lib.efn_row_axpby_norm_strided(X.data_ptr(), stride_x, a.data_ptr(),
                                              None if Y is None else Y.data_ptr(), stride_y,
                                              None if Y is None else b.data_ptr(), nrow, m, n,
                                              out.data_ptr(), stride_out, nrm.data_ptr(),
                                              work.data_ptr(), work.numel(), stream_ptr()),
               "row_axpby_norm_strided")
    return nrm.real


_POOL = None
_POOL_LOCK = threading.Lock()


def _pool():
    """Thread pool for the host's small FP64 eigensolves."""
    global _POOL
    with _POOL_LOCK:  # serving engines on several host threads may get here together
        if _POOL is None:
            from concurrent.futures import ThreadPoolExecutor
            _POOL = ThreadPoolExecutor(max_workers=max(1, min(16, os.cpu_count() or 1)))
    return _POOL


_BLAS_LOCK = threading.Lock()
_BLAS_USERS = 0
_BLAS_LIMITER = None
_BLAS_CONTROLLER = None   # threadpoolctl's scan of the loaded BLAS libraries, made once


@contextlib.contextmanager
def one_blas_thread():
    """Host BLAS at one thread while a pooled batch of small eigensolves runs:
    these matrices are ≤ 96 × 96, where threaded BLAS inside a pool of threaded
    calls oversubscribes the cores (measured 48 ms for 21 eigensolves of
    72 × 72 with 8 BLAS threads per call, 5–13 ms with one).  The BLAS thread
    setting is process-wide, so concurrent batches share one limit, taken by
    the first and restored to the caller's setting by the last.  A serving
    loop holds it for its whole run: restoring a larger thread count makes
    OpenBLAS start its threads again, which cost the engines' host phases
    30 % when the limit came and went with each batch."""
    global _BLAS_USERS, _BLAS_LIMITER, _BLAS_CONTROLLER
    with _BLAS_LOCK:
        if _BLAS_USERS == 0:
            try:
                if _BLAS_CONTROLLER is None:
                    # scanning the process's libraries costs ms with torch loaded;
                    # the controller keeps the scan, a limit from it costs µs
                    from threadpoolctl import ThreadpoolController
                    _BLAS_CONTROLLER = ThreadpoolController()
                _BLAS_LIMITER = _BLAS_CONTROLLER.limit(limits=1, user_api="blas")
            except ImportError:  # without threadpoolctl the BLAS keeps its threads
                _BLAS_LIMITER = None
        _BLAS_USERS += 1
    try:
        yield
    finally:
        with _BLAS_LOCK:
            _BLAS_USERS -= 1
            if _BLAS_USERS == 0 and _BLAS_LIMITER is not None:
                _BLAS_LIMITER.restore_original_limits()
                _BLAS_LIMITER = None


def eigh_stack_host(a: np.ndarray, lowest: int | None = None):
    """FP64 Hermitian eigendecompositions of a (nprob, k, k) numpy stack, one LAPACK
    call (numpy's zheevd, which releases the GIL) per matrix over a thread pool;
    with `lowest`, the lowest eigenpairs are returned."""
    a = (a + np.conj(np.swapaxes(a, -1, -2))) / 2
    if a.shape[0] == 1:
        res = [np.linalg.eigh(a[0])]
    else:
        with one_blas_thread():
            res = list(_pool().map(np.linalg.eigh, list(a)))
    w = np.stack([r[0] for r in res])
    V = np.stack([r[1] for r in res])
    if lowest is not None and lowest < a.shape[-1]:
        w, V = w[:, :lowest], V[:, :, :lowest]
    return w, V


def eigh_list_host(mats, lowest=None):
    """[(w, V)] of FP64 Hermitian eigendecompositions of a list of numpy
    matrices of any sizes (the thread pool of eigh_stack_host); `lowest`: one
    count for all, or one per matrix, of the lowest eigenpairs to keep."""
    mats = [(a + a.conj().T) / 2 for a in mats]
    if len(mats) == 1:
        res = [np.linalg.eigh(mats[0])]
    else:
        with one_blas_thread():
            res = list(_pool().map(np.linalg.eigh, mats))
    if lowest is None:
        return res
    lows = lowest if isinstance(lowest, (list, tuple)) else [lowest] * len(mats)
    return [(w[:k], V[:, :k]) for (w, V), k in zip(res, lows)]


SMALL_METHODS = ("heevx", "ql")
# The device eigensolver for k ≤ 96 when none is named.  "heevx" measured
# faster on general Hermitian matrices (72², lowest 24: 0.58 ms against
# 1.95 ms) and on LOBPCG's Rayleigh–Ritz matrices (0.81 against 1.95 ms; C1
# as one graph 0.27 s against 0.30 s, C2 6.8 s against 7.2 s), but on some of
# those matrices its lowest-m solve on the device differs from the full
# solve's leading columns in the last bits (up to 4e-14), which breaks the
# graph-replay-equals-eager guarantee; QL keeps it.
SMALL_METHOD = "ql"


def eigh_ragged_device(A: torch.Tensor, sizes, lowest: int | None = None, method: str | None = None):
    """The lowest eigenpairs of each problem's leading sizes[p] × sizes[p]
    block of a (nprob ≤ 64, K, K) complex128 device stack (K ≤ 96), in one
    launch of one of the library's tridiagonal eigensolvers: method "heevx"
    (efn_heevx_batched: bisection and inverse iteration) or "ql"
    (efn_heev_ql_batched: implicit QL); None: SMALL_METHOD.  w (nprob, lowest) ascending, +inf
    past min(lowest, sizes[p]); V (nprob, K, lowest), zero past sizes[p] rows
    and min(lowest, sizes[p]) columns.  Under "ql" a problem whose iteration
    failed gets NaN eigenvalues."""
    import ctypes as Ct
    method = SMALL_METHOD if method is None else method
    if method not in SMALL_METHODS:
        raise ValueError(f"eigh_ragged_device: method must be one of {SMALL_METHODS}, got {method!r}")
    if A.dim() != 3 or A.shape[1] != A.shape[2]:
        raise ValueError(f"eigh_ragged_device: expected a (nprob, K, K) stack, got {tuple(A.shape)}")
    A = A.to(torch.complex128).contiguous()
    nprob, K = int(A.shape[0]), int(A.shape[1])
    if sizes is not None and len(sizes) != nprob:
        raise ValueError(f"eigh_ragged_device: {len(sizes)} sizes for {nprob} problems")
    want = K if lowest is None else int(lowest)
    low = min(want, K)  # more columns than the stack has: padding only
    w = torch.empty((nprob, low), dtype=torch.float64, device=A.device)
    V = torch.empty((nprob, K, low), dtype=torch.complex128, device=A.device)
    info = torch.empty(nprob, dtype=torch.int32, device=A.device)
    sz = None if sizes is None else (Ct.c_int * nprob)(*[int(x) for x in sizes])
    fn = _lib.load().efn_heevx_batched if method == "heevx" else _lib.load().efn_heev_ql_batched
    _lib.check(fn(nprob, K, sz, A.data_ptr(), low, w.data_ptr(), V.data_ptr(), info.data_ptr(), stream_ptr()),
               f"{method} eigensolve")
    if want > low:
        w = torch.cat([w, torch.full((nprob, want - low), float("inf"), dtype=w.dtype, device=w.device)], 1)
        V = torch.cat([V, torch.zeros((nprob, K, want - low), dtype=V.dtype, device=V.device)], 2)
    return w, V


QL_MAX_N = 96  # largest matrix of the library's tridiagonal eigensolvers (their LDS)


def eigh_batched_device(A: torch.Tensor, lowest: int | None = None, max_sweeps: int = 0,
                        method: str | None = None):
    """(w, V) of a (nprob, k, k) complex128 device stack on the device, FP64;
    ascending, V[:, :, j] the vector of w[:, j]; with `lowest`, the lowest
    eigenpairs.  No host round trip.  method "heevx": the library's
    bisection + inverse-iteration kernel (efn_heevx_batched, k ≤ 96, no
    workspace, so a graph can capture it); "ql": its tridiagonal QL kernel
    (efn_heev_ql_batched, k ≤ 96, also capturable; 64 × 24² in 0.36 ms, 72²
    in 2.3 ms); "dc": rocSOLVER's batched divide and conquer
    (efn_heevd_batched, any k; 2.5 ms at 72², 4.1 ms for 64 × 24²); "jacobi":
    the library's cyclic Jacobi kernel (efn_heev_batched, k ≤ 96).  None:
    SMALL_METHOD up to 96, "dc" above."""
    if method is None:
        method = SMALL_METHOD if int(A.shape[1]) <= QL_MAX_N else "dc"
    if method in SMALL_METHODS:
        return eigh_ragged_device(A, None, lowest, method)
    A = A.to(torch.complex128).contiguous()
    nprob, k = int(A.shape[0]), int(A.shape[1])
    w = torch.empty((nprob, k), dtype=torch.float64, device=A.device)
    V = torch.empty_like(A)
    lib = _lib.load()
    if method == "dc":
        wb = int(lib.efn_heevd_work_bytes(nprob, k))
        work = torch.empty(max(wb, 16), dtype=torch.uint8, device=A.device)
        _lib.check(lib.efn_heevd_batched(nprob, k, A.data_ptr(), w.data_ptr(), V.data_ptr(),
                                         work.data_ptr(), wb, stream_ptr()), "heevd_batched")
    else:
        _lib.check(lib.efn_heev_batched(nprob, k, A.data_ptr(), w.data_ptr(), V.data_ptr(),
                                        int(max_sweeps), stream_ptr()), "heev_batched")
    if lowest is not None and lowest < k:
        return w[:, :lowest], V[:, :, :lowest]
    return w, V


def eigh_batched_host(A: torch.Tensor, lowest: int | None = None):
    """FP64 Hermitian eigendecomposition of a (nprob, k, k) stack on the host (LAPACK)."""
    return eigh_stack_host(A.detach().cpu().numpy(), lowest)
