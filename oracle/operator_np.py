"""Oracle: transverse plane-wave basis and the Maxwell operator Θ = ∇×ε⁻¹∇× (NumPy).

Restates eigenfns/operator.py:
  make_basis                 operator.py:44-79
  theta_block                operator.py:82-95   (Θ·H, six FFTs)
  precondition_simple_block  operator.py:98-101
  precondition_block         operator.py:104-127 (MPB transverse-projection preconditioner)
  MaxwellOperatorNP          operator.py:130-184
  energy_density             scripts/run_modes.py:116-122 (ε|E|² per band)

Vectors are (m, 2, G, G, G) complex, C order, z fastest; frequencies in
np.fft.fftfreq order; ``ifftn`` carries 1/G³ and ``fftn`` is unnormalised.
FFTs run in complex128 (the FP64-FFT oracle) unless ``fft_dtype`` says
otherwise.

Pinned by (tests/test_oracle_operator.py): the orthonormal right-handed frame
and the exact integer swap rule at Γ; the homogeneous known answer
Θ·H = (|k+G|²/ε₀)·H per coefficient; the analytic homogeneous and Bloch
spectra of the dense matrix (tests/test_operator.py:27-65 of the reference);
Hermiticity/PSD on a disordered ε; a brute-force DFT at G=4; for the
preconditioner the homogeneous known answer (ε₀/knr²)·H; for the real-space
fields the plane-wave closed forms (|H(r)|² = 1/G⁶, ε|E(r)|² = 1/G⁶ for a unit
coefficient) and, on a disordered ε, Parseval: Σ_r|H(r)|² = ‖H‖²/G³ and
Σ_r ε|E(r)|² = Re⟨H, ΘH⟩/(G³λ) for any H and λ.
"""
from __future__ import annotations

import os

import numpy as np
import scipy.fft as sfft

_WORKERS = os.cpu_count() or 1


def set_workers(n: int) -> None:
    global _WORKERS
    _WORKERS = max(1, int(n))


def make_basis(grid_size: int, box_size: float, k_frac=(0.0, 0.0, 0.0)):
    """Return (kn, t, K) in FP64: kn (G,G,G), t (2,3,G,G,G), K (3,G,G,G).

    Follows operator.py:44-79 operation for operation (FP64): reference axis ẑ,
    switched to x̂ where |K_z| > 0.9·|K|; t̂₁ = K×ref/|·|, t̂₂ = K×t̂₁/|·|; both
    zeroed where K = 0.
    """
    G, L = int(grid_size), float(box_size)
    g1 = 2 * np.pi * np.fft.fftfreq(G, d=L / G)
    k0 = 2 * np.pi * np.asarray(k_frac, np.float64) / L
    KX, KY, KZ = np.meshgrid(g1 + k0[0], g1 + k0[1], g1 + k0[2], indexing="ij")
    K = np.stack([KX, KY, KZ])
    kn = np.sqrt((K ** 2).sum(0))
    ref = np.zeros_like(K)
    ref[2] = 1.0
    swap = np.abs(K[2]) > 0.9 * np.maximum(kn, 1e-300)
    ref[2][swap] = 0.0
    ref[0][swap] = 1.0
    t1 = np.cross(K, ref, axis=0)
    n1 = np.linalg.norm(t1, axis=0)
    sing = n1 == 0
    n1[sing] = 1.0
    t1 /= n1
    t2 = np.cross(K, t1, axis=0)
    n2 = np.linalg.norm(t2, axis=0)
    n2[n2 == 0] = 1.0
    t2 /= n2
    t1[:, sing] = 0.0
    t2[:, sing] = 0.0
    return kn, np.stack([t1, t2]), K


def swap_mask(grid_size: int, box_size: float, k_frac=(0.0, 0.0, 0.0)) -> np.ndarray:
    """The FP64 reference-axis switch of operator.py:57 as a boolean (G,G,G) array."""
    G, L = int(grid_size), float(box_size)
    g1 = 2 * np.pi * np.fft.fftfreq(G, d=L / G)
    k0 = 2 * np.pi * np.asarray(k_frac, np.float64) / L
    KX, KY, KZ = np.meshgrid(g1 + k0[0], g1 + k0[1], g1 + k0[2], indexing="ij")
    K = np.stack([KX, KY, KZ])
    kn = np.sqrt((K ** 2).sum(0))
    return np.abs(K[2]) > 0.9 * np.maximum(kn, 1e-300)


def _ifftn(x, fft_dtype):
    return sfft.ifftn(x.astype(fft_dtype, copy=False), axes=(-3, -2, -1), workers=_WORKERS)


def _fftn(x, fft_dtype):
    return sfft.fftn(x.astype(fft_dtype, copy=False), axes=(-3, -2, -1), workers=_WORKERS)


def theta_block(H, kn, t, inv_eps, fft_dtype=np.complex128):
    """Θ·H for a block H (m,2,G,G,G): operator.py:82-95."""
    a, b = H[:, 0], H[:, 1]
    c1 = -1j * kn * b
    c2 = 1j * kn * a
    Ec = c1[:, None] * t[0][None] + c2[:, None] * t[1][None]
    E = _ifftn(Ec, fft_dtype)
    D = inv_eps * E
    Dc = _fftn(D, fft_dtype)
    f1 = (t[0][None] * Dc).sum(1)
    f2 = (t[1][None] * Dc).sum(1)
    return np.stack([-1j * kn * f2, 1j * kn * f1], axis=1)


def precondition_simple_block(R, kn, scale):
    """Diagonal kinetic preconditioner R/(|k+G|² + s): operator.py:98-101."""
    return R / (kn ** 2 + scale)


def precondition_block(R, kn, t, eps, floor, fft_dtype=np.complex128):
    """MPB transverse-projection preconditioner: operator.py:104-127."""
    a, b = R[:, 0], R[:, 1]
    knr = np.maximum(kn, floor)
    c1 = 1j * b / knr
    c2 = -1j * a / knr
    Ec = c1[:, None] * t[0][None] + c2[:, None] * t[1][None]
    E = _ifftn(Ec, fft_dtype)
    D = eps * E
    Dc = _fftn(D, fft_dtype)
    f1 = (t[0][None] * Dc).sum(1)
    f2 = (t[1][None] * Dc).sum(1)
    return np.stack([1j * f2 / knr, -1j * f1 / knr], axis=1)


class MaxwellOperatorNP:
    """Matrix-free Θ on the CPU (operator.py:130-184), FP64 throughout.

    ``out_dtype`` is the dtype of returned blocks (complex128 by default; the
    FP32 restatement used as a CPU baseline passes complex64 and
    ``fft_dtype=np.complex64``).
    """

    def __init__(self, eps, box_size, k_frac=(0.0, 0.0, 0.0),
                 fft_dtype=np.complex128, out_dtype=np.complex128):
        eps = np.asarray(eps)
        if eps.ndim != 3 or len(set(eps.shape)) != 1:
            raise ValueError(f"eps must be a cube, got {eps.shape}")
        self.grid_size = eps.shape[0]
        self.box_size = float(box_size)
        self.k_frac = tuple(float(x) for x in k_frac)
        self.kn, self.t, self.K = make_basis(self.grid_size, self.box_size, k_frac)
        self.rdtype = np.float64 if fft_dtype == np.complex128 else np.float32
        self.kn_r = self.kn.astype(self.rdtype)
        self.t_r = self.t.astype(self.rdtype)
        self.inv_eps = (1.0 / eps).astype(self.rdtype)
        self.eps_grid = np.asarray(eps, self.rdtype)
        self.eps_mean_inv = float(np.mean(1.0 / eps))
        # operator.py:149-151: the floor is taken from the FP32-cast |k+G|
        kn32 = self.kn.astype(np.float32)
        self.kn_floor = float(0.5 * kn32[kn32 > 0].min())
        self.fft_dtype = fft_dtype
        self.out_dtype = out_dtype

    @property
    def mask(self):
        return (self.kn > 0).astype(np.float64)

    def theta(self, H):
        return theta_block(np.asarray(H), self.kn_r, self.t_r, self.inv_eps,
                           self.fft_dtype).astype(self.out_dtype, copy=False)

    def precondition(self, R, target: float = 0.0):
        floor = self.rdtype(self.kn_floor)
        return precondition_block(np.asarray(R), self.kn_r, self.t_r, self.eps_grid,
                                  floor, self.fft_dtype).astype(self.out_dtype, copy=False)

    def precondition_simple(self, R, target: float = 0.0):
        scale = np.float32(max(target / self.eps_mean_inv, 1e-6) * self.eps_mean_inv + 1e-6)
        return precondition_simple_block(np.asarray(R), self.kn_r, scale).astype(
            self.out_dtype, copy=False)

    def to_cartesian_spectral(self, H):
        t = self.t_r
        return H[:, 0:1] * t[0][None] + H[:, 1:2] * t[1][None]

    def h_realspace(self, H):
        return _ifftn(self.to_cartesian_spectral(np.asarray(H)), self.fft_dtype)

    def e_realspace(self, H, eigvals):
        kn, t = self.kn_r, self.t_r
        H = np.asarray(H)
        a, b = H[:, 0], H[:, 1]
        Ec = (-1j * kn * b)[:, None] * t[0][None] + (1j * kn * a)[:, None] * t[1][None]
        E = self.inv_eps * _ifftn(Ec, self.fft_dtype)
        return E / np.sqrt(np.asarray(eigvals))[:, None, None, None, None]

    def energy_density(self, H, eigvals):
        """ε·Σ_c|E_c(r)|² per band (scripts/run_modes.py:116-122)."""
        E = self.e_realspace(H, eigvals)
        return self.eps_grid[None] * (np.abs(E) ** 2).sum(1)
