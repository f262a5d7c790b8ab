"""The drop-in surface: ``install_alias()`` makes the reference's import names
resolve to this package, and every public function of the hot path keeps the
reference's parameter names, order and defaults (extra parameters may only
follow them, with defaults).  The expected signatures are transcribed from the
reference sources cited beside each entry."""
import importlib
import inspect
import sys

import pytest

# (module, name, [(parameter, default or REQUIRED)], reference file:line)
REQ = object()
VAR = "*"
SIGS = [
    ("operator", "make_basis", [("grid_size", REQ), ("box_size", REQ), ("k_frac", (0.0, 0.0, 0.0)),
                                ("rdtype", "float32")], "eigenfns/operator.py:44"),
    ("operator", "MaxwellOperator", [("eps", REQ), ("box_size", REQ), ("k_frac", (0.0, 0.0, 0.0)),
                                     ("dtype", "complex64")], "eigenfns/operator.py:136"),
    ("solver", "gram", [("A", REQ), ("B", REQ)], "eigenfns/solver.py:35"),
    ("solver", "combine", [("C", REQ), ("blocks", VAR)], "eigenfns/solver.py:49"),
    ("solver", "apply_chunked", [("fn", REQ), ("X", REQ), ("rows", REQ)], "eigenfns/solver.py:74"),
    ("solver", "deflate_chunk_rows", [("grid_size", REQ), ("target_bytes", 0.75e9)],
     "eigenfns/solver.py:85"),
    ("solver", "deflate", [("X", REQ), ("locked", REQ), ("chunk_rows", 128)], "eigenfns/solver.py:95"),
    ("solver", "ortho_block", [("X", REQ), ("passes", 2), ("drop", 1e-5)], "eigenfns/solver.py:113"),
    ("solver", "lobpcg_blocks",
     [("op", REQ), ("nev", REQ), ("m", 96), ("guard", 48), ("tol", 1e-4), ("maxit", 400), ("seed", 0),
      ("precond_shift", "median"), ("verbose", True), ("log_every", 0), ("initial_locked", None),
      ("initial_vals", None), ("on_block", None), ("locked_storage", "auto"), ("theta_chunk", 0)],
     "eigenfns/solver.py:136-152"),
    ("solver", "lobpcg_blocks_resumable", [("op", REQ), ("nev", REQ), ("checkpointer", None),
                                           ("resume", False)], "eigenfns/solver.py:385"),
    ("chebyshev", "lanczos_lambda_max", [("op", REQ), ("iters", 24), ("seed", 0), ("n_seeds", 2)],
     "eigenfns/chebyshev.py:28"),
    ("chebyshev", "chebsi_bottom_up",
     [("op", REQ), ("nev", REQ), ("guard", 48), ("degree", 200), ("tol", 1e-4), ("max_outer", 30),
      ("chunk", 64), ("seed", 0), ("lam_max", None), ("verbose", True)], "eigenfns/chebyshev.py:103-114"),
    ("chebyshev", "kpm_count_below",
     [("op", REQ), ("lam_b", REQ), ("lam_max", REQ), ("degree", 500), ("n_probe", 30), ("seed", 1),
      ("locked", None)], "eigenfns/chebyshev.py:213"),
    ("interior", "rr_extract", [("op", REQ), ("X", REQ), ("theta_chunk", 8)], "eigenfns/interior.py:37"),
    ("interior", "_stream_rotate", [("C", REQ), ("Xh", REQ), ("rows", 8)], "eigenfns/interior.py:93"),
    ("interior", "gram_hosted", [("Ah", REQ), ("Bh", None), ("rows", 8)], "eigenfns/interior.py:112"),
    ("interior", "svqb_hosted", [("Xh", REQ), ("passes", 2), ("rows", 8)], "eigenfns/interior.py:129"),
    ("interior", "rr_extract_hosted", [("op", REQ), ("Xh", REQ), ("theta_chunk", 8)],
     "eigenfns/interior.py:141"),
    ("interior", "bandpass_subspace_hosted",
     [("op", REQ), ("lam_lo", REQ), ("lam_hi", REQ), ("lam_max", REQ), ("m", 104), ("degree", 3000),
      ("max_outer", 2), ("res_tol", 1e-4), ("chunk", 8), ("theta_chunk", 8), ("seed", 0), ("X0h", None),
      ("verbose", True)], "eigenfns/interior.py:181-186"),
    ("interior", "_jackson", [("p", REQ)], "eigenfns/interior.py:312"),
    ("interior", "_step_coeffs", [("xb", REQ), ("p", REQ)], "eigenfns/interior.py:319"),
    ("interior", "bandpass_coeffs", [("lam_lo", REQ), ("lam_hi", REQ), ("lam_max", REQ), ("degree", REQ)],
     "eigenfns/interior.py:328"),
    ("interior", "_apply_cheb_poly", [("op", REQ), ("X", REQ), ("coef", REQ), ("lam_max", REQ)],
     "eigenfns/interior.py:336"),
    ("interior", "bandpass_subspace",
     [("op", REQ), ("lam_lo", REQ), ("lam_hi", REQ), ("lam_max", REQ), ("m", 64), ("degree", 2000),
      ("max_outer", 8), ("res_tol", 1e-4), ("target_count", None), ("chunk", 16), ("theta_chunk", 8),
      ("seed", 0), ("X0", None), ("verbose", True)], "eigenfns/interior.py:357-362"),
    ("interior", "gram_chunked", [("A", REQ), ("B", REQ), ("rows", 8)], "eigenfns/interior.py:432"),
    ("interior", "ortho_block_lowmem", [("X", REQ), ("passes", 2), ("rows", 8)], "eigenfns/interior.py:444"),
    ("io", "BlockCheckpointer", [("directory", REQ), ("tag", REQ), ("meta", None)], "eigenfns/io.py:19-25"),
]


@pytest.fixture(scope="module")
def aliased():
    import eigenfns_amd
    saved = {k: v for k, v in sys.modules.items() if k == "eigenfns" or k.startswith("eigenfns.")}
    for k in saved:
        del sys.modules[k]
    eigenfns_amd.install_alias()
    yield
    for k in [k for k in sys.modules if k == "eigenfns" or k.startswith("eigenfns.")]:
        del sys.modules[k]
    sys.modules.update(saved)


def test_alias_resolves_the_reference_import_names(aliased):
    import eigenfns_amd
    for sub in ("operator", "solver", "chebyshev", "interior", "io"):
        mod = importlib.import_module(f"eigenfns.{sub}")
        assert mod is importlib.import_module(f"eigenfns_amd.{sub}")
    from eigenfns.solver import lobpcg_blocks  # noqa: F401  the reference's own import lines
    from eigenfns.operator import MaxwellOperator  # noqa: F401
    assert importlib.import_module("eigenfns") is eigenfns_amd


def _default_matches(got, want):
    if want is REQ:
        return got is inspect.Parameter.empty
    if isinstance(want, str) and want in ("float32", "complex64"):
        return want in str(got)
    if isinstance(want, float):
        return isinstance(got, (int, float)) and got == pytest.approx(want, rel=1e-12)
    return got == want


@pytest.mark.parametrize("mod,name,params,cite", SIGS, ids=[f"{m}.{n}" for m, n, _, _ in SIGS])
def test_signature_matches_reference(aliased, mod, name, params, cite):
    obj = getattr(importlib.import_module(f"eigenfns.{mod}"), name)
    sig = inspect.signature(obj)
    got = [p for p in sig.parameters.values()]
    for i, (pname, default) in enumerate(params):
        assert i < len(got), f"{name}: missing parameter {pname} ({cite})"
        p = got[i]
        assert p.name == pname, f"{name}: parameter {i} is {p.name}, reference has {pname} ({cite})"
        if default is VAR:
            assert p.kind is inspect.Parameter.VAR_POSITIONAL
        else:
            assert _default_matches(p.default, default), \
                f"{name}.{pname}: default {p.default!r}, reference {default!r} ({cite})"
    for p in got[len(params):]:
        assert p.default is not inspect.Parameter.empty or p.kind in (
            inspect.Parameter.VAR_KEYWORD, inspect.Parameter.VAR_POSITIONAL), \
            f"{name}: extra parameter {p.name} without a default"
