"""The C ABI library loads on the CPU and exports every symbol that
include/gm_capi.h declares (no compute calls: there is no GPU here)."""
import os
import re

from goma_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "gm_capi.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(gm_[a-z0-9_]+)\s*\(", src)))


def test_header_and_binding_agree():
    assert _declared() == sorted(capi.EXPORTS)


def test_library_exports_every_symbol():
    L = capi.lib()
    missing = [s for s in _declared() if not hasattr(L, s)]
    assert not missing, missing


def test_default_options_are_gomas():
    o = capi.solver_opts()
    assert o.eps_resid == 1e-10 and o.eps_corr == 1e10 and o.min_damp == 0.05
    assert o.lin_tol == 1e-6 and o.max_its == 500 and o.line_search == 0
    # additions without a Goma counterpart are off by default: Goma's fixed linear
    # tolerance, re-assembled coarse operators, recalc preconditioners
    assert o.lin_tol_adapt == 0 and o.mg_galerkin == 0 and o.pc_reuse == 0


def test_second_multigrid_attach_is_refused():
    """The library keeps pointers to the attached coarse levels, so a second
    attach (which would drop and free them) is refused before any library call."""
    import pytest

    from goma_amd import problems

    ctx = capi.Context.__new__(capi.Context)
    ctx.h = None
    ctx.prob = problems.lid_cavity(n=4, Re=1.0)
    ctx._coarse = [object()]
    with pytest.raises(RuntimeError, match="already attached"):
        ctx.attach_multigrid(2)
    ctx._coarse = []
