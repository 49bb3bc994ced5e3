"""The C-ABI library loads and exports every symbol include/boxzy.h declares
(no GPU work is launched)."""
import os
import re

from conftest import ROOT


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "boxzy.h")).read()
    # every function declaration at the start of a line, whatever its return type
    return sorted(set(re.findall(r"^(?:[A-Za-z_][A-Za-z_0-9]*\s*\**\s+)+\**(bx_[a-z_0-9]+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ("bx_create", "bx_state", "bx_velocity", "bx_courant", "bx_source", "bx_flux",
              "bx_set_ghost_cells", "bx_avisc", "bx_step", "bx_last_error", "bx_eos_defaults"):
        assert s in syms


def test_library_exports_all_declared_symbols():
    from boxzyhydro_amd import native as nv

    L = nv.lib()
    for s in declared_symbols():
        assert hasattr(L, s), s
    assert set(declared_symbols()) <= set(nv.EXPORTED)
