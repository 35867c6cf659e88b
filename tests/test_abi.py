"""The C-ABI library loads and exports every symbol include/bdstep.h declares.
CPU only: nothing here calls into the device."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "bdstep.h")


def header_symbols():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(bd_[a-z_0-9]+)\s*\(", txt)))


def test_header_declares_the_boundary():
    syms = header_symbols()
    for s in ("bd_ctx_create", "bd_rpy_assemble", "bd_brownian", "bd_forces", "bd_mob_symv", "bd_run"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    from brownpak_amd import bdstep
    lib = bdstep.load_library()
    missing = [s for s in header_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    assert sorted(bdstep.EXPORTS) == header_symbols()


def test_library_is_built_for_gfx950():
    from brownpak_amd import bdstep
    blob = open(bdstep.LIB_PATH, "rb").read()
    assert b"gfx950" in blob


def test_opts_struct_matches_header():
    from brownpak_amd.bdstep import BdOpts
    txt = open(HEADER).read()
    body = txt[txt.index("typedef struct {", txt.index("BD_PT_DIR")):txt.index("} bd_opts;")]
    fields = re.findall(r"\b(?:int|double)\s+([a-z_]+);", body)
    assert [f for f, _ in BdOpts._fields_] == fields


def test_oracle_is_not_imported_by_the_product():
    pkg = os.path.join(ROOT, "brownpak_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".h", ".cpp", ".f90")):
                src = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in src.replace("Oracle", ""), f
