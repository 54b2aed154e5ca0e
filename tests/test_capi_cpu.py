"""CPU-only checks of the C ABI: the gfx950 library loads, exports every
symbol include/lfdem_amd.h declares, and its host-side helpers behave
(no GPU compute calls here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lfdem_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lfd_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from lf_dem_amd import capi
    L = capi.lib()
    names = declared_functions()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # the ctypes signature table covers the whole header
    assert set(names) == set(capi.SIGNATURES)


def test_library_is_gfx950_code_object():
    from lf_dem_amd import capi
    data = open(capi.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_default_params_match_oracle_defaults(oracle):
    from lf_dem_amd import capi
    p = capi.default_params()
    q = oracle.default_params()
    for name, _ in capi.Params._fields_:
        assert getattr(p, name) == getattr(q, name), name
    o = capi.default_opts()
    assert o.tol == 1e-12 and o.kind == 0 and o.lanes_per_row in (1, 2, 4, 8)


def test_last_error_is_a_string():
    from lf_dem_amd import capi
    assert isinstance(capi.lib().lfd_last_error(), bytes)


def test_oracle_not_imported_by_product():
    """The product package never loads the oracle."""
    pkg = os.path.join(ROOT, "lf_dem_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".hpp", ".h")):
                src = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in src.lower(), f


def _header_struct_fields(name):
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    body = re.search(r"typedef struct %s \{(.*?)\} %s;" % (name, name), text, flags=re.S).group(1)
    fields = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        names = decl.split(None, 1)[1]
        fields += [n.strip().lstrip("*").split("[")[0] for n in names.split(",")]
    return fields


def test_solve_opts_layout_matches_header():
    """ctypes mirror of lfd_solve_opts: same fields in the same order."""
    from lf_dem_amd import capi
    assert [n for n, _ in capi.SolveOpts._fields_] == _header_struct_fields("lfd_solve_opts")
    o = capi.default_opts()
    assert o.launch == 0 and o.precond == 1 and o.regroup_interval >= 1
