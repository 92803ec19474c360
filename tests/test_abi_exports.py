"""The C-ABI library loads without a GPU and exports every function the
public headers declare (no device calls are made here)."""
import glob
import os
import re

import pytest

from conftest import ROOT


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"typedef[^;]*;", "", text)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", text,
                             flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_the_reference_link_set():
    # the 13 LASPack symbols the reference fails to link without
    # (CGIter, GMRESIter, SSORPrecond, Q_*, V_*, RTC accessors), plus the
    # ones its comment names (src/tools.cpp:347-349)
    need = {"CGIter", "GMRESIter", "SSORPrecond", "Q_Constr", "Q_Destr", "Q_SetLen",
            "Q_SetEntry", "V_Constr", "V_Destr", "V_SetAllCmp", "SetRTCAccuracy",
            "GetLastNoIter", "GetLastAccuracy", "BiCGSTABIter", "JacobiPrecond", "ILUPrecond"}
    assert need <= set(_declared())


@pytest.mark.parametrize("name", _declared())
def test_library_exports(name):
    import sysu_cfd_heat_amd as P
    L = P.lib()
    assert hasattr(L, name), f"{name} declared in include/ but not exported"


def test_library_without_build_fails_loudly(tmp_path, monkeypatch):
    import sysu_cfd_heat_amd as P
    monkeypatch.setattr(P, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(P, "_lib", None)
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        P.lib()
