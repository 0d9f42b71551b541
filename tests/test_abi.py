"""The C-ABI library loads and exports every symbol include/sunset_amd.h
declares (no device needed: nothing is computed)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(sv_[a-z0-9_]+)\s*\(", txt))
    return names


def test_header_declares_entry_points():
    names = _declared()
    assert {"sv_create", "sv_find_neighbours", "sv_calc_labf_weights", "sv_run",
            "sv_calc_gradient"} <= names


def test_library_exports_every_declared_symbol():
    import sunset_amd
    try:
        L = sunset_amd.lib()
    except sunset_amd.SunsetError as e:
        pytest.skip(str(e))
    missing = [n for n in _declared() if not hasattr(L, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", sunset_amd.core.LIB_PATH],
                         capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (sv_[a-z0-9_]+)", out))
    assert _declared() <= exported, _declared() - exported
    assert L.sv_version().startswith(b"sunset_amd")


def test_library_is_gfx950_code_object():
    """The device code object embedded in our build targets gfx950 only."""
    import sunset_amd
    path = sunset_amd.core.LIB_PATH
    if not os.path.exists(path):
        pytest.skip("library not built")
    data = open(path, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in data
    targets = set(re.findall(rb"amdgcn-amd-amdhsa--(gfx[0-9a-f]+)", data))
    assert targets == {b"gfx950"}, targets


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    import sunset_amd.core as core
    monkeypatch.setattr(core, "_lib", None)
    monkeypatch.setattr(core, "LIB_PATH", str(tmp_path / "nope.so"))
    with pytest.raises(core.SunsetError, match="not built"):
        core.lib()


@pytest.mark.parametrize("morder,forder,form,what", [
    (12, 12, 1, "order 12"), (6, 12, 1, "order 12"), (5, 5, 0, "order 5"),
    (6, 6, 7, "form 7")])
def test_create_refuses_unsupported_orders_and_forms(morder, forder, form, what):
    """sv_create checks the form and the LABFM orders before it touches a
    device (so this runs without one) and leaves the message where
    sv_last_error(NULL) finds it."""
    import ctypes as C
    import sunset_amd
    try:
        L = sunset_amd.lib()
    except sunset_amd.SunsetError as e:
        pytest.skip(str(e))
    h = C.c_void_p()
    cfg = sunset_amd.core.Config(morder, forder, form, 0)
    assert L.sv_create(C.byref(cfg), C.byref(h)) == -1
    assert not h.value
    assert what in L.sv_last_error(None).decode()
