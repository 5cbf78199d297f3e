"""The C-ABI library loads and exports every entry point of include/bsam_gpu.h."""

import ctypes
import os
import re

from bsam_amd import engine

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "bsam_gpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bsam_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_all_declared_symbols():
    lib = ctypes.CDLL(engine.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_errors_are_reported_not_fatal():
    from bsam_amd.layout import Layout
    lay = Layout(2, (8, 8), 1.0 / 8)
    eng = engine.Engine(lay, dt=1e-3, eps=0.03, device=-1)
    # host-only contexts refuse compute calls with an error code and message
    rc = engine.lib().bsam_set_ghost(eng.ctx, 0, 0)
    assert rc < 0
    assert b"host-only" in engine.lib().bsam_last_error(eng.ctx)
    cfg = engine.Cfg(2, 2, 2, 0, 1e-3, 1e-3, 1, 0.5, 1, 4, 2, 10, 1e-8, 0, -1)
    assert not engine.lib().bsam_ctx_create(ctypes.byref(cfg))
    assert b"mbc" in engine.lib().bsam_last_error(None)
    # ErrorAFAS knows errortype 1 and 2 and, like the reference
    # (lib/afasroutines.f90:143-146), refuses every other value
    for bad in (0, 3, -1):
        cfg = engine.Cfg(2, 2, 1, 0, 1e-3, 1e-3, 1, 0.5, 1, 4, 2, 10, 1e-8, 0, -1, bad)
        assert not engine.lib().bsam_ctx_create(ctypes.byref(cfg))
        assert b"errortype" in engine.lib().bsam_last_error(None)
    for et in (1, 2):
        e = engine.Engine(lay, dt=1e-3, eps=0.03, device=-1, errortype=et)
        e.close()


def test_hierarchy_validation():
    from bsam_amd.layout import Layout
    import pytest
    lay = Layout(2, (16, 16), 1.0 / 16)
    lay.add_child(0, ((1, 8), (1, 8)))
    lay.add_child(0, ((5, 12), (5, 12)))   # overlaps the first child
    with pytest.raises(engine.BsamError, match="overlapping"):
        engine.Engine(lay, dt=1e-3, eps=0.03, device=-1)
