"""The C-ABI library loads, exports every symbol include/wamp.h declares, and
rejects bad arguments with a message (no GPU needed: argument checks run
before any HIP call)."""
import ctypes as C
import os
import re

import pytest

from wamp import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "wamp.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(wamp3?_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    syms = declared_symbols()
    assert len(syms) >= 14
    for s in syms:
        assert hasattr(lib, s), s
    # and the ctypes table binds exactly the declared set
    assert sorted(n for n, _, _ in _lib.SIGNATURES) == syms


def _grid(**kw):
    g = dict(nx=16, ny=8, g=5, ld=18, dx=0.1, dy=0.1)
    g.update(kw)
    return _lib.WampGrid(g["nx"], g["ny"], g["g"], g["ld"], g["dx"], g["dy"], (C.c_int32 * 4)(0, 0, 1, 1))


@pytest.mark.parametrize("kw, msg", [
    (dict(nx=0), "nx, ny"),
    (dict(g=3), "g must be >= 4"),
    (dict(ld=10), "ld"),
    (dict(dx=-1.0), "dx, dy"),
])
def test_ctx_create_rejects_bad_grids(kw, msg):
    lib = _lib.load()
    ctx = C.c_void_p()
    eos = _lib.WampEos(4.4, 6000.0, 1.4, 0.0)
    sch = _lib.WampScheme(0, 1, 0, 0, 1, 0.4)
    rc = lib.wamp_ctx_create(C.byref(_grid(**kw)), C.byref(eos), C.byref(sch), 0, C.byref(ctx))
    assert rc < 0 and not ctx.value
    assert msg in lib.wamp_last_error().decode()


def test_ctx_create_rejects_bad_scheme_and_precision():
    lib = _lib.load()
    ctx = C.c_void_p()
    eos = _lib.WampEos(4.4, 6000.0, 1.4, 0.0)
    rc = lib.wamp_ctx_create(C.byref(_grid()), C.byref(eos),
                             C.byref(_lib.WampScheme(2, 1, 0, 0, 1, 0.4)), 0, C.byref(ctx))
    assert rc < 0 and "base" in lib.wamp_last_error().decode()
    rc = lib.wamp_ctx_create(C.byref(_grid()), C.byref(eos),
                             C.byref(_lib.WampScheme(0, 1, 0, 0, 1, 0.4)), 7, C.byref(ctx))
    assert rc < 0 and "precision" in lib.wamp_last_error().decode()


def test_null_context_is_an_error():
    lib = _lib.load()
    rc = lib.wamp_step(None, _lib.ptrs6([0] * 6), None, None, None)
    assert rc < 0 and "context" in lib.wamp_last_error().decode()
    with pytest.raises(_lib.WampError):
        _lib.check(rc, "wamp_step")


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(_lib.WampError, match="not found"):
        _lib.load(str(tmp_path / "nope.so"))


def _grid3(**kw):
    g = dict(nx=8, ny=6, nz=4, g=5, ldz=14, dx=0.1, dy=0.1, dz=0.1)
    g.update(kw)
    return _lib.WampGrid3(g["nx"], g["ny"], g["nz"], g["g"], g["ldz"], g["dx"], g["dy"], g["dz"],
                          (C.c_int32 * 6)(0, 0, 1, 1, 1, 1))


@pytest.mark.parametrize("kw, scheme, msg", [
    (dict(nz=0), (0, 1, 0, 0), "nx, ny, nz"),
    (dict(g=3), (0, 1, 0, 0), "g must be >= 4"),
    (dict(ldz=13), (0, 1, 0, 0), "ldz"),
    (dict(dz=0.0), (0, 1, 0, 0), "dx, dy, dz"),
    ({}, (1, 0, 0, 0), "not built"),  # MP5 without central6: no 3D instantiation
])
def test_ctx3_create_rejects_bad_grids_and_schemes(kw, scheme, msg):
    lib = _lib.load()
    ctx = C.c_void_p()
    eos = _lib.WampEos(4.4, 6000.0, 1.4, 0.0)
    sch = _lib.WampScheme(*scheme, 1, 0.4)
    rc = lib.wamp3_ctx_create(C.byref(_grid3(**kw)), C.byref(eos), C.byref(sch), 0, C.byref(ctx))
    assert rc < 0 and not ctx.value
    assert msg in lib.wamp_last_error().decode()


def test_3d_entry_points_reject_null_contexts():
    lib = _lib.load()
    U = _lib.PtrArr7(*([None] * 7))
    assert lib.wamp3_step(None, U, None, None, None) < 0
    assert lib.wamp3_get_dt(None, U, None, None) < 0
    assert lib.wamp3_prepare(None, U, None) < 0
    assert lib.wamp3_ctx_scratch_bytes(None) == -1
    assert lib.wamp3_ctx_destroy(None) == 0


def test_ctx3_create_rejects_an_unknown_precision():
    lib = _lib.load()
    ctx = C.c_void_p()
    eos = _lib.WampEos(4.4, 6000.0, 1.4, 0.0)
    sch = _lib.WampScheme(0, 0, 1, 1, 1, 0.4)
    rc = lib.wamp3_ctx_create(C.byref(_grid3()), C.byref(eos), C.byref(sch), 2, C.byref(ctx))
    assert rc < 0 and not ctx.value and "precision" in lib.wamp_last_error().decode()
