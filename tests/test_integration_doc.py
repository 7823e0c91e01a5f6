"""The ctypes binding printed in the repository's integration guide (the stub a
maintainer would add beside the reference script, `wamp_binding.py`) loads
libwamp.so and declares every entry point it uses -- executed here on the CPU
(no GPU call is made)."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "wave-multiphase_amd", "csrc", "libwamp.so")


GUIDE = os.path.join(ROOT, "INTEGRATION.md")


def _binding_block():
    text = open(GUIDE).read()
    blocks = re.findall(r"```python\n(.*?)```", text, re.S)
    stub = [b for b in blocks if "wamp_binding.py" in b]
    assert stub, "the integration guide lost its wamp_binding.py block"
    return stub[0]


@pytest.mark.skipif(not os.path.exists(LIB), reason="libwamp.so not built (run __graft_entry__.build())")
def test_binding_block_executes_against_the_library(monkeypatch):
    monkeypatch.chdir(ROOT)
    ns = {}
    exec(compile(_binding_block(), GUIDE, "exec"), ns)
    L = ns["L"]
    for name in ("wamp_ctx_create", "wamp_apply_bc", "wamp_enforce_admissible", "wamp_cons_to_prim",
                 "wamp_prepare", "wamp_compute_fluxes", "wamp_limit_update", "wamp_get_dt", "wamp_step"):
        assert getattr(L, name).argtypes, name
    # the structs match the header's sizes (two int32 x4 + two doubles + int32 x4, ...)
    import ctypes as C
    assert C.sizeof(ns["Grid"]) == 48
    assert C.sizeof(ns["Eos"]) == 32
    assert C.sizeof(ns["Scheme"]) == 32
