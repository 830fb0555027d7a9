"""The drop-in boundary: the C-ABI library loads without a GPU and exports every function include/sf.h declares."""
import ctypes

import numpy as np

from shatterforge_amd import _native


def test_library_loads_and_exports_every_declared_symbol():
    L = _native.lib()
    names = _native.declared_symbols()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the binding covers them all
    assert set(names) <= set(_native._PROTOS), set(names) - set(_native._PROTOS)


def test_errors_are_returned_not_raised():
    L = _native.lib()
    # bad arguments come back as negative codes with a message, never an abort
    rc = L.sf_scene_pile(-1, None, None, None, None)
    assert rc < 0 and L.sf_last_error() < 0
    assert b"n < 0" in L.sf_last_error_message()
    assert L.sf_world_step(None, ctypes.c_float(1 / 60)) < 0
    assert L.sf_load_solve(None, 0.0, 1e-6, 10) < 0


def test_no_gpu_world_creation_fails_cleanly():
    L = _native.lib()
    if L.sf_device_count() > 0:
        return
    w = L.sf_world_create(10, 10, None)
    assert not w
    assert L.sf_last_error() == -4


def test_rust_binding_names_only_exported_symbols():
    """The Rust crate (shatterforge_amd/rust/sf-hip, source only here) declares every function it binds with a name
    the header declares and the library exports, and binds every legacy, serving and load entry point."""
    import os
    import re
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = open(os.path.join(root, "shatterforge_amd", "rust", "sf-hip", "src", "lib.rs")).read()
    block = src[src.index('extern "C" {'):src.index("/// A failed call")]
    names = re.findall(r"pub fn (sf_\w+)\(", block)
    declared = set(_native.declared_symbols())
    L = _native.lib()
    assert names and all(n in declared and hasattr(L, n) for n in names), [n for n in names if n not in declared]
    want = {n for n in declared if n.startswith(("sf_legacy_", "sf_serve_s", "sf_serve_t", "sf_serve_p", "sf_load_solve"))}
    assert want <= set(names), want - set(names)
