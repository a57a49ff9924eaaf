"""The C ABI library (CPU-side checks; no compute calls without a GPU).

* libtgrav.so loads and exports every entry point include/tgrav.h declares;
* argument errors are reported through return codes and tgrav_last_error before any
  device work (AMUSE's ERROR_CODE convention);
* the Python mirror exposes the reference's method names with the reference's argument
  order (interface.py:221-243, 548-559).
"""
import ctypes
import json
import inspect
import os
import re

import pytest

from torch_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tgrav.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(tgrav_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    bound = {n for n, _, _ in _lib.SIGNATURES}
    assert set(names) == bound, set(names) ^ bound


def test_problem_struct_matches_header():
    src = open(HEADER).read()
    body = src[src.index("typedef struct tgrav_problem"):src.index("} tgrav_problem;")]
    fields = re.findall(r"\b([a-z][a-z0-9_]*)\s*[,;]", body.split("{", 1)[1])
    assert [f for f, _ in _lib.Problem._fields_] == fields
    assert ctypes.sizeof(_lib.Problem) == 8 * len(fields)


def test_version_and_variants():
    lib = _lib.load()
    assert lib.tgrav_version() >= 1
    for mode in (_lib.PLAIN_FP64, _lib.MIXED):
        n = lib.tgrav_num_variants(mode)
        assert n >= 4
        k, nt, ts = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        for v in range(n):
            assert lib.tgrav_variant_info(mode, v, ctypes.byref(k), ctypes.byref(nt), ctypes.byref(ts)) == 0
            assert k.value >= 1 and nt.value % 64 == 0 and ts.value % nt.value == 0
        assert lib.tgrav_variant_info(mode, n, None, None, None) == _lib.EINVAL
    assert lib.tgrav_num_variants(7) == _lib.EINVAL
    assert b"bad mode" in lib.tgrav_last_error()


def test_null_context_and_arguments_are_rejected():
    lib = _lib.load()
    assert lib.tgrav_create(None, 0, None) == _lib.EINVAL
    assert b"out is NULL" in lib.tgrav_last_error()
    assert lib.tgrav_set_sources(None, 1, None, None, None, None, None, 0) == _lib.EINVAL
    assert lib.tgrav_eval(None, 1, None, None, None, None, None, None, None, None, 0, 0) == _lib.EINVAL
    assert lib.tgrav_set_constant(None, 1.0) == _lib.EINVAL
    assert lib.tgrav_destroy(None) == 0
    with pytest.raises(_lib.TgravError) as ei:
        _lib.check(lib.tgrav_set_validation(None, 1), "tgrav_set_validation")
    assert ei.value.code == _lib.EINVAL and "context is NULL" in str(ei.value)


def test_create_without_a_device_fails_cleanly():
    """Here (no GPU) tgrav_create must return an error code, not crash."""
    lib = _lib.load()
    ctx = ctypes.c_void_p()
    rc = lib.tgrav_create(ctypes.byref(ctx), 0, None)
    if rc == 0:
        lib.tgrav_destroy(ctx)
        pytest.skip("a GPU is present")
    assert rc in (_lib.EHIP, _lib.EINVAL)
    assert ctx.value is None
    assert lib.tgrav_last_error()


def test_python_mirror_keeps_reference_method_names():
    from torch_amd import GasGravityField, StellarGravityField
    for cls, meth in ((StellarGravityField, "get_gravity_at_point"),
                      (StellarGravityField, "get_potential_at_point"),
                      (GasGravityField, "get_accel_gas_on_particles"),
                      (GasGravityField, "get_potential_at_point")):
        params = list(inspect.signature(getattr(cls, meth)).parameters)
        assert params == ["self", "eps", "x", "y", "z"], (cls, meth, params)


def test_tune_cache_parser(tmp_path):
    """The library's own reader of the autotuner cache (used by tgrav_create) accepts the
    committed cache and rejects malformed files; no device needed."""
    lib = _lib.load()
    cache = os.path.join(os.path.dirname(HEADER), "..", "torch_amd", "tune_cache.json")
    with open(cache) as fh:
        n = len(json.load(fh))
    assert lib.tgrav_tune_cache_entries(os.fsencode(cache)) == n > 0
    good = tmp_path / "t.json"
    good.write_text('{"a|eval|mixed|s1|t2": {"variant": 3, "nsplit": 8, "ms": 1.5, "note": "x",'
                    ' "list": [1, {"y": 2}]}, "b": {}}')
    assert lib.tgrav_tune_cache_entries(os.fsencode(str(good))) == 2
    for bad in ('{"a": 3}', '{"a": {"variant": }}', '[1, 2]', '{"a": {"variant": 1}', ''):
        f = tmp_path / "bad.json"
        f.write_text(bad)
        assert lib.tgrav_tune_cache_entries(os.fsencode(str(f))) == _lib.EINVAL, bad
    assert lib.tgrav_tune_cache_entries(os.fsencode(str(tmp_path / "missing.json"))) == _lib.EINVAL


def test_tune_cache_parser_fuzz(tmp_path):
    """Random, truncated and deeply nested inputs: the reader answers with an entry count or
    TGRAV_EINVAL and never crashes or reads past the text (hypothesis, derandomized)."""
    from hypothesis import HealthCheck, given, settings
    from hypothesis import strategies as st

    lib = _lib.load()
    valid = open(os.path.join(os.path.dirname(HEADER), "..", "torch_amd", "tune_cache.json")).read()
    path = tmp_path / "fuzz.json"

    def check(text):
        path.write_bytes(text if isinstance(text, bytes) else text.encode("utf-8", "surrogatepass"))
        rc = lib.tgrav_tune_cache_entries(os.fsencode(str(path)))
        assert rc == _lib.EINVAL or rc >= 0

    json_values = st.recursive(
        st.none() | st.booleans() | st.integers() | st.floats(allow_nan=False) | st.text(max_size=8),
        lambda inner: st.lists(inner, max_size=4) | st.dictionaries(st.text(max_size=6), inner, max_size=4),
        max_leaves=20)

    @settings(max_examples=300, derandomize=True, deadline=None, database=None,
              suppress_health_check=list(HealthCheck))
    @given(cut=st.integers(0, len(valid)), junk=st.binary(max_size=64), doc=json_values,
           depth=st.integers(0, 3000))
    def run(cut, junk, doc, depth):
        check(valid[:cut])                     # every truncation of the committed cache
        check(valid[:cut].encode() + junk)     # with trailing garbage
        check(json.dumps(doc))                 # arbitrary JSON documents
        check("[" * depth + "]" * depth)       # deep nesting
        check('{"a": ' * depth)                # unterminated nesting

    run()
