"""The C-ABI library loads and exports every function that include/*.h
declares (no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

from conftest import REPO

HEADERS = [os.path.join(REPO, "include", h) for h in sorted(os.listdir(os.path.join(REPO, "include")))
           if h.endswith(".h")]


def declared_functions():
    names = []
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        for m in re.finditer(r"\b(galop_[a-z0-9_]+)\s*\(", text):
            if not text[:m.start()].rstrip().endswith("#define"):
                names.append(m.group(1))
    return sorted(set(names))


def test_headers_declare_functions():
    names = declared_functions()
    assert "galop_eval_dense" in names and "galop_eval_csr_device" in names
    assert len(names) >= 15


def test_library_exports_every_declared_symbol():
    from galop_amd import _lib
    lib = _lib.load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"
    # the ctypes signature table covers every declared symbol
    untyped = [n for n in declared_functions() if n not in _lib.SIGNATURES]
    assert not untyped, untyped


def test_abi_version_without_device():
    from galop_amd import _lib
    assert _lib.load().galop_abi_version() == 5


def test_null_arguments_rejected_without_device():
    from galop_amd import _lib
    lib = _lib.load()
    assert lib.galop_ctx_create(0, None) == _lib.GALOP_ERR_INVALID
    assert lib.galop_dataset_create(None, None, None, None, 0, 0, None) == _lib.GALOP_ERR_INVALID
    assert lib.galop_eval_csr(None, None, None, None, 0, None, None) == _lib.GALOP_ERR_INVALID
    assert b"NULL" in lib.galop_last_error(None)
    assert lib.galop_fill_random_words_device(None, 5, 8, 0, None) == _lib.GALOP_ERR_INVALID
    assert lib.galop_ctx_phase_seconds(None, None, None, None, None, None) == _lib.GALOP_ERR_INVALID
    assert lib.galop_ctx_last_footprint(None, None, None) == _lib.GALOP_ERR_INVALID
    assert lib.galop_device_count() == 0  # no device visible in this container


def test_oracle_is_not_reachable_from_the_product():
    """The product package never imports, links or loads the oracle."""
    pkg = os.path.join(REPO, "galop_amd")
    for root, _, files in os.walk(pkg):
        for fn in files:
            if fn.endswith((".py", ".hip", ".h", ".cpp")):
                text = open(os.path.join(root, fn)).read()
                assert "oracle_py" not in text and "galop_oracle" not in text, fn
