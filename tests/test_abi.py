"""The C-ABI library loads and exports every symbol include/pokerrust_hip.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os

import pytest

from pokerrust_amd import _lib

DROPIN = ["init_builder", "upload_c", "download_c", "build_river_cuda", "transfer_flop_eval_cuda",
          "free_eval_cuda", "evaluate_cuda"]


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(_lib.LIB_PATH):
        from pokerrust_amd import build
        build.build()
    return ctypes.CDLL(_lib.LIB_PATH)


def test_header_declares_the_seven_dropin_symbols():
    syms = _lib.header_symbols()
    for s in DROPIN:
        assert s in syms


def test_library_exports_every_header_symbol(lib):
    missing = [s for s in _lib.header_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_loader_fails_loudly_without_library(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "missing.so"))
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        _lib.load()
