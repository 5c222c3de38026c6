"""The native libraries load and export every entry point their headers
declare (no device calls)."""
import ctypes
import os
import re

import care_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(care_[a-z_0-9]+)\s*\(", text)))


def test_engine_exports_all_declared_symbols():
    lib = ctypes.CDLL(care_amd.ENGINE_LIB)
    for name in _declared("care_amd.h"):
        assert hasattr(lib, name), name


def test_host_exports_all_declared_symbols():
    care_amd.engine()
    lib = ctypes.CDLL(care_amd.HOST_LIB)
    for name in _declared("care_host.h"):
        assert hasattr(lib, name), name


def test_bindings_cover_headers():
    assert set(_declared("care_amd.h")) <= set(care_amd.ENGINE_SYMBOLS)
    assert set(_declared("care_host.h")) <= set(care_amd.HOST_SYMBOLS)


def test_options_defaults():
    opt = care_amd.default_options()
    assert opt.kmer_length == 20 and opt.num_hash_maps == 48
    assert abs(opt.error_factor - 0.06) < 1e-7 and abs(opt.coverage_factor - 0.6) < 1e-7
    assert opt.min_overlap == 30 and abs(opt.max_mismatch_ratio - 0.2) < 1e-7
