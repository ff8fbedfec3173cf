"""CPU checks of the C ABI library and the host mirror (no GPU calls)."""
import ctypes
import os

import pytest

import maglua_amd as M
from maglua_amd import workloads as W


def test_library_exports_every_declared_symbol():
    L = M.lib()
    syms = M.declared_symbols()
    assert len(syms) > 60
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing


def test_every_declared_symbol_has_a_ctypes_signature():
    """A call without argtypes would pass 64-bit handles as C ints."""
    from maglua_amd import _lib
    unsigned = [s for s in M.declared_symbols() if s not in _lib._SIGS]
    assert not unsigned


def test_library_is_gfx950_code_object():
    data = open(M._lib.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_last_error_is_empty_string_initially():
    assert M.lib().msd_last_error() in (b"",)


def test_null_handles_rejected_without_gpu():
    L = M.lib()
    assert L.msd_system_destroy(None) == 0
    assert L.msd_exchange_num_paths(None, ctypes.byref(ctypes.c_int())) == -1
    assert b"NULL" in L.msd_last_error()


def test_lua_style_argument_errors():
    with pytest.raises(M.MagLuaError):
        M.LLG.new("Quaternion")
    with pytest.raises(M.MagLuaError):
        M.maglua._index((5, 1, 1), (4, 4, 1))
    assert M.maglua._index((4, 4), (4, 4, 1)) == 15          # 1-based {x,y}


def test_nn_paths_counts():
    f, t = W.nn_paths(4, 3, 2, (1, 1, 0))
    assert len(f) == 4 * 3 * 2 * 4 + 4 * 3 * 2          # 4 in-plane + 1 z neighbour each
    f1, _ = W.nn_paths(32, 32, 1, (1, 1, 0))
    assert len(f1) == 4 * 1024
    assert (f[:5] == 0).all()                           # site-major order


def test_unseeded_random_draws_distinct_nonzero_seeds():
    """Random.new() and Random.new(0) take a seed from /dev/urandom, as RNG::luaInit does
    (common/random/random.cpp:29-62); an explicit seed is kept."""
    from maglua_amd.maglua import Random
    seeds = {Random().seed, Random(0).seed, Random.new().seed, Random.new(0).seed}
    assert len(seeds) == 4 and 0 not in seeds
    assert Random(1234).seed == 1234
