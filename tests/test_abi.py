"""The C ABI: every declared symbol is exported, struct layouts agree between
include/barnacle_hip.h, the ctypes mirror and the oracle's independent copy."""
import ctypes as C
import os
import re

from barnacle_amd import abi, native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "barnacle_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bn_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = native.load()
    names = declared_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n


def test_struct_sizes_match_ctypes():
    lib = native.load()
    buf = (C.c_int32 * 32)()
    n = lib.bn_struct_sizes(buf, 32)
    assert n == len(abi.STRUCT_ORDER)
    assert list(buf)[:n] == [C.sizeof(t) for t in abi.STRUCT_ORDER]


def test_oracle_struct_sizes_match(orc):
    buf = (C.c_int32 * 32)()
    n = orc.lib.orc_struct_sizes(buf, 32)
    assert list(buf)[:n] == [C.sizeof(t) for t in abi.STRUCT_ORDER[:n]]


def test_error_code_and_message_without_gpu():
    # bn_bvh_build is host-only: invalid input returns BN_E_INVALID with a message.
    lib = native.load()
    nodes = (abi.BVHNode * 1)()
    nn = C.c_int32(1)
    perm = (C.c_int32 * 1)()
    rc = lib.bn_bvh_build(None, 0, nodes, C.byref(nn), perm)
    assert rc == abi.BN_E_INVALID
    assert b"bad arguments" in lib.bn_last_error()
