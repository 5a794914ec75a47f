"""The drop-in boundary: built libraries load and export every C-ABI symbol
declared in include/*.h (no device calls).  CPU only."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "bohrium_amd", "lib")
INCLUDE = os.path.join(ROOT, "include")

HEADER_LIBS = {
    "bhhip.h": "libbh_ve_hip.so",
    "bhc.h": "libbhc.so",
    "bh_ve_hip.h": "libbh_ve_hip.so",
    "bh_costmodel.h": "libbh_ve_hip.so",
    "bh_mem_signal.h": "libbh.so",
    "bh_blas_hip.h": "libbh_blas_hip.so",
}


def declared_functions(header):
    text = open(os.path.join(INCLUDE, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"^typedef[^;]*;", "", text, flags=re.M)
    names = re.findall(r"^[A-Za-z_][\w \*]*?\b(\w+)\s*\([^;{]*\)\s*;", text, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "while")))


def exported(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", os.path.join(LIB, lib)], stdout=subprocess.PIPE,
                         text=True, check=True).stdout
    return set(line.split()[-1] for line in out.splitlines() if line.strip())


@pytest.mark.parametrize("header", sorted(HEADER_LIBS))
def test_every_declared_symbol_is_exported(header):
    names = declared_functions(header)
    assert names, header
    syms = exported(HEADER_LIBS[header])
    missing = [n for n in names if n not in syms]
    assert not missing, "%s declares symbols %s missing from %s" % (header, missing, HEADER_LIBS[header])


@pytest.mark.parametrize("lib", ["libbh.so", "libbhc.so", "libbh_vem_node.so", "libbh_ve_hip.so",
                                 "libbh_blas_hip.so"])
def test_libraries_load(lib):
    ctypes.CDLL(os.path.join(LIB, lib), mode=ctypes.RTLD_GLOBAL)


def test_engine_component_symbols():
    """The parent resolves exactly `create` and `destroy` (reference core/bh_component.cpp:30-61)."""
    syms = exported("libbh_ve_hip.so")
    assert "create" in syms and "destroy" in syms
    node = exported("libbh_vem_node.so")
    assert "create" in node and "destroy" in node


def test_opcode_ids_match_reference_table():
    """A few fixed wire ids of the reference opcode table (core/codegen/opcodes.json)."""
    from bohrium_amd import opcodes as o
    assert (o.ADD, o.IDENTITY, o.FREE, o.ADD_REDUCE, o.RANDOM, o.ADD_ACCUMULATE, o.GATHER, o.CONJ) == \
        (1, 53, 55, 59, 71, 75, 79, 84)
