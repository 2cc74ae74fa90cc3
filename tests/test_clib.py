"""The C-ABI library loads and exports every entry point its headers declare.

No compute call is made here (CPU container); the GPU tests exercise them.
"""
import ctypes
import glob
import os
import re

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(REPO, "gsvit_copy_amd", "lib", "libgsvit_amd.so")
DECL = re.compile(r"^\s*(?:[A-Za-z_][\w\s\*]*?[\s\*])(sv_[a-z_0-9]+|gsv_[a-z_0-9]+|clear_settings|parse_settings|init_and_make_plan|is_output)\s*\(", re.M)

# host-side helpers declared in the headers that are not exported by the library
NOT_EXPORTED = set()


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(REPO, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in DECL.finditer(text):
            names.add(m.group(1))
    return sorted(names - NOT_EXPORTED)


def test_library_built():
    assert os.path.exists(LIB), "run gsvit_copy_amd/build.py"


def test_every_declared_symbol_is_exported():
    lib = ctypes.CDLL(LIB)
    names = declared_functions()
    assert "sv_pool_gpu_ystep_h" in names and "gsv_open" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_drop_in_entry_points_present():
    """the names GSvit's driver calls, src3d/pool.h:98-124,167"""
    lib = ctypes.CDLL(LIB)
    for n in ["sv_gpu_new", "sv_pool_gpu_init", "sv_pool_gpu_copyto", "sv_pool_gpu_copyfrom", "sv_pool_gpu_ystep_h",
              "sv_pool_gpu_ystep_e", "sv_pool_gpu_getsum", "sv_pool_gpu_boundary_copy", "sv_pool_gpu_boundary_hstep",
              "sv_pool_gpu_boundary_estep", "sv_pool_gpu_apply_source_hstep", "sv_pool_gpu_apply_source_estep",
              "sv_pool_gpu_apply_farfield", "sv_pool_gpu_apply_pfarfield", "sv_pool_gpu_getpoints",
              "sv_pool_gpu_getforce"]:
        assert hasattr(lib, n), n


def test_product_does_not_link_the_oracle():
    """the product library must not depend on the checker"""
    data = open(LIB, "rb").read()
    assert b"liboracle" not in data and b"oracle_step" not in data
