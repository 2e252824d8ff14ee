"""libmarvl_align loads without a GPU and exports every function the
headers in include/ declare."""

import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "marvel_amd", "libmarvl_align.so")


def declared_functions():
    names = set()
    for h in ("marvl_align.h", "marvl_filter.h"):
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", src, flags=re.M):
            if not m.group(0).startswith("typedef"):
                names.add(m.group(1))
    return names


def declared_variables():
    src = open(os.path.join(ROOT, "include", "marvl_filter.h")).read()
    return set(re.findall(r"^extern\s+[\w ]+?\s(\w+);", src, flags=re.M))


def test_every_declared_symbol_is_exported():
    lib = ctypes.CDLL(LIB)
    fns = declared_functions()
    assert {"Local_Alignment", "Sort_Kmers", "Match_Filter", "Write_Overlap_Buffer", "mi_la_batch_host"} <= fns
    missing = [n for n in sorted(fns | declared_variables()) if not hasattr(lib, n)]
    assert not missing, missing


def test_filter_globals_have_reference_defaults():
    lib = ctypes.CDLL(LIB)
    assert ctypes.c_int.in_dll(lib, "NTHREADS").value == 4
    assert ctypes.c_int.in_dll(lib, "NSHIFT").value == 2
    assert ctypes.c_int.in_dll(lib, "SYMMETRIC").value == 1


def test_set_filter_params_bounds():
    lib = ctypes.CDLL(LIB)
    assert lib.Set_Filter_Params(1, 6, 0, 35) == 1     # kmer <= 1 (dalign/filter.c:121-122)
    assert lib.Set_Filter_Params(32, 6, 0, 35) == 1    # 1 << 2k undefined from k = 32 (dalign/filter.c:129-131)
    assert lib.Set_Filter_Params(31, 6, 0, 35) == 0    # a block's keys may still not fit: Sort_Kmers says so
    assert lib.Set_Filter_Params(14, 6, 0, 35) == 0
