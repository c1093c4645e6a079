"""The C ABI: the shared library loads without a GPU and exports every
function and object that include/*.h declares."""
import ctypes
import glob
import os
import re

import decomp_mpi_amd as dm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols(text):
    """Function and extern object names declared at file scope of a C header."""
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    text = re.sub(r"//[^\n]*", " ", text)
    text = re.sub(r"^\s*#[^\n]*", " ", text, flags=re.M)
    text = text.replace('extern "C" {', " ")
    while True:  # drop struct / enum bodies
        t = re.sub(r"\{[^{}]*\}", " ", text)
        if t == text:
            break
        text = t
    text = text.replace("}", " ")
    funcs, objs = set(), set()
    for stmt in text.split(";"):
        s = " ".join(stmt.split())
        if not s or s.startswith("typedef") or s.startswith("enum"):
            continue
        if s.startswith("extern ") and "(" not in s:
            for part in s[len("extern "):].split(","):
                objs.add(re.findall(r"[A-Za-z_]\w*", part)[-1])
            continue
        m = re.match(r"^[\w\s\*]*?\b([A-Za-z_]\w*)\s*\(", s)
        if m and "(*" not in s.split("(")[0] + "(" + s.split("(")[1][:1]:
            funcs.add(m.group(1))
    return funcs, objs


def all_declared():
    funcs, objs = set(), set()
    for h in sorted(glob.glob(os.path.join(ROOT, "include", "*.h"))):
        f, o = declared_symbols(open(h).read())
        funcs |= f
        objs |= o
    return funcs, objs


def test_parser_sees_the_api():
    funcs, objs = all_declared()
    for name in ("makeReshapePlan", "runReshapePlan", "runSplitPlan", "makeBlockDecomp", "getTransformed",
                 "makeProcessorGrid", "dm_malloc", "buildParallelPlanf", "timer", "logger"):
        assert name in funcs
    assert "dm_type_double" in objs
    assert "rank_offset" not in funcs  # struct fields are not symbols


def test_every_declared_symbol_is_exported():
    lib = ctypes.CDLL(dm.LIB_PATH)
    funcs, objs = all_declared()
    missing = sorted(n for n in funcs | objs if not hasattr(lib, n))
    assert not missing, f"declared in include/*.h but not exported: {missing}"


def test_loads_without_device():
    C = dm.lib()
    assert C.dm_device_available() in (0, 1)
    comm = C.dm_world_create(3)
    n = ctypes.c_int()
    C.dm_comm_size(comm, ctypes.byref(n))
    assert n.value == 3
    C.dm_world_free(comm)
