"""CPU tests of the C-ABI library: it loads, exports every function that
include/cudabvh.h declares (and every symbol of the reference boundary
layer/cuda_bvh_backend.h:12-141), and its host-only entry points behave as
the reference's. No GPU work is issued."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cudabvh.h")

# The 21 functions of the reference boundary (layer/cuda_bvh_backend.h:21-141).
REFERENCE_SYMBOLS = [
    "cudaBVH_build", "cudaBVH_destroy", "cudaBVH_tracePrimary", "cudaBVH_traceDiffuse", "cudaBVH_traceToRGBA",
    "cudaBVH_traceToGPUPtr", "cudaBVH_traceToGPUAsync", "cudaBVH_traceToHostPtr", "cudaBVH_importBufferFd",
    "cudaBVH_getBounds", "cudaBVH_getNumTris", "cudaBVH_getNumBVH4Nodes", "cudaBVH_getNumCWBVHNodes",
    "cudaBVH_getNodeData", "cudaBVH_getTriData", "cudaBVH_getStacklessBVH2", "cudaBVH_getPackedTris",
    "cudaBVH_resetDevice", "cudaBVH_buildFromAABBs", "cudaBVH_buildTLASFast", "cudaBVH_readGPUMemoryFd",
]


def header_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(cudaBVH_\w+)\s*\(", txt)))


@pytest.fixture(scope="module")
def lib(pkg):
    return pkg.load_library()


def test_exports_every_declared_symbol(lib, pkg):
    declared = header_functions()
    assert set(REFERENCE_SYMBOLS) <= set(declared)
    for name in declared:
        assert hasattr(lib, name), name
        assert name in pkg.SIGNATURES, f"{name} not bound in the Python mirror"
    assert set(pkg.SIGNATURES) == set(declared)


def test_header_compiles_as_c_and_cpp(tmp_path):
    import subprocess
    for comp, ext in (("gcc", "c"), ("g++", "cpp")):
        src = tmp_path / f"t.{ext}"
        src.write_text('#include "cudabvh.h"\nint main(void){ CudaTri t; (void)t; return 0; }\n')
        subprocess.run([comp, "-I", os.path.join(ROOT, "include"), "-c", str(src), "-o", str(tmp_path / "t.o")],
                       check=True)


def test_null_handle_conventions(lib, pkg):
    assert lib.cudaBVH_getNumTris(None) == 0
    assert lib.cudaBVH_getNumBVH4Nodes(None) == 0
    assert lib.cudaBVH_getNumCWBVHNodes(None) == 0
    assert lib.cudaBVH_getNodeData(None) is None
    lib.cudaBVH_destroy(None)                              # accepts NULL (backend :745)
    out = C.c_void_p(1)
    assert lib.cudaBVH_getStacklessBVH2(None, C.byref(out)) == 0 and not out.value
    assert lib.cudaBVH_getPackedTris(None, C.byref(out)) == 0 and not out.value
    ptrs = (C.c_void_p * 9)(*([1] * 9))
    lib.cudaBVH_getTriData(None, ptrs)
    assert all(p is None for p in ptrs)
    assert lib.cudaBVH_tracePrimary(None, 16, 0, 0, 0, None) == 0.0
    assert "NULL handle" in pkg.last_error()
    assert lib.cudaBVH_build(None, 10) is None
    assert lib.cudaBVH_buildFromAABBs(None, 0) is None
    assert lib.cudaBVH_buildTLASFast(None, 0, None, None) == 0


def test_fd_calls_reject_bad_arguments_on_the_host(lib, pkg):
    """importBufferFd / readGPUMemoryFd refuse a negative fd, an empty
    import and a range outside the allocation before any HIP call."""
    assert lib.cudaBVH_importBufferFd(-1, 4096) is None
    assert "fd -1" in pkg.last_error()
    assert lib.cudaBVH_importBufferFd(3, 0) is None
    host = np.zeros(4, np.float32)
    assert lib.cudaBVH_readGPUMemoryFd(-1, 4096, 0, 16, host.ctypes.data) == -1
    assert "outside" in pkg.last_error()
    assert lib.cudaBVH_readGPUMemoryFd(3, 4096, 4090, 16, host.ctypes.data) == -1
    assert lib.cudaBVH_readGPUMemoryFd(3, 4096, 8192, 0, host.ctypes.data) == -1
    assert lib.cudaBVH_readGPUMemoryFd(3, 4096, 0, 16, None) == -1
