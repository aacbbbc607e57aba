"""cudaBVH_importBufferFd and cudaBVH_readGPUMemoryFd on device memory handed
over as a file descriptor (layer/cuda_bvh_backend.cu:1337-1370 and
:1509-1545: the Vulkan layer passes the opaque fd of a VkDeviceMemory). On
this driver an exportable fd of device memory is a dma-buf; HSA's
hsa_amd_portable_export_dmabuf makes one from a torch allocation in the
same process, which stands in for the Vulkan driver's export.

* vertex and index buffers imported through their fds build exactly the
  tree (node bytes, exports, hits) that cudaBVH_build makes from the same
  triangles as a host soup;
* readGPUMemoryFd returns the bytes at an offset of the allocation;

(tests/test_abi.py checks the host-side rejection of bad fds and ranges.)

Each call gets a dup of the fd, and the dups are left open: whether the
import takes ownership of its fd is the driver's business (CUDA's does),
so closing them here could close an unrelated descriptor."""
import ctypes as C
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _hsa():
    hsa = C.CDLL("libhsa-runtime64.so.1")          # the runtime HIP already loaded
    hsa.hsa_amd_portable_export_dmabuf.argtypes = [C.c_void_p, C.c_size_t, C.POINTER(C.c_int), C.POINTER(C.c_uint64)]
    hsa.hsa_amd_portable_export_dmabuf.restype = C.c_int
    hsa.hsa_amd_portable_close_dmabuf.argtypes = [C.c_int]
    hsa.hsa_amd_portable_close_dmabuf.restype = C.c_int
    return hsa


def _export(hsa, t):
    """(fd, offset of t in its allocation, allocation bytes) of a device tensor."""
    fd, off = C.c_int(-1), C.c_uint64(0)
    st = hsa.hsa_amd_portable_export_dmabuf(C.c_void_p(t.data_ptr()), t.numel() * t.element_size(), C.byref(fd),
                                            C.byref(off))
    if st != 0:
        pytest.skip(f"this driver exports no dma-buf for device memory (HSA status {st:#x})")
    size = os.lseek(fd.value, 0, os.SEEK_END)
    os.lseek(fd.value, 0, os.SEEK_SET)
    return fd.value, off.value, size


def _grid(n):
    xs = np.linspace(-1.0, 1.0, n + 1, dtype=np.float32)
    X, Z = np.meshgrid(xs, xs, indexing="ij")
    Y = (0.25 * np.sin(3 * X) * np.cos(2 * Z)).astype(np.float32)
    verts = np.stack([X, Y, Z], axis=-1).reshape(-1, 3).astype(np.float32)
    i = np.arange(n)[:, None] * (n + 1) + np.arange(n)[None, :]
    a, b, c, d = i, i + 1, i + (n + 1), i + (n + 2)
    idx = np.concatenate([np.stack([a, b, c], -1).reshape(-1, 3), np.stack([b, d, c], -1).reshape(-1, 3)])
    return verts, idx.astype(np.int32), verts[idx].reshape(-1, 9)


def test_import_fd_build_and_read(pkg, dev):
    import torch
    print("importBufferFd / readGPUMemoryFd: 40x40 grid mesh", flush=True)
    lib = pkg.load_library()
    hsa = _hsa()
    verts, idx, soup = _grid(40)
    V = torch.from_numpy(verts.reshape(-1)).to(dev)
    I = torch.from_numpy(idx.reshape(-1)).to(dev)
    torch.cuda.synchronize()
    vfd, voff, vsize = _export(hsa, V)
    ifd, ioff, isize = _export(hsa, I)
    try:
        vbytes, ibytes = V.numel() * 4, I.numel() * 4
        assert voff + vbytes <= vsize and ioff + ibytes <= isize
        pv = lib.cudaBVH_importBufferFd(os.dup(vfd), voff + vbytes)
        pi = lib.cudaBVH_importBufferFd(os.dup(ifd), ioff + ibytes)
        assert pv and pi, pkg.last_error()
        h = lib.cudaBVH_buildDevice(C.c_void_p(pv + voff), verts.shape[0], 12, C.c_void_p(pi + ioff), 4, idx.shape[0])
        assert h, pkg.last_error()
        a = pkg.BVH(h, idx.shape[0])
        with a, pkg.BVH.build(soup) as b:
            assert a.num_bvh4_nodes == b.num_bvh4_nodes and a.depth == b.depth
            assert np.array_equal(a.node_data(), b.node_data())
            assert np.array_equal(a.tri_data(), b.tri_data())
            assert np.array_equal(a.packed_tris(), b.packed_tris())
        # bytes at an offset inside the allocation: vertices 100 .. 399
        host = np.zeros(900, np.float32)
        rc = lib.cudaBVH_readGPUMemoryFd(os.dup(vfd), vsize, voff + 100 * 12, host.nbytes, host.ctypes.data)
        assert rc == 0, pkg.last_error()
        assert np.array_equal(host, verts[100:400].reshape(-1))
        hidx = np.zeros(idx.size, np.int32)
        assert lib.cudaBVH_readGPUMemoryFd(os.dup(ifd), isize, ioff, hidx.nbytes, hidx.ctypes.data) == 0
        assert np.array_equal(hidx, idx.reshape(-1))
    finally:
        hsa.hsa_amd_portable_close_dmabuf(vfd)
        hsa.hsa_amd_portable_close_dmabuf(ifd)

