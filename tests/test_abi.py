"""The C ABI: every entry point include/bsam_gpu.h declares is exported by
the built library, and the ctypes mirrors have the C layouts.  No compute
call is made (no GPU needed)."""
import ctypes as C
import os
import re
import subprocess
import tempfile

from bsam_amd import abi, gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "bsam_gpu.h")


def declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int32_t|int64_t|const char\*)\s+(bsam_\w+)\s*\(", txt, re.M)))


def test_header_and_mirror_agree():
    assert declared() == sorted(abi.ENTRY_POINTS)


def test_library_exports_every_symbol():
    assert os.path.exists(gpu.LIB_PATH), "build the library first (__graft_entry__.build())"
    out = subprocess.run(["nm", "-D", "--defined-only", gpu.LIB_PATH], capture_output=True, text=True).stdout
    syms = set(l.split()[-1] for l in out.splitlines() if l.strip())
    missing = [s for s in declared() if s not in syms]
    assert not missing, missing


def test_struct_layouts():
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "bsam_gpu.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu\n", sizeof(bsam_patch_desc), offsetof(bsam_patch_desc, dx),
         offsetof(bsam_patch_desc, xlower), sizeof(bsam_params), offsetof(bsam_params, dt),
         offsetof(bsam_params, model_params));
  return 0;
}
'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        open(c, "w").write(src)
        exe = os.path.join(d, "t")
        subprocess.run(["gcc", "-I", os.path.dirname(HEADER), c, "-o", exe], check=True)
        vals = list(map(int, subprocess.run([exe], capture_output=True, text=True).stdout.split()))
    assert vals == [C.sizeof(abi.PatchDesc), abi.PatchDesc.dx.offset, abi.PatchDesc.xlower.offset,
                    C.sizeof(abi.Params), abi.Params.dt.offset, abi.Params.model_params.offset]


def test_create_without_gpu_fails_cleanly():
    """On a machine without a GPU, bsam_create returns BSAM_ENODEV (it never
    falls back to the CPU)."""
    import pytest
    try:
        import torch
        if torch.cuda.device_count() > 0:
            pytest.skip("a GPU is visible")
    except ImportError:
        pass
    lib = gpu.load()
    ctx = C.c_void_p()
    p = abi.Params()
    p.ndims, p.nrvars, p.mbc, p.model, p.nsmoothingpasses, p.maxvcycles, p.errortype = 2, 2, 1, 1, 2, 5, 1
    rc = lib.bsam_create(C.byref(p), C.byref(ctx))
    assert rc == abi.BSAM_ENODEV
    assert b"device" in lib.bsam_last_error()


def test_mixed_precision_model_check():
    """Mixed precision is built for the CH and tumour models; asking for it
    with the BioGrowth model is refused as an invalid argument before any
    device is touched."""
    lib = gpu.load()
    ctx = C.c_void_p()
    p = abi.Params()
    p.ndims, p.nrvars, p.mbc, p.model, p.nsmoothingpasses, p.maxvcycles, p.errortype = 2, 2, 1, abi.MODEL_BG, 2, 5, 1
    p.precision = abi.PREC_MIXED
    rc = lib.bsam_create(C.byref(p), C.byref(ctx))
    assert rc == abi.BSAM_EINVAL
    assert b"mixed precision" in lib.bsam_last_error()
