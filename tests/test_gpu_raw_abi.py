"""The C ABI with device memory from the HIP runtime itself, no PyTorch.

A caller in another host language binds buffers it allocated with
hipMalloc and lets the library run on the default stream.  This test does
exactly that through ctypes (hipMalloc / hipMemcpy / hipFree of
libamdhip64): it creates a context, binds one hipMalloc'd buffer per
registry field (uploaded from the seeded case; fields the case leaves out
start at zero, as a Domain's do), steps twice with wrfdc_step, copies the
prognostic fields back with hipMemcpy and compares them bit for bit with
the same two steps through ``wrfdc.device.Domain`` (torch-allocated buffers
on torch's stream).  The checksums must agree as well.
"""

import ctypes
import ctypes.util
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def _hip():
    for cand in (ctypes.util.find_library("amdhip64"), "/opt/rocm/lib/libamdhip64.so", "libamdhip64.so"):
        if cand:
            try:
                lib = ctypes.CDLL(cand)
                break
            except OSError:
                continue
    else:
        pytest.skip("libamdhip64 not found")
    lib.hipMalloc.restype = ctypes.c_int
    lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    lib.hipFree.restype = ctypes.c_int
    lib.hipFree.argtypes = [ctypes.c_void_p]
    lib.hipMemcpy.restype = ctypes.c_int
    lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    lib.hipMemset.restype = ctypes.c_int
    lib.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    lib.hipDeviceSynchronize.restype = ctypes.c_int
    return lib


def test_raw_hip_buffers_match_domain(gpu):
    from wrfdc import _lib
    from wrfdc.cases import make_case
    from wrfdc.device import NCHECKSUMS, Domain
    from wrfdc.fields import FIELD_SPECS, Dims

    print("[raw abi] 40x33x20, 8 species, 2 steps: hipMalloc'd buffers vs Domain", flush=True)
    case = make_case(40, 33, 20, nmoist=6, nscalar=2, hydrometeors=True)
    cfg = case.config
    # reference: the torch-backed Domain
    dom = Domain(cfg, gpu)
    dom.upload(case.arrays)
    dom.step(2, 0.0)
    names = ["u_2", "v_2", "w_2", "t_2", "ph_2", "mu_2", "p", "al", "moist", "scalar"]
    want = dom.download(names)
    want_ck = dom.checksums()
    dom.close()

    hip = _hip()
    lib = _lib.load()
    dims = Dims(nx=cfg.nx, ny=cfg.ny, nz=cfg.nz, nmoist=cfg.nmoist, nscalar=cfg.nscalar, spec_bdy_width=cfg.spec_bdy_width)
    st = cfg.as_struct()
    ctx = ctypes.c_void_p()
    _lib.raise_for(lib.wrfdc_create(ctypes.byref(st), ctypes.byref(ctx)), lib, None, "wrfdc_create")
    ptrs = {}
    try:
        for name, kind in FIELD_SPECS:
            shape = dims.shape(kind)
            nbytes = 4 * math.prod(shape)
            if nbytes == 0:
                _lib.raise_for(lib.wrfdc_bind(ctx, name.encode(), None), lib, ctx, "wrfdc_bind")
                continue
            p = ctypes.c_void_p()
            assert hip.hipMalloc(ctypes.byref(p), nbytes) == 0, name
            ptrs[name] = (p, shape, nbytes)
            assert hip.hipMemset(p, 0, nbytes) == 0, name
            if name in case.arrays:
                src = np.ascontiguousarray(np.asarray(case.arrays[name], dtype=np.float32))
                assert tuple(src.shape) == tuple(shape), name
                assert hip.hipMemcpy(p, src.ctypes.data, nbytes, H2D) == 0, name
            _lib.raise_for(lib.wrfdc_bind(ctx, name.encode(), p), lib, ctx, f"wrfdc_bind({name})")
        _lib.raise_for(lib.wrfdc_step(ctx, 2, 0.0), lib, ctx, "wrfdc_step")
        assert hip.hipDeviceSynchronize() == 0
        out = (ctypes.c_double * NCHECKSUMS)()
        _lib.raise_for(lib.wrfdc_checksums(ctx, out), lib, ctx, "wrfdc_checksums")
        got_ck = np.array(list(out))
        got = {}
        for name in names:
            p, shape, nbytes = ptrs[name]
            dst = np.empty(shape, dtype=np.float32)
            assert hip.hipMemcpy(dst.ctypes.data, p, nbytes, D2H) == 0, name
            got[name] = dst
    finally:
        lib.wrfdc_destroy(ctx)
        for p, _, _ in ptrs.values():
            hip.hipFree(p)
    diff = {n: int(np.count_nonzero(got[n] != want[n])) for n in names}
    print(f"[raw abi] differing elements: {diff}", flush=True)
    assert all(v == 0 for v in diff.values()), diff
    assert np.array_equal(got_ck, want_ck), (got_ck, want_ck)
