"""wrfdc_checksums_noalloc (no context, no allocation, no host wait) against
wrfdc_checksums on the same device state.

Both run the same two checksum kernels with the same number of blocks, so
the FP64 sums must agree to the last bit.  Two calls in a row, the second on
a smaller domain, check that the shared partial-sum array is fully
rewritten by each call; the rejections (a missing field, a configuration
wrfdc_create refuses) report through wrfdc_last_error(NULL).
"""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pointer_table(dom):
    from wrfdc.fields import FIELD_SPECS

    ptrs = (ctypes.c_void_p * len(FIELD_SPECS))()
    for i, (name, _) in enumerate(FIELD_SPECS):
        t = dom.buffers[name]
        ptrs[i] = t.data_ptr() if t.numel() else None
    return ptrs


def _noalloc(dom, ptrs, out_dev):
    from wrfdc import _lib

    rc = dom.lib.wrfdc_checksums_noalloc(ctypes.byref(dom._cfg), ptrs, ctypes.c_void_p(out_dev.data_ptr()),
                                         ctypes.c_void_p(dom.stream.cuda_stream))
    _lib.raise_for(rc, dom.lib, None, "wrfdc_checksums_noalloc")
    dom.stream.synchronize()
    return out_dev.cpu().numpy().copy()


def test_noalloc_checksums_match_context_checksums(gpu):
    import torch

    from wrfdc.cases import make_case
    from wrfdc.device import NCHECKSUMS, Domain

    for shape in ((40, 33, 20), (20, 16, 8)):
        print(f"[no-alloc checksums] shape {shape}", flush=True)
        case = make_case(*shape, nmoist=6, nscalar=2, hydrometeors=True, hill_height=500.0)
        dom = Domain(case.config, gpu)
        dom.upload(case.arrays)
        dom.step(1, 0.0)
        ref = dom.checksums()
        out = torch.full((NCHECKSUMS,), float("nan"), dtype=torch.float64, device=gpu)
        got = _noalloc(dom, _pointer_table(dom), out)
        dom.close()
        assert got.shape == ref.shape
        assert np.array_equal(got, ref), (shape, got, ref)


def test_noalloc_checksums_reject_bad_input(gpu):
    import torch

    from wrfdc.cases import make_case
    from wrfdc.device import NCHECKSUMS, Domain
    from wrfdc.fields import FIELD_SPECS

    case = make_case(20, 16, 8, nmoist=1, nscalar=0)
    dom = Domain(case.config, gpu)
    dom.upload(case.arrays)
    out = torch.zeros((NCHECKSUMS,), dtype=torch.float64, device=gpu)
    ptrs = _pointer_table(dom)
    ptrs[[n for n, _ in FIELD_SPECS].index("phb")] = None
    with pytest.raises(KeyError, match="phb"):
        _noalloc(dom, ptrs, out)
    bad = case.config.as_struct()
    bad.nx = 8
    rc = dom.lib.wrfdc_checksums_noalloc(ctypes.byref(bad), _pointer_table(dom), ctypes.c_void_p(out.data_ptr()),
                                         ctypes.c_void_p(dom.stream.cuda_stream))
    assert rc == -1 and b"domain too small" in dom.lib.wrfdc_last_error(None)
    dom.close()
