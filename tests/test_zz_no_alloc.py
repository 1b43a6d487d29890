"""arw_step_noalloc (the step with its scratch in a library-held device
array, for callers that cannot allocate) against arw_step, both through
the C-ABI: bit-identical on every state field, on two domains in a row,
the second smaller than the first (the shared array keeps the first
call's contents), in FP32 and FP64; a domain larger than the array is
refused before any device work."""
import ctypes

import pytest
import torch

from arwen_amd import _lib, cases
from arwen_amd.dycore import _stream, params
from arwen_amd.state import DomainState

pytestmark = pytest.mark.gpu

FIELDS = ("u", "v", "w", "thp", "php", "mup", "p", "al", "alt")


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_noalloc_step_matches_step(dtype):
    lib = _lib.load()
    pre = "arw64_" if dtype == torch.float64 else "arw_"
    step, noalloc = getattr(lib, pre + "step"), getattr(lib,
                                                        pre + "step_noalloc")
    for nx, ny, nz, seed in ((96, 80, 40, 5), (40, 24, 16, 6)):
        print(f"step_noalloc {dtype} {nx}x{ny}x{nz}", flush=True)
        case = cases.ensemble_member(nx, ny, nz, seed)
        a = DomainState.from_host(case, dtype=dtype)
        b = DomainState.from_host(case, dtype=dtype)
        p = params(case.cfg)
        da, db = a.domain_struct(), b.domain_struct()
        db.ww = db.gam = db.dmdt_rf = db.mup_next = None
        for _ in range(2):
            _lib.check(step(da, p, _stream()), pre + "step")
            _lib.check(noalloc(db, p, _stream()), pre + "step_noalloc")
        torch.cuda.synchronize()
        for n in FIELDS:
            assert torch.equal(getattr(a, n), getattr(b, n)), (nx, n)


def test_noalloc_refuses_a_domain_beyond_its_array():
    lib = _lib.load()
    d = _lib.ArwDomain()
    d.nz, d.ny, d.nx = 50, 1000, 1000
    p = params(cases.ensemble_member(8, 8, 8, 1).cfg)
    rc = lib.arw_step_noalloc(d, p, _stream())
    assert rc == -1
    assert b"bytes of step scratch" in lib.arw_last_error()
