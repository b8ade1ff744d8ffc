"""pvhip_active_compact_noalloc (the compaction with the library's own
workspace, for callers that cannot allocate) against pvhip_active_compact
with a caller workspace: entries, counts and position starts bit-identical on
two calls in a row, the second with a smaller input; a request beyond the
workspace is refused before any launch."""
import ctypes as C

import numpy as np
import pytest
import torch

from openpv_amd import abi

pytestmark = pytest.mark.gpu


def _compact(loc, A, noalloc):
    H = abi.hip()
    n = loc.nbatch * loc.num_ext
    entries = torch.full((n, 2), -7, dtype=torch.int32, device="cuda")
    num = torch.full((loc.nbatch,), -1, dtype=torch.int64, device="cuda")
    pos = torch.full((loc.nbatch * (loc.nx_ext * loc.ny_ext + 1),), 0xFFFF, dtype=torch.int32, device="cuda")
    p = lambda t: C.c_void_p(t.data_ptr())
    if noalloc:
        st = H.pvhip_active_compact_noalloc(C.byref(loc), p(A), p(entries), p(num), p(pos), None)
        abi.check(st, "pvhip_active_compact_noalloc")
    else:
        ws = torch.empty(max(1, H.pvhip_active_compact_scratch_bytes(C.byref(loc)) // 4), dtype=torch.int32,
                         device="cuda")
        abi.check(H.pvhip_active_compact(C.byref(loc), p(A), p(entries), p(num), p(pos), p(ws), None),
                  "pvhip_active_compact")
    torch.cuda.synchronize()
    k = num.cpu().numpy()
    e = entries.cpu().numpy().reshape(loc.nbatch, -1, 2)
    return k, [e[b, :k[b]].copy() for b in range(loc.nbatch)], pos.cpu().numpy()


def _activity(loc, density, seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    a = torch.zeros(loc.nbatch, loc.ny_ext, loc.nx_ext, loc.nf, device="cuda")
    h = loc.halo
    v = torch.randn(loc.nbatch, loc.ny, loc.nx, loc.nf, device="cuda", generator=g)
    v = torch.where(torch.rand(v.shape, device="cuda", generator=g) < density, v, torch.zeros_like(v))
    a[:, h.up:h.up + loc.ny, h.lt:h.lt + loc.nx] = v
    return a.reshape(-1).contiguous()


def test_noalloc_compaction_equals_the_workspace_form_twice():
    # the LCA workload's V1 (64 x 32x32x512, halo 2: 2,624 chunk counts), then a
    # smaller layer on the same workspace
    for shape, density, seed in (((64, 32, 32, 512, 2), 0.03, 1), ((3, 16, 12, 24, 1), 0.2, 2)):
        loc = abi.Loc.make(*shape[:4], halo=shape[4])
        A = _activity(loc, density, seed)
        print("noalloc compaction", shape, flush=True)
        k0, e0, p0 = _compact(loc, A, False)
        k1, e1, p1 = _compact(loc, A, True)
        assert np.array_equal(k0, k1) and (k0 > 0).all()
        for b in range(loc.nbatch):
            assert np.array_equal(e0[b], e1[b])
        assert np.array_equal(p0, p1)


def test_noalloc_compaction_refuses_beyond_its_workspace():
    # 8 x 132x132x1024 extended: 8,712 chunk counts fit; 64 such images
    # (69,696 counts) do not
    H = abi.hip()
    ok = abi.Loc.make(8, 128, 128, 1024, halo=2)
    assert H.pvhip_active_compact_scratch_bytes(C.byref(ok)) <= 262144
    big = abi.Loc.make(64, 128, 128, 1024, halo=2)
    assert H.pvhip_active_compact_scratch_bytes(C.byref(big)) > 262144
    fake = C.c_void_p(4096)   # never dereferenced: the size check comes first
    st = H.pvhip_active_compact_noalloc(C.byref(big), fake, fake, fake, None, None)
    assert st == 1 and b"workspace" in H.pvhip_last_error()
