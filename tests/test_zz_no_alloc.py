"""rtps_sort_pairs_noalloc: the sort of rtps_sort_pairs with its scratch in a
fixed device array instead of a caller workspace.  Two calls in a row (the
second smaller, so stale contents of the first are in the array) give keys
and values bit-identical to rtps_sort_pairs on the same inputs; sizes above
the maximum are refused."""
import numpy as np
import pytest

from enjaparticles_amd import _native as N

pytestmark = pytest.mark.gpu


def _sorted_both(n, bits, seed):
    import torch
    L = N.lib()
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(seed)
    k = rng.integers(0, 1 << bits, size=n, dtype=np.uint64).astype(np.uint32)
    v = rng.permutation(n).astype(np.uint32)
    ka = torch.from_numpy(k.view(np.int32)).to(dev)
    va = torch.from_numpy(v.view(np.int32)).to(dev)
    kb, vb = ka.clone(), va.clone()
    st = torch.cuda.current_stream().cuda_stream
    tb = L.rtps_sort_temp_bytes(n, bits)
    tmp = torch.empty(tb, dtype=torch.uint8, device=dev)
    N.check(L.rtps_sort_pairs(n, bits, ka.data_ptr(), va.data_ptr(), tmp.data_ptr(), tb, st))
    N.check(L.rtps_sort_pairs_noalloc(n, bits, kb.data_ptr(), vb.data_ptr(), st))
    torch.cuda.synchronize()
    a = (ka.cpu().numpy().view(np.uint32), va.cpu().numpy().view(np.uint32))
    b = (kb.cpu().numpy().view(np.uint32), vb.cpu().numpy().view(np.uint32))
    return k, v, a, b


def test_noalloc_sort_equals_workspace_sort_twice(cuda):
    for n, bits, seed in ((1 << 20, 24, 1), (70001, 13, 2)):
        print(f"noalloc sort: n={n} key_bits={bits}", flush=True)
        k, v, a, b = _sorted_both(n, bits, seed)
        np.testing.assert_array_equal(b[0], a[0])
        np.testing.assert_array_equal(b[1], a[1])
        # and it is the stable ascending sort of the inputs
        o = np.argsort(k, kind="stable")
        np.testing.assert_array_equal(b[0], k[o])
        np.testing.assert_array_equal(b[1], v[o])


def test_noalloc_sort_refuses_oversize(cuda):
    L = N.lib()
    assert L.rtps_sort_pairs_noalloc((1 << 20) + 1, 20, 1, 1, None) == N.RTPS_ERR_ARG
    assert b"maximum" in L.rtps_last_error()
    assert L.rtps_sort_pairs_noalloc(10, 33, 1, 1, None) == N.RTPS_ERR_ARG
