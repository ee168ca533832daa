"""sph_radix_sort_pairs_noalloc -- the radix sort whose temporary buffer is the
library's static device array (no allocation) -- against sph_radix_sort_pairs
with a caller-allocated buffer: bit-identical sorted keys and values on two
calls in a row (the second with fewer pairs, so it runs on the first call's
leftover array contents), at the maximum size, and equal to a stable host
sort.  Sizes beyond the maximum are rejected before any device work."""
import numpy as np
import pytest

from sph2s_final_amd import _capi as C


def _dev(a):
    import torch
    return torch.from_numpy(a.view(np.int32)).cuda()


def _sort(keys, vals, bits, noalloc):
    import torch
    n = keys.size
    k_in = _dev(keys)
    v_in = None if vals is None else _dev(vals)
    k_out = torch.empty(n, dtype=torch.int32, device="cuda")
    v_out = torch.empty(n, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    vp = None if v_in is None else v_in.data_ptr()
    if noalloc:
        C.check(C.lib().sph_radix_sort_pairs_noalloc(k_in.data_ptr(), vp, k_out.data_ptr(),
                                                     v_out.data_ptr(), n, bits, None))
    else:
        tb = int(C.lib().sph_radix_sort_temp_bytes(n))
        tmp = torch.empty(tb, dtype=torch.uint8, device="cuda")
        C.check(C.lib().sph_radix_sort_pairs(k_in.data_ptr(), vp, k_out.data_ptr(),
                                             v_out.data_ptr(), n, bits, tmp.data_ptr(), tb, None))
    torch.cuda.synchronize()
    return k_out.cpu().numpy().view(np.uint32), v_out.cpu().numpy().view(np.uint32)


def _case(n, bits, seed, with_values):
    rng = np.random.default_rng(seed)
    keys = rng.integers(0, 1 << bits, n, dtype=np.uint64).astype(np.uint32)
    vals = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32) if with_values else None
    return keys, vals


def _check(keys, vals, bits):
    print(f"case noalloc sort n={keys.size} bits={bits} values={vals is not None}", flush=True)
    ks, vs = _sort(keys, vals, bits, noalloc=True)
    ka, va = _sort(keys, vals, bits, noalloc=False)
    assert np.array_equal(ks, ka) and np.array_equal(vs, va)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(ks, keys[order])
    assert np.array_equal(vs, order.astype(np.uint32) if vals is None else vals[order])


@pytest.mark.gpu
def test_noalloc_sort_matches_the_allocating_sort_on_two_calls():
    # The second call is smaller and carries explicit values.
    _check(*_case(1_500_000, 19, 1, False), 19)
    _check(*_case(70_001, 24, 2, True), 24)


@pytest.mark.gpu
def test_noalloc_sort_at_its_maximum_size():
    _check(*_case(1 << 21, 21, 3, False), 21)


def test_noalloc_sort_rejects_sizes_beyond_its_maximum():
    rc = C.lib().sph_radix_sort_pairs_noalloc(None, None, None, None, (1 << 21) + 1, 19, None)
    assert rc == C.SPH_ERR_CAPACITY
    assert b"maximum" in C.lib().sph_last_error()
    assert C.lib().sph_radix_sort_pairs_noalloc(None, None, None, None, 0, 19, None) == C.SPH_OK
