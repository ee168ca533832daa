"""GPU radix sort (sph_radix_sort_pairs) vs a stable host sort: bit-exact keys and
permutations, including empty and ragged sizes, all-equal keys, maximum keys,
nearly-sorted keys and every key width."""
import numpy as np
import pytest
import torch

from sph2s_final_amd import _capi as C

pytestmark = pytest.mark.gpu


def _gpu_sort(keys: np.ndarray, bits: int, vals: np.ndarray | None = None):
    n = keys.size
    dev = torch.device("cuda")
    k_in = torch.from_numpy(keys.view(np.int32)).to(dev)
    v_in = None if vals is None else torch.from_numpy(vals.view(np.int32)).to(dev)
    k_out = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
    v_out = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
    tb = int(C.lib().sph_radix_sort_temp_bytes(n))
    tmp = torch.empty(max(tb, 1), dtype=torch.uint8, device=dev)
    torch.cuda.synchronize()
    C.check(C.lib().sph_radix_sort_pairs(k_in.data_ptr(), None if v_in is None else v_in.data_ptr(),
                                         k_out.data_ptr(), v_out.data_ptr(), n, bits,
                                         tmp.data_ptr(), tb, None))
    torch.cuda.synchronize()
    return (k_out[:n].cpu().numpy().view(np.uint32), v_out[:n].cpu().numpy().view(np.uint32))


def _check(keys, bits, vals=None):
    print(f"case n={keys.size} bits={bits}", flush=True)
    ks, vs = _gpu_sort(keys, bits, vals)
    masked = keys & np.uint32((1 << bits) - 1) if bits < 32 else keys
    order = np.argsort(masked, kind="stable")
    assert np.array_equal(ks, keys[order])
    expect_v = order.astype(np.uint32) if vals is None else vals[order]
    assert np.array_equal(vs, expect_v)


@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 1000, 4095, 4096, 4097, 12289, 100_000,
                               1_000_003])
def test_random_keys_sizes(n):
    rng = np.random.default_rng(n)
    _check(rng.integers(0, 1 << 19, n, dtype=np.uint32), 19)


@pytest.mark.parametrize("bits", [1, 3, 7, 8, 9, 10, 11, 15, 16, 17, 20, 21, 24, 30, 31, 32])
def test_key_widths(bits):
    rng = np.random.default_rng(bits)
    hi = (1 << bits) if bits < 32 else (1 << 32)
    _check(rng.integers(0, hi, 50_000, dtype=np.uint64).astype(np.uint32), bits)


def test_all_equal_keys():
    _check(np.full(70_000, 12345, np.uint32), 19)


def test_maximum_keys():
    keys = np.full(9000, 0xFFFFFFFF, np.uint32)
    keys[::7] = 0
    _check(keys, 32)


def test_nearly_sorted_keys():
    rng = np.random.default_rng(0)
    keys = np.sort(rng.integers(0, 300_000, 500_000, dtype=np.uint32))
    swaps = rng.integers(0, keys.size - 1, 20_000)
    keys[swaps], keys[swaps + 1] = keys[swaps + 1].copy(), keys[swaps].copy()
    _check(keys, 19)


def test_explicit_values():
    rng = np.random.default_rng(3)
    keys = rng.integers(0, 1000, 30_000, dtype=np.uint32)
    vals = rng.integers(0, 1 << 32, 30_000, dtype=np.uint64).astype(np.uint32)
    _check(keys, 10, vals)


def test_empty_input_is_a_no_op():
    assert int(C.lib().sph_radix_sort_pairs(None, None, None, None, 0, 19, None, 0, None)) == 0


def test_rejects_bad_arguments():
    assert C.lib().sph_radix_sort_pairs(None, None, None, None, 10, 0, None, 0, None) == \
        C.SPH_ERR_INVALID_ARGUMENT
    assert C.lib().sph_radix_sort_pairs(None, None, None, None, 10, 33, None, 0, None) == \
        C.SPH_ERR_INVALID_ARGUMENT
