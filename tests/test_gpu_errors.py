"""Error paths on the device, forced from the host so that every kernel still
exits normally: a bounded in-kernel wait for a flag that nothing sets must
give up at its wall-clock bound and surface as TL_ETIMEOUT (through the call
and through tl_sync), and clean calls leave nothing to report."""
import pytest
import torch

import tl_amd

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("wait_ms", [20, 200])
def test_bounded_wait_gives_up_at_its_bound(wait_ms, dev):
    print(f"[test] selftest_timeout {wait_ms} ms", flush=True)
    st, ms = tl_amd.selftest_timeout(wait_ms)
    assert st == tl_amd.TL_ETIMEOUT, st
    # the bound is wall-clock time on the device: the kernel ends near it
    assert 0.8 * wait_ms <= ms <= 5 * wait_ms + 50, ms
    # the report cleared the sticky word
    tl_amd.sync()


def test_untimed_calls_report_through_sync(dev):
    n = 1 << 22
    A = torch.arange(n, dtype=torch.int64, device=dev)
    for _ in range(3):
        tl_amd.permute_btree(A, 16, kind="u64", timed=False)
        tl_amd.unpermute("btree", A, b=16, kind="u64", timed=False)
    tl_amd.sync()   # no wait gave up
    assert torch.equal(A, torch.arange(n, dtype=torch.int64, device=dev))


def test_bad_arguments_are_rejected_before_any_launch(dev):
    A = torch.arange(100, dtype=torch.int64, device=dev)
    with pytest.raises(ValueError):
        tl_amd.permute_btree(A[::2], 16, kind="u64")          # strided view
    with pytest.raises(ValueError):
        tl_amd.permute_btree(A.cpu(), 16, kind="u64")         # host tensor
    with pytest.raises(ValueError):
        tl_amd.permute_btree(A, 16, kind="u32")               # wrong width
    Q = torch.arange(10, dtype=torch.int64, device=dev)
    with pytest.raises(ValueError):
        tl_amd.query_btree(A, 16, Q, torch.empty(5, dtype=torch.int64, device=dev), kind="u64")
    with pytest.raises(tl_amd.TLError):
        tl_amd.permute_btree(A, 0, kind="u64")                # b = 0
    st, _ = tl_amd.selftest_timeout(0)                        # wait_ms out of range
    assert st == tl_amd.TL_EINVAL
