"""dsmc_op_sort_static, the sort operator without a workspace argument,
against dsmc_op_sort with a caller workspace: same permutation and cell
offsets, bit for bit, on two calls in a row (the second smaller, so stale
contents of the shared device array would show), and the size limit is
refused on the host.
"""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np
import pytest

from cudadsmc_amd import _capi

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def note(msg: str) -> None:
    print(msg, flush=True)
    sys.stdout.flush()


def ptr(t) -> C.c_void_p:
    return C.c_void_p(t.data_ptr())


def dev(a: np.ndarray):
    return torch.from_numpy(np.ascontiguousarray(a).view(np.int32)).cuda()


def sort_with_workspace(tk, n: int, n_cells: int):
    perm = torch.zeros(max(n, 1), dtype=torch.int32, device="cuda")
    cs = torch.zeros(n_cells + 1, dtype=torch.int32, device="cuda")
    work = torch.zeros(int(_capi.lib().dsmc_op_sort_work_words(n, n_cells)), dtype=torch.int32, device="cuda")
    _capi.check(_capi.lib().dsmc_op_sort(ptr(tk), n, n_cells, ptr(perm), ptr(cs), ptr(work), None))
    torch.cuda.synchronize()
    return perm.cpu().numpy().view(np.uint32)[:n], cs.cpu().numpy().view(np.uint32)


def sort_static(tk, n: int, n_cells: int):
    perm = torch.zeros(max(n, 1), dtype=torch.int32, device="cuda")
    cs = torch.zeros(n_cells + 1, dtype=torch.int32, device="cuda")
    _capi.check(_capi.lib().dsmc_op_sort_static(ptr(tk), n, n_cells, ptr(perm), ptr(cs), None))
    torch.cuda.synchronize()
    return perm.cpu().numpy().view(np.uint32)[:n], cs.cpu().numpy().view(np.uint32)


def test_static_sort_matches_workspace_sort_twice():
    rng = np.random.default_rng(5)
    for n, n_cells in ((1 << 20, 1 << 20), (100000, 3000)):
        note(f"static sort: {n} keys over {n_cells} cells")
        keys = rng.integers(0, n_cells, n).astype(np.uint32)
        keys[rng.random(n) < 0.05] = 0xFFFFFFFF
        tk = dev(keys)
        want_p, want_c = sort_with_workspace(tk, n, n_cells)
        got_p, got_c = sort_static(tk, n, n_cells)
        assert np.array_equal(got_c, want_c)
        assert np.array_equal(got_p, want_p)
        valid = keys < n_cells
        assert int(got_c[-1]) == int(valid.sum())
        assert np.array_equal(got_p[: got_c[-1]], np.argsort(np.where(valid, keys, n_cells), kind="stable")[: got_c[-1]])


def test_static_sort_refuses_sizes_beyond_its_array():
    tk = torch.zeros(4, dtype=torch.int32, device="cuda")
    perm = torch.zeros(4, dtype=torch.int32, device="cuda")
    cs = torch.zeros(4, dtype=torch.int32, device="cuda")
    rc = _capi.lib().dsmc_op_sort_static(ptr(tk), (1 << 20) + 1, 3, ptr(perm), ptr(cs), None)
    assert rc == -5 and b"at most" in _capi.lib().dsmc_global_last_error()
    rc = _capi.lib().dsmc_op_sort_static(ptr(tk), 3, (1 << 20) + 1, ptr(perm), ptr(cs), None)
    assert rc == -5
