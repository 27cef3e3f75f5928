"""The no-allocation connectivities entry point against the workspace one, through the C ABI:
bit-identical CSR, sigma and rho on two calls in a row (the second smaller, so stale scratch
from the first would show), and the size limit rejected before any device work."""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _samples(seed, sizes, d=20, k=15):
    import torch
    from metalsinglecell import _hip
    rng = np.random.default_rng(seed)
    parts = []
    for m in sizes:
        c = int(rng.integers(3, 9))
        centers = rng.normal(0.0, 5.0, (c, d))
        parts.append((centers[rng.integers(0, c, m)] + rng.normal(0.0, 1.0, (m, d))).astype(np.float32))
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    idx, dist = _hip.knn_batched(np.concatenate(parts), off, k, precision="bf16", splits=1)
    return idx, dist, torch.from_numpy(off).to(idx.device), len(sizes)


def _outputs(n, k, dev):
    import torch
    cap = 2 * n * (k - 1)
    return (torch.empty(n + 1, dtype=torch.int64, device=dev), torch.empty(cap, dtype=torch.int32, device=dev),
            torch.empty(cap, dtype=torch.float32, device=dev), torch.empty(n, dtype=torch.float32, device=dev),
            torch.empty(n, dtype=torch.float32, device=dev), cap)


def _run(noalloc, idx, dist, goff, groups):
    import torch
    from metalsinglecell import _hip
    lib = _hip.load_library()
    n, k = idx.shape
    ip, ind, dat, sig, rho, cap = _outputs(n, k, idx.device)
    P = _hip.ptr
    if noalloc:
        _hip.check(lib.msc_fuzzy_connectivities_batched_noalloc(
            P(idx), P(dist), n, k, P(goff), groups, P(ip), P(ind), P(dat), cap, P(sig), P(rho), _hip.stream_ptr()))
    else:
        ws = _hip.workspace(_hip.fuzzy_batched_workspace_bytes(n, k, groups))
        _hip.check(lib.msc_fuzzy_connectivities_batched(
            P(idx), P(dist), n, k, P(goff), groups, P(ip), P(ind), P(dat), cap, P(sig), P(rho), P(ws), ws.numel(),
            _hip.stream_ptr()))
    torch.cuda.synchronize()
    nnz = int(ip[-1].item())
    return ip.cpu().numpy(), ind[:nnz].cpu().numpy(), dat[:nnz].cpu().numpy(), sig.cpu().numpy(), rho.cpu().numpy()


def test_noalloc_equals_workspace_entry_point_twice():
    first = _samples(1, [900, 3000, 250, 1700])
    second = _samples(2, [400, 800, 120])                 # smaller: stale scratch would show
    for inputs in (first, second):
        got = _run(True, *inputs)
        ref = _run(False, *inputs)
        for g, r in zip(got, ref):
            assert np.array_equal(g, r)
        assert got[1].size > 0


def test_noalloc_rejects_sizes_beyond_its_scratch():
    from metalsinglecell import _hip
    lib = _hip.load_library()
    n, k = 3_000_000, 15
    nb = ctypes.c_size_t(0)
    _hip.check(lib.msc_fuzzy_batched_workspace_size(n, k, 1, ctypes.byref(nb)))
    assert nb.value > 64 << 20
    z = ctypes.c_void_p(16)       # never dereferenced: the size check fails first
    rc = lib.msc_fuzzy_connectivities_batched_noalloc(z, z, n, k, z, 1, z, z, z, 2 * n * (k - 1), z, z, None)
    assert rc == 1 and b"static scratch" in lib.msc_last_error()
