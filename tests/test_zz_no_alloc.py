"""smi_sort_exchange_noalloc (the sort/exchange on the library's static
device scratch) against smi_sort_exchange (the same work on scratch it
allocates): bit-identical outputs on two calls in a row, the second one
smaller (the static array then holds the first call's contents), and a
call that needs more scratch than the array holds is refused before any
work is done."""
import ctypes

import numpy as np
import pytest

from smilei_amd import _lib

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def I3(v):
    return (ctypes.c_int * 3)(*[int(a) for a in v])


def D3(v):
    return (ctypes.c_double * 3)(*[float(a) for a in v])


def ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def particles(nd, n_cells, patch, dx, n, seed):
    """Positions in the box, a tenth pushed up to a patch length outside it,
    random momenta, weights and charges."""
    rng = np.random.default_rng(seed)
    xmax = [n_cells[d] * dx[d] for d in range(nd)]
    pos = [rng.random(n) * xmax[d] for d in range(nd)]
    out = rng.random(n) < 0.1
    for d in range(nd):
        shift = rng.choice([-1.0, 1.0], n) * rng.random(n) * patch[d] * dx[d]
        pos[d] = np.where(out, pos[d] + shift, pos[d])
    mom = [rng.standard_normal(n) for _ in range(3)]
    w = rng.random(n) + 0.5
    q = rng.choice(np.array([-1, 1], np.int16), n)
    return pos + ([None] if nd == 2 else []), mom, w, q


def run(fn, nd, n_cells, patch, dx, data, npatch, clpp):
    pos, mom, w, q = data
    n = len(w)
    dev = [torch.from_numpy(a).to("cuda") if a is not None else None for a in pos + mom + [w, q]]
    key = torch.full((max(n, 1),), -7, dtype=torch.int32, device="cuda")
    counts = torch.full((npatch,), -7, dtype=torch.int32, device="cuda")
    bins = torch.full((npatch * clpp,), -7, dtype=torch.int32, device="cuda")
    nc = list(n_cells) + [1] * (3 - nd)
    ps = list(patch) + [1] * (3 - nd)
    cl = list(dx) + [1.0] * (3 - nd)
    rc = fn(nd, I3(nc), I3(ps), D3(cl), n, *[ptr(t) for t in dev], ptr(key), ptr(counts), ptr(bins), None)
    _lib.check(rc)
    torch.cuda.synchronize()
    return [t.cpu().numpy() for t in dev if t is not None] + [key.cpu().numpy()[:n], counts.cpu().numpy(),
                                                              bins.cpu().numpy()]


@pytest.mark.parametrize("nd,n_cells,patch,dx,sizes", [
    (3, (32, 16, 16), (8, 8, 8), (0.3, 0.2, 0.25), (200000, 30000)),
    (2, (64, 32), (16, 8), (0.3, 0.2), (120000, 5000)),
])
def test_noalloc_matches_allocating_sort_exchange(nd, n_cells, patch, dx, sizes):
    L = _lib.load()
    npatch = int(np.prod([n_cells[d] // patch[d] for d in range(nd)]))
    clpp = int(np.prod([patch[d] for d in range(nd)])) // (64 if nd == 3 else 16)
    for call, n in enumerate(sizes):
        print("noalloc: nd=%d n=%d call %d" % (nd, n, call), flush=True)
        data = particles(nd, n_cells, patch, dx, n, 500 + 10 * nd + call)
        want = run(L.smi_sort_exchange, nd, n_cells, patch, dx, data, npatch, clpp)
        got = run(L.smi_sort_exchange_noalloc, nd, n_cells, patch, dx, data, npatch, clpp)
        for a, b in zip(got, want):
            np.testing.assert_array_equal(a, b)


def test_noalloc_refuses_oversized_call():
    """A call whose scratch exceeds the static array returns
    SMI_ERR_CAPACITY (-4) before touching the (null) particle arrays'
    memory: only the host-side size check runs."""
    L = _lib.load()
    maxw = L.smi_sort_exchange_noalloc_max_words()
    assert maxw == 16 << 20
    n = maxw // 4  # needs about 5.1 words per particle
    fake = ctypes.c_void_p(16)  # never dereferenced: the call is refused first
    rc = L.smi_sort_exchange_noalloc(3, I3((16, 16, 16)), I3((16, 16, 16)), D3((1.0, 1.0, 1.0)), n,
                                     *[fake] * 7, None, None, None, None, None)
    assert rc == -4
    assert b"static buffer" in L.smi_last_error()
