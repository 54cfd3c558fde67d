"""The no-allocation entry point (rnng_l2_eps_pairs_f32_noalloc) against the
allocating one-shot entry point it is a variant of (rnng_l2_eps_pairs_f32):
identical sorted pairs and weights, bit for bit, on two calls in a row (the
second smaller, so stale contents of the shared scratch array would show),
plus its capacity, scratch-limit and argument errors.  Inputs and outputs are
device buffers that the caller allocates: this test does so through the HIP
runtime the library itself is linked to (one runtime in the process)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_H2D, _D2H = 1, 2


def _hip():
    from rnng_amd import _lib
    _lib.lib()   # loads the library and, as its dependency, the HIP runtime
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return h


class _DevArray:
    """Device buffer of the caller, filled from / copied to numpy."""

    def __init__(self, hip, host):
        self.hip, self.shape, self.dtype = hip, host.shape, host.dtype
        self.nbytes = max(host.nbytes, 16)
        p = C.c_void_p()
        assert hip.hipMalloc(C.byref(p), self.nbytes) == 0
        self.ptr = p.value
        host = np.ascontiguousarray(host)
        assert hip.hipMemcpy(self.ptr, host.ctypes.data, host.nbytes, _H2D) == 0

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, _D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


def _call(pts_dev, n, d, eps, ij_dev, w_dev, cap):
    from rnng_amd import Stats, _lib
    needed = C.c_int64(-1)
    st = Stats()
    rc = _lib.lib().rnng_l2_eps_pairs_f32_noalloc(
        C.c_void_p(pts_dev), n, d, np.float32(eps), C.c_void_p(ij_dev), C.c_void_p(w_dev), cap,
        C.byref(needed), C.byref(st))
    return rc, needed.value, st


def _run(hip, pts, eps, cap, fill=-7):
    n, d = pts.shape
    dev = _DevArray(hip, np.asarray(pts, np.float32))
    ij = _DevArray(hip, np.full((max(cap, 1), 2), fill, np.int64))
    w = _DevArray(hip, np.full(max(cap, 1), fill, np.float32))
    # the library runs on its own stream: the copies above must have landed
    assert hip.hipDeviceSynchronize() == 0
    try:
        rc, needed, st = _call(dev.ptr, n, d, eps, ij.ptr, w.ptr, cap)
        return rc, needed, st, ij.numpy(), w.numpy()
    finally:
        for a in (dev, ij, w):
            a.free()


def test_noalloc_equals_allocating_entry_on_two_calls():
    hip = _hip()
    from rnng_amd import _lib, l2, synth
    pts, eps = synth.points_gmm_small(n=9000)
    for m in (9000, 2500):   # the second call is smaller
        sub = pts[:m]
        ref_ij, ref_w = l2.eps_pairs_oneshot(sub, eps, "plain")
        cap = ref_ij.shape[0] + 1000
        print(f"noalloc call: n={m} d={sub.shape[1]} eps={float(eps)} edges={ref_ij.shape[0]}", flush=True)
        rc, needed, st, ij, w = _run(hip, sub, eps, cap)
        assert rc == _lib.RNNG_OK, _lib.lib().rnng_last_error()
        assert needed == ref_ij.shape[0]
        assert st.edges == needed and st.pair_evals == m * (m - 1) // 2
        assert np.array_equal(ij[:needed], ref_ij)
        assert w[:needed].tobytes() == ref_w.tobytes()
        # nothing is written past the edge count
        assert (ij[needed:] == -7).all() and (w[needed:] == -7.0).all()
    # the MFMA path of the allocating entry point gives the same list
    ref2_ij, ref2_w = l2.eps_pairs_oneshot(pts[:2500], eps, "mfma")
    assert np.array_equal(ij[:needed], ref2_ij) and w[:needed].tobytes() == ref2_w.tobytes()


def test_noalloc_errors():
    hip = _hip()
    from rnng_amd import _lib, l2, synth
    pts, eps = synth.points_gmm_small(n=3000)
    ref_ij, _ = l2.eps_pairs_oneshot(pts, eps, "plain")
    ne = ref_ij.shape[0]
    # capacity below the count: the count comes back, nothing is written
    print("noalloc call: capacity error", flush=True)
    rc, needed, _, ij, w = _run(hip, pts, eps, ne - 1)
    assert rc == _lib.RNNG_ERR_CAPACITY and needed == ne
    assert (ij == -7).all() and (w == -7.0).all()
    # more edges than the scratch array holds: 2,100 equal points give
    # 2,203,950 pairs at distance 0, above the 2^21 limit
    print("noalloc call: scratch limit", flush=True)
    same = np.zeros((2100, 4), np.float32)
    rc, needed, _, _, _ = _run(hip, same, 0.0, 16)
    assert rc == _lib.RNNG_ERR_UNSUPPORTED and needed == 2100 * 2099 // 2
    assert b"scratch" in _lib.lib().rnng_last_error()
    # a host pointer is refused before any kernel sees it
    print("noalloc call: host pointer", flush=True)
    host = np.ascontiguousarray(pts)
    out = _DevArray(hip, np.zeros((ne, 2), np.int64))
    rc, _, _ = _call(host.ctypes.data, host.shape[0], host.shape[1], eps, out.ptr, 0, ne)
    out.free()
    assert rc == _lib.RNNG_ERR_ARG
    # after the errors the next call is still exact
    rc, needed, _, ij, w = _run(hip, pts, eps, ne)
    assert rc == _lib.RNNG_OK and needed == ne and np.array_equal(ij, ref_ij)
