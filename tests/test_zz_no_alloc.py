"""sfh_rhs_noalloc (the right-hand side without device allocation) against sfh_rhs of a fresh
SFH_SOLVE_BATCHED FP64 handle of the same surface, through the C ABI.

Device arrays come from hipMalloc of the HIP runtime the library itself is bound to (called
through ctypes).  Two calls in a row share the library's static
scratch array: a 1024-node water surface, then a smaller 256-node helium film, so the second call
runs on what the first left there.  Bar: 1e-12 relative to 1 + max|reference| (the two paths run
the same kernels; only the twiddle tables are computed differently, on the device here and in long
double on the host for the handle).
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _surface(n, seed, amp=0.04):
    rng = np.random.default_rng(seed)
    x = 2 * np.pi * np.arange(n) / n
    a = rng.uniform(0, amp / np.arange(1, 9))
    th = rng.uniform(0, 2 * np.pi, 8)
    return x + 1j * sum(a[k] * np.cos((k + 1) * x + th[k]) for k in range(8))


def _close(a, b, tol=1e-12):
    err = np.abs(a - b).max()
    assert err <= tol * (1.0 + np.abs(b).max()), f"max err {err:.3e}"
    return err


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the HIP runtime the library is bound to: opening the runtime
    by its SONAME returns the copy loaded first, which is the one the library's DT_NEEDED resolved to
    (a process that imported torch can hold a second copy of the runtime under another path)."""

    def __init__(self):
        self.h = ctypes.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.h.hipFree.argtypes = [ctypes.c_void_p]
        self.h.hipDeviceSynchronize.argtypes = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(p), nbytes) == 0
        return p

    def put(self, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        p = self.alloc(a.nbytes)
        assert self.h.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, 1) == 0   # host to device
        return p

    def get(self, p, n):
        a = np.empty(n)
        assert self.h.hipMemcpy(a.ctypes.data_as(ctypes.c_void_p), p, a.nbytes, 2) == 0   # device to host
        return a


def test_rhs_noalloc_matches_batched_handle_on_two_calls():
    import cusuperhelium as c
    lib = c.lib()
    assert lib.sfh_rhs_noalloc_scratch_bytes() >= 8 * 1024 * 1024
    cases = [(1024, dict(), 3), (256, dict(kind=1, depth=1.0), 4)]
    for n, kw, seed in cases:
        Z = _surface(n, seed)
        P = 0.01 * np.sin(2 * np.pi * np.arange(n) / n + 0.2)
        props = c.Props(**kw)
        with c.Ensemble([n], props, solver=c.BATCHED, tol=1e-13, max_iter=32) as e:
            e.set_state(Z, P)
            q_ref, dp_ref, d = e.rhs(diagnostics=True)
        hip = _Hip()
        x, y, p = hip.put(Z.real), hip.put(Z.imag), hip.put(P)
        vx, vy, dp, mu = (hip.put(np.full(n, np.nan)) for _ in range(4))
        cp = props.c()
        print(f"noalloc n={n} {kw}", flush=True)
        rc = lib.sfh_rhs_noalloc(ctypes.byref(cp), n, x, y, p, vx, vy, dp, mu, 1e-13, 32, None)
        assert rc == 0, c._lib.last_error()
        assert hip.h.hipDeviceSynchronize() == 0
        q = hip.get(vx, n) + 1j * hip.get(vy, n)
        eq = _close(q, q_ref)
        ep = _close(hip.get(dp, n), dp_ref)
        em = _close(hip.get(mu, n), d["mu"])
        for buf in (x, y, p, vx, vy, dp, mu):
            hip.h.hipFree(buf)
        print(f"  |dq| {eq:.2e} |ddphi| {ep:.2e} |dmu| {em:.2e}", flush=True)


def test_rhs_noalloc_rejects_sizes_beyond_the_scratch():
    import cusuperhelium as c
    lib = c.lib()
    cp = c.Props().c()
    for n in (2048, 1000, 0):
        rc = lib.sfh_rhs_noalloc(ctypes.byref(cp), n, None, None, None, None, None, None, None, 1e-12, 16, None)
        assert rc == -2, (n, rc)
    assert "unsupported node count" in c._lib.last_error()
