"""pimd_forces_all_beads_noalloc (no context, no device allocation: every
intermediate array in the library's static per-GPU workspace) against
pimd_forces_all_beads_dev on a context, on two calls in a row, the second with
a smaller system and on a stream of its own.  The two entry points run the
same kernels with the same launch shapes, so the outputs must be the same bits.
Device memory and streams come from the HIP runtime the library loaded.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("Upott", "dRRt", "virt", "virialct", "dip_momIt", "dip_momEt", "chg", "iters")


class Hip:
    """Device memory and streams from the HIP runtime the library itself
    loaded (one runtime in the process)."""

    def __init__(self):
        import ctypes as C
        from pimd_amd import ttm
        ttm.load_library()
        self.C = C
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.L.hipStreamDestroy.argtypes = [C.c_void_p]
        self.L.hipStreamSynchronize.argtypes = [C.c_void_p]
        self.L.hipDeviceSynchronize.argtypes = []
        self.held = []

    def ok(self, rc, what):
        assert rc == 0, f"{what} failed: hipError {rc}"

    def upload(self, a):
        a = np.ascontiguousarray(a)
        ptr = self.C.c_void_p()
        self.ok(self.L.hipMalloc(self.C.byref(ptr), max(a.nbytes, 1)), "hipMalloc")
        self.held.append(ptr.value)
        self.ok(self.L.hipMemcpy(ptr, a.ctypes.data, a.nbytes, 1), "hipMemcpy H2D")
        return ptr.value

    def download(self, ptr, like):
        out = np.empty_like(like)
        self.ok(self.L.hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2), "hipMemcpy D2H")
        return out

    def stream(self):
        s = self.C.c_void_p()
        self.ok(self.L.hipStreamCreate(self.C.byref(s)), "hipStreamCreate")
        return s.value

    def close(self):
        self.L.hipDeviceSynchronize()
        for ptr in self.held:
            self.L.hipFree(ptr)
        self.held = []


def _templates(p):
    P, na, N = p.nbeads, p.natoms, p.nwaters
    nan = np.nan
    return {"Upott": np.full(P, nan), "dRRt": np.full((P, na, 3), nan),
            "virt": np.full((P, 9), nan), "virialct": np.full(P, nan),
            "dip_momIt": np.full((P, N, 3), nan), "dip_momEt": np.full((P, N, 3), nan),
            "chg": np.full((P, na), nan), "iters": np.full(P, -999, dtype=np.int32)}


def _case(model, N, P, rc, eps, precision, seed, stream):
    from pimd_amd import ttm
    from pimd_amd.params import TTMParams
    from pimd_amd.systems import make_system
    print("case", model, N, P, rc, eps, precision, "stream" if stream else "default", flush=True)
    RRt, _, RRc, L = make_system(N, P, seed=seed)
    p = TTMParams(pot_model=model, nwaters=N, nbeads=P, box=(L, L, L), rc=rc, eps_ewald=eps,
                  precision=precision).resolved()
    hip = Hip()
    try:
        R, Rc = hip.upload(RRt), hip.upload(RRc)
        tmpl = _templates(p)
        ref = {k: hip.upload(v) for k, v in tmpl.items()}
        got = {k: hip.upload(v) for k, v in tmpl.items()}
        ff = ttm.TTMForceField(p)
        lib = ttm.load_library()
        ttm._check(lib.pimd_forces_all_beads_dev(ff.handle, R, Rc, 1, *[ref[k] for k in KEYS]))
        ttm._check(lib.pimd_sync(ff.handle))
        s = hip.stream() if stream else None
        ttm.forces_noalloc(p, R, Rc, 1, got, stream=s)
        hip.ok(hip.L.hipStreamSynchronize(s), "hipStreamSynchronize")
        if s is not None:
            hip.L.hipStreamDestroy(s)
        ff.close()
        for k in KEYS:
            a, b = hip.download(got[k], tmpl[k]), hip.download(ref[k], tmpl[k])
            assert np.all(np.isfinite(b)), k
            assert np.array_equal(a, b), (k, float(np.max(np.abs(a - b))))
    finally:
        hip.close()
    print("identical", flush=True)


def test_noalloc_matches_context_call_twice_smaller_second():
    _case(3, 128, 8, -1.0, 1e-6, 0, 71, stream=False)
    _case(2, 32, 4, 4.9, 1e-8, 0, 72, stream=True)


def test_noalloc_mixed_and_fixed_charge():
    _case(3, 64, 8, -1.0, 1e-6, 1, 73, stream=True)
    _case(4, 32, 2, -1.0, 1e-6, 0, 74, stream=False)


def test_noalloc_refuses_what_does_not_fit():
    from pimd_amd import ttm
    from pimd_amd.params import TTMParams
    from pimd_amd.systems import make_system
    _, _, _, L = make_system(512, 1, seed=1)
    p = TTMParams(pot_model=3, nwaters=512, nbeads=64, box=(L, L, L))
    with pytest.raises(ttm.PimdError, match="workspace"):
        ttm.forces_noalloc(p, 1, 1)
    q = TTMParams(pot_model=3, nwaters=32, nbeads=2, box=(L / 2, L / 2, L / 2), guess_initdip=True)
    with pytest.raises(ttm.PimdError, match="guess_initdip"):
        ttm.forces_noalloc(q, 1, 1)
