"""mqct_hip_cross_sections_noalloc, the entry point that allocates nothing
(its working memory is the library's fixed device array), against
mqct_hip_cross_sections on the same inputs: bit-identical sigma and
opacities on two calls in a row, the second smaller, with device buffers the
test allocates itself through the HIP runtime; an input beyond the fixed
working memory is refused before anything runs."""
import ctypes

import numpy as np
import pytest

from mqct_amd import _lib as L
from mqct_amd import workloads as W
from mqct_amd.propagate import gpu_cross_sections

pytestmark = pytest.mark.gpu


class DeviceArray:
    """A device copy of a numpy array through the HIP runtime (the caller's
    own allocation, as a no-alloc caller would hold it)."""
    hip = None

    def __init__(self, a=None, nbytes=None):
        if DeviceArray.hip is None:
            DeviceArray.hip = ctypes.CDLL("libamdhip64.so")
        self.nbytes = a.nbytes if a is not None else nbytes
        self.p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(self.p), ctypes.c_size_t(max(self.nbytes, 8))) == 0
        if a is not None:
            a = np.ascontiguousarray(a)
            assert self.hip.hipMemcpy(self.p, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(self.nbytes),
                                      1) == 0  # hipMemcpyHostToDevice

    def get(self, shape, dtype=np.float64):
        out = np.zeros(shape, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.p, ctypes.c_size_t(out.nbytes),
                                  2) == 0  # hipMemcpyDeviceToHost
        return out

    def free(self):
        if self.p:
            self.hip.hipFree(self.p)
            self.p = None


def run_noalloc(dev, q, b, Wt, opacity):
    T, ldq = q.shape
    C, G = b.n_channels, Wt.shape[0]
    bufs = [DeviceArray(np.ascontiguousarray(q)), DeviceArray(L.as_i32(b.indx_chann)),
            DeviceArray(np.ascontiguousarray(Wt)), DeviceArray(nbytes=G * C * 8)]
    op = DeviceArray(nbytes=T * C * 8) if opacity else None
    try:
        L.check(dev.lib.mqct_hip_cross_sections_noalloc(dev.h, T, b.n_states, ldq, bufs[0].p, C, bufs[1].p, G,
                                                        bufs[2].p, bufs[3].p, op.p if op else None),
                "cross_sections_noalloc")
        sig = bufs[3].get((G, C))
        P = op.get((T, C)) if op else None
    finally:
        for x in bufs + ([op] if op else []):
            x.free()
    return sig, P


def test_noalloc_matches_allocating_entry_point(dev):
    rng = np.random.default_rng(23)
    big = W.asym_top_atom(jmax=8, n_r=4).basis     # 969 states, 81 channels
    small = W.asym_top_atom(jmax=3, n_r=4).basis   # 84 states, 16 channels
    for b, T, G, opacity in ((big, 1500, 4, True), (small, 300, 2, False)):
        print(f"noalloc call: {b.n_states} states, {T} trajectories, {G} groups, opacity {opacity}", flush=True)
        q = rng.standard_normal((T, 2 * b.n_states + 8))
        Wt = rng.uniform(0.0, 2.0, (G, T))
        sig_ref, P_ref = gpu_cross_sections(dev, q, b, Wt, opacity=True)
        sig, P = run_noalloc(dev, q, b, Wt, opacity)
        assert np.array_equal(sig, sig_ref)
        if opacity:
            assert np.array_equal(P, P_ref)


def test_noalloc_refuses_inputs_beyond_its_scratch(dev):
    b = W.asym_top_atom(jmax=3, n_r=4).basis
    T, G = 1 << 20, 1                               # P alone would need 128 MiB
    rc = dev.lib.mqct_hip_cross_sections_noalloc(dev.h, T, b.n_states, 2 * b.n_states + 8, ctypes.c_void_p(8),
                                                 b.n_channels, ctypes.c_void_p(8), G, ctypes.c_void_p(8),
                                                 ctypes.c_void_p(8), None)
    assert rc == -1
    assert "fixed scratch" in L.last_error()
