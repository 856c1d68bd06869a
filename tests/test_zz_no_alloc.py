"""fp_build_grid_noalloc -- the reference grid (src/solver.js:174-184) over
device arrays the caller holds, with every intermediate in the library's
static device array -- gives the same cellStart / cellEntries, bit for bit,
as fp_build_grid + fp_read_grid on the same positions, on two calls in a row
(the second with a smaller scene and table), with outputs pre-filled with
garbage so nothing stale can pass.

Device memory for the inputs and outputs comes from the HIP runtime the
library itself is linked against (ctypes on /opt/rocm's libamdhip64); the
entry point under test allocates nothing."""
import ctypes

import numpy as np
import pytest

from tests._gpu import log

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0


class _Hip:
    def __init__(self):
        from fluidphysics_amd import _abi
        self.lib = _abi.load_library()  # loads /opt/rocm/lib/libamdhip64.so.7 first
        self.hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.hip.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
        self.live = []

    def check(self, rc, what):
        assert rc == 0, "%s failed: %d" % (what, rc)

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        self.check(self.hip.hipMalloc(ctypes.byref(p), max(int(nbytes), 4)), "hipMalloc")
        self.live.append(p.value)
        return p.value

    def upload(self, a):
        a = np.ascontiguousarray(a)
        d = self.alloc(a.nbytes)
        self.check(self.hip.hipMemcpy(d, a.ctypes.data, a.nbytes, 1), "hipMemcpy H2D")
        return d

    def download(self, d, count, dtype):
        out = np.empty(count, dtype)
        self.check(self.hip.hipMemcpy(out.ctypes.data, d, out.nbytes, 2), "hipMemcpy D2H")
        return out

    def garbage(self, nbytes):
        d = self.alloc(nbytes)
        self.check(self.hip.hipMemset(d, 0x7f, max(int(nbytes), 4)), "hipMemset")
        return d

    def sync(self):
        self.check(self.hip.hipDeviceSynchronize(), "hipDeviceSynchronize")

    def free(self):
        for p in self.live:
            self.hip.hipFree(p)
        self.live = []


def _reference_grid(s):
    s.build_grid()
    cs, ce = s.read_grid()
    return cs, ce


def _noalloc_grid(h, s):
    st = s.read_state()
    n = s.capacity
    T = s.info().table_size
    pos = np.ascontiguousarray(st["pos"].reshape(-1), dtype=np.float32)
    act = np.ascontiguousarray(st["active"], dtype=np.uint8)
    assert pos.size == 3 * n and act.size == n
    d_pos, d_act = h.upload(pos), h.upload(act)
    d_cs, d_ce = h.garbage(4 * (T + 1)), h.garbage(4 * n)
    rc = h.lib.fp_build_grid_noalloc(d_pos, d_act, n, T, float(s.cellSize), d_cs, d_ce, None)
    assert rc == 0, h.lib.fp_last_error(None)
    h.sync()
    return h.download(d_cs, T + 1, np.int32), h.download(d_ce, n, np.int32)


def _check(label, got, want):
    cs, ce = got
    rcs, rce = want
    m = int(rcs[-1])
    log("  %s: %d entries in %d buckets" % (label, m, len(rcs) - 1))
    assert np.array_equal(cs, rcs), "%s: cellStart differs" % label
    assert np.array_equal(ce[:m], rce[:m]), "%s: cellEntries differ" % label


def test_noalloc_grid_equals_fp_build_grid_on_two_calls():
    from fluidphysics_amd import from_scene, scenes
    h = _Hip()
    big = from_scene(scenes.single_gpu_scene())
    small = from_scene(scenes.plain_cpu_scene())
    try:
        for s in (big, small):
            s.step(DT, 5)
        assert small.capacity < big.capacity and small.info().table_size < big.info().table_size
        log("case noalloc grid: call 1, capacity %d" % big.capacity)
        _check("call 1", _noalloc_grid(h, big), _reference_grid(big))
        log("case noalloc grid: call 2 (smaller), capacity %d" % small.capacity)
        _check("call 2", _noalloc_grid(h, small), _reference_grid(small))
        # sizes beyond the static array are refused
        assert h.lib.fp_build_grid_noalloc(h.live[0], h.live[1], (1 << 20) + 1, 1 << 20, 0.22,
                                           h.live[2], h.live[3], None) == -1
    finally:
        h.sync()
        h.free()
        big.close()
        small.close()
