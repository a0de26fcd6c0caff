"""ribo_ledger_noalloc, the entry point that allocates nothing, against
ribo_world_ledger on the same grids: its partial sums live in the library's
fixed device scratch, shared by every call on the GPU, so two calls in a row
(the second on a smaller grid, whose totals must not pick up the first
call's partials) both equal the world's ledger and a host count."""
import ctypes

import numpy as np
import pytest

import ribossome_amd as ra

pytestmark = pytest.mark.gpu


class DeviceGrid:
    """The caller's own device copy of a grid, through the HIP runtime the
    library links (hipMalloc / hipMemcpy / hipFree by ctypes)."""

    def __init__(self, words: np.ndarray):
        ra.lib()  # loads libamdhip64.so.7 with the library
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.ptr = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(self.ptr), words.nbytes) == 0
        a = np.ascontiguousarray(words)
        assert self.hip.hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, 1) == 0  # host to device

    def close(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()


def grid_and_world_ledger(side, seed, lam):
    cfg = ra.WorldConfig(side, side // 2)
    gamma = ra.gen_gamma(cfg, seed, 0.1, True)
    chem = ra.gen_chem(cfg, seed, lam, gamma)
    with ra.World(cfg, seed) as w:
        w.upload(ra.GAMMA, gamma)
        w.upload(ra.CHEM, chem)
        return chem, w.ledger_blocking()


def test_noalloc_ledger_matches_world_ledger_twice_then_smaller():
    print("case noalloc ledger", flush=True)
    for side, seed, lam in ((1024, 61, 6.0), (96, 62, 3.0)):
        chem, want = grid_and_world_ledger(side, seed, lam)
        dev = DeviceGrid(chem)
        try:
            got = ra.ledger_noalloc(dev.ptr.value, side * side)
            print(f"side {side}: noalloc {got.act} {got.spent}; world {want.act} {want.spent}", flush=True)
            assert got == want
            assert got == ra.Ledger.from_cells(chem)
            # a prefix of the grid: the totals of those cells only
            part = ra.ledger_noalloc(dev.ptr.value, side * side // 3)
            assert part == ra.Ledger.from_cells(chem[: 4 * (side * side // 3)])
        finally:
            dev.close()


def test_noalloc_ledger_limits():
    print("case noalloc limits", flush=True)
    z = ra.ledger_noalloc(0, 0)
    assert z.act == (0, 0, 0, 0) and z.spent == (0, 0, 0, 0)
    with pytest.raises(ra.RiboError, match="exceed the maximum"):
        ra.ledger_noalloc(16, (1 << 30) + 1)
    with pytest.raises(ra.RiboError, match="null"):
        ra.ledger_noalloc(0, 5)
