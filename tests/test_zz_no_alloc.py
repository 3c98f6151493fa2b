"""The allocation-free entry point (gapx_calc_device_noalloc) against the entry point
it is a variant of (gapx_calc_device): the same results, bit for bit, on two calls in
a row, the second with a smaller structure, in the FP64 and bf16x3 forms (bf16x3 also
under the stream-K schedule of its fused GP kernel); a request
beyond its device block is refused with GAPX_E_CAPACITY."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from quip_amd import _lib as L
from quip_amd import configs
from quip_amd.gap_xml import GPCoordinate, write_gap_xml
from quip_amd.potential import Potential

pytestmark = pytest.mark.gpu

GAPX_E_CAPACITY = 7
WANT_ALL = 1 | 2 | 4 | 8 | 16


def _desc_fn(d, at):
    return O.soap_descriptors(GPCoordinate("t", d, 2, 1, 0, 4, np.zeros((1, d.dim)), np.zeros(1), np.ones(1)), at)[0]


@pytest.fixture(scope="module")
def model_path(tmp_path_factory):
    train = [configs.perturb(configs.diamond(5.43, (1, 1, 1)), 0.1, configs.rng_for(0, 950 + k)) for k in range(8)]
    m = configs.soap_model(_desc_fn, train, 64, 0, [configs.soap_string()], "NA", alpha_std=0.1,
                           d2b=(5.0, 1.0, 12, 1.8, 5.0))
    path = str(tmp_path_factory.mktemp("na") / "na.xml")
    write_gap_xml(path, m)
    return path


class _Dev:
    """hipMalloc'd buffers for one call (the test allocates; the entry point must not)."""

    def __init__(self, N):
        self.hip = C.CDLL("libamdhip64.so")
        self.N = N
        self.p = {}
        for k, n in (("pos", 3 * N), ("e", 1), ("le", N), ("f", 3 * N), ("v", 9), ("lv", 9 * N)):
            q = C.c_void_p()
            assert self.hip.hipMalloc(C.byref(q), C.c_size_t(8 * max(n, 1))) == 0
            self.p[k] = q

    def put(self, pos):
        a = L.f64(pos)
        assert self.hip.hipMemcpy(self.p["pos"], L.ptr(a), C.c_size_t(a.nbytes), 1) == 0

    def get(self):
        out = {}
        for k, n in (("e", 1), ("le", self.N), ("f", 3 * self.N), ("v", 9), ("lv", 9 * self.N)):
            out[k] = np.zeros(n)
            assert self.hip.hipMemcpy(L.ptr(out[k]), self.p[k], C.c_size_t(8 * n), 2) == 0
        return out

    def free(self):
        for q in self.p.values():
            self.hip.hipFree(q)


def _call(fn, pot, at, dev):
    dev.put(at.positions)
    st = fn(pot.handle, len(at), dev.p["pos"], L.ptr(L.i32(at.numbers)), L.ptr(L.f64(at.cell)), L.ptr(L.i32(at.pbc)),
            0.0, WANT_ALL, dev.p["e"], dev.p["le"], dev.p["f"], dev.p["v"], dev.p["lv"])
    return st


@pytest.mark.parametrize("precision", ["fp64", "bf16x3", "bf16x3-sk"])
def test_noalloc_equals_device_entry(model_path, precision):
    """bf16x3: the allocating call tunes the fused split kernel and the allocation-free
    call sizes its slots for that tuning; bf16x3-sk: both under the stream-K schedule."""
    pot = Potential(param_filename=model_path, precision=precision.split("-")[0])
    lib = L.lib()
    if precision.endswith("-sk"):
        L.check(lib.gapx_model_set_fused_schedule(pot.handle, -1), "set_fused_schedule")
    big = configs.perturb(configs.diamond(5.43, (4, 4, 4)), 0.1, np.random.default_rng(1))     # 512 atoms
    small = configs.perturb(configs.diamond(5.43, (2, 2, 2)), 0.1, np.random.default_rng(2))   # 64 atoms
    for at in (big, small):
        print("case noalloc", precision, len(at), flush=True)
        ref_dev, na_dev = _Dev(len(at)), _Dev(len(at))
        try:
            L.check(_call(lib.gapx_calc_device, pot, at, ref_dev), "gapx_calc_device")
            L.check(lib.gapx_sync(pot.handle), "gapx_sync")
            ref = ref_dev.get()
            L.check(_call(lib.gapx_calc_device_noalloc, pot, at, na_dev), "gapx_calc_device_noalloc")
            got = na_dev.get()
        finally:
            ref_dev.free()
            na_dev.free()
        for k in ref:
            assert np.array_equal(got[k], ref[k]), (precision, len(at), k)
        assert np.isfinite(got["e"][0]) and got["e"][0] != 0.0


def test_noalloc_refuses_beyond_its_block(model_path):
    pot = Potential(param_filename=model_path)
    at = configs.perturb(configs.diamond(5.43, (25, 25, 25)), 0.05, np.random.default_rng(3))   # 125000 atoms
    dev = _Dev(len(at))
    try:
        st = _call(L.lib().gapx_calc_device_noalloc, pot, at, dev)
    finally:
        dev.free()
    assert st == GAPX_E_CAPACITY and "allocation-free" in L.last_error()
