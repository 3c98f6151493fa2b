"""The batching form's launch sequence without a host synchronisation inside a call.

After a first (exact) build, the device entry points size the neighbour tables from
capacities kept from earlier builds and queue the whole call, neighbour list
included, as one launch sequence (one graph with graph replay).  A device guard
empties the lists of a call whose slots or tables overflow; gapx_sync reads the
flags and redoes such calls exactly, in order.  These tests check that every result
equals the exact build's bit for bit: repeated calls, eager and graph, a call that
overflows alone or among others queued, and slab batches whose lat_eff (calc_lat_eff,
Connection.f95:740-763) is formed on the device from the call's own extents.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from quip_amd import _lib as L
from quip_amd import configs
from quip_amd.gap_xml import GPCoordinate, write_gap_xml
from quip_amd.potential import Potential

pytestmark = pytest.mark.gpu

WANT = 1 | 2 | 4


def _desc_fn(d, at):
    return O.soap_descriptors(GPCoordinate("t", d, 2, 1, 0, 4, np.zeros((1, d.dim)), np.zeros(1), np.ones(1)), at)[0]


@pytest.fixture(scope="module")
def path(tmp_path_factory):
    train = [configs.perturb(configs.diamond(5.43, (1, 1, 1)), 0.1, configs.rng_for(0, 960 + k)) for k in range(6)]
    m = configs.soap_model(_desc_fn, train, 48, 0, [configs.soap_string()], "B", alpha_std=0.1,
                           d2b=(5.0, 1.0, 12, 1.8, 5.0))
    p = str(tmp_path_factory.mktemp("b") / "b.xml")
    write_gap_xml(p, m)
    return p


class _Buf:
    def __init__(self, S, N):
        self.hip = C.CDLL("libamdhip64.so")
        self.S, self.N = S, N
        self.p = {}
        for k, n in (("pos", 3 * N), ("e", S), ("f", 3 * N), ("v", 9 * S)):
            q = C.c_void_p()
            assert self.hip.hipMalloc(C.byref(q), C.c_size_t(8 * max(n, 1))) == 0
            self.p[k] = q

    def put(self, pos):
        a = L.f64(pos)
        assert self.hip.hipMemcpy(self.p["pos"], L.ptr(a), C.c_size_t(a.nbytes), 1) == 0

    def get(self):
        out = {}
        for k, n in (("e", self.S), ("f", 3 * self.N), ("v", 9 * self.S)):
            out[k] = np.zeros(n)
            assert self.hip.hipMemcpy(L.ptr(out[k]), self.p[k], C.c_size_t(8 * n), 2) == 0
        return out

    def free(self):
        for q in self.p.values():
            self.hip.hipFree(q)


def _calc_dev(pot, at, buf):
    L.check(L.lib().gapx_calc_device(pot.handle, len(at), buf.p["pos"], L.ptr(L.i32(at.numbers)),
                                     L.ptr(L.f64(at.cell)), L.ptr(L.i32(at.pbc)), 0.0, WANT, buf.p["e"], None,
                                     buf.p["f"], buf.p["v"], None), "gapx_calc_device")


def _exact(path, at):
    """A fresh model's first call: the exact (counting) build."""
    r = Potential(param_filename=path).calc(at.copy(), force=True, virial=True)
    return r["energy"], r["force"]


@pytest.mark.parametrize("graphs", [False, True])
def test_repeated_device_calls_equal_exact_build(path, graphs):
    at = configs.perturb(configs.diamond(5.43, (3, 3, 3)), 0.1, np.random.default_rng(4))
    e_ref, f_ref = _exact(path, at)
    pot = Potential(param_filename=path)
    pot.set_graphs(graphs)
    buf = _Buf(1, len(at))
    try:
        buf.put(at.positions)
        for _ in range(6):
            _calc_dev(pot, at, buf)
        L.check(L.lib().gapx_sync(pot.handle), "sync")
        out = buf.get()
    finally:
        buf.free()
    assert out["e"][0] == e_ref
    assert np.array_equal(out["f"].reshape(-1, 3), f_ref)


def test_overflowing_call_is_redone(path):
    """Capacities from a sparse structure, then a denser one of the same size: its slots
    overflow, the guard empties its lists, and gapx_sync redoes it exactly -- also when
    other calls with their own outputs are queued around it."""
    loose = configs.perturb(configs.diamond(6.2, (3, 3, 3)), 0.05, np.random.default_rng(5))
    dense = configs.perturb(configs.diamond(5.0, (3, 3, 3)), 0.05, np.random.default_rng(6))
    mid = configs.perturb(configs.diamond(5.43, (3, 3, 3)), 0.05, np.random.default_rng(7))
    refs = {k: _exact(path, a) for k, a in (("loose", loose), ("dense", dense), ("mid", mid))}
    pot = Potential(param_filename=path)
    bufs = {k: _Buf(1, len(loose)) for k in ("loose", "dense", "mid")}
    try:
        for k, a in (("loose", loose), ("dense", dense), ("mid", mid)):
            bufs[k].put(a.positions)
        _calc_dev(pot, loose, bufs["loose"])     # exact: capacities of the loose structure
        _calc_dev(pot, loose, bufs["loose"])     # speculative, fits
        _calc_dev(pot, dense, bufs["dense"])     # speculative, overflows
        _calc_dev(pot, mid, bufs["mid"])         # speculative, queued after the overflow
        L.check(L.lib().gapx_sync(pot.handle), "sync")
        outs = {k: b.get() for k, b in bufs.items()}
        _calc_dev(pot, dense, bufs["dense"])     # with the grown capacities: fits
        L.check(L.lib().gapx_sync(pot.handle), "sync")
        again = bufs["dense"].get()
    finally:
        for b in bufs.values():
            b.free()
    for k in ("loose", "dense", "mid"):
        assert outs[k]["e"][0] == refs[k][0], k
        assert np.array_equal(outs[k]["f"].reshape(-1, 3), refs[k][1]), k
    assert again["e"][0] == refs["dense"][0]


@pytest.mark.parametrize("graphs", [False, True])
def test_slab_batch_device_geometry(path, graphs):
    """Slabs (non-periodic z) in a batch on device positions: lat_eff comes from the
    device on each call's own extents; positions move between calls."""
    rng = np.random.default_rng(8)
    frames = [configs.perturb(configs.slab111(5.43, 4 + 2 * (k % 3), 10.0), 0.1, rng) for k in range(6)]
    frames += [configs.perturb(configs.diamond(5.43, (2, 2, 2)), 0.1, rng) for k in range(2)]
    S = len(frames)
    off = np.zeros(S + 1, np.int32)
    off[1:] = np.cumsum([len(a) for a in frames])
    N = int(off[-1])
    Z = L.i32(np.concatenate([a.numbers for a in frames]))
    lats = L.f64(np.stack([a.cell for a in frames]))
    pbcs = L.i32(np.stack([a.pbc for a in frames]))
    pot = Potential(param_filename=path)
    pot.set_graphs(graphs)
    buf = _Buf(S, N)
    try:
        for step in range(4):
            pos = np.vstack([a.positions for a in frames]) + rng.normal(0, 0.02, (N, 3)) * (step > 0)
            buf.put(pos)
            L.check(L.lib().gapx_calc_batch_device(pot.handle, S, L.ptr(off), buf.p["pos"], L.ptr(Z), L.ptr(lats),
                                                   L.ptr(pbcs), 0.0, WANT, buf.p["e"], None, buf.p["f"], buf.p["v"],
                                                   None), "gapx_calc_batch_device")
            L.check(L.lib().gapx_sync(pot.handle), "sync")
            out = buf.get()
            moved = [a.copy() for a in frames]
            for s_, a in enumerate(moved):
                a.positions = pos[off[s_]:off[s_ + 1]]
            ref = Potential(param_filename=path).calc_batch(moved)
            assert np.array_equal(out["e"], np.array([r["energy"] for r in ref])), step
            assert np.array_equal(out["f"].reshape(-1, 3), np.vstack([r["force"] for r in ref])), step
    finally:
        buf.free()


def test_split_tuning_is_listed(path):
    """The split-form autotuner times the fused kernel's sparse-point parts for the
    centre-count bucket of a call and lists its choice; the tuned call's results stay
    within the split form's budget of the FP64 ones."""
    at = configs.perturb(configs.diamond(5.43, (4, 4, 4)), 0.1, np.random.default_rng(9))
    pot = Potential(param_filename=path, precision="bf16x3")
    r = pot.calc(at.copy(), force=True)
    n = C.c_int()
    prec, bucket, fparts, g1, g2 = (np.zeros(8, np.int32) for _ in range(5))
    ms1, ms2 = np.zeros(8), np.zeros(8)
    L.check(L.lib().gapx_model_tuning_split(pot.handle, 0, 8, L.ptr(prec), L.ptr(bucket), L.ptr(fparts), L.ptr(g1),
                                            L.ptr(g2), L.ptr(ms1), L.ptr(ms2), C.byref(n)), "tuning_split")
    assert n.value == 1
    assert prec[0] == 1 and bucket[0] == 9                      # 512 centres: 2^8 < 512 <= 2^9
    assert fparts[0] in (-1, 1, 2, 4, 8) and g1[0] == -1 and g2[0] == -1 and ms1[0] > 0
    r64 = Potential(param_filename=path).calc(at.copy(), force=True)
    assert abs(r["energy"] - r64["energy"]) / len(at) < 1e-3
    assert np.sqrt(np.mean((r["force"] - r64["force"]) ** 2)) < 1e-2


@pytest.mark.parametrize("prec", ["bf16x3", "fp16x3"])
def test_split_form_device_calls(path, prec):
    """Repeated device calls of a split form (the speculative launch sequence) give the
    split form's own results -- those of its first (exact) host call bit for bit -- and
    they differ from the FP64 form's, within the split form's budget."""
    at = configs.perturb(configs.diamond(5.43, (4, 4, 4)), 0.1, np.random.default_rng(10))
    r64 = Potential(param_filename=path).calc(at.copy(), force=True)
    rs = Potential(param_filename=path, precision=prec).calc(at.copy(), force=True)
    pot = Potential(param_filename=path, precision=prec)
    buf = _Buf(1, len(at))
    try:
        buf.put(at.positions)
        for _ in range(5):
            _calc_dev(pot, at, buf)
        L.check(L.lib().gapx_sync(pot.handle), "sync")
        out = buf.get()
    finally:
        buf.free()
    f = out["f"].reshape(-1, 3)
    print(prec, "dE/atom vs fp64", abs(out["e"][0] - r64["energy"]) / len(at),
          "F rmse", np.sqrt(np.mean((f - r64["force"]) ** 2)))
    assert out["e"][0] == rs["energy"] and np.array_equal(f, rs["force"])
    assert out["e"][0] != r64["energy"] and not np.array_equal(f, r64["force"])
    assert abs(out["e"][0] - r64["energy"]) / len(at) < 1e-3
    assert np.sqrt(np.mean((f - r64["force"]) ** 2)) < 1e-2


def test_forms_in_turn_on_shared_buffers(path):
    """The bench's pattern: an FP64 model and then a split-form model write the same
    device output buffers, each over many timed device calls; each form's results are
    its own (the split form's differ from FP64's and equal its first exact call's)."""
    at = configs.perturb(configs.diamond(5.43, (4, 4, 4)), 0.1, np.random.default_rng(11))
    buf = _Buf(1, len(at))
    outs = {}
    try:
        buf.put(at.positions)
        for prec in ("fp64", "bf16x3"):
            pot = Potential(param_filename=path, precision=prec)
            for _ in range(3):
                _calc_dev(pot, at, buf)
            L.check(L.lib().gapx_sync(pot.handle), "sync")
            pot.set_timing(True)
            for _ in range(20):
                _calc_dev(pot, at, buf)
            L.check(L.lib().gapx_sync(pot.handle), "sync")
            pot.timing_totals()
            pot.set_timing(False)
            outs[prec] = buf.get()
            del pot
    finally:
        buf.free()
    refs = {p: Potential(param_filename=path, precision=p).calc(at.copy(), force=True) for p in ("fp64", "bf16x3")}
    for p in ("fp64", "bf16x3"):
        print(p, outs[p]["e"][0], refs[p]["energy"])
        assert outs[p]["e"][0] == refs[p]["energy"], p
        assert np.array_equal(outs[p]["f"].reshape(-1, 3), refs[p]["force"]), p
    assert outs["fp64"]["e"][0] != outs["bf16x3"]["e"][0]


@pytest.fixture(scope="module")
def path300(tmp_path_factory):
    """A SOAP model with 300 sparse points: ten chunks of 32 per row block of the fused
    split kernel, so that the stream-K schedule splits row blocks in many ways."""
    train = [configs.perturb(configs.diamond(5.43, (1, 1, 1)), 0.1, configs.rng_for(0, 1200 + k)) for k in range(40)]
    m = configs.soap_model(_desc_fn, train, 300, 0, [configs.soap_string()], "K", alpha_std=0.1)
    p = str(tmp_path_factory.mktemp("k") / "k.xml")
    write_gap_xml(p, m)
    return p


def _run_schedule(path, frames, prec, schedule):
    pot = Potential(param_filename=path, precision=prec)
    L.check(L.lib().gapx_model_set_fused_schedule(pot.handle, schedule), "set_fused_schedule")
    return [pot.calc_batch(frames) for _ in range(2)]


@pytest.mark.parametrize("prec", ["bf16x3", "fp16x3"])
@pytest.mark.parametrize("n_frames", [1, 70])
def test_stream_k_schedule_matches_parts(path300, prec, n_frames):
    """The fused split GP kernel under the stream-K schedule against one sparse-point
    part (G written directly): one 512-atom cell (four row blocks over ten chunks: every
    row block shared by several blocks of the grid) and 70 of them (35,840 centres, more
    units per block than a row block has: whole and shared row blocks).  Energies agree
    to the rounding of FP64 partial sums, forces and virials to that of the FP32 G
    partials; two calls of each schedule are bitwise equal; the results stay within the
    split form's budget of the FP64 ones."""
    frames = [configs.perturb(configs.diamond(5.43, (4, 4, 4)), 0.1, configs.rng_for(3, 50 + k))
              for k in range(n_frames)]
    sk = _run_schedule(path300, frames, prec, -1)
    p1 = _run_schedule(path300, frames, prec, 1)
    for runs in (sk, p1):
        for a, b in zip(*runs):
            assert a["energy"] == b["energy"] and np.array_equal(a["force"], b["force"])
    dE = max(abs(a["energy"] - b["energy"]) / len(f) for a, b, f in zip(sk[0], p1[0], frames))
    dF = max(np.abs(a["force"] - b["force"]).max() for a, b in zip(sk[0], p1[0]))
    dV = max(np.abs(a["virial"] - b["virial"]).max() for a, b in zip(sk[0], p1[0]))
    print(prec, n_frames, "stream-K vs one part: dE/atom", dE, "max dF", dF, "max dV", dV)
    assert dE < 1e-12 and dF < 1e-5 and dV < 1e-3
    r64 = Potential(param_filename=path300).calc(frames[0].copy(), force=True)
    assert abs(sk[0][0]["energy"] - r64["energy"]) / len(frames[0]) < 1e-3
    assert np.sqrt(np.mean((sk[0][0]["force"] - r64["force"]) ** 2)) < 1e-2


def test_fused_schedule_argument_checked(path300):
    pot = Potential(param_filename=path300, precision="bf16x3")
    assert L.lib().gapx_model_set_fused_schedule(pot.handle, -2) != 0
    assert L.lib().gapx_model_set_fused_schedule(pot.handle, 9) != 0
    assert L.lib().gapx_model_set_fused_schedule(pot.handle, 0) == 0
