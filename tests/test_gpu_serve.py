"""Serving loop (continuous batching) and batching on the GPU.

Each served result is compared with the same structure evaluated alone
(gapx_calc).  FP64: within 1e-12 eV/atom and 1e-11 eV/A (batching changes only
the split-K factor of the FP64 G GEMM, i.e. rounding).  Split forms: within
1e-12 eV/atom and 1e-11 eV/A (their GEMMs have no split-K, so a row's result
does not depend on the other rows).  Two served FP64 results are also checked
against the oracle to the plain-form tolerance (1e-10 eV/atom, 1e-8 eV/A).
"""
import threading
import time

import numpy as np
import pytest

from oracle import oracle as O
from quip_amd import configs
from quip_amd import _lib as L
from quip_amd.gap_xml import GPCoordinate, write_gap_xml
from quip_amd.potential import Potential
from quip_amd.serve import Server

pytestmark = pytest.mark.gpu


def _desc_fn(d, at):
    return O.soap_descriptors(GPCoordinate("t", d, 2, 1, 0, 4, np.zeros((1, d.dim)), np.zeros(1), np.ones(1)), at)[0]


@pytest.fixture(scope="module")
def sic_model(tmp_path_factory):
    """Two-species model with the serving hyperparameters (n_max 12, l_max 8) and few sparse points."""
    train = [configs.antisite_swaps(configs.perturb(configs.diamond(4.36, (1, 1, 1), species=(14, 6)), 0.1,
                                                    configs.rng_for(4, 500 + k)), 1, configs.rng_for(4, 600 + k))
             for k in range(6)]
    strs = [configs.soap_string(n_max=12, l_max=8, species=(14, 6), centre=14),
            configs.soap_string(n_max=12, l_max=8, species=(14, 6), centre=6)]
    m = configs.soap_model(_desc_fn, train, 24, 4, strs, "GAP_sic_serve", alpha_std=1e-2)
    path = str(tmp_path_factory.mktemp("serve") / "sic.xml")
    write_gap_xml(path, m)
    return m, path


def _submit_mixed(srv, frames, rng):
    """Two client threads submit interleaved halves with random gaps."""
    tickets = [None] * len(frames)

    def client(idx):
        for i in idx:
            tickets[i] = srv.submit(frames[i])
            time.sleep(float(rng.uniform(0, 2e-3)))

    th = [threading.Thread(target=client, args=(list(range(k, len(frames), 2)),)) for k in (0, 1)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    return tickets


@pytest.mark.parametrize("prec", ["fp64", "bf16x3"])
def test_served_results_equal_alone(sic_model, prec):
    m, path = sic_model
    frames = configs.sic_requests(32, n_max=700)
    pot = Potential(param_filename=path, precision=prec)
    with Server(pot, n_slots=2, hbm_budget=8 << 30, max_batch_atoms=6000) as srv:
        tickets = _submit_mixed(srv, frames, np.random.default_rng(3))
        res = [srv.wait(t) for t in tickets]
        st = srv.stats()
    assert st["requests"] == len(frames)
    assert st["batches"] < len(frames), "no request was batched with another"
    alone = Potential(param_filename=path, precision=prec)
    for at, r in zip(frames, res):
        ra = alone.calc(at.copy(), force=True, virial=True)
        n = len(at)
        assert abs(r["energy"] - ra["energy"]) / n < 1e-12
        assert np.abs(r["force"] - ra["force"]).max() < 1e-11
        assert np.abs(r["virial"] - ra["virial"]).max() < 1e-10
        assert r["latency_s"] > 0
    if prec == "fp64":
        for k in (0, int(np.argmax([len(a) for a in frames]))):
            ro = O.gap_calc(m, frames[k])
            assert abs(res[k]["energy"] - ro["energy"]) / len(frames[k]) < 1e-10
            assert np.abs(res[k]["force"] - ro["forces"]).max() < 1e-8


def test_server_admission_by_footprint(sic_model):
    m, path = sic_model
    pot = Potential(param_filename=path)
    big = configs.diamond(4.36, (4, 4, 4), species=(14, 6))
    with Server(pot, n_slots=1, hbm_budget=64 << 20, max_batch_atoms=100000) as srv:
        with pytest.raises(L.GapxError) as e:
            srv.submit(big)
        assert e.value.code == 7   # GAPX_E_CAPACITY
        small = configs.perturb(configs.diamond(4.36, (1, 1, 1), species=(14, 6)), 0.05, configs.rng_for(4, 9))
        r = srv.wait(srv.submit(small))
        ra = Potential(param_filename=path).calc(small.copy(), force=True, virial=True)
        assert abs(r["energy"] - ra["energy"]) < 1e-12 * len(small)


def test_server_reports_errors_per_request(sic_model):
    """A request that fails at evaluation fails alone; the others of its batch complete."""
    m, path = sic_model
    pot = Potential(param_filename=path)
    singular = configs.diamond(4.36, (1, 1, 1), species=(14, 6))
    singular.cell = np.zeros((3, 3))
    huge = configs.diamond(4.36, (1, 1, 1), species=(14, 6))
    huge.cell = np.eye(3) * 1.0e4                 # 2000^3 cells: rejected by the neighbour list
    good = [configs.perturb(configs.diamond(4.36, (2, 2, 2), species=(14, 6)), 0.05, configs.rng_for(4, 20 + k))
            for k in range(3)]
    with Server(pot, n_slots=1, hbm_budget=1 << 30, max_batch_atoms=100000) as srv:
        with pytest.raises(L.GapxError) as e:
            srv.submit(singular)
        assert e.value.code == 6   # GAPX_E_GEOMETRY at submission
        tickets = [srv.submit(a) for a in (good[0], huge, good[1], good[2])]
        with pytest.raises(L.GapxError) as e:
            srv.wait(tickets[1])
        assert e.value.code == 6
        alone = Potential(param_filename=path)
        for t, at in zip([tickets[0], tickets[2], tickets[3]], good):
            r = srv.wait(t)
            assert abs(r["energy"] - alone.calc(at.copy())["energy"]) < 1e-12 * len(at)


def test_session_keeps_neighbour_list_and_results(sic_model):
    """A session resubmitting a moving structure keeps its neighbour list within the skin
    (calc_connect with cutoff_skin, Connection.f95:1085-1127): its results equal, bit for
    bit, those of one model evaluating the same trajectory with the same list cutoff
    (the plain entry point's skin reuse), and fresh builds to rounding -- a reused list's
    distances come from calc_dists (Connection.f95:1393-1395), whose operation order
    differs from the build's, in the reference too.  A bitwise identical resubmission
    takes the resident results; a move beyond half the skin rebuilds the list."""
    m, path = sic_model
    skin = 0.5
    at0 = configs.perturb(configs.diamond(4.36, (2, 2, 2), species=(14, 6)), 0.05, configs.rng_for(4, 31))
    rng = np.random.default_rng(12)
    traj = [at0]
    for _ in range(5):
        a = traj[-1].copy()
        a.positions = a.positions + rng.normal(0.0, 2e-3, a.positions.shape)
        traj.append(a)
    far = traj[-1].copy()
    far.positions = far.positions + rng.normal(0.0, 0.3, far.positions.shape)
    pot = Potential(param_filename=path)
    rc = pot.cutoff() + skin
    with Server(pot, n_slots=2, hbm_budget=8 << 30, max_batch_atoms=6000) as srv:
        with srv.session(skin=skin) as s:
            res = [srv.wait(s.submit(a)) for a in traj]
            again = srv.wait(s.submit(traj[-1].copy()))
            st = s.stats()
            moved = srv.wait(s.submit(far))
            st2 = s.stats()
    n = len(traj)
    assert st["requests"] == n + 1
    assert st["nl_reused"] == n - 1 and st["results_reused"] == 1
    assert st2["nl_reused"] == st["nl_reused"]                     # a move beyond half the skin rebuilt it
    ref = Potential(param_filename=path)
    for k, (a, r) in enumerate(zip(traj + [far], res + [moved])):
        rr = ref.calc(a.copy(), force=True, virial=True, cutoff=rc)
        assert ref.nl_rebuilt() == (k == 0 or k == n)
        assert r["energy"] == rr["energy"], k
        assert np.array_equal(r["force"], rr["force"]), k
        assert np.array_equal(r["virial"], rr["virial"]), k
        fresh = Potential(param_filename=path).calc(a.copy(), force=True, cutoff=rc)
        assert abs(r["energy"] - fresh["energy"]) < 1e-12 * len(a), k
        assert np.abs(r["force"] - fresh["force"]).max() < 1e-11, k
    assert again["energy"] == res[-1]["energy"] and np.array_equal(again["force"], res[-1]["force"])


def test_submit_while_another_thread_waits(sic_model):
    """Tickets are erased by key after a wait: submissions from one thread while another
    waits (which can rehash the ticket table) leave every wait valid."""
    m, path = sic_model
    pot = Potential(param_filename=path)
    frames = [configs.perturb(configs.diamond(4.36, (1, 1, 1), species=(14, 6)), 0.05, configs.rng_for(4, 50 + k))
              for k in range(40)]
    with Server(pot, n_slots=2, hbm_budget=4 << 30, max_batch_atoms=4000) as srv:
        first = srv.submit(frames[0])
        out = {}
        th = threading.Thread(target=lambda: out.setdefault("r", srv.wait(first)))
        th.start()
        tickets = [srv.submit(a) for a in frames[1:]]       # many inserts while the waiter blocks
        th.join()
        rest = [srv.wait(t) for t in tickets]
    ref = Potential(param_filename=path)
    for a, r in zip(frames, [out["r"]] + rest):
        assert abs(r["energy"] - ref.calc(a.copy())["energy"]) < 1e-12 * len(a)


@pytest.fixture(scope="module")
def sic_full(tmp_path_factory):
    """The served model at full size: Si and C centres, n_max 12, l_max 8 (d = 2701),
    2 x 2000 sparse points from GPU descriptors of 64 seeded training cells."""
    tmp = tmp_path_factory.mktemp("sicfull")
    cache = {}

    def fn(d, at):
        if d.string not in cache:
            from quip_amd.gap_xml import GAPModel
            mm = GAPModel("D", 1527075646, np.zeros(119),
                          [GPCoordinate("D1", d, 2, 1.0, 0.0, 1.0, np.zeros((1, d.dim)), np.zeros(1), np.ones(1))])
            p = str(tmp / ("d%d.xml" % len(cache)))
            write_gap_xml(p, mm)
            cache[d.string] = Potential(param_filename=p)
        return cache[d.string].descriptors(at)[0]

    m = configs.model_sic(fn)
    path = str(tmp / "sic.xml")
    write_gap_xml(path, m)
    return m, path


@pytest.mark.parametrize("prec", ["fp64", "bf16x3"])
def test_cfg4_scale_served_equals_alone(sic_full, prec):
    """Requests of the serving workload (Si/C, N log-uniform on [16, 4096], bulk / slab /
    vacancies) arrive as a Poisson stream on the full model; each served result equals
    the structure evaluated alone (FP64 to 1e-12 eV/atom and 1e-11 eV/A; the split form's
    large descriptors run the split GEMMs, whose element results do not depend on the
    batch, to the same bounds)."""
    m, path = sic_full
    frames = configs.sic_requests(20)
    assert max(len(a) for a in frames) > 1000
    pot = Potential(param_filename=path, precision=prec)
    rng = np.random.default_rng(17)
    gaps = rng.exponential(2e-3, len(frames))
    with Server(pot, n_slots=2, hbm_budget=64 << 30, max_batch_atoms=16384) as srv:
        tickets = []
        for a, g in zip(frames, gaps):
            time.sleep(float(g))
            tickets.append(srv.submit(a))
        res = [srv.wait(t) for t in tickets]
        st = srv.stats()
    print("cfg4", prec, "requests", st["requests"], "batches", st["batches"], flush=True)
    alone = Potential(param_filename=path, precision=prec)
    for a, r in zip(frames, res):
        ra = alone.calc(a.copy(), force=True, virial=True)
        assert abs(r["energy"] - ra["energy"]) / len(a) < 1e-12
        assert np.abs(r["force"] - ra["force"]).max() < 1e-11
        assert np.abs(r["virial"] - ra["virial"]).max() < 1e-9
