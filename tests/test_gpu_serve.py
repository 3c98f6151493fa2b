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
