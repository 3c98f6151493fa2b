"""HIP path (libgapx.so through its C-ABI) against the oracle and the reference fixtures.

Tolerances: FP64 plain form 1e-10 eV/atom in energy and 1e-8 eV/A in every
force component; bf16x3 / fp16x3 forms 1 meV/atom and 10 meV/A force RMSE.
Neighbour tables are compared bit for bit (integers and stored distances).
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from quip_amd import configs
from quip_amd import _lib as L
from quip_amd.atoms import Atoms, read_xyz
from quip_amd.gap_xml import GAPModel, GPCoordinate, Descriptor, read_gap_xml, write_gap_xml
from quip_amd.potential import Potential, connect

pytestmark = pytest.mark.gpu

E_TOL = 1e-10      # eV/atom
F_TOL = 1e-8       # eV/A


def _desc_fn(d, at):
    return O.soap_descriptors(GPCoordinate("t", d, 2, 1, 0, 4, np.zeros((1, d.dim)), np.zeros(1), np.ones(1)), at)[0]


@pytest.fixture(scope="module")
def models(tmp_path_factory):
    out = tmp_path_factory.mktemp("models")
    res = {}
    for name, builder in (("si64", configs.model_si64), ("si4096", configs.model_si4096)):
        m = builder(_desc_fn)
        path = str(out / (name + ".xml"))
        write_gap_xml(path, m)
        res[name] = (m, path)
    # two-species model with the SiC hyperparameters (n_max 12, l_max 8) and few sparse points
    train = [configs.antisite_swaps(configs.perturb(configs.diamond(4.36, (1, 1, 1), species=(14, 6)), 0.1,
                                                    configs.rng_for(2, 500 + k)), 1, configs.rng_for(2, 600 + k))
             for k in range(6)]
    strs = [configs.soap_string(n_max=12, l_max=8, species=(14, 6), centre=14),
            configs.soap_string(n_max=12, l_max=8, species=(14, 6), centre=6)]
    m = configs.soap_model(_desc_fn, train, 20, 2, strs, "GAP_sic_small", alpha_std=1e-2)
    path = str(out / "sic_small.xml")
    write_gap_xml(path, m)
    res["sic_small"] = (m, path)
    return res


def _geoms():
    rng = np.random.default_rng(5)
    yield "si64", configs.si64(), 5.0
    yield "small_cell", Atoms([14, 14], [[0, 0, 0], [1.1, 1.2, 0.9]], [[2.6, 0.1, 0], [0.3, 2.4, 0], [0.2, -0.1, 2.9]]), 4.0
    yield "skew", Atoms(np.full(12, 14), rng.uniform(0, 6, (12, 3)), [[6.0, 0, 0], [2.5, 5.5, 0], [1.0, 1.5, 6.5]]), 3.5
    yield "slab_nonperiodic", Atoms(np.full(10, 14), rng.uniform(0, 5, (10, 3)), np.diag([5.0, 5.0, 5.0]),
                                    (True, True, False)), 3.0
    yield "offset", Atoms(np.full(6, 14), rng.uniform(0, 5, (6, 3)) + np.array([40.0, -17.0, 3.0]),
                          np.diag([5.0, 5.0, 5.0])), 3.0
    yield "si4096", configs.si4096(), 5.0
    yield "si4096_skin", configs.si4096(), 5.5
    # about 50 neighbour2 entries per atom: the ranking takes its path beyond 32 entries
    yield "si64_rc8", configs.si64(), 8.0
    # more than 16384 atoms: the one-block count scans pass through LDS in two chunks
    yield "si17576", configs.perturb(configs.diamond(5.43, (13, 13, 13)), 0.1, configs.rng_for(0, 77)), 5.0
    # more than 65536 atoms: the count scans take the tile-sum path
    yield "si74088", configs.perturb(configs.diamond(5.43, (21, 21, 21)), 0.1, configs.rng_for(0, 78)), 5.0
    yield "empty", Atoms(np.zeros(0, np.int32), np.zeros((0, 3)), np.eye(3) * 5.0), 3.0
    yield "single_atom", Atoms([14], [[0.3, 0.2, 0.1]], np.eye(3) * 2.2), 5.0


@pytest.mark.parametrize("case", [g[0] for g in _geoms()])
def test_neighbour_tables_bit_exact(case):
    name, at, rc = [g for g in _geoms() if g[0] == case][0]
    print("case", name, len(at), rc, flush=True)
    a = connect(at, rc)
    b = O.neighbour_list(at, rc)
    for k in ("n1_off", "n1_j", "n1_shift", "n2_off", "n2_j", "n2_idx", "map_shift", "cells"):
        assert np.array_equal(a[k], b[k]), k
    assert np.array_equal(a["n1_r"].view(np.int64), b["n1_r"].view(np.int64))


def test_soap_vectors_match_reference_sparse_points(golden, tmp_path):
    g = np.load(os.path.join(golden, "si_soap_cur_points.npz"))
    d = Descriptor.from_string("soap atom_sigma=0.5 central_weight=1.0 cutoff=4.0 cutoff_transition_width=1.0 "
                               "l_max=8 n_max=8")
    m = GAPModel("T", 1527075646, np.zeros(119),
                 [GPCoordinate("T1", d, 2, 1.0, 0.0, 4.0, g["sparsex"][:2], np.zeros(2), np.ones(2))])
    write_gap_xml(str(tmp_path / "t.xml"), m)
    pot = Potential(param_filename=str(tmp_path / "t.xml"))
    X, off = [], 0
    for f, n in enumerate(g["natoms"]):
        at = Atoms(g["numbers"][off:off + n], g["positions"][off:off + n], g["lattices"][f])
        off += n
        X.append(pot.descriptors(at)[0])
    X = np.vstack(X)
    assert np.abs(X[g["index"] - 1] - g["sparsex"]).max() < 1e-12


def test_distance_2b_water_reference(golden):
    pot = Potential(param_filename=os.path.join(golden, "gap_water", "GAP.xml"))
    at = read_xyz(os.path.join(golden, "gap_water", "gap_sample.xyz"), 0)
    e_ref, f_ref = at.info["energy"], at.arrays["force"].copy()
    r = pot.calc(at, force=True, virial=True)
    assert abs(r["energy"] - e_ref) < 1e-7          # reference test tolerance 1e-5
    assert np.abs(r["force"] - f_ref).max() < 1e-7  # reference test tolerance 1e-6
    ro = O.gap_calc(read_gap_xml(os.path.join(golden, "gap_water", "GAP.xml")), at)
    assert np.abs(r["virial"] - ro["virial"]).max() < 1e-7


def test_cells_smaller_than_cutoff(golden):
    ref = json.load(open(os.path.join(golden, "potential_cell.json")))
    pot = Potential(param_filename=os.path.join(golden, "gap_water", "GAP.xml"))
    for c, e in zip(ref["cell_sizes"], ref["energies"]):
        at = Atoms(ref["numbers"], ref["positions"], np.eye(3) * c)
        assert abs(pot.calc(at)["energy"] - e) < 1e-8


def _compare(r, ro, N, e_tol=E_TOL, f_tol=F_TOL):
    assert abs(r["energy"] - ro["energy"]) / N < e_tol
    assert np.abs(r["force"] - ro["forces"]).max() < f_tol
    assert np.abs(r["virial"] - ro["virial"]).max() < f_tol * 10
    assert np.abs(r["local_energy"] - ro["local_energy"]).max() < e_tol * 10
    assert np.abs(r["local_virial"] - ro["local_virial"]).max() < f_tol * 10


@pytest.mark.parametrize("name", ["si64", "si4096", "sic_small"])
def test_fp64_matches_oracle(models, name):
    m, path = models[name]
    at = {"si64": configs.si64(), "si4096": configs.si4096(),
          "sic_small": configs.antisite_swaps(configs.perturb(configs.diamond(4.36, (2, 2, 2), species=(14, 6)), 0.1,
                                                              np.random.default_rng(9)), 3, np.random.default_rng(10))}[name]
    print("case", name, len(at), flush=True)
    pot = Potential(param_filename=path)
    r = pot.calc(at.copy(), force=True, virial=True, local_energy=True, local_virial=True)
    ro = O.gap_calc(m, at)
    _compare(r, ro, len(at))


def test_batch_of_si4096_sized_structures(models):
    """The batching form on two si4096-sized structures (one seed each, the first si4096):
    FP64 against the oracle structure by structure and against the plain form's calls;
    bf16x3 within the split forms' bound of FP64."""
    m, path = models["si4096"]
    frames = configs.si4096_batch(2)
    print("case si4096x2", [len(a) for a in frames], flush=True)
    rb = Potential(param_filename=path).calc_batch([a.copy() for a in frames])
    rs = Potential(param_filename=path, precision="bf16x3").calc_batch([a.copy() for a in frames])
    plain = Potential(param_filename=path)
    for a, r, s in zip(frames, rb, rs):
        N = len(a)
        ro = O.gap_calc(m, a)
        assert abs(r["energy"] - ro["energy"]) / N < E_TOL
        assert np.abs(r["force"] - ro["forces"]).max() < F_TOL
        rp = plain.calc(a.copy(), force=True, virial=True)
        assert abs(r["energy"] - rp["energy"]) / N < 1e-13
        assert np.abs(r["force"] - rp["force"]).max() < 1e-12
        assert np.allclose(np.asarray(r["virial"]).reshape(3, 3), np.asarray(rp["virial"]).reshape(3, 3),
                           rtol=1e-11, atol=1e-10)
        assert abs(s["energy"] - r["energy"]) / N < 1e-5
        assert np.sqrt(np.mean((s["force"] - r["force"]) ** 2)) < 1e-4


@pytest.mark.parametrize("nl", [(4, 4), (5, 4), (6, 4), (8, 4), (6, 6), (10, 6), (12, 6), (8, 8), (10, 8)])
def test_descriptor_variants_match_oracle(nl, tmp_path):
    """Every (n_max, l_max) the specialised kernels cover, on a cell with an odd number of
    centres (the forward's two-centres-per-wave form leaves its last wave half empty)."""
    n, lm = nl
    train = [configs.perturb(configs.diamond(5.43, (1, 1, 1)), 0.1, configs.rng_for(3, 700 + k)) for k in range(3)]
    m = configs.soap_model(_desc_fn, train, 16, 3, [configs.soap_string(n_max=n, l_max=lm)], "GAP_nl%d_%d" % nl,
                           alpha_std=1e-2)
    path = str(tmp_path / "m.xml")
    write_gap_xml(path, m)
    full = configs.si64()
    at = Atoms(full.numbers[1:], full.positions[1:], full.cell)   # 63 atoms
    print("case", nl, len(at), flush=True)
    r = Potential(param_filename=path).calc(at.copy(), force=True, virial=True, local_energy=True,
                                            local_virial=True)
    _compare(r, O.gap_calc(m, at), len(at))


def test_skin_and_nonperiodic(models):
    """A neighbour cutoff above the model cutoff (QUIP's cutoff_skin) and a non-periodic axis."""
    m, path = models["si64"]
    pot = Potential(param_filename=path)
    at = configs.si64()
    r = pot.calc(at.copy(), force=True, virial=True, local_energy=True, local_virial=True, cutoff=5.7)
    _compare(r, O.gap_calc(m, at, nl_cutoff=5.7), len(at))
    slab = configs.si64(3)
    slab.pbc = np.array([True, True, False])
    slab.cell[2, 2] = 30.0
    r = pot.calc(slab.copy(), force=True, virial=True, local_energy=True, local_virial=True)
    _compare(r, O.gap_calc(m, slab), len(slab))


@pytest.mark.parametrize("prec", ["bf16x3", "fp16x3"])
@pytest.mark.parametrize("nl", [(6, 6), (8, 4), (5, 4), (3, 3), (4, 4)])
def test_split_padded_descriptor_widths(nl, prec, tmp_path):
    """The fused split GP kernel on descriptors whose 32-padded width is below the width the
    kernel is built for (d 148 -> 160 of 256, 181 -> 192 of 256, 76 -> 96 of 128, 25 -> 32 of
    64), and at d 51 -> 64 of 64: its staging lanes past the padded width read in-bounds values
    that must meet zero X components, and its rows past d give G columns that are not
    written.  70 sparse points make three chunks, the last one partial.  The split error on
    these models is about 1e-7; a padding lane that reached a result would be far above the
    bounds checked here."""
    n, lm = nl
    train = [configs.perturb(configs.diamond(5.43, (1, 1, 1)), 0.1, configs.rng_for(3, 800 + k)) for k in range(10)]
    m = configs.soap_model(_desc_fn, train, 70, 3, [configs.soap_string(n_max=n, l_max=lm)], "GAP_pad%d_%d" % nl,
                           alpha_std=1e-2)
    path = str(tmp_path / "m.xml")
    write_gap_xml(path, m)
    at = configs.perturb(configs.diamond(5.43, (4, 4, 4)), 0.1, configs.rng_for(3, 900))
    print("case", nl, prec, len(at), flush=True)
    r = Potential(param_filename=path, precision=prec).calc(at.copy(), force=True, virial=True, local_energy=True)
    ro = O.gap_calc(m, at)
    de = abs(r["energy"] - ro["energy"]) / len(at)
    frmse = np.sqrt(np.mean((r["force"] - ro["forces"]) ** 2))
    print(nl, prec, "dE/atom", de, "F rmse", frmse)
    assert de < 1e-5 and frmse < 1e-4
    assert np.abs(r["local_energy"] - ro["local_energy"]).max() < 1e-4


@pytest.mark.parametrize("prec", ["bf16x3", "fp16x3"])
def test_split_precision_within_budget(models, prec):
    m, path = models["si4096"]
    at = configs.si4096()
    pot = Potential(param_filename=path, precision=prec)
    r = pot.calc(at.copy(), force=True, virial=True, local_energy=True)
    ro = O.gap_calc(m, at)
    de = abs(r["energy"] - ro["energy"]) / len(at)
    frmse = np.sqrt(np.mean((r["force"] - ro["forces"]) ** 2))
    print(prec, "dE/atom", de, "F rmse", frmse)
    assert de < 1e-3 and frmse < 1e-2
    # per-atom energies sit on the right atoms (catches accumulator layout errors)
    assert np.abs(r["local_energy"] - ro["local_energy"]).max() < 1e-4


def test_bitwise_repeatable(models):
    m, path = models["si4096"]
    pot = Potential(param_filename=path)
    at = configs.si4096()
    a = pot.calc(at.copy(), force=True, virial=True, local_virial=True)
    b = pot.calc(at.copy(), force=True, virial=True, local_virial=True)
    assert a["energy"] == b["energy"]
    assert np.array_equal(a["force"], b["force"]) and np.array_equal(a["virial"], b["virial"])
    assert np.array_equal(a["local_virial"], b["local_virial"])


def test_one_pass_neighbour_search(models):
    """The neighbour search runs once into per-atom slots sized by the previous build.
    A first build (exact count and fill), a repeat (one pass), a denser structure
    that overflows the slots (second pass fills the exact tables) and a repeat of
    that one all give the results of a fresh model bit for bit."""
    m, path = models["si64"]
    pot = Potential(param_filename=path)
    sparse = configs.si64()
    sparse.positions = sparse.positions * 1.3
    sparse.cell = sparse.cell * 1.3
    dense = configs.si64()

    def fresh(at):
        return Potential(param_filename=path).calc(at.copy(), force=True, virial=True, local_virial=True)

    def same(a, b):
        assert a["energy"] == b["energy"]
        assert np.array_equal(a["force"], b["force"]) and np.array_equal(a["virial"], b["virial"])
        assert np.array_equal(a["local_virial"], b["local_virial"])

    for at in (sparse, sparse, dense, dense):
        print("one-pass search", len(at), at.cell[0, 0], flush=True)
        same(pot.calc(at.copy(), force=True, virial=True, local_virial=True), fresh(at))


def _same(rb, r, N):
    # a batch changes GEMM 2's split-K factor (chosen from the centre count), so
    # a structure alone and in a batch agree to rounding, not bit for bit
    assert abs(rb["energy"] - r["energy"]) / max(N, 1) < 1e-12
    assert np.abs(rb["force"] - r["force"]).max() < 1e-11
    assert np.abs(rb["virial"] - r["virial"]).max() < 1e-10


def test_batch_matches_single(models):
    m, path = models["si4096"]
    pot = Potential(param_filename=path)
    frames = configs.batch512(n_struct=12)
    assert any(not all(a.pbc) for a in frames)
    res = pot.calc_batch(frames, want=("energy", "force", "virial", "local_energy", "local_virial"))
    for at, rb in zip(frames, res):
        r = pot.calc(at.copy(), force=True, virial=True, local_energy=True, local_virial=True)
        _same(rb, r, len(at))
        assert np.abs(rb["local_energy"] - r["local_energy"]).max() < 1e-12
        assert np.abs(rb["local_virial"] - r["local_virial"]).max() < 1e-10
        assert abs(rb["local_energy"].sum() - rb["energy"]) < 1e-9


def test_batch512_full_size(models):
    """The full 512-structure batch (136k atoms) in one pass: sampled structures
    against single calls and the oracle; size-independent checks on all of them."""
    m, path = models["si4096"]
    pot = Potential(param_filename=path)
    frames = configs.batch512()
    res = pot.calc_batch(frames, want=("energy", "force", "virial", "local_energy", "local_virial"))
    rng = np.random.default_rng(11)
    for s in rng.choice(len(frames), 12, replace=False):
        at = frames[s]
        _same(res[s], pot.calc(at.copy(), force=True, virial=True), len(at))
    for s in (0, int(np.argmax([len(a) for a in frames])), [k for k, a in enumerate(frames) if not all(a.pbc)][0]):
        at = frames[s]
        ro = O.gap_calc(m, at)
        _compare(res[s], ro, len(at))
    for at, r in zip(frames, res):
        assert np.isfinite(r["energy"]) and np.isfinite(r["force"]).all()
        assert np.abs(r["force"].sum(axis=0)).max() < 1e-9 * len(at)   # no net force
        assert np.abs(r["virial"] - r["virial"].T).max() < 1e-8         # rotation invariance
    # split precision over the whole batch (one sparse-point part per centre block)
    # against the FP64 batch: 1 meV/atom, 10 meV/A force RMSE
    mixed = Potential(param_filename=path, precision="bf16x3").calc_batch(frames)
    de = max(abs(a["energy"] - b["energy"]) / len(at) for at, a, b in zip(frames, mixed, res))
    fr = np.sqrt(np.mean(np.concatenate([(a["force"] - b["force"]).ravel() for a, b in zip(mixed, res)]) ** 2))
    print("batch512 bf16x3 max dE/atom", de, "force rmse", fr)
    assert de < 1e-3 and fr < 1e-2


@pytest.fixture(scope="module")
def sic_model(tmp_path_factory):
    """The two-species precision-form model: SOAP n_max 12, l_max 8 (d = 2701), 2 x 2000 sparse points."""
    m = configs.model_sic(_desc_fn)
    path = str(tmp_path_factory.mktemp("sic") / "sic.xml")
    write_gap_xml(path, m)
    return m, path


def test_sic_model_matches_oracle(sic_model):
    """Full two-species model on a 512-atom SiC cell with antisite swaps, FP64 against the oracle."""
    m, path = sic_model
    rng = configs.rng_for(2, 31)
    at = configs.antisite_swaps(configs.perturb(configs.diamond(4.36, (4, 4, 4), species=(14, 6)), 0.1, rng), 6, rng)
    r = Potential(param_filename=path).calc(at.copy(), force=True, virial=True, local_energy=True, local_virial=True)
    _compare(r, O.gap_calc(m, at), len(at))


def test_sic32768_precision_forms(sic_model):
    """The 32768-atom SiC precision form: split forms against FP64 within 1 meV/atom and
    10 meV/A force RMSE; size-independent invariants of the FP64 result."""
    m, path = sic_model
    at = configs.sic32768()
    r64 = Potential(param_filename=path).calc(at.copy(), force=True, virial=True)
    assert np.abs(r64["force"].sum(axis=0)).max() < 1e-8 * len(at)
    assert np.abs(r64["virial"] - r64["virial"].T).max() < 1e-7
    shifted = at.copy()
    shifted.positions = shifted.positions + at.cell[0]          # a whole lattice vector: same structure
    rs = Potential(param_filename=path).calc(shifted, force=True)
    assert abs(rs["energy"] - r64["energy"]) / len(at) < 1e-12
    assert np.abs(rs["force"] - r64["force"]).max() < 1e-10
    for prec in ("bf16x3", "fp16x3"):
        r = Potential(param_filename=path, precision=prec).calc(at.copy(), force=True)
        de = abs(r["energy"] - r64["energy"]) / len(at)
        fr = float(np.sqrt(np.mean((r["force"] - r64["force"]) ** 2)))
        print("sic32768", prec, "dE/atom", de, "force rmse", fr)
        assert de < 1e-3 and fr < 1e-2


@pytest.mark.parametrize("prec", ["bf16x3", "fp16x3"])
def test_split_gemm_tile_shapes_agree(sic_model, prec):
    """A 64-atom SiC cell alone (its split GEMMs run on 64 x 64 tiles) and batched with
    an 8192-atom one (the batch's GEMMs run on 128 x 128 LDS tiles): both kernels issue
    the MFMA passes of an element in the same order, so the small cell's results agree
    to the rounding of the FP64 sums."""
    m, path = sic_model
    small = configs.antisite_swaps(configs.perturb(configs.diamond(4.36, (2, 2, 2), species=(14, 6)), 0.1,
                                                   configs.rng_for(2, 41)), 3, configs.rng_for(2, 42))
    big = configs.perturb(configs.diamond(4.36, (8, 8, 16), species=(14, 6)), 0.1, configs.rng_for(2, 43))
    pot = Potential(param_filename=path, precision=prec)
    alone = pot.calc(small.copy(), force=True, virial=True)
    rb = pot.calc_batch([small, big])[0]
    print("tile shapes", prec, abs(rb["energy"] - alone["energy"]), np.abs(rb["force"] - alone["force"]).max())
    assert abs(rb["energy"] - alone["energy"]) / len(small) < 1e-12
    assert np.abs(rb["force"] - alone["force"]).max() < 1e-11
    assert np.abs(rb["virial"] - alone["virial"]).max() < 1e-10


def test_skin_reuse(models):
    """calc_connect with a skin: small moves keep the list (distances recomputed as
    calc_dists does) and match the oracle's fresh build; a move of half the skin
    or more rebuilds it."""
    m, path = models["si64"]
    pot = Potential(param_filename=path)
    at = configs.si64(3)
    rc = pot.cutoff() + 0.6
    rng = np.random.default_rng(21)
    pot.calc(at.copy(), force=True, cutoff=rc)
    assert pot.nl_rebuilt()
    for step in range(4):
        at.positions = at.positions + rng.normal(0.0, 0.01, at.positions.shape)
        r = pot.calc(at.copy(), force=True, virial=True, local_energy=True, local_virial=True, cutoff=rc)
        assert not pot.nl_rebuilt()
        _compare(r, O.gap_calc(m, at, nl_cutoff=rc), len(at))
    at.positions[5] += np.array([0.4, 0.0, 0.0])
    r = pot.calc(at.copy(), force=True, virial=True, local_energy=True, local_virial=True, cutoff=rc)
    assert pot.nl_rebuilt()
    _compare(r, O.gap_calc(m, at, nl_cutoff=rc), len(at))
    # no skin: every call builds
    pot.calc(at.copy(), force=True)
    assert pot.nl_rebuilt()
    pot.calc(at.copy(), force=True)
    assert pot.nl_rebuilt()


def test_quip_cli_trajectory(models, tmp_path):
    """The quip-style front end over a three-frame trajectory (E F V L with the
    default 0.5 A skin): every printed energy and every AT frame's forces and
    local energies match the oracle."""
    import io
    from quip_amd import quip as Q
    from quip_amd.atoms import write_xyz
    m, path = models["si64"]
    rng = np.random.default_rng(31)
    frames = [configs.si64(3)]
    for k in range(2):
        nxt = frames[-1].copy()
        nxt.positions = nxt.positions + rng.normal(0.0, 0.01, nxt.positions.shape)
        frames.append(nxt)
    src = str(tmp_path / "traj.xyz")
    write_xyz(src, frames)
    out = io.StringIO()
    Q.run(["atoms_filename=" + src, "param_filename=" + path, "E", "F", "V", "L"], stdout=out)
    text = out.getvalue().splitlines()
    energies = [float(ln.split("=")[1]) for ln in text if ln.startswith("Energy=")]
    body = "\n".join(ln[3:] for ln in text if ln.startswith("AT "))
    back = read_xyz(io.StringIO(body))
    assert len(energies) == 3 and len(back) == 3
    for at, e, b in zip(frames, energies, back):
        ro = O.gap_calc(m, at)
        assert abs(e - ro["energy"]) / len(at) < E_TOL
        assert np.abs(b.arrays["force"] - ro["forces"]).max() < 1e-7      # printed with %16.8f
        assert np.abs(b.arrays["local_energy"] - ro["local_energy"]).max() < 1e-7


def _lammps_view(at, rc, rng):
    """A periodic cell as a LAMMPS pair style sees it: the cell's atoms as local
    atoms, every periodic image within rc of one of them as a ghost, and each local
    atom's full neighbour list (local and ghost atoms within rc, 1-based), with the
    local atoms listed in a shuffled order."""
    import itertools
    n = len(at)
    gpos, gZ, gown = [], [], []
    for sh in itertools.product((-1, 0, 1), repeat=3):
        if sh == (0, 0, 0):
            continue
        img = at.positions + np.array(sh, float) @ at.cell
        dmin = np.linalg.norm(img[:, None, :] - at.positions[None, :, :], axis=2).min(axis=1)
        keep = np.nonzero(dmin < rc)[0]
        gpos.append(img[keep]); gZ.append(at.numbers[keep]); gown.append(keep)
    x = np.vstack([at.positions] + gpos)
    Z = np.concatenate([at.numbers] + gZ)
    owner = np.concatenate([np.arange(n)] + gown)
    ilist = rng.permutation(n)
    num, neigh = [], []
    for i in ilist:
        d = np.linalg.norm(x - x[i], axis=1)
        js = np.nonzero((d < rc) & (np.arange(len(x)) != i))[0]
        num.append(len(js))
        neigh.extend(rng.permutation(js) + 1)
    return x, Z, owner, ilist, np.array(num), np.array(neigh)


def test_lammps_entry_matches_periodic(models):
    """gapx_calc_lammps on the ghost-atom view of a periodic cell (SOAP + distance_2b
    model): with the ghosts' forces, local energies and local virials folded onto
    their owners (LAMMPS's reverse communication) the results are the periodic
    ones of the oracle; a vacuum region returns zeros; overlapping atoms and bad
    indices are errors."""
    m, path = models["si64"]
    pot = Potential(param_filename=path)
    at = configs.si64()
    rc = pot.cutoff() + 0.5            # the pair style's cutoff plus its neighbour skin
    rng = np.random.default_rng(51)
    x, Z, owner, ilist, num, neigh = _lammps_view(at, rc, rng)
    nl, ng = len(at), len(x) - len(at)
    print("lammps view", nl, ng, len(neigh), flush=True)
    r = pot.calc_lammps(nl, ng, Z, ilist, num, neigh, x, lattice=at.cell)
    ro = O.gap_calc(m, at)
    F = np.zeros((nl, 3)); np.add.at(F, owner, r["force"])
    LE = np.zeros(nl); np.add.at(LE, owner, r["local_energy"])
    LV = np.zeros((nl, 9)); np.add.at(LV, owner, r["local_virial"])
    assert abs(r["energy"] - ro["energy"]) / nl < E_TOL
    assert abs(r["local_energy"].sum() - r["energy"]) < 1e-9
    assert np.abs(F - ro["forces"]).max() < F_TOL
    assert np.abs(r["virial"] - ro["virial"]).max() < F_TOL * 10
    assert np.abs(LE - ro["local_energy"]).max() < E_TOL * 10
    assert np.abs(LV - ro["local_virial"]).max() < F_TOL * 10
    assert np.abs(r["local_energy"][nl:]).max() > 0     # ghosts carry their share of the pair terms
    # the calc_connect path afterwards is unaffected
    rp = pot.calc(at.copy(), force=True)
    assert abs(rp["energy"] - ro["energy"]) / nl < E_TOL
    # vacuum region
    rv = pot.calc_lammps(0, 3, Z[:3], [], [], [], x[:3])
    assert rv["energy"] == 0.0 and not rv["force"].any()
    # errors: two atoms on top of each other, a neighbour index out of range
    x2 = x.copy(); x2[ilist[0]] = x2[neigh[0] - 1]
    with pytest.raises(L.GapxError) as e:
        pot.calc_lammps(nl, ng, Z, ilist, num, neigh, x2)
    assert e.value.code == 6
    bad = neigh.copy(); bad[3] = len(x) + 1
    with pytest.raises(L.GapxError) as e:
        pot.calc_lammps(nl, ng, Z, ilist, num, bad, x)
    assert e.value.code == 1


def _eq(a, b):
    assert a["energy"] == b["energy"]
    assert np.array_equal(a["force"], b["force"])
    assert np.array_equal(a["virial"], b["virial"])


@pytest.mark.parametrize("prec", ["fp64", "bf16x3"])
def test_graph_replay_matches_eager(models, prec):
    """The captured launch sequence, replayed several times with eager work and
    synchronisation in between, gives the eager results bit for bit: identical
    inputs (replay), small displacements (replay on new data), alternating
    structures (re-capture) and a batch."""
    m, path = models["si4096"]
    eager = Potential(param_filename=path, precision=prec)
    graph = Potential(param_filename=path, precision=prec)
    graph.set_graphs(True)
    for p in (eager, graph):   # two models: fixed tile rules, so that neither tunes differently
        L.check(L.lib().gapx_model_set_autotune(p.handle, 0), "set_autotune")
    a = configs.si64(1)
    b = configs.si64(2)
    rng = np.random.default_rng(8)
    for rep in range(4):
        at = a.copy()
        at.positions = at.positions + rng.normal(0.0, 1e-4, at.positions.shape) * (rep > 1)
        ref = eager.calc(at.copy(), force=True, virial=True)
        _eq(graph.calc(at.copy(), force=True, virial=True), ref)
        _eq(graph.calc(at.copy(), force=True, virial=True), ref)        # replay
        _eq(graph.calc(b.copy(), force=True, virial=True), eager.calc(b.copy(), force=True, virial=True))
    frames = configs.batch512(n_struct=6)
    ref = eager.calc_batch(frames)
    for rep in range(3):
        for x, y in zip(graph.calc_batch(frames), ref):
            _eq(x, y)


def test_finite_differences_on_gpu(models):
    m, path = models["si64"]
    pot = Potential(param_filename=path)
    at = configs.si64(5)
    r = pot.calc(at.copy(), force=True)
    h = 1e-5
    for i, k in ((0, 0), (7, 1), (33, 2)):
        p, q = at.copy(), at.copy()
        p.positions[i, k] += h
        q.positions[i, k] -= h
        fd = -(pot.calc(p)["energy"] - pot.calc(q)["energy"]) / (2 * h)
        assert abs(fd - r["force"][i, k]) < 1e-7


def test_rotation_translation_invariance(models):
    m, path = models["si64"]
    pot = Potential(param_filename=path)
    at = configs.si64(2)
    e0 = pot.calc(at.copy(), force=True)
    R = np.array([[0.36, 0.48, -0.8], [-0.8, 0.6, 0.0], [0.48, 0.64, 0.6]])
    rot = Atoms(at.numbers, at.positions @ R.T + np.array([1.0, -2.0, 0.5]), at.cell @ R.T)
    e1 = pot.calc(rot, force=True)
    assert abs(e1["energy"] - e0["energy"]) < 1e-10
    assert np.abs(e1["force"] - e0["force"] @ R.T).max() < 1e-9


def _hip():
    return C.CDLL("libamdhip64.so")


@pytest.mark.parametrize("case", ["si64", "slab"])
def test_device_resident_entry_point(models, case):
    """gapx_calc_device with positions and outputs in HBM equals gapx_calc (for the
    slab, lat_eff's extents come from a device reduction instead of the host)."""
    m, path = models["si64"]
    pot = Potential(param_filename=path)
    at = configs.si64() if case == "si64" else configs.perturb(configs.slab111(5.43, 6, 10.0), 0.1,
                                                               configs.rng_for(9, 1))
    ref = pot.calc(at.copy(), force=True, virial=True)
    hip = _hip()
    N = len(at)
    ptrs = {}
    for k, nb in (("pos", N * 24), ("e", 8), ("f", N * 24), ("v", 72)):
        p = C.c_void_p()
        assert hip.hipMalloc(C.byref(p), C.c_size_t(nb)) == 0
        ptrs[k] = p
    try:
        pos = L.f64(at.positions)
        assert hip.hipMemcpy(ptrs["pos"], L.ptr(pos), C.c_size_t(N * 24), 1) == 0
        lib = L.lib()
        L.check(lib.gapx_calc_device(pot.handle, N, ptrs["pos"], L.ptr(L.i32(at.numbers)), L.ptr(L.f64(at.cell)),
                                     L.ptr(L.i32(at.pbc)), 0.0, L.WANT_E | L.WANT_F | L.WANT_V, ptrs["e"], None,
                                     ptrs["f"], ptrs["v"], None), "gapx_calc_device")
        L.check(lib.gapx_sync(pot.handle), "gapx_sync")
        e, f, v = np.zeros(1), np.zeros((N, 3)), np.zeros(9)
        for host, k, nb in ((e, "e", 8), (f, "f", N * 24), (v, "v", 72)):
            assert hip.hipMemcpy(L.ptr(host), ptrs[k], C.c_size_t(nb), 2) == 0
        assert e[0] == ref["energy"]
        assert np.array_equal(f, ref["force"])
        assert np.array_equal(v.reshape(3, 3).T, ref["virial"])
    finally:
        for p in ptrs.values():
            hip.hipFree(p)


def test_errors_are_status_codes(tmp_path, golden):
    lib = L.lib()
    h = C.c_void_p()
    bad = b"<GAP><GAP_params label='x'><gpSparse>"
    st = lib.gapx_model_load(bad, len(bad), None, b".", 0, C.byref(h))
    assert st == 2 and "XML" in L.last_error()
    xml = open(os.path.join(golden, "gap_water", "GAP.xml")).read().encode()
    st = lib.gapx_model_load(xml, len(xml), None, str(tmp_path).encode(), 0, C.byref(h))
    assert st == 3 and "sparseX" in L.last_error()
    st = lib.gapx_model_load(xml, len(xml), b"no_such_label", os.path.join(golden, "gap_water").encode(), 0,
                             C.byref(h))
    assert st == 2
    pot = Potential(param_filename=os.path.join(golden, "gap_water", "GAP.xml"))
    with pytest.raises(L.GapxError):
        pot.calc(Atoms([1, 1], [[0, 0, 0], [1, 1, 1]], np.zeros((3, 3))))      # singular lattice
    with pytest.raises(L.GapxError):
        pot.calc(Atoms([1, 1], [[0, 0, 0], [1, 1, 1]], np.eye(3) * 5), cutoff=1.0)   # below model cutoff
    with pytest.raises(ValueError):
        pot.calc(Atoms([1, 1], [[0, 0, 0], [1, 1, 1]], np.eye(3) * 5), args_str="E_scale=2")


def test_fortran_host_module(models, tmp_path):
    """The amdflang IPModel_GAPX module over the C-ABI gives the Python host's numbers."""
    import subprocess
    drv = os.path.join(os.path.dirname(L.LIB_PATH), "fortran", "gapx_calc_driver")
    if not os.path.exists(drv):
        pytest.skip("Fortran driver not built (amdflang absent)")
    m, path = models["si64"]
    at = configs.si64(4)
    inp = tmp_path / "in.txt"
    with open(inp, "w") as fh:
        fh.write(path + "\n%d\n" % len(at))
        fh.write(" ".join("%.17e" % x for x in at.cell.ravel()) + "\n1 1 1\n")
        for z, p in zip(at.numbers, at.positions):
            fh.write("%d %.17e %.17e %.17e\n" % (z, p[0], p[1], p[2]))
    out = subprocess.run([drv, str(inp)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    e = float([ln.split()[1] for ln in out.stdout.splitlines() if ln.startswith("ENERGY")][0])
    f = np.array([[float(x) for x in ln.split()[2:5]] for ln in out.stdout.splitlines() if ln.startswith("FORCE")])
    v = np.array([[float(x) for x in ln.split()[1:4]] for ln in out.stdout.splitlines() if ln.startswith("VIRIAL")])
    r = Potential(param_filename=path).calc(at.copy(), force=True, virial=True)
    assert abs(e - r["energy"]) < 1e-12
    assert np.abs(f - r["force"]).max() < 1e-14
    assert np.abs(v - r["virial"]).max() < 1e-13
