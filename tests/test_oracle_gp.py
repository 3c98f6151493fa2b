"""The oracle's gp_predict and the IPModel_GAP_Calc options, pinned by what the
mathematics fixes.

No reference fixture holds a SOAP-GAP prediction, so the dot_product branch of
gp_predict is pinned here by closed forms on the reference's own descriptor
vectors (tests/Si.soap.cur_points.json), chosen so that a plausible slip fails:
delta instead of delta^2, an ignored sparseCutoff, a wrong power zeta, a lost
E_scale or e0.  The calc options (atom_mask_name, only_descriptor,
energy_per_coordinate; IPModel_GAP.f95:324-335, 398-403, 462, 568, 576-590) are
pinned by complementarity and sums; the GAP variance (local_gap_variance,
IPModel_GAP.f95:330-332, 464-470, 484-488, 558-567) by closed forms, its
non-negativity, its smallness at a sparse point and finite differences of its
gradient.  These run without a GPU.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
from quip_amd import configs
from quip_amd.atoms import Atoms
from quip_amd.gap_xml import Descriptor, GAPModel, GPCoordinate

SOAP8 = "soap atom_sigma=0.5 central_weight=1.0 cutoff=4.0 cutoff_transition_width=1.0 l_max=8 n_max=8"


@pytest.fixture(scope="module")
def xs(golden):
    """The reference's 100 unit-norm SOAP vectors (d = 325)."""
    X = np.load(os.path.join(golden, "si_soap_cur_points.npz"))["sparsex"]
    assert np.allclose(np.linalg.norm(X, axis=1), 1.0, rtol=0, atol=1e-14)
    return X


def _dot_coord(X, alpha, delta=1.0, zeta=4.0, sc=None, desc=SOAP8):
    X = np.atleast_2d(np.asarray(X, dtype=float))
    sc = np.ones(len(X)) if sc is None else np.asarray(sc, dtype=float)
    return GPCoordinate("c", Descriptor.from_string(desc), 2, delta, 0.0, zeta, X, np.asarray(alpha, float), sc)


# ---------------------------------------------------------------- dot_product gp_predict

def test_self_covariance_is_signal_variance_squared(xs):
    """k(x, x) = delta^2 (x.x)^zeta = delta^2 for a unit descriptor: one sparse point equal
    to x with alpha = 1 predicts delta^2 exactly (delta = 3 as test_gapfit.py:95 fits)."""
    for q in (0, 17, 99):
        for zeta in (1.0, 2.0, 4.0):
            e = O.gp_predict(_dot_coord(xs[q], [1.0], delta=3.0, zeta=zeta), xs[q])
            assert abs(e - 9.0) < 1e-13, (q, zeta, e)


def test_sparse_cutoff_and_alpha_scale_each_point(xs):
    """The prediction is linear in alpha_s and in sparseCutoff_s of every sparse point."""
    x = xs[3]
    pts = xs[[10, 20, 30]]
    alpha = np.array([0.7, -1.3, 2.1])
    single = np.array([O.gp_predict(_dot_coord(p, [1.0], delta=1.5), x) for p in pts])
    for sc in ([1.0, 1.0, 1.0], [0.25, 1.0, 3.0], [0.0, 0.5, 1.0]):
        e = O.gp_predict(_dot_coord(pts, alpha, delta=1.5, sc=sc), x)
        assert abs(e - np.sum(alpha * np.array(sc) * single)) < 1e-13 * max(1.0, abs(e))
    e2 = O.gp_predict(_dot_coord(pts, alpha, delta=1.5, sc=[2.0, 2.0, 2.0]), x)
    e1 = O.gp_predict(_dot_coord(pts, alpha, delta=1.5), x)
    assert abs(e2 - 2.0 * e1) < 1e-13 * abs(e1)


def test_zeta_is_the_power_of_the_dot_product(xs):
    """Two one-point models with X_s = x2 and x3: the ratio of their predictions at x is
    ((x.x2)/(x.x3))^zeta, and delta enters squared."""
    x, x2, x3 = xs[5], xs[40], xs[77]
    c2, c3 = x @ x2, x @ x3
    assert c2 > 0.5 and c3 > 0.5
    for zeta in (2.0, 3.0, 4.0):
        e2 = O.gp_predict(_dot_coord(x2, [1.0], delta=2.0, zeta=zeta), x)
        e3 = O.gp_predict(_dot_coord(x3, [1.0], delta=2.0, zeta=zeta), x)
        assert abs(e2 / e3 - (c2 / c3) ** zeta) < 1e-13
        assert abs(e2 - 4.0 * c2 ** zeta) < 1e-13


def test_gradient_of_prediction_matches_finite_differences(xs):
    c = _dot_coord(xs[[1, 2, 3, 4]], [0.3, -0.2, 0.5, 0.1], delta=1.7, zeta=4.0, sc=[1.0, 0.5, 2.0, 1.0])
    x = xs[60].copy()
    e, g = O.gp_predict(c, x, grad=True)
    h = 1e-6
    for q in (0, 7, 100, 324):
        p, m = x.copy(), x.copy()
        p[q] += h
        m[q] -= h
        fd = (O.gp_predict(c, p) - O.gp_predict(c, m)) / (2 * h)
        assert abs(fd - g[q]) < 1e-8


# ---------------------------------------------------------------- whole calc, closed forms

def _perfect_diamond(reps=(1, 1, 1)):
    return configs.diamond(5.43, reps)


def _one_point_model(at, desc_string, alpha, delta, sc, zeta, e0_si):
    d = Descriptor.from_string(desc_string)
    x0 = O.soap_descriptors(_dot_coord(np.zeros(d.dim), [0.0], desc=desc_string), at)[0][0]
    c = GPCoordinate("s", d, 2, delta, 0.0, zeta, x0[None, :], np.array([alpha]), np.array([sc]))
    e0 = np.zeros(119)
    e0[14] = e0_si
    return GAPModel("T", 1527075646, e0, [c])


def test_perfect_crystal_energy_closed_form():
    """Every atom of a perfect diamond crystal has the same unit descriptor x0.  A model
    whose one sparse point is x0 gives each atom alpha delta^2 sparseCutoff (x0.x0)^zeta
    + e0, times E_scale, and no force.  delta = 3, sparseCutoff 0.5, e0 = -1.25, E_scale 2."""
    at = _perfect_diamond()
    ds = configs.soap_string(cutoff=4.0, n_max=6, l_max=4)
    m = _one_point_model(at, ds, alpha=0.7, delta=3.0, sc=0.5, zeta=4.0, e0_si=-1.25)
    r = O.gap_calc_ex(m, at, E_scale=2.0)
    per_atom = 2.0 * (0.7 * 9.0 * 0.5 - 1.25)
    assert abs(r["energy"] - len(at) * per_atom) < 1e-11
    assert np.abs(r["local_energy"] - per_atom).max() < 1e-12
    assert np.abs(r["forces"]).max() < 1e-10
    assert abs(r["energy_per_coordinate"][0] - len(at) * 0.7 * 9.0 * 0.5) < 1e-11


# ---------------------------------------------------------------- calc options

@pytest.fixture(scope="module")
def two_coord_model():
    """SOAP (n 5, l 4) + distance_2b, with delta != 1, sparseCutoff != 1 and e0 != 0."""
    train = [configs.perturb(configs.diamond(5.43, (1, 1, 1)), 0.1, configs.rng_for(0, 400 + k)) for k in range(6)]
    fn = lambda d, at: O.soap_descriptors(_dot_coord(np.zeros(d.dim), [0.0], desc=d.string), at)[0]
    m = configs.soap_model(fn, train, 24, 0, [configs.soap_string(cutoff=4.0, n_max=5, l_max=4)], "T",
                           alpha_std=1.0, delta=1.8, d2b=(4.0, 1.0, 8, 1.8, 4.0))
    rng = np.random.default_rng(11)
    for c in m.coords:
        c.sparse_cutoff = rng.uniform(0.3, 1.5, c.n_sparse)
    m.e0[14] = -0.37
    return m


def _cell(seed):
    return configs.perturb(configs.diamond(5.43, (1, 1, 1)), 0.12, np.random.default_rng(seed))


def test_atom_mask_complements_add_up(two_coord_model):
    """atom_mask_name restricts descriptor centres, the first atom of pair descriptors and
    e0 to the masked atoms, so a mask and its complement add up to the unmasked calc; a
    SOAP-only model gives unmasked atoms no local energy."""
    at = _cell(21)
    mask = np.zeros(len(at), bool)
    mask[[0, 2, 3, 6]] = True
    full = O.gap_calc_ex(two_coord_model, at)
    a = O.gap_calc_ex(two_coord_model, at, atom_mask=mask)
    b = O.gap_calc_ex(two_coord_model, at, atom_mask=~mask)
    assert abs(a["energy"] + b["energy"] - full["energy"]) < 1e-11
    for k in ("forces", "virial", "local_energy", "local_virial"):
        assert np.abs(a[k] + b[k] - full[k]).max() < 1e-11, k
    soap_only = GAPModel("T", 1527075646, two_coord_model.e0, two_coord_model.coords[:1])
    s = O.gap_calc_ex(soap_only, at, atom_mask=mask)
    s_all = O.gap_calc_ex(soap_only, at)
    assert np.abs(s["local_energy"][~mask]).max() == 0.0
    assert np.abs(s["local_energy"][mask] - s_all["local_energy"][mask]).max() < 1e-12


def test_only_descriptor_and_energy_per_coordinate(two_coord_model):
    """only_descriptor=k evaluates coordinate k alone (e0 still added); the coordinates'
    energies add up to the whole, and energy_per_coordinate is each coordinate's share
    before e0 and E_scale."""
    at = _cell(22)
    e0sum = len(at) * two_coord_model.e0[14]
    full = O.gap_calc_ex(two_coord_model, at, E_scale=1.5)
    parts = [O.gap_calc_ex(two_coord_model, at, E_scale=1.5, only_descriptor=k) for k in (1, 2)]
    assert abs(parts[0]["energy"] + parts[1]["energy"] - full["energy"] - 1.5 * e0sum) < 1e-11
    assert np.abs(parts[0]["forces"] + parts[1]["forces"] - full["forces"]).max() < 1e-11
    epc = full["energy_per_coordinate"]
    assert abs(1.5 * (epc.sum() + e0sum) - full["energy"]) < 1e-11
    for k in (0, 1):
        assert abs(1.5 * (epc[k] + e0sum) - parts[k]["energy"]) < 1e-11
    none = O.gap_calc_ex(two_coord_model, at, E_scale=1.5, only_descriptor=3)   # no such coordinate
    assert abs(none["energy"] - 1.5 * e0sum) < 1e-12


# ---------------------------------------------------------------- GAP variance

def test_variance_one_point_closed_forms(xs):
    """One sparse point: var = k(x,x) - k^2/(k(X,X) + reg).  dot_product on a unit
    descriptor: delta^2 - delta^4 c^(2 zeta)/(delta^2 + reg), c = x.X; ard_se on a
    dimer at r inside the transition-free range: delta^2 - k^2/(delta^2 + reg) with
    k = delta^2 exp(-(r-X)^2/(2 theta^2)), and each atom of the pair gets the variance
    of both directed instances halved."""
    at = _perfect_diamond()
    ds = configs.soap_string(cutoff=4.0, n_max=6, l_max=4)
    d = Descriptor.from_string(ds)
    X = O.soap_descriptors(_dot_coord(np.zeros(d.dim), [0.0], desc=ds), _cell(5))[0]
    x_at = O.soap_descriptors(_dot_coord(np.zeros(d.dim), [0.0], desc=ds), at)[0]
    c = GPCoordinate("s", d, 2, 1.7, 0.0, 3.0, X[2][None, :], np.array([0.4]), np.ones(1))
    m = GAPModel("T", 1527075646, np.zeros(119), [c])
    reg = 0.01
    r = O.gap_calc_ex(m, at, variance=True, variance_regularisation=reg, max_instances=64)
    cc = x_at @ X[2]
    want = 1.7 ** 2 - 1.7 ** 4 * cc ** 6 / (1.7 ** 2 + reg)
    assert np.abs(r["local_gap_variance"] - want).max() < 1e-12
    assert np.abs(r["instance_variance"] - want).max() < 1e-12
    assert np.abs(r["gap_variance_gradient"]).max() < 1e-10            # symmetric crystal
    # distance_2b dimer, r = 2.3 with cutoff 4, transition width 0.5
    dd = Descriptor.from_string("distance_2b cutoff=4.0 cutoff_transition_width=0.5 Z1=14 Z2=14")
    c2 = GPCoordinate("d", dd, 1, 0.9, 0.0, 1.0, np.array([[2.0]]), np.array([0.3]), np.ones(1),
                      theta=np.array([0.8]))
    dimer = Atoms([14, 14], [[0, 0, 0], [2.3, 0, 0]], np.eye(3) * 20.0, pbc=(False, False, False))
    r2 = O.gap_calc_ex(GAPModel("D", 1527075646, np.zeros(119), [c2]), dimer, variance=True,
                       variance_regularisation=reg, max_instances=8)
    k = 0.81 * np.exp(-0.5 * (0.3 / 0.8) ** 2)
    want2 = 0.81 - k * k / (0.81 + reg)
    assert len(r2["instance_variance"]) == 2
    assert np.abs(r2["instance_variance"] - want2).max() < 1e-14
    assert np.abs(r2["local_gap_variance"] - want2).max() < 1e-14
    assert np.array_equal(r2["instance_atoms"], [[0, 1], [1, 0]])


def test_variance_nonnegative_and_small_at_sparse_points(two_coord_model):
    """The estimate is a Schur complement, so it is never negative; at a structure whose
    descriptors are all sparse points it vanishes as the regularisation does."""
    at = _cell(23)
    r = O.gap_calc_ex(two_coord_model, at, variance=True, max_instances=4096)
    assert r["instance_variance"].min() > -1e-12
    assert r["local_gap_variance"].min() > -1e-12
    d = two_coord_model.coords[0].descriptor
    X = O.soap_descriptors(two_coord_model.coords[0], at)[0]
    c = GPCoordinate("s", d, 2, 1.3, 0.0, 4.0, X, np.ones(len(X)), np.ones(len(X)))
    own = GAPModel("T", 1527075646, np.zeros(119), [c])
    small = O.gap_calc_ex(own, at, variance=True, variance_regularisation=1e-9)
    big = O.gap_calc_ex(own, at, variance=True, variance_regularisation=1e-3)
    assert small["local_gap_variance"].max() < 1e-8
    assert big["local_gap_variance"].max() > 10 * small["local_gap_variance"].max()


def test_variance_gradient_matches_finite_differences(two_coord_model):
    """gap_variance_gradient(:,j) = d/dr_j of the summed local_gap_variance (with its
    plus sign: IPModel_GAP.f95:484-488 adds, where the force subtracts)."""
    at = _cell(24)
    r = O.gap_calc_ex(two_coord_model, at, variance=True)
    g = r["gap_variance_gradient"]
    h = 1e-5
    for i in (0, 5):
        for k in range(3):
            p, m = at.copy(), at.copy()
            p.positions[i, k] += h
            m.positions[i, k] -= h
            vp = O.gap_calc_ex(two_coord_model, p, variance=True)["local_gap_variance"].sum()
            vm = O.gap_calc_ex(two_coord_model, m, variance=True)["local_gap_variance"].sum()
            fd = (vp - vm) / (2 * h)
            assert abs(fd - g[i, k]) < 1e-7 * max(1.0, abs(fd)), (i, k, fd, g[i, k])
