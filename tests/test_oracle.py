"""The CPU oracle against the reference's own fixtures and against what the
mathematics fixes (brute force, finite differences, invariances).

These run without a GPU.  They pin every function of oracle/oracle.c that the
HIP path is compared with.
"""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from quip_amd.atoms import Atoms, read_xyz
from quip_amd.gap_xml import Descriptor, GPCoordinate, read_gap_xml
from quip_amd import configs


def _coord(desc_string, d=None):
    desc = Descriptor.from_string(desc_string)
    d = desc.dim
    return GPCoordinate("x", desc, 2, 1.0, 0.0, 4.0, np.zeros((1, d)), np.zeros(1), np.ones(1))


# ---------------------------------------------------------------- fixtures

def test_soap_vectors_match_gap_fit_sparse_points(golden):
    """tests/Si.soap.cur_points.json: 100 SOAP vectors, tolerance 1e-8 (test_gapfit.py:58)."""
    g = np.load(os.path.join(golden, "si_soap_cur_points.npz"))
    c = _coord("soap atom_sigma=0.5 central_weight=1.0 cutoff=4.0 cutoff_transition_width=1.0 l_max=8 n_max=8")
    X, off = [], 0
    for f, n in enumerate(g["natoms"]):
        at = Atoms(g["numbers"][off:off + n], g["positions"][off:off + n], g["lattices"][f])
        off += n
        X.append(O.soap_descriptors(c, at)[0])
    X = np.vstack(X)
    err = np.abs(X[g["index"] - 1] - g["sparsex"]).max()
    assert err < 1e-12, err


def test_soap_gradients_c2h(golden):
    """tests/test_descriptor.py:80-190: gradient index order and gradient values (tol 1e-7)."""
    g = np.load(os.path.join(golden, "c2h_soap_grad.npz"))
    at = Atoms(g["numbers"], g["positions"], g["cell"])
    c = _coord(str(g["descriptor"]))
    pairs = []
    for a in range(3):
        x, grad, ii = O.soap_grad_atom(c, at, a, nl_cutoff=float(g["nl_cutoff"]))
        pairs += [(a, j) for j in ii]
        if a == 0:
            grad0 = grad
    assert np.array_equal(np.array(pairs), g["grad_index"])
    assert np.abs(grad0[:2] - g["grad"]).max() < 1e-7


def test_distance_2b_water_energy_forces(golden):
    """tests/test_gappot.py:44-48 (E 1e-5, F 1e-6)."""
    m = read_gap_xml(os.path.join(golden, "gap_water", "GAP.xml"))
    at = read_xyz(os.path.join(golden, "gap_water", "gap_sample.xyz"), 0)
    r = O.gap_calc(m, at)
    assert abs(r["energy"] - at.info["energy"]) < 1e-7
    assert np.abs(r["forces"] - at.arrays["force"]).max() < 1e-7


def test_cells_smaller_than_cutoff(golden):
    """tests/test_potential_cell.py:37-53: periodic self-images inside the cutoff."""
    ref = json.load(open(os.path.join(golden, "potential_cell.json")))
    m = read_gap_xml(os.path.join(golden, "gap_water", "GAP.xml"))
    for c, e in zip(ref["cell_sizes"], ref["energies"]):
        at = Atoms(ref["numbers"], ref["positions"], np.eye(3) * c)
        assert abs(O.gap_calc(m, at)["energy"] - e) < 1e-8


# ---------------------------------------------------------------- neighbour list

def _brute_pairs(at, rc):
    inv = np.linalg.inv(at.cell)
    nimg = [int(np.ceil(rc * np.linalg.norm(inv[:, k]))) + 1 if at.pbc[k] else 0 for k in range(3)]
    out = set()
    for i in range(len(at)):
        for j in range(len(at)):
            for a in range(-nimg[0], nimg[0] + 1):
                for b in range(-nimg[1], nimg[1] + 1):
                    for c in range(-nimg[2], nimg[2] + 1):
                        d = at.positions[j] - at.positions[i] + np.array([a, b, c]) @ at.cell
                        r = np.linalg.norm(d)
                        if r < rc and not (i == j and a == b == c == 0):
                            out.add((i, j, a, b, c))
    return out


def _geometries():
    rng = np.random.default_rng(5)
    yield configs.si64(), 5.0
    small = Atoms([14, 14], [[0, 0, 0], [1.1, 1.2, 0.9]], [[2.6, 0.1, 0], [0.3, 2.4, 0], [0.2, -0.1, 2.9]])
    yield small, 4.0                                     # cell smaller than the cutoff
    skew = Atoms(np.full(12, 14), rng.uniform(0, 6, (12, 3)), [[6.0, 0, 0], [2.5, 5.5, 0], [1.0, 1.5, 6.5]])
    yield skew, 3.5
    slab = Atoms(np.full(10, 14), rng.uniform(0, 5, (10, 3)), np.diag([5.0, 5.0, 5.0]), (True, True, False))
    yield slab, 3.0                                      # non-periodic axis
    off = Atoms(np.full(6, 14), rng.uniform(0, 5, (6, 3)) + np.array([40.0, -17.0, 3.0]), np.diag([5.0, 5.0, 5.0]))
    yield off, 3.0                                       # positions far outside the cell


@pytest.mark.parametrize("case", range(5))
def test_neighbour_list_against_brute_force(case):
    at, rc = list(_geometries())[case]
    nl = O.neighbour_list(at, rc)
    got = set()
    N = len(at)
    for i in range(N):
        for k in range(nl["n1_off"][i], nl["n1_off"][i + 1]):
            j = nl["n1_j"][k]
            s = nl["n1_shift"][k]
            assert j >= i
            got.add((i, j, *s))
            got.add((j, i, *(-s)))
            d = at.positions[j] - at.positions[i] + s @ at.cell
            assert abs(np.linalg.norm(d) - nl["n1_r"][k]) < 1e-12
        for k in range(nl["n2_off"][i], nl["n2_off"][i + 1]):
            j, idx = nl["n2_j"][k], nl["n2_idx"][k]
            assert j < i and nl["n1_j"][nl["n1_off"][j] + idx] == i
    want = _brute_pairs(at, rc)
    # pbc=F axes: brute force has no images there, the NL must not either
    assert got == want


# ---------------------------------------------------------------- SOAP-GAP physics

@pytest.fixture(scope="module")
def small_model():
    """A small SOAP(+distance_2b) GAP model built from the si64 recipe with fewer sparse points."""
    train = [configs.perturb(configs.diamond(5.43, (1, 1, 1)), 0.1, configs.rng_for(0, 300 + k)) for k in range(8)]
    fn = lambda d, at: O.soap_descriptors(GPCoordinate("t", d, 2, 1, 0, 4, np.zeros((1, d.dim)), np.zeros(1),
                                                       np.ones(1)), at)[0]
    return configs.soap_model(fn, train, 40, 0, [configs.soap_string(cutoff=4.0, n_max=5, l_max=4)], "T",
                              alpha_std=1.0, d2b=(4.0, 1.0, 8, 1.8, 4.0))


def _cell8(seed=3):
    return configs.perturb(configs.diamond(5.43, (1, 1, 1)), 0.12, np.random.default_rng(seed))


def test_forces_match_finite_differences(small_model):
    at = _cell8()
    r = O.gap_calc(small_model, at)
    h = 1e-5
    for i in (0, 5):
        for k in range(3):
            p, m = at.copy(), at.copy()
            p.positions[i, k] += h
            m.positions[i, k] -= h
            fd = -(O.gap_calc(small_model, p)["energy"] - O.gap_calc(small_model, m)["energy"]) / (2 * h)
            assert abs(fd - r["forces"][i, k]) < 1e-7 * max(1.0, abs(fd))


def test_virial_matches_strain_derivative(small_model):
    """virial = -dE/d(strain), the QUIP convention (stress = -virial/volume, potential.py:281)."""
    at = _cell8(4)
    r = O.gap_calc(small_model, at)
    h = 1e-6
    for a in range(3):
        for b in range(3):
            eps = np.zeros((3, 3))
            eps[a, b] = h
            p, m = at.copy(), at.copy()
            for s, x in ((1, p), (-1, m)):
                F = np.eye(3) + s * eps
                x.positions = at.positions @ F.T
                x.cell = at.cell @ F.T
            fd = -(O.gap_calc(small_model, p)["energy"] - O.gap_calc(small_model, m)["energy"]) / (2 * h)
            assert abs(fd - r["virial"][a, b]) < 1e-6 * max(1.0, abs(fd))
    assert np.allclose(r["local_virial"].sum(0).reshape(3, 3, order="F"), r["virial"], atol=1e-12)
    assert np.isclose(r["local_energy"].sum(), r["energy"], atol=1e-12)
    assert np.abs(r["forces"].sum(0)).max() < 1e-10


def test_rotation_translation_permutation_invariance(small_model):
    at = _cell8(6)
    e0 = O.gap_calc(small_model, at)
    th = 0.7
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1.0]]) @ \
        np.array([[1, 0, 0], [0, np.cos(0.3), -np.sin(0.3)], [0, np.sin(0.3), np.cos(0.3)]])
    rot = Atoms(at.numbers, at.positions @ R.T + np.array([0.3, -1.1, 2.0]), at.cell @ R.T)
    e1 = O.gap_calc(small_model, rot)
    assert abs(e1["energy"] - e0["energy"]) < 1e-10
    assert np.abs(e1["forces"] - e0["forces"] @ R.T).max() < 1e-9
    perm = np.random.default_rng(0).permutation(len(at))
    pa = Atoms(at.numbers[perm], at.positions[perm], at.cell)
    e2 = O.gap_calc(small_model, pa)
    assert abs(e2["energy"] - e0["energy"]) < 1e-10
    assert np.abs(e2["forces"] - e0["forces"][perm]).max() < 1e-10


def test_soap_basis_is_orthonormal_projection():
    """T = C^-1 L maps grid values h_a to coefficients in an orthonormal basis of the
    Gaussian interpolants h(r) = sum_a w_a exp(-alpha (r - r_a)^2), w = C^-1 h:
    (h T).(h' T) = 8 sqrt(2) alpha^(5/2) * int_0^inf r^2 h(r) h'(r) dr (the overlap
    formula's normalisation).  The centre Gaussian (grid values exp(-alpha r_a^2))
    lands in channel 0 only, with value U00."""
    cutoff, sigma, n = 4.0, 0.5, 6
    rb, T, u00 = O.soap_basis(cutoff, sigma, n)
    alpha = 0.5 / sigma ** 2
    assert np.allclose(rb, np.arange(n) * (cutoff + sigma * np.sqrt(20 * np.log(10))) / n, rtol=0, atol=1e-14)
    C = np.exp(-alpha * (rb[:, None] - rb[None, :]) ** 2)
    g = np.exp(-alpha * rb ** 2) @ T
    assert abs(g[0] - u00) < 1e-12 and np.abs(g[1:]).max() < 1e-10
    rng = np.random.default_rng(1)
    h1, h2 = rng.standard_normal(n), rng.standard_normal(n)
    r = np.linspace(0.0, 14.0, 140001)
    phi = np.exp(-alpha * (r[:, None] - rb[None, :]) ** 2)
    i1, i2 = phi @ np.linalg.solve(C, h1), phi @ np.linalg.solve(C, h2)
    w = np.full(len(r), r[1] - r[0]); w[0] = w[-1] = 0.5 * w[0]
    integral = np.sum(w * r ** 2 * i1 * i2) * 8 * np.sqrt(2) * alpha ** 2.5
    assert abs((h1 @ T) @ (h2 @ T) - integral) < 1e-8 * max(1.0, abs(integral))
