"""HIP path against the oracle for the IPModel_GAP_Calc options and the model shapes
that the fast kernels do not cover.

- dot_product conventions: delta != 1, sparseCutoff != 1, e0 != 0 and E_scale != 1
  (IPModel_GAP.f95:445-460, 576-596) against the oracle, whose gp_predict is pinned
  by closed forms (tests/test_oracle_gp.py), and a closed form on the GPU itself;
- atom_mask_name, only_descriptor, energy_per_coordinate (:324-335, 398-403, 462, 568);
- local_gap_variance / print_gap_variance and the variance gradient (:330-332,
  464-470, 484-488, 506-516, 558-567);
- the generic SOAP kernels (n_max, l_max outside the specialised shapes) and a
  two-species n 4, l 4 model on C/H, shaped like tests/test_descriptor.py:61-79.

Tolerances: FP64 1e-10 eV/atom and 1e-8 eV/A as the plain form; the variance to
1e-10 relative to delta^2 (it is formed in FP64 in every precision form).
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from quip_amd import _lib as L
from quip_amd import configs
from quip_amd.atoms import Atoms
from quip_amd.gap_xml import Descriptor, GAPModel, GPCoordinate, write_gap_xml
from quip_amd.potential import Potential

pytestmark = pytest.mark.gpu

E_TOL = 1e-10
F_TOL = 1e-8


def _desc_fn(d, at):
    return O.soap_descriptors(GPCoordinate("t", d, 2, 1, 0, 4, np.zeros((1, d.dim)), np.zeros(1), np.ones(1)), at)[0]


def _write(m, tmp_path, name):
    path = str(tmp_path / (name + ".xml"))
    write_gap_xml(path, m)
    return path


@pytest.fixture(scope="module")
def conv_model(tmp_path_factory):
    """SOAP (n 6, l 4) + distance_2b with delta = 1.8, sparseCutoff in [0.3, 1.5], e0(Si) = -0.37."""
    train = [configs.perturb(configs.diamond(5.43, (1, 1, 1)), 0.1, configs.rng_for(0, 700 + k)) for k in range(6)]
    m = configs.soap_model(_desc_fn, train, 32, 0, [configs.soap_string(cutoff=4.5, n_max=6, l_max=4)], "CONV",
                           alpha_std=0.5, delta=1.8, d2b=(4.5, 1.0, 10, 1.8, 4.5))
    rng = np.random.default_rng(31)
    for c in m.coords:
        c.sparse_cutoff = rng.uniform(0.3, 1.5, c.n_sparse)
    m.e0[14] = -0.37
    return m, _write(m, tmp_path_factory.mktemp("conv"), "conv")


@pytest.fixture(scope="module")
def var_model(tmp_path_factory):
    """As conv_model with every sparseCutoff 1 (the variance is defined for those)."""
    train = [configs.perturb(configs.diamond(5.43, (1, 1, 1)), 0.1, configs.rng_for(0, 800 + k)) for k in range(6)]
    m = configs.soap_model(_desc_fn, train, 40, 0, [configs.soap_string(cutoff=4.5, n_max=6, l_max=4)], "VAR",
                           alpha_std=0.5, delta=1.3, d2b=(4.5, 1.0, 10, 1.8, 4.5))
    m.e0[14] = 0.11
    return m, _write(m, tmp_path_factory.mktemp("var"), "var")


def _cell(seed, reps=(2, 2, 2), sigma=0.12):
    return configs.perturb(configs.diamond(5.43, reps), sigma, np.random.default_rng(seed))


def _cmp(r, ro, N):
    assert abs(r["energy"] - ro["energy"]) / N < E_TOL
    assert np.abs(r["force"] - ro["forces"]).max() < F_TOL
    assert np.abs(r["virial"] - ro["virial"]).max() < 10 * F_TOL
    assert np.abs(r["local_energy"] - ro["local_energy"]).max() < 10 * E_TOL
    assert np.abs(r["local_virial"] - ro["local_virial"]).max() < 10 * F_TOL


ALL = dict(force=True, virial=True, local_energy=True, local_virial=True)


# ---------------------------------------------------------------- dot_product conventions

@pytest.mark.parametrize("precision", ["fp64", "bf16x3"])
def test_delta_sparse_cutoff_e0_e_scale_match_oracle(conv_model, precision):
    m, path = conv_model
    at = _cell(41)
    print("case conv", precision, len(at), flush=True)
    pot = Potential("IP GAP E_scale=1.5", param_filename=path, precision=precision)
    r = pot.calc(at.copy(), **ALL)
    ro = O.gap_calc_ex(m, at, E_scale=1.5)
    if precision == "fp64":
        _cmp(r, ro, len(at))
    else:
        assert abs(r["energy"] - ro["energy"]) / len(at) < 1e-3
        assert np.sqrt(np.mean((r["force"] - ro["forces"]) ** 2)) < 1e-2


def test_perfect_crystal_closed_form_on_gpu(tmp_path):
    """One sparse point equal to the common descriptor of a perfect crystal: every atom
    gets E_scale (alpha delta^2 sparseCutoff + e0), forces vanish."""
    at = configs.diamond(5.43, (2, 2, 2))
    ds = configs.soap_string(cutoff=4.0, n_max=6, l_max=4)
    d = Descriptor.from_string(ds)
    x0 = _desc_fn(d, at)[0]
    c = GPCoordinate("s", d, 2, 3.0, 0.0, 4.0, x0[None, :], np.array([0.7]), np.array([0.5]))
    e0 = np.zeros(119)
    e0[14] = -1.25
    path = _write(GAPModel("P", 1527075646, e0, [c]), tmp_path, "p")
    r = Potential("IP GAP E_scale=2.0", param_filename=path).calc(at.copy(), force=True, local_energy=True)
    per_atom = 2.0 * (0.7 * 9.0 * 0.5 - 1.25)
    assert abs(r["energy"] - len(at) * per_atom) < 1e-10
    assert np.abs(r["local_energy"] - per_atom).max() < 1e-12
    assert np.abs(r["force"]).max() < 1e-10


# ---------------------------------------------------------------- calc options

def test_atom_mask_matches_oracle(conv_model):
    m, path = conv_model
    at = _cell(42)
    mask = np.random.default_rng(3).random(len(at)) < 0.4
    at.arrays["local"] = mask
    pot = Potential(param_filename=path)
    r = pot.calc(at.copy(), args_str="atom_mask_name=local", **ALL)
    ro = O.gap_calc_ex(m, at, atom_mask=mask)
    _cmp(r, ro, len(at))
    rn = pot.calc(at.copy(), args_str="atom_mask_name=local", **ALL)      # repeat: the mask is re-uploaded
    assert rn["energy"] == r["energy"]
    full = pot.calc(at.copy(), **ALL)
    at.arrays["other"] = ~mask
    comp = pot.calc(at.copy(), args_str="atom_mask_name=other", **ALL)
    assert abs(r["energy"] + comp["energy"] - full["energy"]) < 1e-10
    assert np.abs(r["force"] + comp["force"] - full["force"]).max() < 1e-10
    with pytest.raises(ValueError):
        pot.calc(at.copy(), args_str="atom_mask_name=missing")


def test_only_descriptor_and_energy_per_coordinate_match_oracle(conv_model):
    m, path = conv_model
    at = _cell(43)
    pot = Potential("IP GAP E_scale=1.5", param_filename=path)
    r = pot.calc(at, args_str="energy_per_coordinate=epc", **ALL)
    ro = O.gap_calc_ex(m, at, E_scale=1.5)
    _cmp(r, ro, len(at))
    assert np.abs(at.info["epc"] - ro["energy_per_coordinate"]).max() < 1e-10 * len(at)
    for k in (1, 2):
        rk = pot.calc(at.copy(), args_str="only_descriptor=%d energy_per_coordinate=epc" % k, **ALL)
        rok = O.gap_calc_ex(m, at, E_scale=1.5, only_descriptor=k)
        _cmp(rk, rok, len(at))
        assert rk["energy_per_coordinate"][2 - k] == 0.0
        assert abs(rk["energy_per_coordinate"][k - 1] - ro["energy_per_coordinate"][k - 1]) < 1e-10 * len(at)


def test_device_entry_with_options_equals_host_entry(var_model):
    """gapx_calc_device_ex gives gapx_calc_ex's results bit for bit."""
    m, path = var_model
    at = _cell(44)
    N = len(at)
    mask = (np.arange(N) % 3 != 0).astype(np.uint8)
    at.arrays["m"] = mask.astype(bool)
    pot = Potential(param_filename=path)
    r = pot.calc(at.copy(), args_str="atom_mask_name=m energy_per_coordinate=e local_gap_variance=v", **ALL)
    hip = C.CDLL("libamdhip64.so")
    sizes = (("pos", 3 * N), ("e", 1), ("le", N), ("f", 3 * N), ("v", 9), ("lv", 9 * N), ("epc", 2), ("lgv", N),
             ("gvg", 3 * N))
    ptrs = {}
    try:
        for k, n in sizes:
            p = C.c_void_p()
            assert hip.hipMalloc(C.byref(p), C.c_size_t(8 * n)) == 0
            ptrs[k] = p
        pos = L.f64(at.positions)
        assert hip.hipMemcpy(ptrs["pos"], L.ptr(pos), C.c_size_t(24 * N), 1) == 0
        Z, lat, pbc = L.i32(at.numbers), L.f64(at.cell), L.i32(at.pbc)
        want = L.WANT_E | L.WANT_F | L.WANT_V | L.WANT_LOCAL_E | L.WANT_LOCAL_V
        L.check(L.lib().gapx_calc_device_ex(pot.handle, N, ptrs["pos"], L.ptr(Z), L.ptr(lat), L.ptr(pbc), 0.0, want,
                                            L.ptr(mask), 0, 1e-3,
                                            *[ptrs[k] for k in ("e", "le", "f", "v", "lv", "epc", "lgv", "gvg")]),
                "gapx_calc_device_ex")
        L.check(L.lib().gapx_sync(pot.handle), "sync")
        o = {}
        for k, n in sizes[1:]:
            o[k] = np.zeros(n)
            assert hip.hipMemcpy(L.ptr(o[k]), ptrs[k], C.c_size_t(8 * n), 2) == 0
    finally:
        for p in ptrs.values():
            hip.hipFree(p)
    assert o["e"][0] == r["energy"]
    assert np.array_equal(o["f"].reshape(N, 3), r["force"])
    assert np.array_equal(o["epc"], r["energy_per_coordinate"])
    assert np.array_equal(o["lgv"], r["local_gap_variance"])
    assert np.array_equal(o["gvg"].reshape(N, 3), r["gap_variance_gradient"])


# ---------------------------------------------------------------- GAP variance

@pytest.mark.parametrize("precision", ["fp64", "bf16x3"])
def test_variance_matches_oracle(var_model, precision, capsys):
    m, path = var_model
    at = _cell(45)
    print("case variance", precision, len(at), flush=True)
    pot = Potential(param_filename=path, precision=precision)
    r = pot.calc(at, args_str="local_gap_variance=var print_gap_variance=T gap_variance_regularisation=0.002",
                 force=True)
    ro = O.gap_calc_ex(m, at, variance=True, variance_regularisation=0.002, max_instances=20000)
    scale = 1.3 ** 2
    assert np.abs(at.arrays["var"] - ro["local_gap_variance"]).max() < 1e-10 * scale
    assert np.abs(at.arrays["gap_variance_gradient"] - ro["gap_variance_gradient"]).max() < 1e-8 * scale
    # per-instance values in the reference's order (soap centres, then distance_2b pairs)
    got = [pot.descriptor_variances(k) for k in range(2)]
    nsoap = len(got[0][0])
    assert nsoap == len(at)
    iv = np.concatenate([got[0][0], got[1][0]])
    ic = np.concatenate([got[0][1], got[1][1]])
    ia = np.concatenate([got[0][2], got[1][2]])
    assert len(iv) == len(ro["instance_variance"])
    assert np.array_equal(ia, ro["instance_atoms"])
    assert np.abs(iv - ro["instance_variance"]).max() < 1e-10 * scale
    assert np.abs(ic - ro["instance_cutoff2"]).max() < 1e-14
    out = capsys.readouterr().out
    assert out.count("GAP_VARIANCE potential") == 1 + 2 * len(iv)


def test_variance_refused_with_sparse_cutoff(conv_model):
    m, path = conv_model
    with pytest.raises(L.GapxError) as e:
        Potential(param_filename=path).calc(_cell(46), args_str="local_gap_variance=v")
    assert e.value.code == 5


# ---------------------------------------------------------------- generic SOAP kernels and C/H

GENERIC = [(3, 3), (6, 2), (7, 5), (16, 10), (8, 12)]


@pytest.mark.parametrize("shape", GENERIC)
def test_generic_soap_shapes_match_oracle(shape, tmp_path):
    n, l = shape
    ds = configs.soap_string(cutoff=4.0, n_max=n, l_max=l)
    train = [configs.perturb(configs.diamond(5.43, (1, 1, 1)), 0.1, configs.rng_for(0, 900 + k)) for k in range(3)]
    m = configs.soap_model(_desc_fn, train, 12, 0, [ds], "G", alpha_std=1.0, delta=1.2)
    path = _write(m, tmp_path, "g")
    at = configs.perturb(configs.diamond(5.43, (1, 1, 2)), 0.12, np.random.default_rng(n * 31 + l))
    print("case generic", n, l, len(at), flush=True)
    r = Potential(param_filename=path).calc(at.copy(), **ALL)
    _cmp(r, O.gap_calc(m, at), len(at))


def _ch_cell(seed, n_c=6, n_h=10, box=6.0):
    """Random C/H periodic cell with no two atoms closer than 1.0 A."""
    rng = np.random.default_rng(seed)
    pos = []
    while len(pos) < n_c + n_h:
        p = rng.uniform(0, box, 3)
        ok = True
        for q in pos:
            d = p - q
            d -= box * np.round(d / box)
            if np.linalg.norm(d) < 1.0:
                ok = False
                break
        if ok:
            pos.append(p)
    return Atoms([6] * n_c + [1] * n_h, np.array(pos), np.eye(3) * box)


def test_two_species_ch_model_matches_oracle(tmp_path):
    """n_species 2 (H, C), centres on both, n 4, l 4: the reference's C2H descriptor shape
    (tests/test_descriptor.py:61-79) with two species channels."""
    strs = ["soap cutoff=3.0 cutoff_transition_width=0.5 atom_sigma=0.5 n_max=4 l_max=4 central_weight=1.0 "
            "n_species=2 species_Z={1 6} n_Z=1 Z={%d}" % z for z in (6, 1)]
    train = [_ch_cell(100 + k) for k in range(4)]
    m = configs.soap_model(_desc_fn, train, 16, 0, strs, "CH", alpha_std=1.0, delta=1.1)
    m.e0[1], m.e0[6] = -0.5, -1.5
    path = _write(m, tmp_path, "ch")
    for seed in (7, 8):
        at = _ch_cell(seed)
        print("case ch", seed, flush=True)
        r = Potential(param_filename=path).calc(at.copy(), **ALL)
        _cmp(r, O.gap_calc(m, at), len(at))


def test_descriptor_beyond_lds_is_refused(tmp_path):
    """A shape whose kernels need more LDS than a block has is refused at load, not at launch."""
    ds = ("soap cutoff=4.0 atom_sigma=0.5 n_max=16 l_max=12 n_species=4 species_Z={1 6 8 14} n_Z=1 Z={14}")
    d = Descriptor.from_string(ds)
    c = GPCoordinate("big", d, 2, 1.0, 0.0, 4.0, np.zeros((1, d.dim)), np.ones(1), np.ones(1))
    path = _write(GAPModel("B", 1527075646, np.zeros(119), [c]), tmp_path, "big")
    with pytest.raises(L.GapxError) as e:
        Potential(param_filename=path)
    assert e.value.code == 5 and "LDS" in str(e.value)


# ---------------------------------------------------------------- Fortran host

def test_fortran_host_options(var_model, tmp_path):
    """The amdflang module with atom_mask_name, energy_per_coordinate and local_gap_variance."""
    drv = os.path.join(os.path.dirname(L.LIB_PATH), "fortran", "gapx_calc_driver")
    if not os.path.exists(drv):
        pytest.skip("Fortran driver not built (amdflang absent)")
    m, path = var_model
    at = _cell(47, reps=(1, 1, 1))
    mask = np.array([1, 0, 1, 1, 0, 1, 1, 0], bool)
    inp = tmp_path / "in.txt"
    with open(inp, "w") as fh:
        fh.write(path + "\n%d\n" % len(at))
        fh.write(" ".join("%.17e" % x for x in at.cell.ravel()) + "\n1 1 1\n")
        for z, p, k in zip(at.numbers, at.positions, mask):
            fh.write("%d %.17e %.17e %.17e %d\n" % (z, p[0], p[1], p[2], int(k)))
    args = "atom_mask_name=local energy_per_coordinate=epc local_gap_variance=var"
    out = subprocess.run([drv, str(inp), args], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.splitlines()
    e = float([ln.split()[1] for ln in lines if ln.startswith("ENERGY ")][0])
    epc = np.array([float(ln.split()[2]) for ln in lines if ln.startswith("ENERGY_PER_COORDINATE")])
    lgv = np.array([float(ln.split()[2]) for ln in lines if ln.startswith("LOCAL_GAP_VARIANCE")])
    gvg = np.array([[float(x) for x in ln.split()[2:5]] for ln in lines if ln.startswith("GAP_VARIANCE_GRADIENT")])
    ro = O.gap_calc_ex(m, at, atom_mask=mask, variance=True)
    assert abs(e - ro["energy"]) < 1e-10 * len(at)
    assert np.abs(epc - ro["energy_per_coordinate"]).max() < 1e-10 * len(at)
    assert np.abs(lgv - ro["local_gap_variance"]).max() < 1e-10
    assert np.abs(gvg - ro["gap_variance_gradient"]).max() < 1e-8
