"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  The product path (quip_amd) never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
MAX_SPECIES = 16


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE, "-B"])


class OrcCoord(C.Structure):
    _fields_ = [("kind", C.c_int), ("cutoff", C.c_double), ("cutoff_transition_width", C.c_double),
                ("atom_sigma", C.c_double), ("central_weight", C.c_double), ("basis_error_exponent", C.c_double),
                ("n_max", C.c_int), ("l_max", C.c_int), ("n_species", C.c_int),
                ("species_Z", C.c_int * MAX_SPECIES), ("n_Z", C.c_int), ("Z", C.c_int * MAX_SPECIES),
                ("Z1", C.c_int), ("Z2", C.c_int), ("covariance", C.c_int), ("delta", C.c_double),
                ("zeta", C.c_double), ("n_sparse", C.c_int), ("dim", C.c_int),
                ("sparseX", C.POINTER(C.c_double)), ("alpha", C.POINTER(C.c_double)),
                ("sparse_cutoff", C.POINTER(C.c_double)), ("theta", C.POINTER(C.c_double))]


class OrcCalcOpts(C.Structure):
    _fields_ = [("atom_mask", C.c_void_p), ("only_descriptor", C.c_int), ("want_variance", C.c_int),
                ("variance_regularisation", C.c_double), ("energy_per_coordinate", C.c_void_p),
                ("local_gap_variance", C.c_void_p), ("gap_variance_gradient", C.c_void_p), ("inst_cap", C.c_int),
                ("inst_var", C.c_void_p), ("inst_cut2", C.c_void_p), ("inst_coord", C.c_void_p),
                ("inst_atoms", C.c_void_p), ("n_inst", C.c_void_p)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = C.CDLL(_LIB)
        _lib.orc_nl_build.restype = C.c_void_p
        _lib.orc_nl_build.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.POINTER(C.c_int)]
        _lib.orc_nl_counts.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_void_p]
        _lib.orc_nl_export.argtypes = [C.c_void_p] + [C.c_void_p] * 8
        _lib.orc_nl_free.argtypes = [C.c_void_p]
        _lib.orc_soap_basis.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int, C.c_void_p, C.c_void_p,
                                        C.POINTER(C.c_double)]
        _lib.orc_soap_descriptors.argtypes = [C.POINTER(OrcCoord), C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                              C.c_void_p, C.c_double, C.c_void_p, C.c_void_p, C.POINTER(C.c_int)]
        _lib.orc_soap_grad_atom.argtypes = [C.POINTER(OrcCoord), C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                            C.c_void_p, C.c_double, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                            C.c_int, C.POINTER(C.c_int)]
        _lib.orc_gap_calc.argtypes = [C.c_int, C.POINTER(OrcCoord), C.c_void_p, C.c_double, C.c_int, C.c_void_p,
                                      C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_int] + [C.c_void_p] * 5
        _lib.orc_gap_calc_ex.argtypes = _lib.orc_gap_calc.argtypes + [C.POINTER(OrcCalcOpts)]
        _lib.orc_gp_predict.restype = C.c_double
        _lib.orc_gp_predict.argtypes = [C.POINTER(OrcCoord), C.c_void_p, C.c_void_p]
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def _geom(at):
    pos = np.ascontiguousarray(at.positions, dtype=np.float64)
    Z = np.ascontiguousarray(at.numbers, dtype=np.int32)
    lat = np.ascontiguousarray(at.cell, dtype=np.float64)
    pbc = np.ascontiguousarray(at.pbc, dtype=np.int32)
    return pos, Z, lat, pbc


def neighbour_list(at, cutoff):
    """QUIP calc_connect tables: dict of CSR arrays (0-based)."""
    L = lib()
    pos, Z, lat, pbc = _geom(at)
    err = C.c_int(0)
    h = L.orc_nl_build(len(at), _p(pos), _p(lat), _p(pbc), cutoff, C.byref(err))
    if not h:
        raise RuntimeError("orc_nl_build failed (%d)" % err.value)
    t1, t2 = C.c_int(), C.c_int()
    cells = np.zeros(3, np.int32)
    L.orc_nl_counts(h, C.byref(t1), C.byref(t2), _p(cells))
    N = len(at)
    out = dict(n1_off=np.zeros(N + 1, np.int32), n1_j=np.zeros(t1.value, np.int32),
               n1_shift=np.zeros((t1.value, 3), np.int32), n1_r=np.zeros(t1.value),
               n2_off=np.zeros(N + 1, np.int32), n2_j=np.zeros(t2.value, np.int32),
               n2_idx=np.zeros(t2.value, np.int32), map_shift=np.zeros((N, 3), np.int32), cells=cells)
    L.orc_nl_export(h, *[_p(out[k]) for k in ("n1_off", "n1_j", "n1_shift", "n1_r", "n2_off", "n2_j",
                                               "n2_idx", "map_shift")])
    L.orc_nl_free(h)
    return out


def make_coords(coords):
    """GAP model coordinates (quip_amd.gap_xml.GPCoordinate list) -> OrcCoord array (keeps refs)."""
    arr = (OrcCoord * len(coords))()
    keep = []
    for k, c in enumerate(coords):
        d = c.descriptor
        o = arr[k]
        o.kind = 1 if d.kind == "soap" else 2
        o.cutoff = d.cutoff
        o.cutoff_transition_width = d.cutoff_transition_width
        o.atom_sigma = d.atom_sigma
        o.central_weight = d.central_weight
        o.basis_error_exponent = d.basis_error_exponent
        o.n_max, o.l_max, o.n_species, o.n_Z = d.n_max, d.l_max, d.n_species, d.n_Z
        for s, z in enumerate(d.species_Z):
            o.species_Z[s] = z
        for s, z in enumerate(d.Z):
            o.Z[s] = z
        o.Z1, o.Z2 = d.Z1, d.Z2
        o.covariance = c.covariance_type
        o.delta, o.zeta = c.delta, c.zeta
        o.n_sparse, o.dim = c.sparseX.shape
        X = np.ascontiguousarray(c.sparseX, dtype=np.float64)
        al = np.ascontiguousarray(c.alpha, dtype=np.float64)
        sc = np.ascontiguousarray(c.sparse_cutoff, dtype=np.float64)
        th = np.ascontiguousarray(c.theta if c.theta is not None else np.ones(c.sparseX.shape[1]), dtype=np.float64)
        keep += [X, al, sc, th]
        o.sparseX = X.ctypes.data_as(C.POINTER(C.c_double))
        o.alpha = al.ctypes.data_as(C.POINTER(C.c_double))
        o.sparse_cutoff = sc.ctypes.data_as(C.POINTER(C.c_double))
        o.theta = th.ctypes.data_as(C.POINTER(C.c_double))
    return arr, keep


def soap_basis(cutoff, atom_sigma, n_max, basis_error_exponent=10.0):
    rb = np.zeros(n_max)
    T = np.zeros((n_max, n_max))
    u00 = C.c_double()
    lib().orc_soap_basis(cutoff, atom_sigma, basis_error_exponent, n_max, _p(rb), _p(T), C.byref(u00))
    return rb, T, u00.value


def soap_descriptors(coord, at, nl_cutoff=None):
    arr, keep = make_coords([coord])
    pos, Z, lat, pbc = _geom(at)
    d = coord.descriptor.dim
    X = np.zeros((len(at), d))
    idx = np.zeros(len(at), np.int32)
    nd = C.c_int()
    rc = nl_cutoff if nl_cutoff is not None else coord.descriptor.cutoff
    st = lib().orc_soap_descriptors(arr, len(at), _p(pos), _p(Z), _p(lat), _p(pbc), rc, _p(X), _p(idx), C.byref(nd))
    if st:
        raise RuntimeError("orc_soap_descriptors failed (%d)" % st)
    return X[:nd.value], idx[:nd.value]


def soap_grad_atom(coord, at, atom, nl_cutoff=None, max_entries=512):
    arr, keep = make_coords([coord])
    pos, Z, lat, pbc = _geom(at)
    d = coord.descriptor.dim
    x = np.zeros(d)
    grad = np.zeros((max_entries, 3, d))
    ii = np.zeros(max_entries, np.int32)
    ne = C.c_int()
    rc = nl_cutoff if nl_cutoff is not None else coord.descriptor.cutoff
    st = lib().orc_soap_grad_atom(arr, len(at), _p(pos), _p(Z), _p(lat), _p(pbc), rc, atom, _p(x), _p(grad),
                                  _p(ii), max_entries, C.byref(ne))
    if st:
        raise RuntimeError("orc_soap_grad_atom failed (%d)" % st)
    return x, grad[:ne.value], ii[:ne.value]


def gap_calc(model, at, nl_cutoff=None, n_threads=0, E_scale=1.0):
    """IPModel_GAP_Calc restatement: dict(energy, local_energy, forces (N,3), virial (3,3), local_virial (N,9))."""
    arr, keep = make_coords(model.coords)
    pos, Z, lat, pbc = _geom(at)
    N = len(at)
    e = np.zeros(1)
    le = np.zeros(N)
    f = np.zeros((N, 3))
    v = np.zeros((3, 3))        # stored column-major -> v[b, a] = virial(a, b)
    lv = np.zeros((N, 9))
    e0 = np.ascontiguousarray(model.e0, dtype=np.float64)
    rc = nl_cutoff if nl_cutoff is not None else model.cutoff
    st = lib().orc_gap_calc(len(model.coords), arr, _p(e0), E_scale, N, _p(pos), _p(Z), _p(lat), _p(pbc), rc,
                            n_threads, _p(e), _p(le), _p(f), _p(v), _p(lv))
    if st:
        raise RuntimeError("orc_gap_calc failed (%d)" % st)
    return dict(energy=e[0], local_energy=le, forces=f, virial=v.T.copy(), local_virial=lv)


def gp_predict(coord, x, grad=False):
    """gp_predict of one descriptor vector: prediction (and its gradient)."""
    arr, keep = make_coords([coord])
    x = np.ascontiguousarray(x, dtype=np.float64)
    g = np.zeros(len(x)) if grad else None
    e = lib().orc_gp_predict(arr, _p(x), _p(g))
    return (e, g) if grad else e


def gap_calc_ex(model, at, nl_cutoff=None, E_scale=1.0, atom_mask=None, only_descriptor=0, variance=False,
                variance_regularisation=1e-3, n_threads=0, max_instances=0):
    """IPModel_GAP_Calc with its args_str options: atom_mask (N booleans), only_descriptor
    (1-based coordinate), energy_per_coordinate, and with `variance` local_gap_variance,
    gap_variance_gradient (N,3) and the per-instance variances (print_gap_variance)."""
    arr, keep = make_coords(model.coords)
    pos, Z, lat, pbc = _geom(at)
    N = len(at)
    e = np.zeros(1)
    le = np.zeros(N)
    f = np.zeros((N, 3))
    v = np.zeros((3, 3))
    lv = np.zeros((N, 9))
    epc = np.zeros(len(model.coords))
    lgv = np.zeros(N)
    gvg = np.zeros((N, 3))
    cap = max_instances
    ivar, icut, icoord, iatoms = np.zeros(cap + 1), np.zeros(cap + 1), np.zeros(cap + 1, np.int32), \
        np.zeros((cap + 1, 2), np.int32)
    ninst = np.zeros(1, np.int32)
    mask = None if atom_mask is None else np.ascontiguousarray(np.asarray(atom_mask, dtype=bool).astype(np.uint8))
    o = OrcCalcOpts()
    o.atom_mask = _p(mask)
    o.only_descriptor = int(only_descriptor)
    o.want_variance = 1 if variance else 0
    o.variance_regularisation = variance_regularisation
    o.energy_per_coordinate = _p(epc)
    o.local_gap_variance = _p(lgv)
    o.gap_variance_gradient = _p(gvg)
    o.inst_cap = cap
    o.inst_var, o.inst_cut2, o.inst_coord, o.inst_atoms = _p(ivar), _p(icut), _p(icoord), _p(iatoms)
    o.n_inst = _p(ninst) if cap > 0 else None
    e0 = np.ascontiguousarray(model.e0, dtype=np.float64)
    rc = nl_cutoff if nl_cutoff is not None else model.cutoff
    st = lib().orc_gap_calc_ex(len(model.coords), arr, _p(e0), E_scale, N, _p(pos), _p(Z), _p(lat), _p(pbc), rc,
                               n_threads, _p(e), _p(le), _p(f), _p(v), _p(lv), C.byref(o))
    if st:
        raise RuntimeError("orc_gap_calc_ex failed (%d)" % st)
    out = dict(energy=e[0], local_energy=le, forces=f, virial=v.T.copy(), local_virial=lv,
               energy_per_coordinate=epc)
    if variance:
        n = min(int(ninst[0]), cap)
        out.update(local_gap_variance=lgv, gap_variance_gradient=gvg, instance_variance=ivar[:n],
                   instance_cutoff2=icut[:n], instance_coord=icoord[:n], instance_atoms=iatoms[:n])
    return out
