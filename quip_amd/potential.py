"""Host-side mirror of QUIP's Potential / IPModel_GAP interface over libgapx.so.

``Potential("IP GAP", param_filename="gp.xml")`` loads a GAP XML model the way
quippy.potential.Potential does (quippy/quippy/potential.py:43-90 ->
Potential_Filename_Initialise, Potential.f95:438-468: the sparseX files are
resolved relative to the XML file).  ``calc(at, args_str=...)`` follows
IPModel_GAP_Calc (IPModel_GAP.f95:233-605): the args string names the outputs
(``energy``, ``force``, ``virial``, ``local_energy``, ``local_virial``), unknown
keys are ignored like ParamReader's ignore_unknown, results are stored on the
Atoms object (info for per-config values, arrays for per-atom values) and also
returned.  The ASE-calculator conventions of quippy are kept: stress =
-virial / volume in Voigt order (potential.py:281-284).

All arithmetic runs in the HIP library on the GPU; there is no CPU path.
"""
import ctypes as C
import os

import numpy as np

from . import _lib as L
from .gap_xml import parse_args_str


class Potential:
    def __init__(self, args_str="IP GAP", param_filename=None, param_str=None, label=None, device=0,
                 precision="fp64", calc_args=""):
        kv = parse_args_str(args_str or "")
        if label is None:
            label = kv.get("label")
        if param_filename is not None:
            with open(param_filename) as fh:
                xml = fh.read()
            xml_dir = os.path.dirname(os.path.abspath(param_filename))
        elif param_str is not None:
            xml, xml_dir = param_str, os.getcwd()
        else:
            raise ValueError("Potential needs param_filename or param_str")
        lib = L.lib()
        if lib.gapx_api_version() != 1:
            raise L.GapxError("libgapx API version %d, expected 1" % lib.gapx_api_version())
        h = C.c_void_p()
        raw = xml.encode()
        L.check(lib.gapx_model_load(raw, len(raw), label.encode() if label else None, xml_dir.encode(), device,
                                    C.byref(h)), "gapx_model_load")
        self._h = h
        self.label = label
        self.calc_args = calc_args
        self.set_precision(precision)
        if "E_scale" in kv:
            L.check(lib.gapx_model_set_e_scale(h, float(kv["E_scale"])), "set_e_scale")
        self.results = {}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and L._lib is not None:
            L._lib.gapx_model_free(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def set_precision(self, precision):
        L.check(L.lib().gapx_model_set_precision(self._h, L.PREC[precision]), "set_precision")
        self.precision = precision

    def cutoff(self):
        c = C.c_double()
        L.check(L.lib().gapx_model_cutoff(self._h, C.byref(c)), "gapx_model_cutoff")
        return c.value

    def n_coordinates(self):
        n = C.c_int()
        L.check(L.lib().gapx_model_info(self._h, C.byref(n), -1, None, None), "gapx_model_info")
        return n.value

    def set_nl_reuse(self, on=True):
        """Keep the neighbour list within its skin (calc_connect with cutoff_skin)."""
        L.check(L.lib().gapx_model_set_nl_reuse(self._h, 1 if on else 0), "set_nl_reuse")

    def nl_rebuilt(self):
        """Whether the last calc rebuilt the neighbour list (calc_connect's did_rebuild)."""
        r = C.c_int()
        L.check(L.lib().gapx_last_nl_rebuilt(self._h, C.byref(r)), "last_nl_rebuilt")
        return bool(r.value)

    def set_graphs(self, on=True):
        """Replay the launch sequence after the neighbour count as a HIP graph."""
        L.check(L.lib().gapx_model_set_graphs(self._h, 1 if on else 0), "set_graphs")

    def set_timing(self, on=True):
        L.check(L.lib().gapx_model_set_timing(self._h, 1 if on else 0), "set_timing")

    STAGES = ("neighbour_list", "soap_forward", "gemm1", "gemm2", "soap_adjoint", "d2b_gather")

    def last_timings(self):
        """Stage times (ms) of the last calc, from HIP events on the model stream."""
        t = np.zeros(6)
        L.check(L.lib().gapx_last_timings(self._h, L.ptr(t)), "last_timings")
        return dict(zip(self.STAGES, t))

    def timing_totals(self):
        """(stage -> summed ms, number of calls) since set_timing(True)."""
        t = np.zeros(6)
        n = C.c_int64()
        L.check(L.lib().gapx_timing_totals(self._h, L.ptr(t), C.byref(n)), "timing_totals")
        return dict(zip(self.STAGES, t)), n.value

    # args_str keys this host reads (IPModel_GAP.f95:324-335 plus the output names); others
    # are ignored as ParamReader's ignore_unknown does
    def calc(self, at, energy=None, force=None, virial=None, local_energy=None, local_virial=None, args_str="",
             cutoff=None):
        """IPModel_GAP_Calc (IPModel_GAP.f95:233-605).

        `energy`, `force`, `virial`, `local_energy`, `local_virial` may be True (the
        result is allocated) or caller-owned float64 arrays that are filled: energy of
        one element, force (N, 3) or QUIP's (3, N), virial (3, 3), local_energy (N,),
        local_virial (N, 9) or (9, N).  The results are computed into fresh buffers and
        copied into the caller's arrays with the layout their shape names.

        args_str keys: the output names above; `atom_mask_name` (a boolean array of
        at.arrays: only those atoms are descriptor centres / first atoms of pair
        descriptors and get e0); `only_descriptor` (1-based coordinate); `energy_per_coordinate`
        (name of the at.info entry that receives one energy per coordinate);
        `local_gap_variance` (name of the at.arrays entry for the per-atom variance; with
        forces or virials also at.arrays["gap_variance_gradient"]); `print_gap_variance`;
        `gap_variance_regularisation` (default 0.001); `cutoff`.  `r_scale` / `E_scale` are
        errors at calc time, as in the reference (:348-350)."""
        kv = parse_args_str((self.calc_args or "") + " " + (args_str or ""))
        if "r_scale" in kv or "E_scale" in kv:
            raise ValueError("IPModel_GAP_Calc: r_scale / E_scale cannot be given at calc time "
                             "(IPModel_GAP.f95:348-350)")
        N = len(at)
        want = L.WANT_E
        shapes = {"energy": (1,), "force": (N, 3), "virial": (3, 3), "local_energy": (N,), "local_virial": (N, 9)}
        flags = {"energy": L.WANT_E, "force": L.WANT_F, "virial": L.WANT_V, "local_energy": L.WANT_LOCAL_E,
                 "local_virial": L.WANT_LOCAL_V}
        given = {"energy": energy, "force": force, "virial": virial, "local_energy": local_energy,
                 "local_virial": local_virial}
        bufs = {}
        for name, val in given.items():
            if (val is not None and val is not False) or name in kv:
                want |= flags[name]
                if isinstance(val, np.ndarray) and val.dtype != np.float64:
                    raise TypeError("%s must be a float64 array (got %s)" % (name, val.dtype))
                bufs[name] = np.zeros(shapes[name])
        mask = None
        if "atom_mask_name" in kv:
            name = kv["atom_mask_name"]
            if name not in at.arrays:
                raise ValueError("IPModel_GAP_Calc did not find %s property in the atoms object" % name)
            mask = np.ascontiguousarray(np.asarray(at.arrays[name]).astype(bool).astype(np.uint8))
            if mask.shape != (N,):
                raise ValueError("atom mask %s has shape %s, expected (%d,)" % (name, mask.shape, N))
        only = int(kv.get("only_descriptor", 0))
        epc_name = kv.get("energy_per_coordinate", "")
        lgv_name = kv.get("local_gap_variance", "")
        print_var = kv.get("print_gap_variance", "F").upper().startswith(("T", ".T"))
        reg = float(kv.get("gap_variance_regularisation", 1e-3))
        grad = bool(want & (L.WANT_F | L.WANT_V | L.WANT_LOCAL_V))
        ncoord = self.n_coordinates()
        epc = np.zeros(ncoord) if epc_name else None
        lgv = np.zeros(N) if lgv_name else None
        gvg = np.zeros((N, 3)) if lgv_name and grad else None
        if print_var:
            want |= L.WANT_VARIANCE
        pos, Z = L.f64(at.positions), L.i32(at.numbers)
        lat, pbc = L.f64(at.cell), L.i32(at.pbc)
        vflat = np.zeros(9) if "virial" in bufs else None
        e = bufs.get("energy", np.zeros(1))
        rc = float(cutoff) if cutoff is not None else float(kv.get("cutoff", 0.0))
        L.check(L.lib().gapx_calc_ex(self._h, N, L.ptr(pos), L.ptr(Z), L.ptr(lat), L.ptr(pbc), rc, want,
                                     L.ptr(mask), only, reg, L.ptr(e), L.ptr(bufs.get("local_energy")),
                                     L.ptr(bufs.get("force")), L.ptr(vflat), L.ptr(bufs.get("local_virial")),
                                     L.ptr(epc), L.ptr(lgv), L.ptr(gvg)), "gapx_calc_ex")
        out = {"energy": float(e[0])}
        at.info["energy"] = out["energy"]
        if "force" in bufs:
            out["force"] = bufs["force"]
            at.arrays["force"] = out["force"]
        if vflat is not None:
            out["virial"] = vflat.reshape(3, 3).T.copy()   # stored column-major: vflat[3*b + a] = virial(a, b)
            at.info["virial"] = out["virial"]
        if "local_energy" in bufs:
            out["local_energy"] = bufs["local_energy"]
            at.arrays["local_energy"] = out["local_energy"]
        if "local_virial" in bufs:
            out["local_virial"] = bufs["local_virial"]
            at.arrays["local_virial"] = out["local_virial"]
        if epc is not None:
            out["energy_per_coordinate"] = epc
            at.info[epc_name] = epc
        if lgv is not None:
            out["local_gap_variance"] = lgv
            at.arrays[lgv_name] = lgv
            if gvg is not None:
                out["gap_variance_gradient"] = gvg
                at.arrays["gap_variance_gradient"] = gvg
        if print_var:
            out["gap_variance"] = self._print_variances(ncoord, only)
        # caller-owned arrays receive the results in the layout their shape names
        for name, val in given.items():
            if isinstance(val, np.ndarray):
                res = out[name] if name != "energy" else np.array([out["energy"]])
                if name in ("force", "local_virial") and val.shape == res.T.shape and val.shape != res.shape:
                    np.copyto(val, res.T)
                elif val.size == res.size:
                    np.copyto(val, res.reshape(val.shape))
                else:
                    raise ValueError("%s has shape %s, expected %s" % (name, val.shape, res.shape))
        self.results = out
        return out

    def descriptor_variances(self, coord):
        """Per-instance GAP variances of coordinate `coord` (0-based) from the last calc
        that formed them: (variance, covariance_cutoff^2, atoms (n, 2); -1 for soap)."""
        n = C.c_int()
        L.check(L.lib().gapx_descriptor_variances(self._h, coord, 0, None, None, None, C.byref(n)),
                "gapx_descriptor_variances")
        v, c2, a = np.zeros(n.value), np.zeros(n.value), np.zeros((n.value, 2), np.int32)
        L.check(L.lib().gapx_descriptor_variances(self._h, coord, n.value, L.ptr(v), L.ptr(c2), L.ptr(a),
                                                  C.byref(n)), "gapx_descriptor_variances")
        return v, c2, a

    def _print_variances(self, ncoord, only):
        """print_gap_variance (IPModel_GAP.f95:401-404, 506-516): one line per instance."""
        label = self.label or ""
        print("GAP_VARIANCE potential %s calculating for %d descriptors" % (label, ncoord))
        res = []
        for k in range(ncoord):
            if only > 0 and ncoord > 1 and k != only - 1:
                continue
            v, c2, a = self.descriptor_variances(k)
            res.append((v, c2, a))
            if len(v) == 0:
                print("GAP_VARIANCE potential %s descriptor %d not found" % (label, k + 1))
            for i in range(len(v)):
                print("GAP_VARIANCE potential %s descriptor %d var( %d ) = %.16g * %.16g cutoff"
                      % (label, k + 1, i + 1, v[i], c2[i]))
                ii = [x + 1 for x in a[i] if x >= 0]
                print("GAP_VARIANCE potential %s descriptor %d ii( %d ) = %s"
                      % (label, k + 1, i + 1, " ".join(str(x) for x in ii)))
        return res

    # --- ASE-calculator style accessors (quippy/quippy/potential.py:177-300) ---
    def get_property(self, name, at):
        r = self.calc(at, force=True, virial=True)
        if name == "energy":
            return r["energy"]
        if name == "forces":
            return r["force"]
        if name == "stress":
            s = -r["virial"] / at.get_volume()
            return np.array([s[0, 0], s[1, 1], s[2, 2], s[1, 2], s[0, 2], s[0, 1]])
        raise KeyError(name)

    def calc_lammps(self, nlocal, nghost, numbers, ilist, num_neigh, neigh, positions, lattice=None, tags=None):
        """The LAMMPS pair style call (quip_lammps_wrapper.f95:30-156, gapx_calc_lammps):
        `numbers` and `positions` (n_all x 3) cover the local atoms followed by the
        ghosts; `ilist` (0-based) the local atoms whose full neighbour lists are
        packed in `neigh` (1-based, `num_neigh[k]` for ilist[k]).  Returns energy,
        local_energy (n_all), virial (3x3), local_virial (n_all x 9) and force
        (n_all x 3), ghost entries included."""
        n = nlocal + nghost
        Z, pos = L.i32(numbers), L.f64(np.asarray(positions).reshape(n, 3))
        il, nn, ng = L.i32(ilist), L.i32(num_neigh), L.i32(neigh)
        lat = L.f64(np.eye(3) if lattice is None else lattice)
        tg = L.i32(np.arange(1, n + 1) if tags is None else tags)
        e, le, v, lv, f = np.zeros(1), np.zeros(n), np.zeros(9), np.zeros((n, 9)), np.zeros((n, 3))
        L.check(L.lib().gapx_calc_lammps(self._h, nlocal, nghost, L.ptr(Z), L.ptr(tg), len(il), len(ng), L.ptr(il),
                                         L.ptr(nn), L.ptr(ng), L.ptr(lat), L.ptr(pos), L.ptr(e), L.ptr(le), L.ptr(v),
                                         L.ptr(lv), L.ptr(f)), "gapx_calc_lammps")
        return {"energy": float(e[0]), "local_energy": le, "virial": v.reshape(3, 3).T, "local_virial": lv,
                "force": f}

    def calc_batch(self, frames, want=("energy", "force", "virial"), cutoff=0.0):
        """Many structures in one pass of the engine (gapx_calc_batch); returns
        one dict per structure with the same keys as calc()."""
        off = np.zeros(len(frames) + 1, np.int32)
        off[1:] = np.cumsum([len(a) for a in frames])
        pos = L.f64(np.vstack([a.positions for a in frames]))
        Z = L.i32(np.concatenate([a.numbers for a in frames]))
        lats = L.f64(np.stack([a.cell for a in frames]))
        pbcs = L.i32(np.stack([a.pbc for a in frames]))
        n = off[-1]
        E = np.zeros(len(frames))
        F = np.zeros((n, 3)) if "force" in want else None
        V = np.zeros((len(frames), 9)) if "virial" in want else None
        LE = np.zeros(n) if "local_energy" in want else None
        LV = np.zeros((n, 9)) if "local_virial" in want else None
        w = (L.WANT_E | (L.WANT_F if F is not None else 0) | (L.WANT_V if V is not None else 0) |
             (L.WANT_LOCAL_E if LE is not None else 0) | (L.WANT_LOCAL_V if LV is not None else 0))
        L.check(L.lib().gapx_calc_batch(self._h, len(frames), L.ptr(off), L.ptr(pos), L.ptr(Z), L.ptr(lats),
                                        L.ptr(pbcs), cutoff, w, L.ptr(E), L.ptr(LE), L.ptr(F), L.ptr(V), L.ptr(LV)),
                "gapx_calc_batch")
        res = []
        for s in range(len(frames)):
            a, b = off[s], off[s + 1]
            r = {"energy": E[s]}
            if F is not None:
                r["force"] = F[a:b]
            if V is not None:
                r["virial"] = V[s].reshape(3, 3).T
            if LE is not None:
                r["local_energy"] = LE[a:b]
            if LV is not None:
                r["local_virial"] = LV[a:b]
            res.append(r)
        return res

    def descriptors(self, at, coord=0, cutoff=0.0):
        """SOAP vectors of coordinate `coord` for every centre atom (descriptor calc)."""
        d = C.c_int()
        L.check(L.lib().gapx_model_info(self._h, None, coord, C.byref(d), None), "gapx_model_info")
        N = len(at)
        X = np.zeros((max(N, 1), d.value))
        idx = np.zeros(max(N, 1), np.int32)
        nd = C.c_int()
        pos, Z, lat, pbc = L.f64(at.positions), L.i32(at.numbers), L.f64(at.cell), L.i32(at.pbc)
        L.check(L.lib().gapx_descriptors(self._h, coord, N, L.ptr(pos), L.ptr(Z), L.ptr(lat), L.ptr(pbc), cutoff,
                                         L.ptr(X), L.ptr(idx), C.byref(nd)), "gapx_descriptors")
        return X[:nd.value], idx[:nd.value]


def connect(at, cutoff, device=0):
    """QUIP calc_connect tables built on the GPU (same layout as the oracle's)."""
    lib = L.lib()
    pos, lat, pbc = L.f64(at.positions), L.f64(at.cell), L.i32(at.pbc)
    h = C.c_void_p()
    L.check(lib.gapx_connect(device, len(at), L.ptr(pos), L.ptr(lat), L.ptr(pbc), cutoff, C.byref(h)), "gapx_connect")
    try:
        t1, t2 = C.c_int(), C.c_int()
        cells = np.zeros(3, np.int32)
        L.check(lib.gapx_nl_counts(h, C.byref(t1), C.byref(t2), L.ptr(cells)), "gapx_nl_counts")
        N = len(at)
        out = dict(n1_off=np.zeros(N + 1, np.int32), n1_j=np.zeros(t1.value, np.int32),
                   n1_shift=np.zeros((t1.value, 3), np.int32), n1_r=np.zeros(t1.value),
                   n2_off=np.zeros(N + 1, np.int32), n2_j=np.zeros(t2.value, np.int32),
                   n2_idx=np.zeros(t2.value, np.int32), map_shift=np.zeros((N, 3), np.int32), cells=cells)
        L.check(lib.gapx_nl_export(h, *[L.ptr(out[k]) for k in ("n1_off", "n1_j", "n1_shift", "n1_r", "n2_off",
                                                                  "n2_j", "n2_idx", "map_shift")]), "gapx_nl_export")
        return out
    finally:
        lib.gapx_nl_free(h)
