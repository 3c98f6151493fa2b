"""GAP XML model files: reader and writer (host side, Python).

Schema, as QUIP reads it (src/Potentials/IPModel_GAP.f95:618-944 for
GAP_params / GAP_data / e0; the gpSparse / gpCoordinates / sparseX stanzas are
what gp_readXML consumes, see tests/GAP.xml:1-170 and
quippy/quippy/gap_tools.py:40-56):

  <ROOT_LABEL>
    <Potential label=.. init_args="IP GAP label=.."/>
    <GAP_params label=.. gap_version=..>
      <GAP_data do_core=..> <e0 Z=.. value=../> ... </GAP_data>
      <gpSparse label=.. n_coordinate=..>
        <gpCoordinates label dimensions signal_variance signal_mean sparsified
                       n_permutations covariance_type [zeta] n_sparseX sparseX_filename>
          <theta>..</theta> <descriptor>..</descriptor> <permutation i=..>..</permutation>
          <sparseX i=.. alpha=.. sparseCutoff=../> ...
        </gpCoordinates> ...
      </gpSparse>
      <command_line><![CDATA[..]]></command_line>
    </GAP_params>
  </ROOT_LABEL>

The sparseX file holds n_sparseX * dimensions values, one per line, point-major.
covariance_type 1 = ard_se, 2 = dot_product.  ``signal_variance`` is delta; the
kernels use delta**2.

The product C-ABI library parses the same XML itself (C++); this module is used
by the Python host mirror, the model generator and the tests.
"""
import os
import re
import xml.etree.ElementTree as ET
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

GAP_VERSION = 1527075646   # the gap_version this reader accepts at most (tests/GAP.xml:3)


def parse_args_str(s):
    """QUIP args string -> dict (ParamReader.f95 syntax: key=value, key={a b}, key)."""
    out = {}
    for m in re.finditer(r'(\w+)\s*=\s*(\{[^}]*\}|"[^"]*"|\S+)|(\w+)', s):
        if m.group(1):
            v = m.group(2)
            if v.startswith("{") or v.startswith('"'):
                v = v[1:-1]
            out[m.group(1)] = v
        else:
            out[m.group(3)] = "T"
    return out


def _fortran_true(v):
    """Fortran list-directed logical: T, .true., t... are true."""
    v = (v or "").strip().lstrip(".")
    return v[:1] in ("T", "t")


# Descriptor options that change the descriptor and are not built: accepted only at their
# default.  Keys GAP does not read are ignored, as its parser does (gap_fit's keys ride in
# the same string).
SOAP_ONLY_DEFAULT = {"normalise": True, "normalize": True, "average": False, "diagonal_radial": False,
                     "central_reference_all_species": False, "covariance_sigma0": 0.0, "cutoff_dexp": 0.0,
                     "cutoff_scale": 1.0, "cutoff_rate": 1.0, "radial_enhancement": 0.0, "nu_R": 2.0, "nu_S": 2.0,
                     "Z_mix": False, "R_mix": False, "sym_mix": False, "compress_file": "", "compress_mode": ""}
D2B_ONLY_DEFAULT = {"only_intra": False, "only_inter": False, "n_exponents": 1.0, "exponents": 1.0,
                    "resid_name": ""}


def _check_defaults(kv, table, s):
    for k, dflt in table.items():
        if k not in kv:
            continue
        v = kv[k]
        if isinstance(dflt, bool):
            ok = _fortran_true(v) == dflt
        elif isinstance(dflt, str):
            ok = v == dflt
        else:
            try:
                ok = float(v) == dflt
            except ValueError:
                ok = False
        if not ok:
            raise NotImplementedError("descriptor option %s=%s is not built (only its default): %r" % (k, v, s))


@dataclass
class Descriptor:
    """Parsed descriptor init string (soap or distance_2b)."""
    kind: str
    cutoff: float
    cutoff_transition_width: float = 0.5
    atom_sigma: float = 0.5
    central_weight: float = 1.0
    basis_error_exponent: float = 10.0
    n_max: int = 0
    l_max: int = 0
    n_species: int = 1
    species_Z: List[int] = field(default_factory=lambda: [0])
    n_Z: int = 1
    Z: List[int] = field(default_factory=lambda: [0])
    Z1: int = 0
    Z2: int = 0
    string: str = ""

    @property
    def dim(self):
        if self.kind == "soap":
            K = self.n_species * self.n_max
            return K * (K + 1) // 2 * (self.l_max + 1) + 1
        return 1

    @classmethod
    def from_string(cls, s):
        toks = s.split()
        kind = toks[0]
        kv = parse_args_str(" ".join(toks[1:]))
        if kind == "soap":
            _check_defaults(kv, SOAP_ONLY_DEFAULT, s)
            n_species = int(kv.get("n_species", 1))
            species = [int(z) for z in kv["species_Z"].split()] if "species_Z" in kv else [0] * n_species
            n_Z = int(kv.get("n_Z", 1))
            Z = [int(z) for z in kv["Z"].split()] if "Z" in kv else [0] * n_Z
            if len(species) != n_species or len(Z) != n_Z:
                raise ValueError("soap descriptor: n_species/species_Z or n_Z/Z mismatch in %r" % s)
            return cls("soap", float(kv["cutoff"]),
                       float(kv.get("cutoff_transition_width", 0.5)),
                       float(kv.get("atom_sigma", kv.get("atom_gaussian_width", 0.5))),
                       float(kv.get("central_weight", 1.0)),
                       float(kv.get("basis_error_exponent", 10.0)),
                       int(kv["n_max"]), int(kv["l_max"]), n_species, species, n_Z, Z, string=s)
        if kind == "distance_2b":
            _check_defaults(kv, D2B_ONLY_DEFAULT, s)
            return cls("distance_2b", float(kv["cutoff"]),
                       float(kv.get("cutoff_transition_width", 0.5)),
                       Z1=int(kv.get("Z1", 0)), Z2=int(kv.get("Z2", 0)), string=s)
        raise ValueError("unsupported descriptor type %r" % kind)


@dataclass
class GPCoordinate:
    label: str
    descriptor: Descriptor
    covariance_type: int            # 1 ard_se, 2 dot_product
    delta: float
    signal_mean: float
    zeta: float
    sparseX: np.ndarray             # (n_sparse, dim)
    alpha: np.ndarray
    sparse_cutoff: np.ndarray
    theta: Optional[np.ndarray] = None

    @property
    def n_sparse(self):
        return self.sparseX.shape[0]


@dataclass
class GAPModel:
    label: str
    gap_version: int
    e0: np.ndarray                  # (119,) indexed by Z
    coords: List[GPCoordinate]
    command_line: str = ""

    @property
    def cutoff(self):
        """IPModel_GAP%cutoff = max over descriptors (IPModel_GAP.f95:182-188)."""
        return max(c.descriptor.cutoff for c in self.coords)


def read_gap_xml(path=None, xml_text=None, label=None, xml_dir=None):
    """Parse a GAP XML potential (label-matched like IPModel_GAP.f95:629-653)."""
    if xml_text is None:
        with open(path) as fh:
            xml_text = fh.read()
        xml_dir = os.path.dirname(os.path.abspath(path))
    xml_dir = xml_dir or "."
    root = ET.fromstring(xml_text)
    params = None
    for gp in root.iter("GAP_params"):
        if label is None or gp.get("label") == label:
            params = gp
            break
    if params is None:
        raise ValueError("GAP_params with label %r not found" % label)
    ver = int(params.get("gap_version", "0"))
    if ver > GAP_VERSION:
        raise ValueError("gap_version %d is newer than supported %d" % (ver, GAP_VERSION))
    e0 = np.zeros(119)
    data = params.find("GAP_data")
    if data is not None:
        if data.get("e0") is not None:          # every species, then per-Z entries (IPModel_GAP.f95:670-674)
            e0[:] = float(data.get("e0").split()[0])
        if _fortran_true(data.get("do_pca", "F")):
            raise NotImplementedError("GAP_data do_pca=T is not built")
        for e in data.findall("e0"):
            e0[int(e.get("Z"))] = float(e.get("value"))
    coords = []
    sparse = params.find("gpSparse")
    for gc in sparse.findall("gpCoordinates"):
        desc = Descriptor.from_string(gc.find("descriptor").text.strip())
        d = int(gc.get("dimensions"))
        n = int(gc.get("n_sparseX"))
        cov = int(gc.get("covariance_type"))
        zeta = float(gc.get("zeta", "1.0"))
        pts = sorted(gc.findall("sparseX"), key=lambda e: int(e.get("i")))
        alpha = np.array([float(e.get("alpha")) for e in pts])
        scut = np.array([float(e.get("sparseCutoff", "1.0")) for e in pts])
        fn = gc.get("sparseX_filename")
        X = np.loadtxt(os.path.join(xml_dir, fn), dtype=np.float64).reshape(n, d)
        theta = None
        th = gc.find("theta")
        if th is not None and th.text and th.text.strip():
            theta = np.array([float(v) for v in th.text.split()])
        if d != desc.dim:
            raise ValueError("gpCoordinates %s: dimensions %d != descriptor dimension %d"
                             % (gc.get("label"), d, desc.dim))
        if float(gc.get("signal_mean", "0")) != 0.0:
            raise NotImplementedError("gpCoordinates %s: signal_mean=%s is not built (only 0)"
                                      % (gc.get("label"), gc.get("signal_mean")))
        if int(gc.get("n_permutations", "1")) > 1:
            raise NotImplementedError("gpCoordinates %s: n_permutations=%s is not built (only 1)"
                                      % (gc.get("label"), gc.get("n_permutations")))
        coords.append(GPCoordinate(gc.get("label"), desc, cov, float(gc.get("signal_variance")),
                                   float(gc.get("signal_mean", "0")), zeta, X, alpha, scut, theta))
    cl = params.find("command_line")
    return GAPModel(params.get("label"), ver, e0, coords, (cl.text or "") if cl is not None else "")


def write_gap_xml(path, model: GAPModel):
    """Write ``model`` in the GAP XML layout plus one sparseX file per coordinate."""
    base = os.path.basename(path)
    lab = model.label
    lines = ["<%s>" % lab,
             '<Potential label="%s" init_args="IP GAP label=%s"/>' % (lab, lab),
             '<GAP_params label="%s" gap_version="%d">' % (lab, model.gap_version),
             '  <GAP_data do_core="F">']
    for z in range(1, 119):
        lines.append('    <e0 Z="%d" value="%.17g"/>' % (z, model.e0[z]))
    lines.append("  </GAP_data>")
    lines.append('  <gpSparse label="%s" n_coordinate="%d">' % (lab, len(model.coords)))
    for k, c in enumerate(model.coords):
        fn = "%s.sparseX.%s" % (base, c.label)
        zeta = ' zeta="%.17g"' % c.zeta if c.covariance_type == 2 else ""
        lines.append('    <gpCoordinates label="%s" dimensions="%d" signal_variance="%.17g" signal_mean="%.17g" '
                     'sparsified="T" n_permutations="1" covariance_type="%d"%s n_sparseX="%d" sparseX_filename="%s">'
                     % (c.label, c.descriptor.dim, c.delta, c.signal_mean, c.covariance_type, zeta, c.n_sparse, fn))
        if c.theta is not None:
            lines.append("      <theta>%s</theta>" % " ".join("%.17g" % t for t in c.theta))
        lines.append("      <descriptor>%s</descriptor>" % c.descriptor.string)
        lines.append('      <permutation i="1">1</permutation>')
        for i in range(c.n_sparse):
            lines.append('      <sparseX i="%d" alpha="%.17g" sparseCutoff="%.17g"/>' % (i + 1, c.alpha[i], c.sparse_cutoff[i]))
        lines.append("    </gpCoordinates>")
        np.savetxt(os.path.join(os.path.dirname(os.path.abspath(path)), fn), c.sparseX.reshape(-1), fmt="%.20e")
    lines.append("  </gpSparse>")
    lines.append("  <command_line><![CDATA[%s]]></command_line>" % model.command_line)
    lines.append("</GAP_params>")
    lines.append("</%s>" % lab)
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")
