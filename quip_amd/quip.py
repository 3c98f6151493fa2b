"""Command-line evaluator with the `quip` program's interface for the GAP path.

    python -m quip_amd.quip atoms_filename=in.xyz param_filename=gp.xml E F V [L] \
        [init_args="IP GAP"] [calc_args=...] [cutoff=-1] [cutoff_skin=0.5] \
        [output_file=stdout] [real_format=%16.8f] [precision=fp64] [device=0]

Follows src/Programs/quip.f95: the parameters are registered at quip.f95:144-234
(bare ``E`` means ``E=T``; ``energy``, ``forces``, ``virial``, ``local`` are the
long names), every frame of ``atoms_filename`` is evaluated in order
(quip.f95:333-346: the neighbour-list cutoff is ``cutoff`` if given, else the
potential's, with ``cutoff_skin`` so that consecutive frames of a trajectory
reuse the list while no atom has moved more than half the skin), the potential
is called with ``calc_args`` plus the requested outputs (quip.f95:698-716), and
the report is printed as quip.f95:723-750 prints it: ``Energy=``, three
``Virial`` rows, three ``Pressure`` rows (eV/A^3 and GPa) and the cell volume,
followed by the frame in extended XYZ with every line prefixed ``AT``
(quip.f95:822).  ``precision`` and ``device`` are this library's additions
(``fp64`` is the plain double-precision form).

Everything physical runs through ``Potential`` and so on the GPU.
"""
import sys

import numpy as np

from .atoms import read_xyz, write_xyz
from .gap_xml import parse_args_str
from .potential import Potential

# ELEM_CHARGE * 1e30 / 1e9: eV/A^3 -> GPa (src/libAtoms/Units.f95:120)
EV_A3_IN_GPA = 1.602176634e-19 * 1.0e30 / 1.0e9

_FLAGS = {"E": "E", "energy": "E", "F": "F", "forces": "F", "V": "V", "virial": "V", "L": "L", "local": "L"}
_DEFAULTS = {"atoms_filename": "stdin", "param_filename": "quip_params.xml", "init_args": "IP GAP",
             "calc_args": "", "cutoff": "-1.0", "cutoff_skin": "0.5", "output_file": "stdout",
             "real_format": "%16.8f", "precision": "fp64", "device": "0"}


def _truth(v):
    return str(v).strip().upper() in ("T", "TRUE", "1", "Y", "YES")


def parse_cli(argv):
    """Command line -> settings dict.  Arguments are joined and read with the
    args-string grammar (key=value, key={a b}, key="a b", bare key = T); an
    unknown key is an error, as param_read_args reports (quip.f95:236-244)."""
    kv = parse_args_str(" ".join('%s="%s"' % tuple(a.split("=", 1)) if "=" in a and " " in a else a
                                 for a in argv))
    out = dict(_DEFAULTS)
    do = {"E": False, "F": False, "V": False, "L": False}
    for k, v in kv.items():
        if k in _FLAGS:
            do[_FLAGS[k]] = _truth(v)
        elif k in _DEFAULTS:
            out[k] = v
        else:
            raise SystemExit("quip: unknown argument '%s'" % k)
    out.update(do)
    return out


def calc_args_for(opt):
    """Output names appended to calc_args (quip.f95:700-707)."""
    extra = []
    if opt["E"]:
        extra.append("energy")
    if opt["F"]:
        extra.append("force")
    if opt["V"]:
        extra.append("virial")
    if opt["L"]:
        if opt["E"]:
            extra.append("local_energy")
        if opt["V"]:
            extra.append("local_virial")
    return (opt["calc_args"] + " " + " ".join(extra)).strip()


def report(at, opt, log):
    """The lines quip.f95:723-750 prints after a calculation, then the AT-prefixed frame."""
    vol = abs(np.linalg.det(at.cell))
    if opt["E"]:
        log.write("Energy=%.16g\n" % at.info["energy"])
        at.info["Energy"] = at.info["energy"]
    if opt["V"]:
        V = np.asarray(at.info["virial"])
        P = V / vol
        for i in range(3):
            log.write("Virial %s\n" % " ".join("%.16g" % x for x in V[i]))
        for i in range(3):
            log.write("Pressure eV/A^3 %s   GPa %s\n" % (" ".join("%.16g" % x for x in P[i]),
                                                         " ".join("%.16g" % x for x in P[i] * EV_A3_IN_GPA)))
    log.write("Cell Volume: %.16g A^3\n" % vol)


class _Prefixed:
    """Text stream that puts a prefix in front of every line (mainlog%prefix)."""

    def __init__(self, out, prefix):
        self.out, self.prefix, self.bol = out, prefix, True

    def write(self, s):
        for piece in s.splitlines(keepends=True):
            if self.bol:
                self.out.write(self.prefix + " ")
            self.out.write(piece)
            self.bol = piece.endswith("\n")


def run(argv, stdin=None, stdout=None):
    opt = parse_cli(argv)
    if not (opt["E"] or opt["F"] or opt["V"] or opt["L"]):
        raise SystemExit("quip: Nothing to be calculated")       # quip.f95:820
    stdin = stdin or sys.stdin
    frames = read_xyz(stdin if opt["atoms_filename"] == "stdin" else opt["atoms_filename"])
    pot = Potential(opt["init_args"], param_filename=opt["param_filename"], precision=opt["precision"],
                    device=int(opt["device"]))
    cut = float(opt["cutoff"])
    rc = (cut if cut >= 0.0 else pot.cutoff()) + float(opt["cutoff_skin"])
    args = calc_args_for(opt)
    close = None
    if opt["output_file"] != "stdout":
        log = close = open(opt["output_file"], "w")
    else:
        log = stdout or sys.stdout
    try:
        for at in frames:
            pot.calc(at, args_str=args, cutoff=rc)
            report(at, opt, log)
            write_xyz(_Prefixed(log, "AT"), at, real_format=opt["real_format"])
    finally:
        if close is not None:
            close.close()
    return frames


def main():
    run(sys.argv[1:])


if __name__ == "__main__":
    main()
