"""Host-side checks that need no GPU: the C-ABI library loads and exports every
entry point include/gapx.h declares, the product path fails loudly without
its HIP library, and the GAP XML writer/reader round-trips."""
import os
import re

import numpy as np
import pytest

from quip_amd import _lib as L
from quip_amd import configs
from quip_amd.gap_xml import read_gap_xml, write_gap_xml

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "gapx.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(gapx_\w+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = L.lib()
    names = _declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(L.EXPORTED)
    assert lib.gapx_api_version() == 1


def test_product_path_fails_loudly_without_library(monkeypatch, tmp_path):
    monkeypatch.setattr(L, "_lib", None)
    monkeypatch.setattr(L, "LIB_PATH", str(tmp_path / "missing.so"))
    from quip_amd.potential import Potential
    with pytest.raises(L.GapxError):
        Potential(param_filename=os.path.join(ROOT, "tests", "golden", "gap_water", "GAP.xml"))


def test_gap_xml_round_trip(tmp_path, golden):
    m = read_gap_xml(os.path.join(golden, "gap_water", "GAP.xml"))
    write_gap_xml(str(tmp_path / "w.xml"), m)
    m2 = read_gap_xml(str(tmp_path / "w.xml"))
    assert len(m2.coords) == len(m.coords)
    for a, b in zip(m.coords, m2.coords):
        assert np.array_equal(a.sparseX, b.sparseX) and np.array_equal(a.alpha, b.alpha)
        assert np.array_equal(a.sparse_cutoff, b.sparse_cutoff) and a.delta == b.delta
        assert a.descriptor.string == b.descriptor.string
    assert np.array_equal(m.e0, m2.e0)


def test_workload_generators_are_seeded():
    a, b = configs.si4096(), configs.si4096()
    assert len(a) == 4055 and np.array_equal(a.positions, b.positions)
    s = configs.sic32768()
    assert len(s) == 32768 and (s.numbers == 14).sum() == 16384
    frames = configs.batch512(n_struct=16)
    assert all(8 <= len(f) <= 512 for f in frames)


def test_quip_cli_arguments_and_report():
    """The quip-style front end: argument grammar, calc_args it builds and the
    printed report (quip.f95:144-234, 698-750, 822) for a frame whose results
    are already on it (no evaluation here)."""
    import io
    from quip_amd import quip as Q
    from quip_amd.atoms import read_xyz
    opt = Q.parse_cli(["atoms_filename=a.xyz", "param_filename=gp.xml", "E", "forces", "V=T", "L",
                       "calc_args=add_noise=F", "cutoff_skin=0.3"])
    assert opt["E"] and opt["F"] and opt["V"] and opt["L"]
    assert opt["atoms_filename"] == "a.xyz" and opt["cutoff_skin"] == "0.3"
    assert Q.calc_args_for(opt) == "add_noise=F energy force virial local_energy local_virial"
    assert Q.calc_args_for(Q.parse_cli(["F"])) == "force"
    with pytest.raises(SystemExit):
        Q.parse_cli(["bogus_key=1"])
    at = configs.diamond(5.43, (1, 1, 1))
    at.info["energy"] = -42.5
    V = np.array([[1.0, 0.5, 0.0], [0.5, 2.0, 0.0], [0.0, 0.0, 3.0]])
    at.info["virial"] = V
    at.arrays["force"] = np.arange(24.0).reshape(8, 3)
    out = io.StringIO()
    Q.report(at, opt, out)
    Q.write_xyz(Q._Prefixed(out, "AT"), at, real_format=opt["real_format"])
    lines = out.getvalue().splitlines()
    vol = 5.43 ** 3
    assert lines[0] == "Energy=-42.5"
    assert [float(x) for x in lines[1].split()[1:]] == [1.0, 0.5, 0.0]
    p = lines[4].split()
    assert p[:2] == ["Pressure", "eV/A^3"] and abs(float(p[2]) - 1.0 / vol) < 1e-15
    assert abs(float(p[6]) - 1.0 / vol * 160.21766339999996) < 1e-9
    assert lines[7].startswith("Cell Volume:") and abs(float(lines[7].split()[2]) - vol) < 1e-9
    at_lines = lines[8:]
    assert all(ln.startswith("AT ") for ln in at_lines) and len(at_lines) == 10
    back = read_xyz(io.StringIO("\n".join(ln[3:] for ln in at_lines)), 0)
    assert np.allclose(back.positions, at.positions, atol=1e-8)
    assert np.allclose(back.arrays["force"], at.arrays["force"])
    assert back.info["Energy"] == -42.5
    assert np.allclose(np.array(back.info["virial"].split(), float).reshape(3, 3, order="F"), V)
