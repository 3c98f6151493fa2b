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


# ---------------------------------------------------------------- model-file refusals

def _water_xml(golden):
    return open(os.path.join(golden, "gap_water", "GAP.xml")).read()


def _load_status(xml, golden):
    """gapx_model_load's status: the XML and sparseX files are read and checked before
    any device is touched, so a refusal shows here without a GPU; a file that passes
    them stops at the device (GAPX_E_DEVICE or GAPX_E_ARG for a missing device)."""
    import ctypes as C
    raw = xml.encode() if isinstance(xml, str) else xml
    h = C.c_void_p()
    st = L.lib().gapx_model_load(raw, len(raw), None, os.path.join(golden, "gap_water").encode(), 0, C.byref(h))
    if st == 0:
        L.lib().gapx_model_free(h)
    return st, L.last_error()


GAPX_E_ARG, GAPX_E_MODEL, GAPX_E_DEVICE, GAPX_E_UNSUPPORTED = 1, 2, 4, 5


@pytest.mark.parametrize("edit,word", [
    (('do_core="F"', 'do_core="F" do_pca="T"'), "do_pca"),
    (('signal_mean=".00000000000000000E+000"', 'signal_mean=".25000000000000000E+000"'), "signal_mean"),
    (('n_permutations="1"', 'n_permutations="2"'), "n_permutations"),
    (("Z1=8 Z2=8</descriptor>", "Z1=8 Z2=8 only_intra=T</descriptor>"), "only_intra"),
    (("Z1=8 Z2=8</descriptor>", "Z1=8 Z2=8 exponents=2</descriptor>"), "exponents"),
])
def test_model_options_not_built_are_refused(golden, edit, word):
    xml = _water_xml(golden)
    assert edit[0] in xml
    st, msg = _load_status(xml.replace(edit[0], edit[1], 1), golden)
    assert st == GAPX_E_UNSUPPORTED and word in msg, (st, msg)
    with pytest.raises(NotImplementedError):
        read_gap_xml(xml_text=xml.replace(edit[0], edit[1], 1), xml_dir=os.path.join(golden, "gap_water"))


def test_soap_options_not_built_are_refused(golden, tmp_path):
    """normalise=F, average=T and friends change the descriptor: refused unless at their
    default; keys GAP's parser does not read (gap_fit's) are ignored, as it ignores them."""
    d = configs.soap_string(cutoff=4.0, n_max=4, l_max=4)
    X = np.zeros((2, 4 * 5 // 2 * 5 + 1))
    from quip_amd.gap_xml import Descriptor, GAPModel, GPCoordinate
    for extra, refused in (("normalise=F", True), ("average=T", True), ("covariance_sigma0=0.5", True),
                           ("normalise=T average=F", False), ("n_sparse=100 sparse_method=cur_points delta=3", False)):
        s = d + " " + extra
        if refused:
            with pytest.raises(NotImplementedError):
                Descriptor.from_string(s)
        c = GPCoordinate("c", Descriptor.from_string(d), 2, 1.0, 0.0, 4.0, X, np.ones(2), np.ones(2))
        c.descriptor.string = s
        path = str(tmp_path / "m.xml")
        write_gap_xml(path, GAPModel("M", 1527075646, np.zeros(119), [c]))
        import ctypes as C
        raw = open(path).read().encode()
        h = C.c_void_p()
        st = L.lib().gapx_model_load(raw, len(raw), None, str(tmp_path).encode(), 0, C.byref(h))
        if st == 0:
            L.lib().gapx_model_free(h)
        if refused:
            assert st == GAPX_E_UNSUPPORTED and extra.split("=")[0] in L.last_error()
        else:
            assert st in (0, GAPX_E_ARG, GAPX_E_DEVICE), (st, L.last_error())


def test_scalar_e0_attribute_sets_every_species(golden):
    """<GAP_data e0=...> sets e0 of every species before the per-Z entries override it
    (IPModel_GAP.f95:670-674); the Python reader mirrors the C reader."""
    xml = _water_xml(golden).replace('do_core="F"', 'do_core="F" e0="-3.5"', 1)
    m = read_gap_xml(xml_text=xml, xml_dir=os.path.join(golden, "gap_water"))
    base = read_gap_xml(os.path.join(golden, "gap_water", "GAP.xml"))
    listed = set()
    import re as _re
    for z in _re.findall(r'<e0 Z="(\d+)"', xml):
        listed.add(int(z))
    for z in range(1, 119):
        assert m.e0[z] == (base.e0[z] if z in listed else -3.5)
    assert _load_status(xml, golden)[0] in (0, GAPX_E_ARG, GAPX_E_DEVICE)


def test_truncated_xml_buffers_are_errors(golden):
    """The reader takes a length-delimited buffer: a comment, CDATA section, tag or
    attribute cut off by the length is an error, and nothing past it is read."""
    for cut in ("<GAP><!-- open comment", "<GAP><c><![CDATA[abc", "<GAP><e0 Z=\"1\" value=\"1.0", "<GAP><e0 Z=",
                "<GAP><", "<GAP><e0 Z=\"1\""):
        raw = cut.encode() + b"-->]]>\"/></GAP>"     # the closing text lies beyond the length passed
        import ctypes as C
        buf = C.create_string_buffer(raw, len(raw))
        h = C.c_void_p()
        st = L.lib().gapx_model_load(buf, len(cut), None, b".", 0, C.byref(h))
        assert st == GAPX_E_MODEL, (cut, st, L.last_error())
