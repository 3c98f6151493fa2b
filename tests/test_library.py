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
