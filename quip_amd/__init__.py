"""quip_amd: MI355X-native SOAP-GAP energy/force/virial engine behind QUIP's IP GAP interface."""
from .atoms import Atoms, read_xyz, write_xyz  # noqa: F401
from .gap_xml import read_gap_xml, write_gap_xml, Descriptor, GPCoordinate, GAPModel  # noqa: F401
