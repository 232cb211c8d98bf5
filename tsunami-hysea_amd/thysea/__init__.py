"""thysea: host-side Python mirror of the MI355X-native Tsunami-HySEA time-stepping path.

The compute path is the HIP library ``libthysea.so`` (tsunami-hysea_amd/csrc), reached
through the C-ABI declared in include/thysea.h.  This package holds the Python side of
the host: problem setup (mirror of src/Problema.cxx), seeded scenario generation and a
ctypes binding of the C-ABI.
"""
from . import problem, scenario  # noqa: F401
