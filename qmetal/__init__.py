"""`import qmetal` for callers of the reference's Python package
(python/qmetal/__init__.py:1-28): the same names -- Circuit, Simulator,
Hamiltonian, QMetalError, version, is_available, device_name, max_qubits --
served by the MI355X build in qmetal_amd, so reference scripts run unchanged.
The library is found the same way (QMETAL_LIBRARY, else the built
libqmetal.so beside qmetal_amd)."""
from qmetal_amd import *  # noqa: F401,F403
from qmetal_amd import __all__  # noqa: F401
