"""MI355X-native DMRG effective-Hamiltonian hot path (device library libmpsd.so).

The product is the C-ABI shared library built from ``csrc/`` (see include/mpsd.h); this
package only builds it (``mpsdyn_amd.build``) and binds it (``mpsdyn_amd.device``).
"""
from . import device  # noqa: F401

__all__ = ["device"]
