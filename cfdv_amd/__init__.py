"""MI355X-native FDV time step (hahnpv/CFDV hot path): element build,
coloring-based EBE GMRES and refinement flags behind a C ABI."""
from .mesh import Mesh, Thermo  # noqa: F401

__all__ = ["Mesh", "Thermo"]
