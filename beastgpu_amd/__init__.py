"""MI355X-native BEM matrix assembly behind BEAST's assemble(op, X, Y)."""
from . import meshes, spaces, operators  # noqa: F401
from .operators import Helmholtz3D, Maxwell3D  # noqa: F401
from .spaces import lagrangec0d1, raviartthomas  # noqa: F401
