"""MI355X-native Roberts boundary-integral RK4 stepper (Python layer over the gfx950 C ABI)."""
from ._lib import LIB_PATH, LibraryMissing, SolverError, lib  # noqa: F401
from ._lib import CGaussLegendreOptions, COptomechanicalVariables  # noqa: F401
from .rhs import CRK4Options, CSimulationProperties, SimulationManager  # noqa: F401
from .stepper import AUTO, BATCHED, FP64, GMRES, HELIUM, LU, MIXED, WATER, Ensemble, Props, autotune, tuned  # noqa: F401
