"""mumaxplus on MI355X: micromagnetic time stepping with hand-written HIP kernels.

The public classes mirror the reference package (mumaxplus/__init__.py of
dirkpetersen/mumax-plus): World, Grid, Ferromagnet, Antiferromagnet, TimeSolver and the
quantity / parameter / variable wrappers.  The compute path is the HIP
library libmumaxplus_amd.so, reached through ``_mumaxpluscpp``.
"""

from . import _mumaxpluscpp as _cpp
from .grid import Grid
from .world import World
from .magnet import Magnet
from .ferromagnet import Ferromagnet
from .antiferromagnet import Antiferromagnet
from .timesolver import TimeSolver, TimeSolverOutput
from .fieldquantity import FieldQuantity
from .scalarquantity import ScalarQuantity
from .parameter import Parameter
from .interparameter import InterParameter
from .poissonsystem import PoissonSystem
from .variable import Variable
from .dmitensor import DmiTensor, DmiTensorGroup
from .strayfield import StrayField
from .batch import Batch
from . import serving
from . import util

__all__ = ["_cpp", "Grid", "World", "Magnet", "Ferromagnet", "Antiferromagnet", "TimeSolver",
           "TimeSolverOutput", "FieldQuantity", "ScalarQuantity", "Parameter", "Variable", "DmiTensor",
           "DmiTensorGroup", "InterParameter", "PoissonSystem", "StrayField", "Batch", "serving", "util"]
