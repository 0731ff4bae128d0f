"""MI355X-native MPGOS: Python mirror of the reference's ProblemSolver over the
gfx950 kernels of libmpgos.so."""
from ._lib import LIB_PATH, MPGOSError, lib
from .enums import *  # noqa: F401,F403  (reference enumerator names)
from .enums import Algorithms, KernelMode, ListOfSolverOptions, ListOfVariables, Precision, Status
from .solver import ProblemSolver, list_systems
from .perblock import PB, PBOption, PBProblemSolver, list_pb_systems
