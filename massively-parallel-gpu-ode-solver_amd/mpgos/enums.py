"""Enumerations of the reference's per-thread interface, same names and order
(SourceCodes/SingleSystem_PerThread_Interface.cuh:33-62), plus the new entries
of include/mpgos.h appended after them."""
import enum


class Algorithms(enum.IntEnum):
    RK4 = 0
    RKCK45 = 1


class Precision(enum.IntEnum):
    DOUBLE = 0
    FLOAT = 1
    FLOAT_KAHAN = 2


class ListOfVariables(enum.IntEnum):
    All = 0
    TimeDomain = 1
    ActualState = 2
    ActualTime = 3
    ControlParameters = 4
    SharedParameters = 5
    IntegerSharedParameters = 6
    Accessories = 7
    IntegerAccessories = 8
    DenseOutput = 9
    DenseIndex = 10
    DenseTime = 11
    DenseState = 12
    # new per-instance outputs
    AcceptedSteps = 13
    RejectedSteps = 14
    Status = 15
    EventCount = 16
    EventLog = 17
    EventLogIndex = 18
    AdjointState = 19
    ParameterGradient = 20


class ListOfSolverOptions(enum.IntEnum):
    ThreadsPerBlock = 0
    InitialTimeStep = 1
    ActiveNumberOfThreads = 2
    MaximumTimeStep = 3
    MinimumTimeStep = 4
    TimeStepGrowLimit = 5
    TimeStepShrinkLimit = 6
    RelativeTolerance = 7
    AbsoluteTolerance = 8
    EventTolerance = 9
    EventDirection = 10
    DenseOutputMinimumTimeStep = 11
    DenseOutputSaveFrequency = 12
    PreferSharedMemory = 13
    # new
    KernelMode = 14
    PersistentWavesPerSimd = 15
    GraphCapture = 16
    EventLogCapacity = 17
    LaunchVariant = 18
    PersistentRoundSteps = 19
    QueueOrder = 20
    PersistentSoloInstances = 21
    RelaunchSteps = 22
    RelaunchLaunches = 23


class KernelMode(enum.IntEnum):
    PLAIN = 0
    PERSISTENT = 1
    RELAUNCH = 2
    SORTED = 3


class Status(enum.IntFlag):
    NONFINITE_RK4 = 1
    NONFINITE_MIN_STEP = 2
    MIN_STEP_REACHED = 4
    EVENT_OMITTED = 8
    CHECKPOINT_OVERFLOW = 16
    STEP_BUDGET = 32


# export the reference's bare enumerator names (RK4, TimeDomain, InitialTimeStep, ...)
globals().update({m.name: m for e in (Algorithms, ListOfVariables, ListOfSolverOptions) for m in e})
