/*
 * mpgos.h -- C ABI of the MI355X-native MPGOS solver library (libmpgos.so).
 *
 * The reference exposes its solver path as two C++ class templates,
 *   ProblemSolver<NT,SD,NCP,NSP,NISP,NE,NA,NIA,NDO,Algorithm,Precision>
 *     (SourceCodes/SingleSystem_PerThread_Interface.cuh:117-208)
 *   ProblemSolver<NS,UPS,UD,TPB,SPB,NC,CBW,CCI,NUP,NSP,NGP,NiGP,NUA,NiUA,NSA,NiSA,NE,NDO,Algorithm,Precision>
 *     (SourceCodes/CoupledSystems_PerBlock_Interface.cuh:135-248)
 * with user callbacks bound at compile time.  Callbacks cannot cross a run-time
 * C ABI, so every system is compiled into the library (see the registry in
 * csrc/systems/) and selected by name at creation; everything else is a plain
 * function over an opaque handle, plain pointers and sizes.  Each entry point
 * names the member function of the reference it replaces.
 *
 * Error convention: the reference calls exit() on CUDA errors, bad indices and
 * bad options (SingleSystem_PerThread_Interface.cuh:13-22, 565-581, 659-663).
 * Here every call returns MPGOS_OK (0) or a negative MPGOS_E_* code and leaves
 * a message for mpgos_last_error(); the C++ drop-in facade turns a failure back
 * into the reference's exit(EXIT_FAILURE).
 *
 * Threading: one host thread per handle, as in the reference.  All transfers
 * and launches are asynchronous on the handle's own HIP stream.
 */
#ifndef MPGOS_H
#define MPGOS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ codes */
enum {
    MPGOS_OK = 0,
    MPGOS_E_ARG = -1,      /* bad index, option or variable (reference: BoundCheck -> exit) */
    MPGOS_E_HIP = -2,      /* HIP runtime error (reference: gpuErrCHK -> exit(1)) */
    MPGOS_E_MEMORY = -3,   /* not enough device / pinned memory (reference: PT-IF:388-396) */
    MPGOS_E_SYSTEM = -4,   /* unknown system name or unsupported algorithm/precision */
    MPGOS_E_STATE = -5     /* call not valid in the current state (e.g. adjoint without forward) */
};

/* Algorithms, same order as enum Algorithms (SingleSystem_PerThread_Interface.cuh:33) */
enum { MPGOS_RK4 = 0, MPGOS_RKCK45 = 1 };

/* Arithmetic mode.  DOUBLE and FLOAT are the reference's Precision template
 * argument; FLOAT_KAHAN is new: FP32 state and stages with compensated (Kahan)
 * accumulation of time and state, FP64 error norm, step-size controller and
 * event step control. */
enum { MPGOS_DOUBLE = 0, MPGOS_FLOAT = 1, MPGOS_FLOAT_KAHAN = 2 };

/* Per-thread ListOfVariables, same names and order as
 * SingleSystem_PerThread_Interface.cuh:35-47; entries from PT_ACCEPTED_STEPS on
 * are new per-instance outputs (counters, status, event log). */
enum {
    MPGOS_PT_ALL = 0,
    MPGOS_PT_TIME_DOMAIN,
    MPGOS_PT_ACTUAL_STATE,
    MPGOS_PT_ACTUAL_TIME,
    MPGOS_PT_CONTROL_PARAMETERS,
    MPGOS_PT_SHARED_PARAMETERS,
    MPGOS_PT_INTEGER_SHARED_PARAMETERS,
    MPGOS_PT_ACCESSORIES,
    MPGOS_PT_INTEGER_ACCESSORIES,
    MPGOS_PT_DENSE_OUTPUT,
    MPGOS_PT_DENSE_INDEX,
    MPGOS_PT_DENSE_TIME,
    MPGOS_PT_DENSE_STATE,
    /* new */
    MPGOS_PT_ACCEPTED_STEPS,   /* int64 [NT]: accepted steps of the last Solve */
    MPGOS_PT_REJECTED_STEPS,   /* int64 [NT]: rejected attempts of the last Solve */
    MPGOS_PT_STATUS,           /* int32 [NT]: MPGOS_STATUS_* bits of the last Solve */
    MPGOS_PT_EVENT_COUNT,      /* int32 [NT]: detected events (all indices) in the last Solve */
    MPGOS_PT_EVENT_LOG,        /* P [NT*(1+SD)*cap]: (t, x_0..x_SD-1) per logged hit, SoA */
    MPGOS_PT_EVENT_LOG_INDEX,  /* int32 [NT*cap]: event function index of each logged hit */
    MPGOS_PT_ADJOINT_STATE,    /* double [SD*NT]: adjoint input dL/dx(T); after mpgos_pt_adjoint_reverse, dL/dx(0) */
    MPGOS_PT_PARAMETER_GRADIENT, /* double [NCP*NT]: dL/dcPAR after mpgos_pt_adjoint_reverse */
    MPGOS_PT_NUM_VARIABLES
};

/* Per-thread ListOfSolverOptions, same names and order as
 * SingleSystem_PerThread_Interface.cuh:49-62; entries from PT_KERNEL_MODE on are new. */
enum {
    MPGOS_PT_THREADS_PER_BLOCK = 0,
    MPGOS_PT_INITIAL_TIME_STEP,
    MPGOS_PT_ACTIVE_NUMBER_OF_THREADS,
    MPGOS_PT_MAXIMUM_TIME_STEP,
    MPGOS_PT_MINIMUM_TIME_STEP,
    MPGOS_PT_TIME_STEP_GROW_LIMIT,
    MPGOS_PT_TIME_STEP_SHRINK_LIMIT,
    MPGOS_PT_RELATIVE_TOLERANCE,
    MPGOS_PT_ABSOLUTE_TOLERANCE,
    MPGOS_PT_EVENT_TOLERANCE,
    MPGOS_PT_EVENT_DIRECTION,
    MPGOS_PT_DENSE_OUTPUT_MINIMUM_TIME_STEP,
    MPGOS_PT_DENSE_OUTPUT_SAVE_FREQUENCY,
    MPGOS_PT_PREFER_SHARED_MEMORY,
    /* new */
    MPGOS_PT_KERNEL_MODE,          /* 0 = one lane per instance (reference form), 1 = persistent work queue with lane
                                      compaction, 2 = one lane per instance relaunched with repacking, 3 = pilot launch
                                      then the instances ordered by estimated remaining work (MPGOS_KERNEL_*) */
    MPGOS_PT_PERSISTENT_WAVES_PER_SIMD, /* persistent grid size in waves per SIMD (1..8), default from occupancy */
    MPGOS_PT_GRAPH_CAPTURE,        /* 1 = Solve() replays a captured hipGraph of its launches */
    MPGOS_PT_EVENT_LOG_CAPACITY,   /* hits logged per instance (allocates EVENT_LOG), default 0 */
    MPGOS_PT_LAUNCH_VARIANT,       /* register-budget variant of the kernel (0 = default, 1..3 = waves-per-EU 2,4,8) */
    MPGOS_PT_PERSISTENT_ROUND_STEPS, /* persistent form: attempted steps between lane compactions (default 32) */
    MPGOS_PT_QUEUE_ORDER,          /* persistent form: order instances are taken from the work queue (MPGOS_QUEUE_*) */
    MPGOS_PT_PERSISTENT_SOLO_INSTANCES, /* persistent form: the first N instances of the queue each run alone in a
                                           workgroup's first lane (no lane divergence on the longest serial chains);
                                           at most one per workgroup, default 0 */
    MPGOS_PT_RELAUNCH_STEPS,       /* relaunch form: attempted steps per instance per launch; sorted form: attempted
                                      steps of the pilot launch (default 300) */
    MPGOS_PT_RELAUNCH_LAUNCHES,    /* relaunch form: budgeted launches before the final one, which runs every instance
                                      to its end (up to 2^31 - 1 attempted steps, MPGOS_STATUS_STEP_BUDGET beyond); default 16 */
    MPGOS_PT_NUM_OPTIONS
};

/* MPGOS_PT_QUEUE_ORDER values.  The order only schedules: results are
 * bit-identical for every order. */
enum {
    MPGOS_QUEUE_INDEX = 0,           /* instance index order (default) */
    MPGOS_QUEUE_LONGEST_FIRST = 1    /* longest integration interval TimeDomain[1] - TimeDomain[0] first
                                        (longest-processing-time-first when steps scale with the interval) */
};

/* MPGOS_PT_KERNEL_MODE values.  RELAUNCH: one lane per instance, relaunched
 * every RELAUNCH_STEPS attempted steps with the unfinished instances packed
 * into the lowest lanes of the whole grid (lane states suspended in HBM); all
 * launches are queued at once (asynchronous, graph-capturable).  SORTED: a
 * pilot launch of RELAUNCH_STEPS attempted steps per instance, then the
 * unfinished instances ordered on the device by their estimated remaining
 * steps (largest first) and run to their ends in one launch, so that each
 * wavefront holds instances of similar length; also asynchronous and
 * graph-capturable.  Every form gives the same bits. */
enum { MPGOS_KERNEL_PLAIN = 0, MPGOS_KERNEL_PERSISTENT = 1, MPGOS_KERNEL_RELAUNCH = 2, MPGOS_KERNEL_SORTED = 3 };

/* Per-instance status bits, replacing the reference's device printf
 * (ErrorControllers.cuh:15,66,73; EventHandling.cuh:38). */
enum {
    MPGOS_STATUS_NONFINITE_RK4 = 1,      /* RK4 state not finite: instance terminated */
    MPGOS_STATUS_NONFINITE_MIN_STEP = 2, /* RKCK45 not finite even at minimum step: terminated */
    MPGOS_STATUS_MIN_STEP_REACHED = 4,   /* step forced at the minimum size, tolerance not guaranteed */
    MPGOS_STATUS_EVENT_OMITTED = 8,      /* event step below the minimum step: detection omitted */
    MPGOS_STATUS_CHECKPOINT_OVERFLOW = 16, /* adjoint record pass: more checkpoints than the capacity */
    MPGOS_STATUS_STEP_BUDGET = 32        /* relaunch form: 2^31 - 1 attempted steps in its last launch; the instance
                                            stopped there (state and time written back as they were) */
};

typedef struct mpgos_solver mpgos_solver;

typedef struct {
    int nt;   /* NT: number of instances */
    int sd;   /* SD: system dimension */
    int ncp;  /* NCP: control parameters per instance */
    int nsp;  /* NSP: shared parameters */
    int nisp; /* NISP: integer shared parameters */
    int ne;   /* NE: event functions */
    int na;   /* NA: accessories per instance */
    int nia;  /* NIA: integer accessories per instance */
    int ndo;  /* NDO: dense-output points per instance */
    int algorithm;
    int precision;
    int has_vjp; /* system provides the right-hand-side vector-Jacobian product (adjoint) */
} mpgos_pt_dims;

/* ---------------------------------------------------------- per-thread module */

/* ProblemSolver::ProblemSolver(int AssociatedDevice)
 *   (SingleSystem_PerThread_Interface.cuh:309-518): allocates pinned host mirrors
 *   and device buffers, creates the stream and event, applies the default options
 *   (dt0 1e-2, dtmax 1e6, dtmin 1e-12, grow 5, shrink 0.1, rtol = atol = 1e-8,
 *   event tolerance 1e-6, direction 0, PreferSharedMemory on).
 *   `system` names a registered system (mpgos_pt_list_systems). */
int mpgos_pt_create(const char* system, int algorithm, int precision, int nt, int ndo,
                    int device, mpgos_solver** out);

/* ProblemSolver::~ProblemSolver() (SingleSystem_PerThread_Interface.cuh:521-562) */
int mpgos_pt_destroy(mpgos_solver* h);

/* The template arguments of the instantiated ProblemSolver (PT-IF:117). */
int mpgos_pt_get_dims(const mpgos_solver* h, mpgos_pt_dims* out);

/* ProblemSolver::SolverOption(ListOfSolverOptions, T) and
 * ProblemSolver::SolverOption(ListOfSolverOptions, int, T)
 *   (SingleSystem_PerThread_Interface.cuh:610-703).  `index` is ignored by the
 *   one-argument options. */
int mpgos_pt_solver_option(mpgos_solver* h, int option, int index, double value);

/* ProblemSolver::SetHost overloads (SingleSystem_PerThread_Interface.cuh:705-830).
 *   problem: ProblemNumber (ignored for global scope), serial: SerialNumber or
 *   component, step: TimeStepNumber (DenseState only).  Integer variables take
 *   the value cast to int, as the reference does. */
int mpgos_pt_set_host(mpgos_solver* h, int variable, int problem, int serial, int step, double value);

/* ProblemSolver::GetHost overloads (SingleSystem_PerThread_Interface.cuh:967-1080). */
int mpgos_pt_get_host(const mpgos_solver* h, int variable, int problem, int serial, int step, double* value);

/* Bulk access to the pinned host mirror of one variable, in the reference's SoA
 * layout (index problem + serial*NT [+ step*NT*SD]); the vectorised form of
 * SetHost/GetHost.  Returns NULL on error; *nbytes gets the size. */
void* mpgos_pt_host_buffer(mpgos_solver* h, int variable, size_t* nbytes);

/* Device address of one variable's buffer (same layout), for callers that keep
 * inputs resident in HBM and fill or read them with their own kernels. */
void* mpgos_pt_device_buffer(mpgos_solver* h, int variable, size_t* nbytes);

/* ProblemSolver::SynchroniseFromHostToDevice / SynchroniseFromDeviceToHost
 *   (SingleSystem_PerThread_Interface.cuh:832-965); asynchronous on the stream. */
int mpgos_pt_sync_h2d(mpgos_solver* h, int variable);
int mpgos_pt_sync_d2h(mpgos_solver* h, int variable);

/* ProblemSolver::Solve() (SingleSystem_PerThread_Interface.cuh:1279-1285):
 *   asynchronous launch of the integration kernel on the handle's stream. */
int mpgos_pt_solve(mpgos_solver* h);
/* New (no reference counterpart): one slice of a Solve.  Every unfinished
 * instance attempts at most `max_attempted_steps` steps in this launch; an
 * instance that has not finished is suspended (its whole lane state is kept in
 * device memory) and resumed by the next call.  The first call of a sequence
 * starts the Solve, the call that returns *unfinished == 0 completes it, and
 * the results are bit-identical to mpgos_pt_solve.  For workloads whose
 * longest instance runs for minutes: the caller regains control between
 * slices (progress, cancellation).  Synchronous; not with GraphCapture. */
int mpgos_pt_solve_slice(mpgos_solver* h, long long max_attempted_steps, int* unfinished);
/* New, no counterpart in the reference: Solve() for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Runs the relaunch form (MPGOS_KERNEL_RELAUNCH,
 * with the handle's ThreadsPerBlock, LaunchVariant, RelaunchSteps and
 * RelaunchLaunches, whatever KernelMode says) over the handle's own buffers,
 * with its lane records, instance lists and counts in the library's static
 * scratch memory of MPGOS_NOALLOC_SCRATCH_BYTES per GPU instead of memory the
 * handle allocates; results are bit-identical to mpgos_pt_solve.  Calls no
 * hipMalloc/hipFree and captures no graph.  Needs
 * NT * (record bytes + 12) + 4 * (RelaunchLaunches + 1) bytes, the record
 * being the instance's lane state rounded up to 16 bytes (160 for Lorenz in
 * FP64, i.e. up to about 390000 instances); larger handles get MPGOS_E_ARG.
 * Every call on one GPU uses that GPU's single scratch block: calls that
 * overlap on one GPU (two handles whose streams run concurrently) race, so
 * order them on one stream or synchronise between them.  Asynchronous on the
 * handle's stream, like mpgos_pt_solve. */
#define MPGOS_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int mpgos_pt_solve_noalloc(mpgos_solver* h);

/* ProblemSolver::SynchroniseDevice / InsertSynchronisationPoint / SynchroniseSolver
 *   (SingleSystem_PerThread_Interface.cuh:1288-1312). */
int mpgos_pt_synchronise_device(mpgos_solver* h);
int mpgos_pt_insert_synchronisation_point(mpgos_solver* h);
int mpgos_pt_synchronise_solver(mpgos_solver* h);

/* ProblemSolver::Print(ListOfVariables) and Print(DenseOutput, ThreadID)
 *   (SingleSystem_PerThread_Interface.cuh:1111-1276): same file names and text
 *   format (width 18, precision 10, scientific), written into `directory`
 *   (NULL = current directory).  thread_id is used for DenseOutput only. */
int mpgos_pt_print(mpgos_solver* h, int variable, int thread_id, const char* directory);

/* Time of the last Solve() kernel(s) in milliseconds, measured with HIP events
 * on the handle's stream (valid after SynchroniseSolver). */
int mpgos_pt_last_solve_ms(const mpgos_solver* h, double* ms);

/* HIP stream of the handle (hipStream_t), for callers that order their own work. */
void* mpgos_pt_stream(mpgos_solver* h);

/* Discrete adjoint (no counterpart in the reference; FP64, systems with a
 * right-hand-side VJP).  The step sequence of the recorded Solve is taken as
 * given (the controller is not differentiated).
 *   mpgos_pt_adjoint_record: a Solve() that also stores (x, t, proposed dt)
 *     before the first and after every `checkpoint_interval`-th accepted step,
 *     at most `checkpoint_capacity` checkpoints per instance (exceeding it sets
 *     MPGOS_STATUS_CHECKPOINT_OVERFLOW for that instance).
 *   mpgos_pt_adjoint_reverse: reads MPGOS_PT_ADJOINT_STATE = dL/dx(T) from the
 *     device buffer, recomputes each checkpoint segment with the same adaptive
 *     loop and propagates the adjoint backwards through the Runge-Kutta stages;
 *     leaves dL/dx(0) in MPGOS_PT_ADJOINT_STATE and dL/dcPAR in
 *     MPGOS_PT_PARAMETER_GRADIENT (device buffers; synchronise to read). */
int mpgos_pt_adjoint_record(mpgos_solver* h, int checkpoint_interval, int checkpoint_capacity);
int mpgos_pt_adjoint_reverse(mpgos_solver* h);

/* Names of the registered per-thread systems, comma separated. */
const char* mpgos_pt_list_systems(void);

/* Message of the last failed call on this thread. */
const char* mpgos_last_error(void);

/* ---------------------------------------------------------- per-block module */

/* Per-block ListOfVariables, same names and order as
 * SourceCodes/CoupledSystems_PerBlock_Interface.cuh:35-53; entries from
 * PB_ACCEPTED_STEPS on are new per-system outputs. */
enum {
    MPGOS_PB_ALL = 0,
    MPGOS_PB_TIME_DOMAIN,
    MPGOS_PB_ACTUAL_TIME,
    MPGOS_PB_ACTUAL_STATE,
    MPGOS_PB_UNIT_PARAMETERS,
    MPGOS_PB_SYSTEM_PARAMETERS,
    MPGOS_PB_GLOBAL_PARAMETERS,
    MPGOS_PB_INTEGER_GLOBAL_PARAMETERS,
    MPGOS_PB_UNIT_ACCESSORIES,
    MPGOS_PB_INTEGER_UNIT_ACCESSORIES,
    MPGOS_PB_SYSTEM_ACCESSORIES,
    MPGOS_PB_INTEGER_SYSTEM_ACCESSORIES,
    MPGOS_PB_COUPLING_MATRIX,
    MPGOS_PB_COUPLING_STRENGTH,
    MPGOS_PB_COUPLING_INDEX,
    MPGOS_PB_DENSE_OUTPUT,
    MPGOS_PB_DENSE_INDEX,
    MPGOS_PB_DENSE_TIME,
    MPGOS_PB_DENSE_STATE,
    /* new */
    MPGOS_PB_ACCEPTED_STEPS,   /* int64 [NS] */
    MPGOS_PB_REJECTED_STEPS,   /* int64 [NS] */
    MPGOS_PB_STATUS,           /* int32 [NS]: MPGOS_STATUS_* bits */
    MPGOS_PB_ADJOINT_STATE,    /* FP64 [UD][NS*UPS]: dL/dx(end) in, dL/dx(start) out (mpgos_pb_adjoint_reverse) */
    MPGOS_PB_UNIT_PARAMETER_GRADIENT,    /* FP64 [NUP][NS*UPS]: dL/d(unit parameters) */
    MPGOS_PB_SYSTEM_PARAMETER_GRADIENT,  /* FP64 [NSP][NS]: dL/d(system parameters) */
    MPGOS_PB_NUM_VARIABLES
};

/* Per-block ListOfSolverOptions, same names and order as
 * SourceCodes/CoupledSystems_PerBlock_Interface.cuh:55-68. */
enum {
    MPGOS_PB_INITIAL_TIME_STEP = 0,
    MPGOS_PB_ACTIVE_SYSTEMS,
    MPGOS_PB_MAXIMUM_TIME_STEP,
    MPGOS_PB_MINIMUM_TIME_STEP,
    MPGOS_PB_TIME_STEP_GROW_LIMIT,
    MPGOS_PB_TIME_STEP_SHRINK_LIMIT,
    MPGOS_PB_RELATIVE_TOLERANCE,
    MPGOS_PB_ABSOLUTE_TOLERANCE,
    MPGOS_PB_EVENT_TOLERANCE,
    MPGOS_PB_EVENT_DIRECTION,
    MPGOS_PB_DENSE_OUTPUT_SAVE_FREQUENCY,
    MPGOS_PB_DENSE_OUTPUT_MINIMUM_TIME_STEP,
    MPGOS_PB_SHARED_GLOBAL_VARIABLES,
    MPGOS_PB_SHARED_COUPLING_MATRICES,
    /* New (no reference counterpart): which coupling kernel runs.  0 automatic
     * (the dense-contraction MFMA kernel for a full FP64 matrix with UPS a
     * multiple of 32 in [32, 256], else lane per pair), 1 lane per pair (any
     * storage form; SPB as created), 2 dense contraction (8 systems per
     * workgroup, 2*UPS threads).  Results are bit-identical between them. */
    MPGOS_PB_COUPLING_KERNEL,
    /* New: units per lane of the dense-contraction kernel, 4 (2*UPS threads,
     * the default) or 2 (4*UPS threads).  A launch shape only. */
    MPGOS_PB_MFMA_UNITS_PER_LANE,
    MPGOS_PB_NUM_OPTIONS
};

typedef struct {
    int ns, ups, ud, spb, nc, cbw, cci;
    int nup, nsp, ngp, nigp, nua, niua, nsa, nisa, ne, ndo;
    int algorithm, precision;
} mpgos_pb_dims;

/* ProblemSolver<NS,UPS,UD,TPB,SPB,NC,CBW,CCI,...>::ProblemSolver(int)
 *   (CoupledSystems_PerBlock_Interface.cuh:353-657).  UD, NC and the parameter
 *   and accessory counts come from the registered system; NS, UPS, SPB
 *   (systems per workgroup, SPB*UPS <= 1024), the coupling storage form
 *   (CBW bandwidth radius, 0 = full; CCI circularity) and NDO are chosen here.
 *   Defaults as the reference: dt0 1e-2, dtmax 1e6, dtmin 1e-12, grow 5,
 *   shrink 0.1, rtol = atol = 1e-8, SharedGlobalVariables on. */
int mpgos_pb_create(const char* system, int algorithm, int precision, int ns, int ups, int spb, int cbw, int cci,
                    int ndo, int device, mpgos_solver** out);
int mpgos_pb_destroy(mpgos_solver* h);
int mpgos_pb_get_dims(const mpgos_solver* h, mpgos_pb_dims* out);
/* SolverOption (CoupledSystems_PerBlock_Interface.cuh:772-865) */
int mpgos_pb_solver_option(mpgos_solver* h, int option, int index, double value);
/* SetHost / GetHost (CoupledSystems_PerBlock_Interface.cuh:868-1140, 1339-1529).
 * Arguments by scope:
 *   unit scope (ActualState, UnitParameters, [Integer]UnitAccessories): a0 system, a1 unit, a2 serial;
 *   system scope (TimeDomain, SystemParameters, [Integer]SystemAccessories, DenseTime, CouplingStrength): a0 system, a2 serial;
 *   global scope ([Integer]GlobalParameters, CouplingIndex): a2 serial;
 *   CouplingMatrix: a0 coupling number, a1 row, a2 column (stored in the form given by CBW/CCI, PB-IF:994-1081);
 *   DenseIndex, ActualTime, AcceptedSteps, RejectedSteps, Status: a0 system;
 *   DenseState: a0 system, a1 unit, a2 component, a3 time-step number. */
int mpgos_pb_set_host(mpgos_solver* h, int variable, int a0, int a1, int a2, int a3, double value);
int mpgos_pb_get_host(const mpgos_solver* h, int variable, int a0, int a1, int a2, int a3, double* value);
void* mpgos_pb_host_buffer(mpgos_solver* h, int variable, size_t* nbytes);
void* mpgos_pb_device_buffer(mpgos_solver* h, int variable, size_t* nbytes);
/* SynchroniseFromHostToDevice / FromDeviceToHost (PB-IF:1143-1337) */
int mpgos_pb_sync_h2d(mpgos_solver* h, int variable);
int mpgos_pb_sync_d2h(mpgos_solver* h, int variable);
/* Solve (PB-IF:1949-1965), synchronisation (PB-IF:1967-1992) */
int mpgos_pb_solve(mpgos_solver* h);
int mpgos_pb_synchronise_device(mpgos_solver* h);
int mpgos_pb_insert_synchronisation_point(mpgos_solver* h);
int mpgos_pb_synchronise_solver(mpgos_solver* h);
int mpgos_pb_last_solve_ms(const mpgos_solver* h, double* ms);

/* Discrete adjoint of the per-block integration (new; no reference counterpart).
 * FP64 systems that provide OdeVJP.  The accepted step sizes are taken as given
 * (frozen), the controller and event step control are not differentiated.
 * mpgos_pb_adjoint_record: a Solve() that also stores the unit states every
 *   `checkpoint_interval` accepted steps (at most `checkpoint_capacity` per
 *   system) and (t, h) of every accepted step (at most `step_capacity` per
 *   system); a system that needs more gets MPGOS_STATUS_CHECKPOINT_OVERFLOW and
 *   NaN gradients.
 * mpgos_pb_adjoint_reverse: with dL/dx(end) in MPGOS_PB_ADJOINT_STATE (device
 *   buffer), computes dL/dx(start) (same buffer), dL/d(unit parameters) and
 *   dL/d(system parameters) by recomputing each checkpoint segment and
 *   back-propagating through the stages and the coupling terms (transposed
 *   coupling product).  Asynchronous on the solver's stream. */
int mpgos_pb_adjoint_record(mpgos_solver* h, int checkpoint_interval, int checkpoint_capacity, int step_capacity);
int mpgos_pb_adjoint_reverse(mpgos_solver* h);
void* mpgos_pb_stream(mpgos_solver* h);
const char* mpgos_pb_list_systems(void);

/* ---------------------------------------------------------- math probes
 * Evaluate the library's elementary functions (0 cos, 1 sin, 2 exp, 3 log,
 * 4 pow(x, y), 5 sqrt, 6 the per-block controller's power pow_ctrl(x, y), 7 the per-thread controller's root
 * ctrl_root(x, y != 0): x^(1/5) when y != 0, else x^(1/4)) on the host, or on the device over device buffers, so tests can
 * show that both sides produce the same bits. */
int mpgos_math_eval_host(int fn, const double* x, const double* y, double* out, long n);
int mpgos_math_eval_device(int fn, const double* x_dev, const double* y_dev, double* out_dev, long n);

#ifdef __cplusplus
}
#endif

#endif /* MPGOS_H */
