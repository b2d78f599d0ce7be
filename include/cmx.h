/*
 * cmx -- MI355X monodomain tissue time-step, C ABI.
 *
 * The library replaces the reference's monodomain hot path behind one opaque
 * context: the cell-ODE sweep (AbstractCardiacTissue::SolveCellSystems), the
 * ICI right-hand side b = M z (MonodomainSolver::SetupLinearSystem) and the
 * CG diffusion solve (LinearSystem::Solve -> KSPSolve).  Plain pointers and
 * sizes only; every call returns 0 on success or a negative CMX_E* code, with
 * the message available from cmx_last_error().  A context is used by one host
 * thread at a time.  Host arrays are caller-owned and copied; device buffers
 * are owned by the context.
 */
#ifndef CMX_H_
#define CMX_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CMX_ABI_VERSION 1

/* Error codes (reference: EXCEPTION / KSPEXCEPT, global/src/Exception.hpp,
 * linalg/src/LinearSystem.cpp:1118-1129). */
enum {
    CMX_OK = 0,
    CMX_E_ARG = -1,          /* invalid argument (size, null, unknown model) */
    CMX_E_HIP = -2,          /* HIP runtime failure */
    CMX_E_NOMEM = -3,        /* device allocation failed */
    CMX_E_DIVERGED = -4,     /* CG hit max iterations (KSP_DIVERGED_ITS) */
    CMX_E_NONFINITE = -5,    /* NaN/Inf in the solution (cell ODE blew up) */
    CMX_E_STATE = -6         /* call out of order */
};

/* Cell models (CellML sources heart/src/odes/cellml/LuoRudy1991.cellml,
 * TenTusscher2006Epi.cellml, HodgkinHuxley1952.cellml, FoxModel2002.cellml,
 * Mahajan2008.cellml, DiFrancescoNoble1985.cellml, FaberRudy2000.cellml,
 * Maleckar2008.cellml, NobleVargheseKohlNoble1998a.cellml; O'Hara-Rudy 2011 endocardial, which the reference tree
 * does not hold, from the published equations, forward-Euler and Rush-Larsen
 * integration only). */
enum { CMX_MODEL_LR91 = 0, CMX_MODEL_TT06_EPI = 1, CMX_MODEL_ORD_ENDO = 2, CMX_MODEL_HH52 = 3,
       CMX_MODEL_FOX2002 = 4, CMX_MODEL_MAHAJAN2008 = 5, CMX_MODEL_DN85 = 6, CMX_MODEL_FR2000 = 7,
       CMX_MODEL_MALECKAR2008 = 8, CMX_MODEL_N98 = 9 };

/* Cell integrators: EulerIvpOdeSolver (ode/src/solver/EulerIvpOdeSolver.cpp:38-57),
 * AbstractRushLarsenCardiacCell (heart/src/odes/AbstractRushLarsenCardiacCell.cpp:57-153),
 * AbstractBackwardEulerCardiacCell (heart/src/odes/AbstractBackwardEulerCardiacCell.hpp:
 * 116-330 with CardiacNewtonSolver.hpp:66-160) and AbstractGeneralizedRushLarsenCardiacCell
 * of first / second order (heart/src/odes/AbstractGeneralizedRushLarsenCardiacCell.cpp:70-170,
 * ode/src/solver/GRL1IvpOdeSolver.cpp, GRL2IvpOdeSolver.cpp).  dt_pde must be a whole number
 * of dt_ode. */
enum { CMX_FORWARD_EULER = 0, CMX_RUSH_LARSEN = 1, CMX_BACKWARD_EULER = 2, CMX_GRL1 = 3, CMX_GRL2 = 4 };

/* Precision: plain FP64, or FP32 cells + FP32 CG on the correction with FP64
 * voltage, FP64 dot reductions and FP64 residual refinement. */
enum { CMX_FP64 = 0, CMX_MIXED = 1 };

/* Sparse format for M, K, A, chosen at create time.  CMX_FMT_STENCIL: one
 * set of column offsets for all rows and per-pattern values (few distinct
 * rows).  CMX_FMT_ELL_OFFSETS: ELL values per row with one byte per row naming
 * its set of column offsets (at most 255 sets; no column indices).
 * CMX_FMT_ELL: ELL with a column index per entry.  CMX_FMT_AUTO takes the first
 * of those three that applies. */
enum { CMX_FMT_AUTO = 0, CMX_FMT_STENCIL = 1, CMX_FMT_ELL = 2, CMX_FMT_ELL_OFFSETS = 3 };

/* Mesh: P1 simplices (replaces AbstractTetrahedralMesh + the FE assemblers,
 * pde/src/solver/AbstractFeVolumeIntegralAssembler.hpp:257-459). */
typedef struct {
    int32_t dim;               /* 1, 2 or 3 */
    int32_t n_nodes;
    const double* coords;      /* [n_nodes][dim], cm */
    int32_t n_elems;
    const int32_t* elems;      /* [n_elems][dim+1] node indices */
    const double* sigma_elem;  /* [n_elems][dim][dim] intracellular conductivity, or NULL */
    double sigma[9];           /* [dim][dim] used when sigma_elem == NULL (mS/cm) */
    int32_t elem_dim;          /* 0 or dim: simplices of the node dimension; 1: line elements
                                  (cables) in 2-D or 3-D space, elems [n_elems][2], conductivity
                                  sigma[0] along the line (MonodomainProblem<1, SPACE_DIM>) */
} cmx_mesh;

/* Stimulus protocol, one row of a table (SimpleStimulus when period <= 0,
 * heart/src/stimulus/SimpleStimulus.cpp:44-76; RegularStimulus otherwise,
 * heart/src/stimulus/RegularStimulus.cpp:58-80).  magnitude in uA/cm^3. */
typedef struct {
    double magnitude, start, duration, period, stop;
} cmx_stimulus;

/* Tissue: the cells the factory creates for every node
 * (AbstractCardiacCellFactory::CreateCardiacCellForTissueNode,
 * heart/src/problem/cell_factories/AbstractCardiacCellFactory.hpp:100): one
 * model and integrator for every node, or, with n_groups > 1, node i gets
 * group_model[group_of[i]] and group_integrator[group_of[i]] (each group is
 * stepped by its own launch of the cell kernel). */
typedef struct {
    int32_t model;             /* CMX_MODEL_* (ignored when n_groups > 1) */
    int32_t integrator;        /* CMX_FORWARD_EULER ... CMX_GRL2 (ignored when n_groups > 1) */
    int32_t n_stimuli;         /* rows in stimuli (<= 255) */
    const cmx_stimulus* stimuli;
    const int32_t* stim_index; /* [n_nodes]: -1 = no stimulus, else row; NULL = none */
    int32_t n_params;          /* per-node parameter planes given (<= model's count; 0 with n_groups > 1) */
    const double* params;      /* [n_params][n_nodes] or NULL (model defaults) */
    int32_t n_groups;          /* 0 or 1: model / integrator everywhere; > 1: per-node cell types
                                  (no per-node parameters, no state-variable interpolation, no adjoint) */
    const int32_t* group_of;   /* [n_nodes] group of each node */
    const int32_t* group_model;       /* [n_groups] CMX_MODEL_* */
    const int32_t* group_integrator;  /* [n_groups] CMX_FORWARD_EULER ... */
} cmx_tissue;

/* Options (HeartConfig subset, heart/src/problem/HeartConfigDefaults.hpp:52-152). */
typedef struct {
    int32_t precision;         /* CMX_FP64 / CMX_MIXED */
    int32_t format;            /* CMX_FMT_* */
    double dt_pde, dt_ode;     /* ms; dt_ode must divide dt_pde */
    double am, cm;             /* surface-area-to-volume (1/cm), capacitance (uF/cm^2) */
    double ksp_rtol, ksp_atol; /* stop when ||D^-1 r|| <= max(rtol ||D^-1 b||, atol) */
    int32_t ksp_max_its;
    int32_t refine_sweeps;     /* mixed: FP64 true-residual checks per step (0 = none) */
    int32_t use_graph;         /* replay steps from captured hipGraphs */
    int32_t mass_lumping;      /* HeartConfig::SetUseMassLumping */
    int32_t extrapolate;       /* CG initial guess instead of the reference's V^n (same
                                  stopping test): 1 = V^n + x_{n-1}, 2 = V^n + 2 x_{n-1} - x_{n-2},
                                  with x_m = V^{m+1} - V^m the step corrections */
    int32_t operator_splitting; /* HeartConfig::SetUseReactionDiffusionOperatorSplitting:
                                  Strang splitting, cells over half steps with V updated and a
                                  diffusion-only solve in between
                                  (OperatorSplittingMonodomainSolver.cpp:40-137) */
    int32_t state_variable_interpolation; /* HeartConfig::SetUseStateVariableInterpolation:
                                  SVI correction of the ICI right-hand side
                                  (MonodomainCorrectionTermAssembler.cpp:48-74); not applied with
                                  operator splitting, whose solver has none */
    int32_t separate_refinement; /* mixed: 0 (default) computes a step's FP64 true residual in the
                                  first solve's last CG iteration whenever the refinement solve needs
                                  no iterations; 1 always in a kernel of its own (same test) */
    int32_t lookup_tables;     /* mixed, Rush-Larsen, one cell model: each gate's steady state and
                                  decay factor exp(-dt_ode/tau) read from a table in V (-250..350 mV,
                                  0.02 mV rows, linear interpolation) instead of evaluated, as the
                                  reference's *Opt cells tabulate V-only expressions
                                  (heart/src/odes/AbstractLookupTableCollection.hpp:49); a voltage
                                  outside the table fails the step (CMX_E_NONFINITE).  0 = evaluate */
} cmx_opts;

typedef struct {
    int64_t steps;             /* PDE steps taken */
    int64_t cg_iterations;     /* total CG iterations */
    int64_t refine_sweeps;     /* mixed: FP64 residual sweeps */
    int32_t last_iterations;   /* CG iterations of the last step */
    int32_t format;            /* CMX_FMT_STENCIL, CMX_FMT_ELL or CMX_FMT_ELL_OFFSETS in use */
    int32_t n_nodes, n_states, row_width, n_patterns;   /* n_patterns: value patterns (stencil) or
                                                               offset sets (ELL_OFFSETS) */
    int64_t nnz;
    double last_residual;      /* ||D^-1 r|| at the end of the last solve */
    double device_bytes;       /* resident device memory (bytes) */
    int64_t graph_launches;    /* captured multi-step graphs replayed */
    int64_t graph_captures;    /* graphs captured and instantiated */
    int32_t graph_mode;        /* 0 plain launches (use_graph = 0), 1 graph replay, 2 plain launches
                                  because a profiler (rocprofv3) is attached */
    int32_t solver_blocks;     /* blocks of the solver kernels (cmx_autotune) */
    int64_t resumes;           /* steps whose final solve the host resumed past its budget */
    int64_t folded_refinements; /* mixed: steps whose FP64 true residual came from the last CG
                                  iteration (no kernel of its own) */
    int32_t adjoint_slots;     /* stored forward states of the last cmx_adjoint (checkpoints, segment
                                  checkpoints and traced steps; each n_states x n_nodes doubles) */
    int32_t adjoint_sub_interval; /* its segment-checkpoint interval (0: one level) */
    int32_t cg_rows_per_thread; /* rows each thread of the CG kernels holds in registers (1 or 2) */
    int32_t graph_steps;       /* steps in the largest captured graph */
    int64_t adjoint_iterations; /* CG iterations of cmx_adjoint's solves (all calls) */
} cmx_stats;

typedef struct cmx_ctx cmx_ctx;

/* Replaces MonodomainProblem::Initialise -> CreateCardiacTissue + the
 * one-time matrix assembly in MonodomainSolver::SetupLinearSystem
 * (heart/src/problem/AbstractCardiacProblem.cpp:117-214,
 * heart/src/solver/electrics/monodomain/MonodomainSolver.cpp:51-76).
 * Cells start from the model's initial state, V from its initial voltage. */
int cmx_create(const cmx_mesh* mesh, const cmx_tissue* tissue, const cmx_opts* opts,
               cmx_ctx** out);
int cmx_destroy(cmx_ctx* ctx);
/* Message for the last failing call on ctx (or of the last failed create when ctx is NULL). */
const char* cmx_last_error(const cmx_ctx* ctx);

/* Advance n_steps PDE steps from time t0: t_k = t0 + k*dt_pde
 * (AbstractDynamicLinearPdeSolver::Solve, pde/src/solver/AbstractDynamicLinearPdeSolver.hpp:328-497). */
int cmx_step(cmx_ctx* ctx, double t0, int32_t n_steps);

/* Launch-shape autotuner (no counterpart in the reference): times n_steps
 * PDE steps from the current state at t0 for each candidate number of
 * solver blocks (candidates == NULL: 128, 192, 256, 320, 384, 512, capped
 * at one block per 1024 rows), then, at the fastest, the CG kernels' rows per
 * thread held in registers (1, 2: the tile shape) and, with graph replay, the
 * steps in the largest captured graph (4 .. 32: the unroll depth of the step
 * loop), keeps the fastest configuration, and restores the state,
 * the initial-guess history, activation times and statistics, so the next
 * cmx_step starts where it would have.  The CG iteration budgets (kernels
 * enqueued per solve, adapted after every batch) are those the fastest trial
 * settled on, as a first batch of those steps would have left them.  The
 * block count changes the reductions' summation order only (results agree to
 * rounding).
 * ms_per_step[n_candidates or 6] (may be NULL) gets each candidate's time. */
int cmx_autotune(cmx_ctx* ctx, double t0, int32_t n_steps, int32_t n_candidates, const int32_t* candidates,
                 int32_t* chosen, double* ms_per_step);
/* Sets the launch configuration the autotuner picks: solver blocks, CG rows
 * per thread (1 or 2) and steps per captured graph (a power of two up to 32);
 * 0 keeps a value.  Read it back with cmx_get_stats. */
int cmx_set_launch_config(cmx_ctx* ctx, int32_t solver_blocks, int32_t cg_rows, int32_t graph_steps);

/* Transmembrane potential (mV): AbstractCardiacProblem::GetSolution /
 * SetInitialCondition (heart/src/problem/AbstractCardiacProblem.cpp:265-292). */
int cmx_get_voltage(cmx_ctx* ctx, double* V);
int cmx_set_voltage(cmx_ctx* ctx, const double* V);
/* Cell states as [n_states][n_nodes] planes in the model's state order; the
 * V plane (index 0) is read from / written to the voltage vector
 * (AbstractCardiacCellInterface::GetStdVecStateVariables / SetStateVariables).
 * With several cell types n_states is the largest state count and node i's
 * column holds its own model's states, NaN past them (ignored by set). */
int cmx_get_state(cmx_ctx* ctx, double* states);
int cmx_set_state(cmx_ctx* ctx, const double* states);
/* Ionic current cache (uA/cm^2) of the last step
 * (AbstractCardiacTissue::rGetIionicCacheReplicated, heart/src/tissue/AbstractCardiacTissue.cpp:730-734). */
int cmx_get_iionic(cmx_ctx* ctx, double* iion);

/* Activation and recovery times, ActivationOutputModifier
 * (heart/src/problem/ActivationOutputModifier.cpp:40-128) driven as
 * AbstractCardiacProblem::Solve drives it (AbstractCardiacProblem.cpp:459-464,
 * 552-555): the voltage at t_now and then after every every_steps-th PDE step
 * of a cmx_step call (counted from its t0; every_steps = printing dt / PDE dt)
 * updates first activation (V > threshold), first recovery (V < threshold),
 * second activation and second recovery, -1 until seen.  A fifth column holds
 * the first upward crossing of the threshold linearly interpolated between
 * PDE steps (checked every step).  every_steps <= 0 stops tracking. */
int cmx_set_activation(cmx_ctx* ctx, double threshold, int32_t every_steps, double t_now);
/* out[5][n_nodes]: first act., first rec., second act., second rec., interpolated first crossing (ms). */
int cmx_get_activation(cmx_ctx* ctx, double* out);
int cmx_get_stats(cmx_ctx* ctx, cmx_stats* stats);

/* Per-kernel timing with HIP events recorded on the context's stream around
 * every launch of the kinds selected by mask (bit k = kind k), in the spirit
 * of the reference's HeartEventHandler events Ode / AssRhs / Ksp
 * (global/src/timing/HeartEventHandler.hpp).  Times accumulate until reset. */
enum { CMX_K_CELL = 0, CMX_K_RESIDUAL = 1, CMX_K_SPMV = 2, CMX_K_UPDATE = 3, CMX_K_PUPDATE = 4,
       CMX_K_FINISH = 5, CMX_K_SVI = 6, CMX_K_COUNT = 7 };
int cmx_set_profiling(cmx_ctx* ctx, int32_t mask);
/* total_ms[CMX_K_COUNT], launches[CMX_K_COUNT]; reset != 0 clears them afterwards. */
int cmx_get_kernel_times(cmx_ctx* ctx, double* total_ms, int64_t* launches, int32_t reset);

/* Loads a cell model built as a plugin (a shared library from a generated
 * CellML model, chaste_amd/csrc/cmx_model_plugin.h; chaste_amd.build.
 * build_model_plugin makes one from a .cellml file) and returns its model id
 * (64 and up), usable wherever a CMX_MODEL_* id is, with every integrator.
 * Replaces DynamicCellModelLoader::CreateCell / CellMLLoader
 * (heart/src/odes/DynamicCellModelLoader.cpp:55-63, heart/src/io/CellMLLoader.cpp):
 * the library stays loaded until the process ends.  Errors: a file that
 * does not load, is not a plugin, or was built against other headers. */
int cmx_load_model(const char* path, int32_t* model_id);

/* Model metadata (CellML order, V first). */
int cmx_model_info(int32_t model, int32_t* n_states, int32_t* n_params);
int cmx_model_state_name(int32_t model, int32_t index, const char** name);
int cmx_model_initial_state(int32_t model, double* y);
/* Parameter `index` of a model: its CellML name (component__variable) and
 * default value.  LR91's parameters are the three conductances its CellML
 * annotates with the metadata the reference's generator turns into modifiable
 * parameters: g_Na, g_Kmax, P_si (in that order). */
int cmx_model_param(int32_t model, int32_t index, const char** name, double* default_value);
/* The reference's name of parameter index (the CellML oxmeta term or cmeta:id
 * that AbstractParameterisedSystem::GetParameterIndex matches,
 * ode/src/common/AbstractParameterisedSystem.cpp:224-237; "" when the CellML
 * annotates none), and the index of a parameter by either name. */
int cmx_model_param_alias(int32_t model, int32_t index, const char** alias);
int cmx_model_find_param(int32_t model, const char* name, int32_t* index);
/* The gate lookup table that cmx_opts.lookup_tables / cmx_cells.lookup_tables
 * use for Rush-Larsen steps of length dt (the tabulated-expression idea of the
 * reference's AbstractLookupTableCollection, heart/src/odes/
 * AbstractLookupTableCollection.hpp:49): row r at V = vmin + r dv holds, for the
 * model's j-th gate, its steady state at column 2j and its decay factor
 * exp(-dt/tau) at column 2j + 1 (width floats per row).  With table NULL only
 * the sizes are returned.  Host code only. */
int cmx_gate_table(int32_t model, double dt, float* table, int32_t* rows, int32_t* width, double* vmin, double* dv);

/* Per-node parameter planes [n_params][n_nodes] in place of those given at
 * create, as AbstractParameterisedSystem::SetParameter on every node's cell
 * (ode/src/common/AbstractParameterisedSystem.cpp:256-275); planes beyond
 * n_params take the model defaults. */
int cmx_set_params(cmx_ctx* ctx, int32_t n_params, const double* params);

/* Gradient of the voltage-mismatch loss
 *   L = sum_{j=1..n_steps/obs_every} sum_o (V_{obs_nodes[o]}(t0 + j obs_every dt) - obs_values[j-1][o])^2
 * over the n_steps PDE steps the context takes from its current state at t0
 * (as cmx_step(ctx, t0, n_steps) takes them) with respect to every per-node
 * cell parameter, grad_params[n_params][n_nodes] (may be NULL), and to the
 * initial state, grad_state0[n_states][n_nodes] (V plane first; may be NULL).
 * The reference has no adjoint: this is the discrete adjoint of its FP64 ICI
 * step (cell-step Jacobian columns by forward-mode differentiation of the
 * generated model, one symmetric diffusion solve per step, FP64 Jacobi-PCG
 * to the context's rtol, 1e-10 when it is 0), recomputing the forward states
 * from checkpoints saved every checkpoint_every steps: a segment between two
 * checkpoints is replayed with every step's state kept (device memory about
 * (n_steps/checkpoint_every + checkpoint_every) x n_states x n_nodes
 * doubles).  When that exceeds the memory bound (cmx_set_adjoint_memory), each
 * segment is replayed once more first to save checkpoints every
 * q ~ sqrt(checkpoint_every) steps and only q steps are kept at a time
 * (n_steps/checkpoint_every + checkpoint_every/q + q states; one forward pass
 * more).  The gradient does not depend on the schedule.  Needs CMX_FP64, forward-Euler or Rush-Larsen cells, no splitting, no SVI,
 * extrapolate = 0, activation tracking off and a cell model with an adjoint
 * (every built-in and plugin model but O'Hara-Rudy and Maleckar 2008).  On return the context holds
 * the state at t0 + n_steps dt. */
int cmx_adjoint(cmx_ctx* ctx, double t0, int32_t n_steps, int32_t checkpoint_every, int32_t n_obs,
                const int32_t* obs_nodes, int32_t obs_every, const double* obs_values, double* loss,
                double* grad_params, double* grad_state0);
/* Bound (bytes) on the forward states cmx_adjoint stores; 0 (default): the
 * device's free memory at the call less 1 GiB.  No reference counterpart. */
int cmx_set_adjoint_memory(cmx_ctx* ctx, double bytes);

/* Reads a Triangle/TetGen mesh <base>.node / <base>.ele, ASCII (items numbered
 * from 0 or 1, '#' comments) or binary (BIN header tag), as
 * TrianglesMeshReader does (mesh/src/reader/TrianglesMeshReader.cpp:593-990);
 * linear simplices of the node dimension only.  Node indices come back from
 * zero.  With coords == NULL only the sizes are returned; elem_attr gets
 * n_elem_attr doubles per element (may be NULL).  Errors: CMX_E_ARG with the
 * message in cmx_io_last_error(). */
int cmx_read_mesh(const char* base, int32_t* dim, int32_t* n_nodes, int32_t* n_elems, int32_t* n_elem_attr,
                  double* coords, int32_t* elems, double* elem_attr);
/* The same for a mesh whose elements have dimension elem_dim (0 = the node
 * dimension, as cmx_read_mesh): line elements (elem_dim 1) of a mesh in 2-D
 * or 3-D space come from <base>.edge, header "n n_attributes [BIN]", as
 * TrianglesMeshReader<1, SPACE_DIM> reads them
 * (mesh/src/reader/TrianglesMeshReader.cpp:690-712); elems [n_elems][elem_dim+1]. */
int cmx_read_mesh_elements(const char* base, int32_t elem_dim, int32_t* dim, int32_t* n_nodes, int32_t* n_elems,
                           int32_t* n_elem_attr, double* coords, int32_t* elems, double* elem_attr);

/* Reads a fibre orientation file (FibreReader, heart/src/io/FibreReader.cpp:
 * 246-335): a header with the number of items (and optionally BIN), then
 * per_item values per element -- dim*dim for .ortho (fibre, sheet, normal
 * directions, one after another), dim for .axi.  out == NULL: count only. */
int cmx_read_fibres(const char* path, int32_t per_item, int32_t* n_items, double* out);

/* Per-element intracellular conductivity tensors [n_elems][dim][dim] from
 * fibre directions: A diag(sigma) A^T with A's columns the fibre, sheet and
 * normal directions (OrthotropicConductivityTensors.cpp:40-130), or
 * sigma[1] I + (sigma[0] - sigma[1]) f f^T when axisymmetric
 * (AxisymmetricConductivityTensors.cpp:100-170). */
int cmx_conductivity_tensors(int32_t dim, int32_t n_elems, int32_t axisymmetric, const double* fibres,
                             const double* sigma, double* out);

const char* cmx_io_last_error(void);

/* HDF5 results file in the layout of the reference's Hdf5DataWriter
 * (io/src/writer/Hdf5DataWriter.cpp:544-663, read by Hdf5DataReader): dataset
 * "Data" [time][node][variable] of doubles with the attributes
 * "Variable Details" (one "name(unit)" string per variable) and
 * "IsDataComplete" = 1, dataset "Time" [time] with attribute "Unit"
 * (time_unit, NULL for "msecs").  Rows are streamed to disk one time step at
 * a time (values[node][variable]); cmx_h5_close writes the time column and the
 * final sizes.  Host code: no device work. */
typedef struct cmx_h5 cmx_h5;
int cmx_h5_open(const char* path, int32_t n_nodes, int32_t n_vars, const char* const* names,
                const char* const* units, const char* time_unit, cmx_h5** out);
int cmx_h5_write_step(cmx_h5* h, double time, const double* values);
int cmx_h5_close(cmx_h5* h);

/* Batched single cells: n_cells independent cells (no tissue), each
 * integrated from t = 0 to its own t_end as AbstractCardiacCell::Compute
 * integrates one cell (heart/src/odes/AbstractCardiacCell.cpp:64-88,
 * ode/src/solver/AbstractOneStepIvpOdeSolver.cpp:93-153,
 * heart/src/odes/AbstractRushLarsenCardiacCell.cpp:57-104): rates at the
 * step's start, V by forward Euler, the rest by the integrator, the cell's
 * own area stimulus (uA/cm^2; Cm = 1) and parameters.  CMX_MIXED computes
 * the rates and the gates in FP32 and accumulates V and the other states
 * (concentrations) in FP64. */
typedef struct {
    int32_t model;               /* CMX_MODEL_* */
    int32_t integrator;          /* CMX_FORWARD_EULER or CMX_RUSH_LARSEN */
    int32_t precision;           /* CMX_FP64 or CMX_MIXED */
    int32_t n_cells;
    double dt;                   /* ms; every t_end a whole number of dt */
    const double* t_end;         /* [n_cells] ms */
    const cmx_stimulus* stimuli; /* [n_cells] one protocol per cell, or NULL */
    int32_t n_params;            /* per-cell parameter planes given (<= model's count) */
    const double* params;        /* [n_params][n_cells] or NULL */
    double threshold;            /* mV, for the action-potential summary */
    int32_t steps_per_launch;    /* steps per kernel launch (0: 2000) */
    int32_t lookup_tables;       /* mixed Rush-Larsen: gate steady states and decay factors over dt
                                    from a table in V (as cmx_opts.lookup_tables); a voltage
                                    outside the table fails the run.  Not with the
                                    allocation-free variant. */
    int32_t lanes_per_cell;      /* 1: one thread steps a cell; 2: two wavefronts share 64 cells, one
                                    advancing the gates and one the rest (the step's critical path
                                    halves when wavefronts are too few to interleave on a SIMD);
                                    0: per launch, 2 while the launch's cells would fill at most half
                                    of the device's SIMDs with one lane each, else 1 (the
                                    allocation-free variant takes 0 as 1).  Lookup tables need 1. */
    int32_t trace_every;         /* > 0: each cell's V sampled at t = 0 and after every trace_every
                                    steps into trace (AbstractCardiacCell::Compute's OdeSolution
                                    with that sampling interval); 0: no trace.  Not with the
                                    allocation-free variant. */
    double* trace;               /* [n_samples][n_cells] host array, n_samples = (the longest
                                    cell's steps) / trace_every + 1; NaN past a cell's own t_end */
} cmx_cells;

typedef struct {
    double kernel_ms;            /* device time of all launches */
    int64_t cell_steps;          /* sum of the cells' step counts */
    int32_t launches;
    int32_t block;               /* threads per block of the one-lane launches */
    int32_t split_launches;      /* launches with two wavefronts per 64 cells */
} cmx_cells_stats;

/* y0: [n_states][n_cells] initial states (NULL: the model's initial state);
 * y_out: [n_states][n_cells] final states (may be NULL); props: [4][n_cells]
 * (may be NULL) the action-potential summary of each cell's voltage trace as
 * CellProperties computes it at the given threshold
 * (heart/src/postprocessing/CellProperties.cpp:54-200): completed action
 * potentials, the last onset (linear interpolation), the peak of the last
 * completed one, and the time (interpolated) at which it fell back below the
 * threshold; -1 where none.  stats may be NULL.  Errors: CMX_E_* with the
 * message in cmx_cells_last_error(). */
int cmx_cells_run(const cmx_cells* cells, const double* y0, double* y_out, double* props, cmx_cells_stats* stats);
const char* cmx_cells_last_error(void);

/* No counterpart in the reference.  cmx_cells_run for callers that cannot
 * allocate device memory (code under hipGraph capture, tight serving loops,
 * harnesses that forbid hipMalloc): it calls no hipMalloc, hipFree or other
 * allocating routine and never synchronises.  Every array is a DEVICE
 * pointer -- cells->t_end, cells->stimuli, cells->params, y0 [n_states][n]
 * (NULL: the model's initial state), y_out [n_states][n] and props [4][n]
 * (either may be NULL) -- and the work is enqueued on `stream` (a
 * hipStream_t; NULL = the null stream); the caller synchronises before
 * reading y_out / props.  The kernels' working buffers (state and parameter
 * planes, step counts, padded stimuli, the CellProperties accumulators) live
 * in one static device array of CMX_CELLS_SCRATCH_BYTES per GPU, held until
 * the process ends; n_cells beyond what it holds (118,144 TT06 or 195,072
 * LR91 cells in FP64) is rejected with CMX_E_ARG.  Results are those
 * of cmx_cells_run on the same inputs (cells are not sorted by run length,
 * and every cell runs in one launch).  A cell whose t_end is not a whole
 * number of dt, or whose regular stimulus lasts longer than its period,
 * comes back as NaN in y_out and props (the host cannot see device inputs
 * to reject them).  Calls that overlap on one GPU share the array and race:
 * serialise them (one stream, or host ordering). */
#define CMX_CELLS_SCRATCH_BYTES (32u << 20)
int cmx_cells_run_noalloc(const cmx_cells* cells, const double* y0, double* y_out, double* props, void* stream);

/* Builds a regular slab mesh exactly as ConstructRegularSlabMesh does
 * (mesh/src/common/AbstractTetrahedralMesh.cpp:277-614): nodes
 * (k*(ey+1)+j)*(ex+1)+i, 2-D staggered triangles, 3-D six tetrahedra per cube.
 * With coords == NULL only the sizes are returned. */
int cmx_slab_mesh(int32_t dim, int32_t ex, int32_t ey, int32_t ez, double h,
                  int32_t* n_nodes, int32_t* n_elems, double* coords, int32_t* elems);

#ifdef __cplusplus
}
#endif
#endif /* CMX_H_ */
