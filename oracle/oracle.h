/*
 * FP64 CPU oracle for the monodomain tissue time-step -- TEST INFRASTRUCTURE.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker (or the timed CPU baseline).
 * The product (chaste_amd) never links, loads or falls back to it.
 *
 * It restates the reference algorithm on the CPU in double precision:
 *   mesh       ConstructRegularSlabMesh            mesh/src/common/AbstractTetrahedralMesh.cpp:277-614
 *   assembly   2nd-order Gauss P1 mass/stiffness   pde/src/solver/AbstractFeVolumeIntegralAssembler.hpp:253-459,
 *                                                  pde/src/solver/MassMatrixAssembler.hpp:76-118,
 *                                                  heart/src/solver/electrics/monodomain/MonodomainAssembler.cpp:42-66,
 *                                                  heart/src/solver/electrics/monodomain/MonodomainStiffnessMatrixAssembler.hpp:68-84
 *   cells      forward Euler / Rush-Larsen, V clamped
 *                                                  heart/src/odes/AbstractCardiacCell.cpp:70-111,
 *                                                  heart/src/odes/AbstractRushLarsenCardiacCell.cpp:57-153,
 *                                                  ode/src/solver/EulerIvpOdeSolver.cpp:38-57,
 *                                                  ode/src/solver/AbstractOneStepIvpOdeSolver.cpp:41-153
 *   LR91 RHS   hand restatement                    heart/dynamic/DynamicallyLoadableLr91.cpp:45-369
 *   TT06 RHS   CellML restatement (generated header heart/src/odes/cellml/TenTusscher2006Epi.cellml)
 *   tissue     SolveCellSystems + caches           heart/src/tissue/AbstractCardiacTissue.cpp:464-734
 *   RHS        b = M (Am Cm V/dt - Am Iion - Istim) heart/src/solver/electrics/monodomain/MonodomainSolver.cpp:41-131
 *   solve      CG, Jacobi PC, warm start, PETSc default preconditioned-norm test
 *                                                  linalg/src/LinearSystem.cpp:674-1130
 *   time       t_k = t0 + k dt                     global/src/TimeStepper.cpp:43-200,
 *                                                  pde/src/solver/AbstractDynamicLinearPdeSolver.hpp:328-497
 *   stimulus   SimpleStimulus / RegularStimulus    heart/src/stimulus/SimpleStimulus.cpp:44-76,
 *                                                  heart/src/stimulus/RegularStimulus.cpp:58-80
 *
 * Pins beyond the reference's golden files (tests/test_oracle_golden.py):
 *   orc_cell_compute_except_voltage   hand and generated LR91 agree, V held,
 *                                     two half intervals equal one
 *   orc_cell_compute_v                equals orc_cell_compute's V at every step
 */
#ifndef CMX_ORACLE_H_
#define CMX_ORACLE_H_

#ifdef __cplusplus
extern "C" {
#endif

enum orc_model { ORC_LR91_HAND = 0, ORC_LR91_CELLML = 1, ORC_TT06_EPI = 2, ORC_ORD_ENDO = 3, ORC_HH52 = 4,
                 ORC_FOX2002 = 5, ORC_MAHAJAN2008 = 6, ORC_DN85 = 7, ORC_FR2000 = 8, ORC_MALECKAR2008 = 9,
                 ORC_N98 = 10 };
/* Cell integrators: forward Euler, Rush-Larsen, backward Euler (the
 * reference's *BackwardEuler classes), generalized Rush-Larsen of first and
 * second order (--grl1 / --grl2 cells); all but FE on generated models only. */
enum orc_integrator { ORC_FE = 0, ORC_RL = 1, ORC_BE = 2, ORC_GRL1 = 3, ORC_GRL2 = 4 };

int orc_model_num_states(int model);

/* Single cell (AbstractCardiacCell::Compute): V integrated by forward Euler,
 * the other states by the integrator (orc_integrator; backward Euler advances
 * V first and steps the rest at the new V).  y is in/out, CellML order
 * (V first).  Stimulus: magnitude for t in [start, start+dur] (SimpleStimulus)
 * or every period if period > 0 (RegularStimulus).  Writes the state every
 * steps_per_sample steps (including t0) into samples[(k)*S + i]; returns the
 * number of samples written, or -1 on error. */
int orc_cell_compute(int model, int integrator, const double* params, double* y,
                     double t0, double t_end, double dt,
                     double stim_mag, double stim_dur, double stim_start, double stim_period,
                     int steps_per_sample, double* samples, int max_samples);

/* orc_cell_compute from t = 0 to t_end keeping V only: v_out[0] = V(0), then
 * V after every step (v_out holds round(t_end/dt) + 1 values); returns the
 * count written, or -1 on error. */
int orc_cell_compute_v(int model, int integrator, const double* params, double* y, double t_end, double dt,
                       double stim_mag, double stim_dur, double stim_start, double stim_period, double* v_out);

/* AbstractCardiacCell::ComputeExceptVoltage over [t0, t_end] in steps of dt
 * (V clamped, the other states by the integrator), then GetIIonic at the
 * result (what SolveCellSystems caches per node).  y is in/out; returns the
 * ionic current, NaN on error. */
double orc_cell_compute_except_voltage(int model, int integrator, const double* params, double* y,
                                       double t0, double t_end, double dt);

/* The model's default initial state (CellML order, V first); returns S. */
int orc_model_initial_state(int model, double* y);

/* GetIIonic at the given state (uA/cm^2). */
double orc_cell_iionic(int model, const double* params, const double* y);

/* Regular slab mesh sizes: elements per axis (ey/ez ignored below dim). */
int orc_slab_sizes(int dim, int ex, int ey, int ez, int* n_nodes, int* n_elems);
/* coords[n*dim], elems[ne*(dim+1)] */
int orc_slab_fill(int dim, int ex, int ey, int ez, double h, double* coords, int* elems);

/* P1 assembly with the reference's quadrature. sigma: dim*dim tensors, one
 * shared by all elements (sigma_per_elem == 0) or one per element
 * (sigma[e*dim*dim ...], AbstractCardiacTissue::rGetIntracellularConductivityTensor).
 * Outputs CSR (row_ptr[n+1], col[nnz], sorted) of M (unscaled mass, row-sum
 * lumped when lumped != 0) and K (stiffness).  Call with col == NULL to get
 * nnz only. */
int orc_assemble(int dim, int n, const double* coords, int ne, const int* elems,
                 const double* sigma, int sigma_per_elem, int lumped, int* nnz_out,
                 int* row_ptr, int* col, double* mass, double* stiff);

/* Jacobi-PCG on CSR (x in/out = initial guess).  Stops when ||D^-1 r||_2 <=
 * max(rtol * ||D^-1 b||_2, atol).  Returns iterations, or -1 if max_its hit. */
int orc_pcg(int n, const int* row_ptr, const int* col, const double* val,
            const double* b, double* x, double rtol, double atol, int max_its);

typedef struct {
    int dim, ex, ey, ez;      /* slab: elements per axis */
    double h;                 /* space step (cm) */
    int model;                /* orc_model */
    int integrator;           /* orc_integrator */
    double sigma[3];          /* diagonal intracellular conductivity, fibres along x (mS/cm) */
    double am, cm;            /* 1/cm, uF/cm^2 */
    double dt_pde, dt_ode;    /* ms */
    double stim_mag, stim_dur, stim_start;  /* uA/cm^3, ms, ms */
    double stim_xlim;         /* stimulated nodes: x < stim_xlim */
    double t_end;             /* ms, starts at 0 */
    double ksp_rtol, ksp_atol;
    int max_its;
    int threads;              /* OpenMP threads (<=0: library default) */
    int n_param;              /* per-node parameters (0: model defaults) */
    const double* params;     /* [n_param][n] node-major planes, or NULL */
    int lumped;               /* HeartConfig::SetUseMassLumping */
    const double* sigma_elem; /* [n_elems][dim*dim] per-element tensors, or NULL for sigma[] */
    double stim_period;       /* > 0: RegularStimulus with this period (ms) */
    double* modifier_out;     /* [4][n] ActivationOutputModifier times, or NULL */
    double modifier_threshold;
    int modifier_every;       /* process the voltage every this many PDE steps (and at t = 0) */
    int split;                /* HeartConfig::SetUseReactionDiffusionOperatorSplitting: Strang
                                 splitting, cells (V updated) over each half step around a
                                 diffusion-only solve */
    int svi;                  /* HeartConfig::SetUseStateVariableInterpolation (ICI step only,
                                 as the reference's splitting solver has no correction) */
    const unsigned char* stim_mask;  /* [n] stimulated nodes, or NULL for x < stim_xlim */
    /* a general mesh instead of the slab (dim = the mesh dimension): */
    int mesh_n_nodes, mesh_n_elems;
    const double* mesh_coords;       /* [n][dim], or NULL for the slab */
    const int* mesh_elems;           /* [ne][dim+1] from zero */
    /* mixed cell types (a factory returning different cells per node,
       AbstractCardiacCellFactory::CreateCardiacCellForTissueNode): per-node
       orc_model / orc_integrator, or NULL for model / integrator everywhere */
    const int* node_model;
    const int* node_integrator;
} orc_mono_cfg;

typedef struct {
    double t_ode, t_rhs, t_ksp, t_total;   /* seconds, reference event names Ode/AssRhs/Ksp */
    long long cg_iterations;
    int steps;
    int n_nodes, nnz;
} orc_mono_stats;

/* Full monodomain run from the model's default state.  Any output may be NULL.
 * V_out[n]; state_out[S*n] (state-major planes, V plane included);
 * act_out[n] first upward crossing of act_threshold (ms, linear interpolation,
 * -1 if none), checked after every PDE step (with splitting, along the
 * polyline through V at t, after the first half step at the midpoint, and at
 * t_next); v_trace[n_trace_nodes * steps]
 * records V at trace_nodes after each step. */
int orc_mono_run(const orc_mono_cfg* cfg, double* V_out, double* state_out,
                 double* act_out, double act_threshold,
                 int n_trace, const int* trace_nodes, double* v_trace,
                 orc_mono_stats* stats);

#ifdef __cplusplus
}
#endif
#endif
