/*
 * C++ host API for the MI355X monodomain path, mirroring the reference's
 * user-facing classes so that code written against them keeps its shape:
 *
 *   HeartConfig                 heart/src/problem/HeartConfig.hpp:203 (singleton;
 *                               defaults of heart/src/problem/HeartConfigDefaults.hpp:52-152)
 *   SimpleStimulus,
 *   RegularStimulus, ZeroStimulus
 *                               heart/src/stimulus/SimpleStimulus.cpp:44-76,
 *                               heart/src/stimulus/RegularStimulus.cpp:41-80
 *   CellLuoRudy1991FromCellML,
 *   CellTenTusscher2006EpiFromCellML[RushLarsen], CellHodgkinHuxley1952FromCellML,
 *   CellFoxModel2002FromCellML, CellMahajan2008FromCellML
 *                               the classes chaste_codegen generates from
 *                               heart/src/odes/cellml/{LuoRudy1991,TenTusscher2006Epi,
 *                               HodgkinHuxley1952,FoxModel2002,Mahajan2008}.cellml
 *   AbstractCardiacCellFactory, PlaneStimulusCellFactory
 *                               heart/src/problem/cell_factories/AbstractCardiacCellFactory.hpp:100-108,
 *                               heart/src/problem/cell_factories/PlaneStimulusCellFactory.hpp:55-86
 *   TetrahedralMesh::ConstructRegularSlabMesh
 *                               mesh/src/common/AbstractTetrahedralMesh.cpp:569-614
 *   MonodomainProblem           heart/src/problem/MonodomainProblem.cpp:44-137 with
 *                               AbstractCardiacProblem::Initialise / Solve
 *                               (heart/src/problem/AbstractCardiacProblem.cpp:117-214, 371-584)
 *
 * A cell factory returns a small descriptor (model, integrator, stimulus)
 * per node instead of a heap-allocated cell object: the whole tissue lives in
 * device memory as structure-of-arrays and is advanced by the gfx950 kernels
 * behind include/cmx.h.  Errors throw chaste_amd::Exception (the reference's
 * EXCEPTION / KSPEXCEPT, global/src/Exception.hpp).
 */
#ifndef CHASTE_AMD_MONODOMAIN_HPP_
#define CHASTE_AMD_MONODOMAIN_HPP_

#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <sys/stat.h>
#include <memory>
#include <stdexcept>
#include <algorithm>
#include <string>
#include <vector>

#include "../cmx.h"

namespace chaste_amd {

class Exception : public std::runtime_error {
public:
    Exception(int code, const std::string& msg) : std::runtime_error(msg), mCode(code) {}
    int GetCode() const { return mCode; }
private:
    int mCode;
};

inline void Check(int rc, const cmx_ctx* ctx)
{
    if (rc != CMX_OK) throw Exception(rc, cmx_last_error(ctx));
}

// ---------------------------------------------------------------------------
// HeartConfig subset (same names, units and defaults)
// ---------------------------------------------------------------------------
class HeartConfig {
public:
    static HeartConfig* Instance()
    {
        static HeartConfig instance;
        return &instance;
    }
    static void Reset() { *Instance() = HeartConfig(); }

    void SetSpaceDimension(unsigned d) { mDim = d; }
    unsigned GetSpaceDimension() const { return mDim; }
    void SetSimulationDuration(double ms) { mDuration = ms; }
    double GetSimulationDuration() const { return mDuration; }
    void SetSlabDimensions(double x, double y, double z, double h) { mDim = 3; mSize[0] = x; mSize[1] = y; mSize[2] = z; mStep = h; }
    // HeartConfig.cpp SetMeshFileName (mesh read by TrianglesMeshReader; the
    // fibre file, when given, is <prefix>.ortho or <prefix>.axi)
    enum FibreOrientation { NoFibreOrientation = 0, Orthotropic = 1, Axisymmetric = 2 };
    // Results file <output directory>/<prefix>.h5 (the directory is taken
    // under $CHASTE_TEST_OUTPUT when that is set), V at t = 0 and after every
    // printing step, as AbstractCardiacProblem::Solve writes it
    // (heart/src/problem/AbstractCardiacProblem.cpp:459-464, 552-555).
    void SetOutputDirectory(const std::string& dir) { mOutputDir = dir; }
    const std::string& GetOutputDirectory() const { return mOutputDir; }
    void SetOutputFilenamePrefix(const std::string& prefix) { mOutputPrefix = prefix; }
    const std::string& GetOutputFilenamePrefix() const { return mOutputPrefix; }
    void SetMeshFileName(const std::string& prefix, FibreOrientation fibres = NoFibreOrientation)
    {
        mMeshFile = prefix;
        mFibres = fibres;
    }
    const std::string& GetMeshName() const { return mMeshFile; }
    FibreOrientation GetConductivityMedia() const { return mFibres; }
    void SetSheetDimensions(double x, double y, double h) { mDim = 2; mSize[0] = x; mSize[1] = y; mSize[2] = 0; mStep = h; }
    void SetFibreLength(double x, double h) { mDim = 1; mSize[0] = x; mSize[1] = mSize[2] = 0; mStep = h; }
    const double* GetSlabSize() const { return mSize; }
    double GetInterNodeSpace() const { return mStep; }

    void SetIntracellularConductivities(const std::vector<double>& s) { mSigma = s; }
    std::vector<double> GetIntracellularConductivities() const { return mSigma; }
    void SetSurfaceAreaToVolumeRatio(double am) { mAm = am; }
    double GetSurfaceAreaToVolumeRatio() const { return mAm; }
    void SetCapacitance(double cm) { mCm = cm; }
    double GetCapacitance() const { return mCm; }

    void SetOdePdeAndPrintingTimeSteps(double ode, double pde, double print)
    {
        if (ode <= 0 || pde <= 0 || print <= 0) throw Exception(CMX_E_ARG, "Time steps must be positive");
        if (ode > pde) throw Exception(CMX_E_ARG, "Ode time-step should not be greater than PDE time-step");
        mOdeDt = ode; mPdeDt = pde; mPrintDt = print;
    }
    void SetOdeTimeStep(double dt) { SetOdePdeAndPrintingTimeSteps(dt, mPdeDt, mPrintDt); }
    void SetPdeTimeStep(double dt) { SetOdePdeAndPrintingTimeSteps(mOdeDt, dt, mPrintDt); }
    void SetPrintingTimeStep(double dt) { SetOdePdeAndPrintingTimeSteps(mOdeDt, mPdeDt, dt); }
    double GetOdeTimeStep() const { return mOdeDt; }
    double GetPdeTimeStep() const { return mPdeDt; }
    double GetPrintingTimeStep() const { return mPrintDt; }

    void SetUseAbsoluteTolerance(double atol) { mUseAbs = true; mAtol = atol; }
    void SetUseRelativeTolerance(double rtol) { mUseAbs = false; mRtol = rtol; }
    bool GetUseAbsoluteTolerance() const { return mUseAbs; }
    bool GetUseRelativeTolerance() const { return !mUseAbs; }
    double GetAbsoluteTolerance() const
    {
        if (!mUseAbs) throw Exception(CMX_E_STATE, "Absolute tolerance is not set in Chaste parameters");
        return mAtol;
    }
    double GetRelativeTolerance() const
    {
        if (mUseAbs) throw Exception(CMX_E_STATE, "Relative tolerance is not set in Chaste parameters");
        return mRtol;
    }
    void SetKSPSolver(const char* s) { if (std::string(s) != "cg") throw Exception(CMX_E_ARG, "only the cg KSP solver is provided"); }
    void SetKSPPreconditioner(const char* pc) { mPc = pc; }  // Jacobi on the device for every PC name
    std::string GetKSPPreconditioner() const { return mPc; }
    void SetUseMassLumping(bool on = true) { mLumping = on; }
    bool GetUseMassLumping() const { return mLumping; }
    // HeartConfig.cpp:2874-2882 (MonodomainProblem.cpp:65-69 picks the splitting solver)
    void SetUseReactionDiffusionOperatorSplitting(bool on = true) { mSplitting = on; }
    bool GetUseReactionDiffusionOperatorSplitting() const { return mSplitting; }
    // HeartConfig.cpp:2756-2779 (MonodomainSolver.cpp:200-209 builds the correction assembler)
    void SetUseStateVariableInterpolation(bool on = true) { mSvi = on; }
    bool GetUseStateVariableInterpolation() const { return mSvi; }

    // MI355X-specific knobs (no reference counterpart).
    void SetGpuPrecision(int precision) { mPrecision = precision; }
    int GetGpuPrecision() const { return mPrecision; }
    void SetGpuRefinementSweeps(int sweeps) { mRefine = sweeps; }
    int GetGpuRefinementSweeps() const { return mRefine; }
    void SetKspMaxIterations(int its) { mMaxIts = its; }
    int GetKspMaxIterations() const { return mMaxIts; }
    // mixed precision, Rush-Larsen cells: gate coefficients from a table in V
    // (the reference selects its *Opt cell classes for this)
    void SetGpuLookupTables(bool on = true) { mLut = on; }
    bool GetGpuLookupTables() const { return mLut; }

private:
    unsigned mDim = 3;
    double mDuration = 10.0;
    double mSize[3] = {0.2, 0.2, 0.2};
    double mStep = 0.01;
    std::vector<double> mSigma = {1.75, 1.75, 1.75};
    double mAm = 1400.0, mCm = 1.0;
    double mOdeDt = 0.01, mPdeDt = 0.01, mPrintDt = 0.01;
    bool mUseAbs = true;
    double mAtol = 2e-4, mRtol = 1e-6;
    std::string mPc = "bjacobi";
    bool mLumping = false;
    bool mSplitting = false;
    std::string mMeshFile;
    std::string mOutputDir, mOutputPrefix;
    FibreOrientation mFibres = NoFibreOrientation;
    bool mSvi = false;
    int mPrecision = CMX_FP64;
    int mRefine = 1;
    int mMaxIts = 1000;
    bool mLut = false;
};

// ---------------------------------------------------------------------------
// Stimuli
// ---------------------------------------------------------------------------
struct AbstractStimulusFunction {
    virtual ~AbstractStimulusFunction() = default;
    virtual cmx_stimulus Describe() const = 0;
};

struct ZeroStimulus : AbstractStimulusFunction {
    cmx_stimulus Describe() const override { return cmx_stimulus{0.0, 0.0, 0.0, -1.0, 0.0}; }
};

struct SimpleStimulus : AbstractStimulusFunction {
    SimpleStimulus(double magnitude, double duration, double timeOfStimulus = 0.0)
        : mMagnitude(magnitude), mDuration(duration), mStart(timeOfStimulus) {}
    cmx_stimulus Describe() const override { return cmx_stimulus{mMagnitude, mStart, mDuration, -1.0, 0.0}; }
    double mMagnitude, mDuration, mStart;
};

struct RegularStimulus : AbstractStimulusFunction {
    RegularStimulus(double magnitude, double duration, double period, double startTime, double stopTime = DBL_MAX)
        : mMagnitude(magnitude), mDuration(duration), mPeriod(period), mStart(startTime), mStop(stopTime) {}
    cmx_stimulus Describe() const override { return cmx_stimulus{mMagnitude, mStart, mDuration, mPeriod, mStop}; }
    double mMagnitude, mDuration, mPeriod, mStart, mStop;
};

// ---------------------------------------------------------------------------
// Cell types (descriptors of the generated models)
// ---------------------------------------------------------------------------
struct CellDescriptor {
    int model = CMX_MODEL_LR91;
    int integrator = CMX_FORWARD_EULER;
    std::shared_ptr<AbstractStimulusFunction> stimulus;
};

struct CellLuoRudy1991FromCellML {
    static constexpr int kModel = CMX_MODEL_LR91;
    static constexpr int kIntegrator = CMX_FORWARD_EULER;
};
struct CellLuoRudy1991FromCellMLRushLarsen {
    static constexpr int kModel = CMX_MODEL_LR91;
    static constexpr int kIntegrator = CMX_RUSH_LARSEN;
};
struct CellTenTusscher2006EpiFromCellML {
    static constexpr int kModel = CMX_MODEL_TT06_EPI;
    static constexpr int kIntegrator = CMX_FORWARD_EULER;
};
struct CellTenTusscher2006EpiFromCellMLRushLarsen {
    static constexpr int kModel = CMX_MODEL_TT06_EPI;
    static constexpr int kIntegrator = CMX_RUSH_LARSEN;
};
// O'Hara-Rudy 2011 endocardial cells (not in the reference tree; written from
// the published equations): forward Euler and Rush-Larsen.
struct CellOHaraRudy2011EndoFromCellML {
    static constexpr int kModel = CMX_MODEL_ORD_ENDO;
    static constexpr int kIntegrator = CMX_FORWARD_EULER;
};
struct CellOHaraRudy2011EndoFromCellMLRushLarsen {
    static constexpr int kModel = CMX_MODEL_ORD_ENDO;
    static constexpr int kIntegrator = CMX_RUSH_LARSEN;
};
// Backward-Euler classes (LuoRudy1991BackwardEuler[Opt].hpp,
// TenTusscher2006EpiBackwardEuler[Opt].hpp; the Opt variants' lookup tables
// only change round-off, so both names map to the same cell).
struct CellLuoRudy1991FromCellMLBackwardEuler {
    static constexpr int kModel = CMX_MODEL_LR91;
    static constexpr int kIntegrator = CMX_BACKWARD_EULER;
};
struct CellLuoRudy1991FromCellMLBackwardEulerOpt : CellLuoRudy1991FromCellMLBackwardEuler {};
struct CellTenTusscher2006EpiFromCellMLBackwardEuler {
    static constexpr int kModel = CMX_MODEL_TT06_EPI;
    static constexpr int kIntegrator = CMX_BACKWARD_EULER;
};
struct CellTenTusscher2006EpiFromCellMLBackwardEulerOpt : CellTenTusscher2006EpiFromCellMLBackwardEuler {};
// Generalized Rush-Larsen cells (chaste_codegen --grl1 / --grl2).
struct CellLuoRudy1991FromCellMLGRL1 {
    static constexpr int kModel = CMX_MODEL_LR91;
    static constexpr int kIntegrator = CMX_GRL1;
};
struct CellLuoRudy1991FromCellMLGRL2 {
    static constexpr int kModel = CMX_MODEL_LR91;
    static constexpr int kIntegrator = CMX_GRL2;
};
struct CellTenTusscher2006EpiFromCellMLGRL1 {
    static constexpr int kModel = CMX_MODEL_TT06_EPI;
    static constexpr int kIntegrator = CMX_GRL1;
};
struct CellTenTusscher2006EpiFromCellMLGRL2 {
    static constexpr int kModel = CMX_MODEL_TT06_EPI;
    static constexpr int kIntegrator = CMX_GRL2;
};

// Further generated models (chaste_codegen classes of
// heart/src/odes/cellml/{HodgkinHuxley1952,FoxModel2002,Mahajan2008,
// DiFrancescoNoble1985,FaberRudy2000,Maleckar2008,NobleVargheseKohlNoble1998a}.cellml).
#define CHASTE_AMD_CELL_TAGS(Base, model)                                  \
    struct Base {                                                          \
        static constexpr int kModel = model;                               \
        static constexpr int kIntegrator = CMX_FORWARD_EULER;              \
    };                                                                     \
    struct Base##RushLarsen {                                              \
        static constexpr int kModel = model;                               \
        static constexpr int kIntegrator = CMX_RUSH_LARSEN;                \
    };                                                                     \
    struct Base##BackwardEuler {                                           \
        static constexpr int kModel = model;                               \
        static constexpr int kIntegrator = CMX_BACKWARD_EULER;             \
    };                                                                     \
    struct Base##BackwardEulerOpt : Base##BackwardEuler {};                \
    struct Base##GRL1 {                                                    \
        static constexpr int kModel = model;                               \
        static constexpr int kIntegrator = CMX_GRL1;                       \
    };                                                                     \
    struct Base##GRL2 {                                                    \
        static constexpr int kModel = model;                               \
        static constexpr int kIntegrator = CMX_GRL2;                       \
    };
CHASTE_AMD_CELL_TAGS(CellHodgkinHuxley1952FromCellML, CMX_MODEL_HH52)
CHASTE_AMD_CELL_TAGS(CellFoxModel2002FromCellML, CMX_MODEL_FOX2002)
CHASTE_AMD_CELL_TAGS(CellMahajan2008FromCellML, CMX_MODEL_MAHAJAN2008)
CHASTE_AMD_CELL_TAGS(CellDiFrancescoNoble1985FromCellML, CMX_MODEL_DN85)
CHASTE_AMD_CELL_TAGS(CellFaberRudy2000FromCellML, CMX_MODEL_FR2000)
CHASTE_AMD_CELL_TAGS(CellMaleckar2008FromCellML, CMX_MODEL_MALECKAR2008)
CHASTE_AMD_CELL_TAGS(CellNobleVargheseKohlNoble1998aFromCellML, CMX_MODEL_N98)
#undef CHASTE_AMD_CELL_TAGS

// A cell model loaded at run time from a plugin built from any CellML file
// (chaste_amd.build.build_model_plugin): the model id for a CellDescriptor,
// with any integrator.  The reference's DynamicCellModelLoader::CreateCell
// (heart/src/odes/DynamicCellModelLoader.cpp:55-63) returns a cell object
// instead; here the id stands for every node's cell of that type.
inline int LoadCellModel(const std::string& pluginPath)
{
    int32_t id = -1;
    Check(cmx_load_model(pluginPath.c_str(), &id), nullptr);
    return id;
}

// ---------------------------------------------------------------------------
// Mesh (regular slabs; general simplices through SetNodesAndElements)
// ---------------------------------------------------------------------------
template <unsigned DIM>
class TetrahedralMesh {
public:
    void ConstructRegularSlabMesh(double spaceStep, double width, double height = 0, double depth = 0)
    {
        // element counts snapped as the reference does (AbstractTetrahedralMesh.cpp:580-600)
        const int ex = (int)((width + 0.5 * spaceStep) / spaceStep);
        const int ey = DIM > 1 ? (int)((height + 0.5 * spaceStep) / spaceStep) : 0;
        const int ez = DIM > 2 ? (int)((depth + 0.5 * spaceStep) / spaceStep) : 0;
        if (std::fabs(ex * spaceStep - width) > DBL_EPSILON * width ||
            (DIM > 1 && std::fabs(ey * spaceStep - height) > DBL_EPSILON * height) ||
            (DIM > 2 && std::fabs(ez * spaceStep - depth) > DBL_EPSILON * depth))
            throw Exception(CMX_E_ARG, "Space step does not divide the size of the mesh");
        ConstructRegularSlabMeshWithDimensions(spaceStep, ex, ey, ez);
    }
    void ConstructRegularSlabMeshWithDimensions(double spaceStep, int ex, int ey = 0, int ez = 0)
    {
        int32_t n = 0, ne = 0;
        Check(cmx_slab_mesh(DIM, ex, ey, ez, spaceStep, &n, &ne, nullptr, nullptr), nullptr);
        mCoords.resize((size_t)n * DIM);
        mElems.resize((size_t)ne * (DIM + 1));
        Check(cmx_slab_mesh(DIM, ex, ey, ez, spaceStep, nullptr, nullptr, mCoords.data(), mElems.data()), nullptr);
    }
    // TetrahedralMesh::ConstructFromMeshReader with a TrianglesMeshReader on
    // <base>.node / <base>.ele.
    void ConstructFromMeshFile(const std::string& base)
    {
        int32_t d = 0, n = 0, ne = 0, na = 0;
        CheckIo(cmx_read_mesh(base.c_str(), &d, &n, &ne, &na, nullptr, nullptr, nullptr));
        if ((unsigned)d != DIM) throw Exception(CMX_E_ARG, "SPACE_DIM  != dimension read from file ");
        mCoords.resize((size_t)n * DIM);
        mElems.resize((size_t)ne * (DIM + 1));
        mAttributes.resize((size_t)ne * na);
        CheckIo(cmx_read_mesh(base.c_str(), &d, &n, &ne, &na, mCoords.data(), mElems.data(),
                              na ? mAttributes.data() : nullptr));
    }
    void SetNodesAndElements(std::vector<double> coords, std::vector<int32_t> elems)
    {
        mCoords = std::move(coords);
        mElems = std::move(elems);
    }
    unsigned GetNumNodes() const { return (unsigned)(mCoords.size() / DIM); }
    unsigned GetNumElements() const { return (unsigned)(mElems.size() / (DIM + 1)); }
    const double* GetNodeLocation(unsigned i) const { return &mCoords[(size_t)i * DIM]; }
    const std::vector<double>& rGetCoords() const { return mCoords; }
    const std::vector<int32_t>& rGetElements() const { return mElems; }

private:
    static void CheckIo(int rc)
    {
        if (rc) throw Exception(rc, cmx_io_last_error());
    }
    std::vector<double> mCoords;
    std::vector<int32_t> mElems;
    std::vector<double> mAttributes;
};

// ---------------------------------------------------------------------------
// Cell factories
// ---------------------------------------------------------------------------
template <unsigned DIM>
class AbstractCardiacCellFactory {
public:
    AbstractCardiacCellFactory() : mpZeroStimulus(new ZeroStimulus) {}
    virtual ~AbstractCardiacCellFactory() = default;
    virtual CellDescriptor CreateCardiacCellForTissueNode(const double* x) = 0;
    void SetMesh(TetrahedralMesh<DIM>* pMesh) { mpMesh = pMesh; }
    TetrahedralMesh<DIM>* GetMesh() { return mpMesh; }
protected:
    template <class CELL>
    CellDescriptor Make(std::shared_ptr<AbstractStimulusFunction> stim) const
    {
        return CellDescriptor{CELL::kModel, CELL::kIntegrator, std::move(stim)};
    }
    std::shared_ptr<AbstractStimulusFunction> mpZeroStimulus;
    TetrahedralMesh<DIM>* mpMesh = nullptr;
};

// Stimulated on the x = 0 plane (x^2 <= 1e-10), zero stimulus elsewhere.
template <class CELL, unsigned DIM>
class PlaneStimulusCellFactory : public AbstractCardiacCellFactory<DIM> {
public:
    explicit PlaneStimulusCellFactory(double stimulusMagnitude = -600, double stimulusDuration = 0.5)
        : mpStimulus(new SimpleStimulus(stimulusMagnitude, stimulusDuration)) {}
    CellDescriptor CreateCardiacCellForTissueNode(const double* x) override
    {
        if (x[0] * x[0] <= 1e-10) return this->template Make<CELL>(mpStimulus);
        return this->template Make<CELL>(this->mpZeroStimulus);
    }
private:
    std::shared_ptr<AbstractStimulusFunction> mpStimulus;
};

// ---------------------------------------------------------------------------
// The PDE solver contract (pde/src/solver/AbstractDynamicLinearPdeSolver.hpp:
// 160-182, 328-497) and the GPU solver behind it
// ---------------------------------------------------------------------------

// The reference's solvers pass PETSc vectors: SetInitialCondition keeps the
// caller's vector without taking ownership, Solve() returns a new one that
// the caller destroys (PetscTools::Destroy).  Here a Vec is a host vector of
// nodal voltages with the same ownership rules (DestroyVec).
typedef std::vector<double>* Vec;
inline void DestroyVec(Vec& v)
{
    delete v;
    v = nullptr;
}

template <unsigned DIM>
class AbstractDynamicLinearPdeSolver {
public:
    virtual ~AbstractDynamicLinearPdeSolver() = default;
    void SetTimes(double tStart, double tEnd)
    {
        mTstart = tStart;
        mTend = tEnd;
        if (mTstart >= mTend) throw Exception(CMX_E_ARG, "Start time has to be less than end time");
        mTimesSet = true;
    }
    virtual void SetTimeStep(double dt)
    {
        if (dt <= 0) throw Exception(CMX_E_ARG, "Time step has to be greater than zero");
        mIdealTimeStep = dt;
    }
    void SetInitialCondition(Vec initialCondition)
    {
        if (!initialCondition) throw Exception(CMX_E_ARG, "null initial condition");
        mInitialCondition = initialCondition;
    }
    virtual Vec Solve() = 0;

protected:
    // AbstractDynamicLinearPdeSolver::Solve's preconditions (:330-343)
    void CheckSetUp() const
    {
        if (!mTimesSet) throw Exception(CMX_E_STATE, "SetTimes() has not been called");
        if (mIdealTimeStep <= 0.0) throw Exception(CMX_E_STATE, "SetTimeStep() has not been called");
        if (!mInitialCondition) throw Exception(CMX_E_STATE, "SetInitialCondition() has not been called");
    }
    // TimeStepper(tStart, tEnd, dt, enforceConstantTimeStep = true) step count
    // (global/src/TimeStepper.cpp:43-104): dt must divide the interval.
    long ConstantSteps() const
    {
        const long k = (long)std::floor((mTend - mTstart) / mIdealTimeStep + 0.5);
        const double eps = mTend > 1.0 ? DBL_EPSILON * mTend : DBL_EPSILON;
        if (k < 1 || std::fabs(mTend - (mTstart + mIdealTimeStep * k)) > eps)
            throw Exception(CMX_E_ARG, "TimeStepper estimates non-constant timesteps will need to be used: check "
                                       "timestep divides (end_time-start_time) (or divides printing timestep)");
        return k;
    }
    double mTstart = 0.0, mTend = 0.0, mIdealTimeStep = -1.0;
    bool mTimesSet = false;
    Vec mInitialCondition = nullptr;
};

// MonodomainSolver (heart/src/solver/electrics/monodomain/MonodomainSolver.cpp)
// on the device: the tissue's cells and the assembled matrices live in the
// cmx context; Solve() advances from the initial condition's voltage over
// [tStart, tEnd] in steps of the PDE time step (cell sweep, ICI right-hand
// side and CG solve per step) and returns the voltage at tEnd.  The matrix
// A = (Am Cm / dt) M + K is assembled for the context's PDE time step, so
// SetTimeStep accepts that step only.  The cell states persist between
// Solve() calls, as the reference's tissue keeps its cells.  The initial
// condition is uploaded unless it is the vector the device already holds
// (the previous Solve()'s result, unchanged).
template <unsigned DIM>
class GpuMonodomainSolver : public AbstractDynamicLinearPdeSolver<DIM> {
public:
    GpuMonodomainSolver(cmx_ctx* pCtx, unsigned numNodes, double pdeTimeStep)
        : mpCtx(pCtx), mNumNodes(numNodes), mPdeTimeStep(pdeTimeStep) {}
    void SetTimeStep(double dt) override
    {
        if (dt > 0 && std::fabs(dt - mPdeTimeStep) > 1e-12 * mPdeTimeStep)
            throw Exception(CMX_E_ARG, "the GPU solver's matrix is assembled for the PDE time step " +
                                           std::to_string(mPdeTimeStep) + " ms");
        AbstractDynamicLinearPdeSolver<DIM>::SetTimeStep(dt);
    }
    Vec Solve() override
    {
        this->CheckSetUp();
        if (this->mInitialCondition->size() != mNumNodes)
            throw Exception(CMX_E_ARG, "the initial condition must hold one value per node");
        const long n_steps = this->ConstantSteps();
        if (*this->mInitialCondition != mDeviceV) {
            Check(cmx_set_voltage(mpCtx, this->mInitialCondition->data()), mpCtx);
            ++mUploads;
        }
        Check(cmx_step(mpCtx, this->mTstart, (int32_t)n_steps), mpCtx);
        Vec out = new std::vector<double>(mNumNodes);
        try {
            Check(cmx_get_voltage(mpCtx, out->data()), mpCtx);
        } catch (...) {
            DestroyVec(out);
            throw;
        }
        mDeviceV = *out;
        return out;
    }
    // Host-to-device uploads of an initial condition so far.
    long GetNumberOfUploads() const { return mUploads; }
    cmx_ctx* GetContext() { return mpCtx; }

private:
    cmx_ctx* mpCtx;
    unsigned mNumNodes;
    double mPdeTimeStep;
    std::vector<double> mDeviceV;   // the voltage the device holds after the last Solve()
    long mUploads = 0;
};

// ---------------------------------------------------------------------------
// MonodomainProblem
// ---------------------------------------------------------------------------
template <unsigned DIM>
class MonodomainProblem {
public:
    explicit MonodomainProblem(AbstractCardiacCellFactory<DIM>* pCellFactory) : mpFactory(pCellFactory) {}
    ~MonodomainProblem()
    {
        if (mpCtx) cmx_destroy(mpCtx);
    }
    MonodomainProblem(const MonodomainProblem&) = delete;
    MonodomainProblem& operator=(const MonodomainProblem&) = delete;

    void SetMesh(TetrahedralMesh<DIM>* pMesh) { mpMesh = pMesh; }
    TetrahedralMesh<DIM>& rGetMesh() { return *mpMesh; }

    // Builds the mesh from HeartConfig when none was set, creates one cell
    // per node through the factory and assembles M, K on the device.
    void Initialise()
    {
        HeartConfig* cfg = HeartConfig::Instance();
        if (!mpMesh) {
            if (!cfg->GetMeshName().empty()) {
                mOwnMesh.ConstructFromMeshFile(cfg->GetMeshName());
            } else {
                if (cfg->GetSpaceDimension() != DIM) throw Exception(CMX_E_ARG, "Space dimension does not match");
                const double* sz = cfg->GetSlabSize();
                mOwnMesh.ConstructRegularSlabMesh(cfg->GetInterNodeSpace(), sz[0], sz[1], sz[2]);
            }
            mpMesh = &mOwnMesh;
        }
        mpFactory->SetMesh(mpMesh);
        const unsigned n = mpMesh->GetNumNodes();
        std::vector<int32_t> stim_index(n, -1);
        std::vector<cmx_stimulus> table;
        std::vector<const AbstractStimulusFunction*> seen;
        // one cell group per distinct (model, integrator) the factory returns
        std::vector<int32_t> group_of(n), group_model, group_integrator;
        for (unsigned i = 0; i < n; ++i) {
            CellDescriptor d = mpFactory->CreateCardiacCellForTissueNode(mpMesh->GetNodeLocation(i));
            int g = -1;
            for (size_t q = 0; q < group_model.size(); ++q)
                if (group_model[q] == d.model && group_integrator[q] == d.integrator) g = (int)q;
            if (g < 0) {
                g = (int)group_model.size();
                group_model.push_back(d.model);
                group_integrator.push_back(d.integrator);
            }
            group_of[i] = g;
            const cmx_stimulus s = d.stimulus ? d.stimulus->Describe() : cmx_stimulus{0, 0, 0, -1, 0};
            if (s.magnitude == 0.0) continue;
            int k = -1;
            for (size_t q = 0; q < seen.size(); ++q)
                if (seen[q] == d.stimulus.get()) k = (int)q;
            if (k < 0) {
                if (table.size() >= 255) throw Exception(CMX_E_ARG, "more than 255 distinct stimuli");
                k = (int)table.size();
                table.push_back(s);
                seen.push_back(d.stimulus.get());
            }
            stim_index[i] = k;
        }
        cmx_mesh mesh{};
        mesh.dim = DIM;
        mesh.n_nodes = (int32_t)n;
        mesh.coords = mpMesh->rGetCoords().data();
        mesh.n_elems = (int32_t)mpMesh->GetNumElements();
        mesh.elems = mpMesh->rGetElements().data();
        const std::vector<double> sig = cfg->GetIntracellularConductivities();
        for (unsigned r = 0; r < DIM; ++r) mesh.sigma[r * DIM + r] = sig[r];
        // Fibre orientation file: one tensor per element
        // (OrthotropicConductivityTensors / AxisymmetricConductivityTensors).
        std::vector<double> tensors;
        if (cfg->GetConductivityMedia() != HeartConfig::NoFibreOrientation && !cfg->GetMeshName().empty()) {
            const bool axi = cfg->GetConductivityMedia() == HeartConfig::Axisymmetric;
            const std::string path = cfg->GetMeshName() + (axi ? ".axi" : ".ortho");
            const int per = axi ? (int)DIM : (int)(DIM * DIM);
            int32_t items = 0;
            if (cmx_read_fibres(path.c_str(), per, &items, nullptr)) throw Exception(CMX_E_ARG, cmx_io_last_error());
            if ((unsigned)items != mpMesh->GetNumElements())
                throw Exception(CMX_E_ARG, "The size of the fibre file does not match the number of elements in the mesh");
            std::vector<double> fib((size_t)items * per);
            if (cmx_read_fibres(path.c_str(), per, &items, fib.data())) throw Exception(CMX_E_ARG, cmx_io_last_error());
            double s3[3] = {sig[0], DIM > 1 ? sig[1] : sig[0], DIM > 2 ? sig[2] : (DIM > 1 ? sig[1] : sig[0])};
            tensors.resize((size_t)items * DIM * DIM);
            if (cmx_conductivity_tensors(DIM, items, axi ? 1 : 0, fib.data(), s3, tensors.data()))
                throw Exception(CMX_E_ARG, cmx_io_last_error());
            mesh.sigma_elem = tensors.data();
        }
        cmx_tissue tissue{};
        mModel = group_model.empty() ? CMX_MODEL_LR91 : group_model[0];
        mCellTypes = (int)group_model.size();
        tissue.model = mModel;
        tissue.integrator = group_integrator.empty() ? CMX_FORWARD_EULER : group_integrator[0];
        if (group_model.size() > 1) {
            tissue.n_groups = (int32_t)group_model.size();
            tissue.group_of = group_of.data();
            tissue.group_model = group_model.data();
            tissue.group_integrator = group_integrator.data();
        }
        tissue.n_stimuli = (int32_t)table.size();
        tissue.stimuli = table.data();
        tissue.stim_index = stim_index.data();
        cmx_opts o{};
        o.precision = cfg->GetGpuPrecision();
        o.format = CMX_FMT_AUTO;
        o.dt_pde = cfg->GetPdeTimeStep();
        o.dt_ode = cfg->GetOdeTimeStep();
        o.am = cfg->GetSurfaceAreaToVolumeRatio();
        o.cm = cfg->GetCapacitance();
        if (cfg->GetUseAbsoluteTolerance()) { o.ksp_atol = cfg->GetAbsoluteTolerance(); o.ksp_rtol = 0.0; }
        else { o.ksp_rtol = cfg->GetRelativeTolerance(); o.ksp_atol = 0.0; }
        o.ksp_max_its = cfg->GetKspMaxIterations();
        o.refine_sweeps = cfg->GetGpuRefinementSweeps();
        o.mass_lumping = cfg->GetUseMassLumping() ? 1 : 0;
        o.use_graph = 1;
        o.extrapolate = 0;
        o.operator_splitting = cfg->GetUseReactionDiffusionOperatorSplitting() ? 1 : 0;
        o.state_variable_interpolation = cfg->GetUseStateVariableInterpolation() ? 1 : 0;
        o.lookup_tables = cfg->GetGpuLookupTables() ? 1 : 0;
        if (mpCtx) { cmx_destroy(mpCtx); mpCtx = nullptr; }
        Check(cmx_create(&mesh, &tissue, &o, &mpCtx), nullptr);
        if (!mParams.empty() && mCellTypes == 1) {   // parameters set before a re-initialisation
            int32_t S = 0, P = 0;
            Check(cmx_model_info(mModel, &S, &P), nullptr);
            Check(cmx_set_params(mpCtx, P, mParams.data()), mpCtx);
        }
        mCurrentTime = 0.0;
        mAm = o.am;
        mCm = o.cm;
    }

    // Advances to the simulation duration in printing steps, as
    // AbstractCardiacProblem::Solve does (heart/src/problem/AbstractCardiacProblem.cpp:
    // 371-584): a TimeStepper(current time, duration, printing dt) whose last
    // interval ends exactly at the duration (global/src/TimeStepper.cpp:43-121),
    // each interval one PDE TimeStepper with a constant dt
    // (pde/src/solver/AbstractDynamicLinearPdeSolver.hpp:367), which must
    // divide the interval.
    void Solve()
    {
        if (!mpCtx) throw Exception(CMX_E_STATE, "Initialise() must be called before Solve()");
        HeartConfig* cfg = HeartConfig::Instance();
        // The reference reads Am and Cm when it assembles, so values set
        // between Initialise() and the first Solve() apply: rebuild then.
        if (cfg->GetSurfaceAreaToVolumeRatio() != mAm || cfg->GetCapacitance() != mCm) {
            if (mCurrentTime != 0.0)
                throw Exception(CMX_E_STATE, "Am and Cm cannot change once the simulation has started");
            Initialise();
        }
        const double end = cfg->GetSimulationDuration(), print = cfg->GetPrintingTimeStep(), dt = cfg->GetPdeTimeStep();
        if (end <= mCurrentTime) throw Exception(CMX_E_ARG, "The simulation duration must be positive");
        const long per = (long)std::floor(print / dt + 0.5);
        if (std::fabs(per * dt - print) > 1e-10 * print)
            throw Exception(CMX_E_ARG, "Printing time-step should be a multiple of PDE time step");
        const double t_start = mCurrentTime;
        const double eps_print = end > 1.0 ? DBL_EPSILON * end : DBL_EPSILON;
        // the solver the problem creates and deletes (AbstractCardiacProblem.cpp:
        // 389-390, 573); the voltage passes through Vecs as in the reference's loop
        std::unique_ptr<GpuMonodomainSolver<DIM>> solver(CreateSolver());
        solver->SetTimeStep(dt);
        cmx_h5* out = OpenResults();
        Vec initial_condition = nullptr;
        try {
            initial_condition = new std::vector<double>(GetSolution());
            if (out && cmx_h5_write_step(out, t_start, initial_condition->data()))
                throw Exception(CMX_E_ARG, cmx_io_last_error());
            double t = t_start;
            for (long k = 0; t < end; ++k) {
                double t_next = t_start + (k + 1) * print;   // printing TimeStepper
                if (end - t_next <= eps_print) t_next = end;
                solver->SetTimes(t, t_next);
                solver->SetInitialCondition(initial_condition);
                Vec solution = solver->Solve();
                DestroyVec(initial_condition);
                initial_condition = solution;
                mCurrentTime = t_next;
                if (out && cmx_h5_write_step(out, t_next, solution->data()))
                    throw Exception(CMX_E_ARG, cmx_io_last_error());
                t = t_next;
            }
        } catch (...) {
            DestroyVec(initial_condition);
            if (out) cmx_h5_close(out);
            throw;
        }
        DestroyVec(initial_condition);
        if (out && cmx_h5_close(out)) throw Exception(CMX_E_ARG, cmx_io_last_error());
        mCurrentTime = end;
    }

    // AbstractCardiacProblem::CreateSolver (heart/src/problem/AbstractCardiacProblem.hpp:452,
    // MonodomainProblem.cpp:52-77): the monodomain PDE solver over this
    // problem's tissue; the caller owns it.  After Initialise().
    GpuMonodomainSolver<DIM>* CreateSolver()
    {
        if (!mpCtx) throw Exception(CMX_E_STATE, "Initialise() must be called before CreateSolver()");
        return new GpuMonodomainSolver<DIM>(mpCtx, mpMesh->GetNumNodes(), HeartConfig::Instance()->GetPdeTimeStep());
    }

    // Path of the results file, or empty when no output directory is set.
    std::string GetResultsPath() const
    {
        HeartConfig* cfg = HeartConfig::Instance();
        if (cfg->GetOutputDirectory().empty()) return "";
        const char* root = std::getenv("CHASTE_TEST_OUTPUT");
        std::string dir = root ? std::string(root) + "/" + cfg->GetOutputDirectory() : cfg->GetOutputDirectory();
        const std::string prefix = cfg->GetOutputFilenamePrefix().empty() ? "results" : cfg->GetOutputFilenamePrefix();
        return dir + "/" + prefix + ".h5";
    }

    std::vector<double> GetSolution() const
    {
        std::vector<double> V(mpMesh->GetNumNodes());
        Check(cmx_get_voltage(mpCtx, V.data()), mpCtx);
        return V;
    }
    void SetInitialCondition(const std::vector<double>& V) { Check(cmx_set_voltage(mpCtx, V.data()), mpCtx); }
    std::vector<double> GetIionicCache() const
    {
        std::vector<double> out(mpMesh->GetNumNodes());
        Check(cmx_get_iionic(mpCtx, out.data()), mpCtx);
        return out;
    }
    cmx_stats GetStats() const
    {
        cmx_stats s{};
        Check(cmx_get_stats(mpCtx, &s), mpCtx);
        return s;
    }
    double GetCurrentTime() const { return mCurrentTime; }
    cmx_ctx* GetContext() { return mpCtx; }

    // The named cell parameter of every node, as the reference sets it on
    // each cell object (AbstractParameterisedSystem::SetParameter,
    // ode/src/common/AbstractParameterisedSystem.cpp:256-275): values[n_nodes].
    // Names are the model's CellML names (cmx_model_param); other parameters
    // keep their values.  After Initialise().
    void SetCellParameter(const std::string& rName, const std::vector<double>& values)
    {
        if (!mpCtx) throw Exception(CMX_E_STATE, "SetCellParameter needs Initialise() first");
        if (mCellTypes > 1) throw Exception(CMX_E_ARG, "SetCellParameter needs a single cell type");
        const unsigned n = mpMesh->GetNumNodes();
        if (values.size() != n) throw Exception(CMX_E_ARG, "one value per node is required");
        int32_t S = 0, P = 0;
        Check(cmx_model_info(mModel, &S, &P), nullptr);
        // the library's name or the reference's (GetParameterIndex)
        int32_t index = -1;
        Check(cmx_model_find_param(mModel, rName.c_str(), &index), nullptr);
        if (mParams.empty()) {
            mParams.resize((size_t)P * n);
            for (int q = 0; q < P; ++q) {
                double def = 0.0;
                Check(cmx_model_param(mModel, q, nullptr, &def), nullptr);
                std::fill(mParams.begin() + (size_t)q * n, mParams.begin() + (size_t)(q + 1) * n, def);
            }
        }
        std::copy(values.begin(), values.end(), mParams.begin() + (size_t)index * n);
        Check(cmx_set_params(mpCtx, P, mParams.data()), mpCtx);
    }

private:
    cmx_h5* OpenResults()
    {
        const std::string path = GetResultsPath();
        if (path.empty()) return nullptr;
        const std::string dir = path.substr(0, path.rfind('/'));
        std::string partial;
        for (size_t p = 0; p <= dir.size(); ++p) {   // mkdir -p
            if (p == dir.size() || dir[p] == '/') {
                partial = dir.substr(0, p);
                if (!partial.empty()) ::mkdir(partial.c_str(), 0755);
            }
        }
        const char* names[1] = {"V"};
        const char* units[1] = {"mV"};
        cmx_h5* h = nullptr;
        if (cmx_h5_open(path.c_str(), (int32_t)mpMesh->GetNumNodes(), 1, names, units, "msecs", &h))
            throw Exception(CMX_E_ARG, cmx_io_last_error());
        return h;
    }
    AbstractCardiacCellFactory<DIM>* mpFactory;
    TetrahedralMesh<DIM>* mpMesh = nullptr;
    TetrahedralMesh<DIM> mOwnMesh;
    cmx_ctx* mpCtx = nullptr;
    double mCurrentTime = 0.0;
    double mAm = 0.0, mCm = 0.0;
    int mModel = CMX_MODEL_LR91;
    int mCellTypes = 1;             // distinct (model, integrator) pairs the factory returned
    std::vector<double> mParams;    // [P][n] planes once a parameter is set
};

}  // namespace chaste_amd

#endif  // CHASTE_AMD_MONODOMAIN_HPP_
