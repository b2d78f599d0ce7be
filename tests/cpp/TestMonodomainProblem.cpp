// C++ API test on the GPU path, in the shape of the reference's
// heart/test/monodomain/TestMonodomainProblem.hpp:784-880: the same
// HeartConfig calls, PlaneStimulusCellFactory<CellLuoRudy1991FromCellML, DIM>
// (-600*5000 uA/cm^3) and the same golden voltages at nodes 1, 3, 5, 7
// (atol 5e-3 mV).  Prints one JSON line per case; exit status 0 iff all pass.
#include <cstdio>
#include <functional>
#include <string>
#include <vector>

#include "chaste_amd/Monodomain.hpp"

using namespace chaste_amd;

static int failures = 0;

template <unsigned DIM>
static void RunCase(const char* name, const double golden[4], double duration, const std::string& out_dir = "",
                    const char* prefix = "")
{
    if (!out_dir.empty()) {
        HeartConfig::Instance()->SetOutputDirectory(out_dir);
        HeartConfig::Instance()->SetOutputFilenamePrefix(prefix);
    }
    PlaneStimulusCellFactory<CellLuoRudy1991FromCellML, DIM> cell_factory(-600 * 5000);
    MonodomainProblem<DIM> monodomain_problem(&cell_factory);
    HeartConfig::Instance()->SetSimulationDuration(duration);
    monodomain_problem.Initialise();
    monodomain_problem.Solve();
    std::vector<double> V = monodomain_problem.GetSolution();
    const unsigned nodes[4] = {1, 3, 5, 7};
    double worst = 0;
    for (int q = 0; q < 4; ++q) worst = std::max(worst, std::fabs(V[nodes[q]] - golden[q]));
    const bool ok = worst < 5e-3;
    if (!ok) ++failures;
    std::printf("{\"case\": \"%s\", \"V\": [%.6f, %.6f, %.6f, %.6f], \"max_err\": %.3e, \"cg_iterations\": %lld, \"pass\": %s}\n",
                name, V[1], V[3], V[5], V[7], worst, (long long)monodomain_problem.GetStats().cg_iterations,
                ok ? "true" : "false");
    std::fflush(stdout);
}

// heart/test/monodomain/TestMonodomainProblem.hpp:226-290 on its mesh file
// (mesh/test/data/1D_0_to_1mm_10_elements): sigma 0.0005, Am = Cm = 1 set
// between Initialise() and Solve(), 2 ms; nodes 1, 3, 5, 7, 9, 10 (atol 5e-3).
static std::vector<double> g_mesh_file_V;   // the single 2 ms run, for the two-halves case

static void RunMeshFileCase(const std::string& meshes, const std::string& out_dir)
{
    HeartConfig::Reset();
    HeartConfig::Instance()->SetIntracellularConductivities({0.0005});
    HeartConfig::Instance()->SetSimulationDuration(2.0);
    HeartConfig::Instance()->SetMeshFileName(meshes + "/1D_0_to_1mm_10_elements");
    if (!out_dir.empty()) {
        HeartConfig::Instance()->SetOutputDirectory(out_dir);
        HeartConfig::Instance()->SetOutputFilenamePrefix("MonodomainLR91_1d");
    }
    PlaneStimulusCellFactory<CellLuoRudy1991FromCellML, 1> cell_factory;
    MonodomainProblem<1> monodomain_problem(&cell_factory);
    monodomain_problem.Initialise();
    HeartConfig::Instance()->SetSurfaceAreaToVolumeRatio(1.0);
    HeartConfig::Instance()->SetCapacitance(1.0);
    monodomain_problem.Solve();
    std::vector<double> V = monodomain_problem.GetSolution();
    const unsigned nodes[6] = {1, 3, 5, 7, 9, 10};
    const double golden[6] = {20.7710232, 21.5319692, 22.9280817, 24.0611303, -0.770330519, -19.2234919};
    double worst = 0;
    for (int q = 0; q < 6; ++q) worst = std::max(worst, std::fabs(V[nodes[q]] - golden[q]));
    const bool ok = worst < 5e-3;
    if (!ok) ++failures;
    g_mesh_file_V = V;
    std::printf("{\"case\": \"mesh_file_1d\", \"V9\": %.6f, \"max_err\": %.3e, \"pass\": %s}\n", V[9], worst,
                ok ? "true" : "false");
    std::fflush(stdout);
}

// heart/test/monodomain/TestMonodomainProblem.hpp:1178-1225: the same problem
// solved to 1 ms, then (SetSimulationDuration(2.0), Solve() again) to 2 ms:
// the goldens (atol 5e-3) and the single run's voltages (5e-11).
static void RunTwoHalvesCase(const std::string& meshes)
{
    HeartConfig::Reset();
    HeartConfig::Instance()->SetIntracellularConductivities({0.0005});
    HeartConfig::Instance()->SetMeshFileName(meshes + "/1D_0_to_1mm_10_elements");
    HeartConfig::Instance()->SetSurfaceAreaToVolumeRatio(1.0);
    HeartConfig::Instance()->SetCapacitance(1.0);
    PlaneStimulusCellFactory<CellLuoRudy1991FromCellML, 1> cell_factory;
    MonodomainProblem<1> monodomain_problem(&cell_factory);
    monodomain_problem.Initialise();
    HeartConfig::Instance()->SetSimulationDuration(1.0);
    monodomain_problem.Solve();
    HeartConfig::Instance()->SetSimulationDuration(2.0);
    monodomain_problem.Solve();
    std::vector<double> V = monodomain_problem.GetSolution();
    const unsigned nodes[6] = {1, 3, 5, 7, 9, 10};
    const double golden[6] = {20.7710232, 21.5319692, 22.9280817, 24.0611303, -0.770330519, -19.2234919};
    double worst = 0, vs_single = 0;
    for (int q = 0; q < 6; ++q) worst = std::max(worst, std::fabs(V[nodes[q]] - golden[q]));
    for (size_t i = 0; i < V.size() && i < g_mesh_file_V.size(); ++i)
        vs_single = std::max(vs_single, std::fabs(V[i] - g_mesh_file_V[i]));
    const bool ok = worst < 5e-3 && g_mesh_file_V.size() == V.size() && vs_single <= 5e-11;
    if (!ok) ++failures;
    std::printf("{\"case\": \"two_halves\", \"max_err\": %.3e, \"vs_single\": %.3e, \"pass\": %s}\n", worst,
                vs_single, ok ? "true" : "false");
    std::fflush(stdout);
}

// heart/test/monodomain/TestMonodomainProblem.hpp:658-691: printing time
// step 0.1 ms over 0.3 ms on the same mesh file writes exactly t = 0, 0.1,
// 0.2, 0.3 (the Python side reads the results file back).
static void RunPrintingTimesCase(const std::string& meshes, const std::string& out_dir)
{
    HeartConfig::Reset();
    HeartConfig::Instance()->SetPrintingTimeStep(0.1);
    HeartConfig::Instance()->SetSimulationDuration(0.3);
    HeartConfig::Instance()->SetMeshFileName(meshes + "/1D_0_to_1mm_10_elements");
    HeartConfig::Instance()->SetOutputDirectory(out_dir);
    HeartConfig::Instance()->SetOutputFilenamePrefix("mono_testPrintTimes");
    PlaneStimulusCellFactory<CellLuoRudy1991FromCellML, 1> cell_factory;
    MonodomainProblem<1> monodomain_problem(&cell_factory);
    monodomain_problem.Initialise();
    monodomain_problem.Solve();
    std::printf("{\"case\": \"printing_times\", \"results\": \"%s\", \"pass\": true}\n",
                monodomain_problem.GetResultsPath().c_str());
    std::fflush(stdout);
}

// A duration that is not a multiple of the printing step: the printing
// TimeStepper's last interval ends exactly at the duration
// (global/src/TimeStepper.cpp:106-121), so printing 0.1 ms over 0.35 ms
// writes t = 0, 0.1, 0.2, 0.3, 0.35 and ends where a run printed every PDE
// step ends (5e-11); a duration the PDE step does not divide throws.
static void RunShortLastIntervalCase(const std::string& meshes, const std::string& out_dir)
{
    std::vector<double> V[2];
    for (int pass = 0; pass < 2; ++pass) {
        HeartConfig::Reset();
        HeartConfig::Instance()->SetIntracellularConductivities({0.0005});
        HeartConfig::Instance()->SetMeshFileName(meshes + "/1D_0_to_1mm_10_elements");
        HeartConfig::Instance()->SetPrintingTimeStep(pass == 0 ? 0.1 : 0.01);
        HeartConfig::Instance()->SetSimulationDuration(0.35);
        if (pass == 0) {
            HeartConfig::Instance()->SetOutputDirectory(out_dir);
            HeartConfig::Instance()->SetOutputFilenamePrefix("mono_shortLastInterval");
        }
        PlaneStimulusCellFactory<CellLuoRudy1991FromCellML, 1> cell_factory;
        MonodomainProblem<1> monodomain_problem(&cell_factory);
        monodomain_problem.Initialise();
        monodomain_problem.Solve();
        V[pass] = monodomain_problem.GetSolution();
    }
    double diff = 0;
    for (size_t i = 0; i < V[0].size(); ++i) diff = std::max(diff, std::fabs(V[0][i] - V[1][i]));
    HeartConfig::Reset();
    HeartConfig::Instance()->SetMeshFileName(meshes + "/1D_0_to_1mm_10_elements");
    HeartConfig::Instance()->SetSimulationDuration(0.355);
    PlaneStimulusCellFactory<CellLuoRudy1991FromCellML, 1> cell_factory;
    MonodomainProblem<1> monodomain_problem(&cell_factory);
    monodomain_problem.Initialise();
    int thrown = 0;
    try { monodomain_problem.Solve(); } catch (const Exception& e) { ++thrown; }
    const bool ok = diff <= 5e-11 && thrown == 1;
    if (!ok) ++failures;
    std::printf("{\"case\": \"short_last_interval\", \"vs_every_step\": %.3e, \"thrown\": %d, \"pass\": %s}\n",
                diff, thrown, ok ? "true" : "false");
    std::fflush(stdout);
}

// The solver contract of AbstractDynamicLinearPdeSolver
// (pde/src/solver/AbstractDynamicLinearPdeSolver.hpp:160-182, 280-343) on the
// solver MonodomainProblem::CreateSolver() returns, driven directly on the
// 1-D mesh-file case: the preconditions throw in the reference's order, the
// initial condition stays the caller's, Solve() returns a new vector, and two
// solves (0-1 ms, 1-2 ms) end where the problem's printed run ends (5e-11).
static void RunSolverContractCase(const std::string& meshes)
{
    HeartConfig::Reset();
    HeartConfig::Instance()->SetIntracellularConductivities({0.0005});
    HeartConfig::Instance()->SetMeshFileName(meshes + "/1D_0_to_1mm_10_elements");
    HeartConfig::Instance()->SetSurfaceAreaToVolumeRatio(1.0);
    HeartConfig::Instance()->SetCapacitance(1.0);
    PlaneStimulusCellFactory<CellLuoRudy1991FromCellML, 1> cell_factory;
    MonodomainProblem<1> problem(&cell_factory);
    problem.Initialise();
    GpuMonodomainSolver<1>* solver = problem.CreateSolver();
    std::vector<std::string> messages;
    auto expect_throw = [&](const std::function<void()>& f) {
        try { f(); messages.push_back("(no exception)"); }
        catch (const Exception& e) { messages.push_back(e.what()); }
    };
    expect_throw([&] { solver->Solve(); });
    expect_throw([&] { solver->SetTimes(1.0, 1.0); });
    solver->SetTimes(0.0, 1.0);
    expect_throw([&] { solver->Solve(); });
    expect_throw([&] { solver->SetTimeStep(0.02); });
    solver->SetTimeStep(0.01);
    expect_throw([&] { solver->Solve(); });
    Vec ic = new std::vector<double>(problem.GetSolution());
    const std::vector<double> ic_copy = *ic;
    solver->SetInitialCondition(ic);
    Vec v1 = solver->Solve();
    solver->SetTimes(1.0, 2.0);
    solver->SetInitialCondition(v1);
    Vec v2 = solver->Solve();
    const long uploads_two_solves = solver->GetNumberOfUploads();
    double vs_problem = 0;
    for (size_t i = 0; i < v2->size() && i < g_mesh_file_V.size(); ++i)
        vs_problem = std::max(vs_problem, std::fabs((*v2)[i] - g_mesh_file_V[i]));
    solver->SetTimes(2.0, 2.005);
    expect_throw([&] { solver->Solve(); });
    Vec shifted = new std::vector<double>(*v2);
    for (double& v : *shifted) v += 1.0;
    solver->SetTimes(2.0, 2.1);
    solver->SetInitialCondition(shifted);
    Vec v3 = solver->Solve();
    const long uploads = solver->GetNumberOfUploads();
    const std::vector<std::string> expected = {"SetTimes() has not been called", "Start time has to be less than end time",
                                               "SetTimeStep() has not been called", "the GPU solver",
                                               "SetInitialCondition() has not been called", "TimeStepper estimates"};
    bool msgs_ok = messages.size() == expected.size();
    for (size_t q = 0; msgs_ok && q < expected.size(); ++q) msgs_ok = messages[q].find(expected[q]) == 0;
    const bool ok = msgs_ok && *ic == ic_copy && v1 != ic && v2 != v1 && v3 != v2 && uploads_two_solves == 1 &&
                    uploads == 2 && g_mesh_file_V.size() == v2->size() && vs_problem <= 5e-11;
    if (!ok) ++failures;
    std::printf("{\"case\": \"solver_contract\", \"messages_ok\": %s, \"uploads\": [%ld, %ld], "
                "\"vs_problem\": %.3e, \"pass\": %s}\n", msgs_ok ? "true" : "false", uploads_two_solves, uploads,
                vs_problem, ok ? "true" : "false");
    if (!msgs_ok)
        for (const std::string& m : messages) std::printf("# %s\n", m.c_str());
    std::fflush(stdout);
    DestroyVec(ic);
    DestroyVec(v1);
    DestroyVec(v2);
    DestroyVec(shifted);
    DestroyVec(v3);
    delete solver;
}

// A factory that returns different cell classes per node (the reference's
// factories may, AbstractCardiacCellFactory.hpp:100): LR91 forward-Euler cells
// for x < 0.15 cm and TT06 Rush-Larsen cells elsewhere on a 0.3 x 0.1 cm
// sheet, stimulated on x = 0.  Prints every node's voltage after 4 ms for the
// Python side to check against the oracle's per-node-type run.
template <unsigned DIM>
class MixedCellFactory : public AbstractCardiacCellFactory<DIM> {
public:
    MixedCellFactory() : mpStimulus(new SimpleStimulus(-600 * 5000, 0.5)) {}
    CellDescriptor CreateCardiacCellForTissueNode(const double* x) override
    {
        auto stim = x[0] * x[0] <= 1e-10 ? mpStimulus : this->mpZeroStimulus;
        if (x[0] < 0.15) return this->template Make<CellLuoRudy1991FromCellML>(stim);
        return this->template Make<CellTenTusscher2006EpiFromCellMLRushLarsen>(stim);
    }
private:
    std::shared_ptr<AbstractStimulusFunction> mpStimulus;
};

static void RunMixedCellTypesCase()
{
    HeartConfig::Reset();
    HeartConfig::Instance()->SetSpaceDimension(2);
    HeartConfig::Instance()->SetSheetDimensions(0.3, 0.1, 0.01);
    HeartConfig::Instance()->SetSimulationDuration(4.0);
    HeartConfig::Instance()->SetUseRelativeTolerance(1e-12);
    MixedCellFactory<2> cell_factory;
    MonodomainProblem<2> monodomain_problem(&cell_factory);
    monodomain_problem.Initialise();
    int thrown = 0;
    const unsigned n = monodomain_problem.rGetMesh().GetNumNodes();
    try { monodomain_problem.SetCellParameter("fast_sodium_current__g_Na", std::vector<double>(n, 1.0)); }
    catch (const Exception& e) { ++thrown; }
    monodomain_problem.Solve();
    const std::vector<double> V = monodomain_problem.GetSolution();
    std::printf("{\"case\": \"mixed_cell_types\", \"thrown\": %d, \"V\": [", thrown);
    for (size_t i = 0; i < V.size(); ++i) std::printf("%s%.17g", i ? ", " : "", V[i]);
    std::printf("], \"pass\": %s}\n", thrown == 1 ? "true" : "false");
    if (thrown != 1) ++failures;
    std::fflush(stdout);
}

// Per-node cell parameters through SetCellParameter (the reference sets them
// on each cell, AbstractParameterisedSystem::SetParameter): the 1-D case of
// RunCase with g_Na = 0 everywhere cannot fire, so node 7 stays below -40 mV
// while the unmodified run reaches 13.7 mV; an unknown name throws.
static void RunParameterCase()
{
    HeartConfig::Reset();
    HeartConfig::Instance()->SetSpaceDimension(1);
    HeartConfig::Instance()->SetFibreLength(1, 0.01);
    HeartConfig::Instance()->SetSimulationDuration(5.0);
    PlaneStimulusCellFactory<CellLuoRudy1991FromCellML, 1> cell_factory(-600 * 5000);
    MonodomainProblem<1> monodomain_problem(&cell_factory);
    monodomain_problem.Initialise();
    const unsigned n = monodomain_problem.rGetMesh().GetNumNodes();
    monodomain_problem.SetCellParameter("fast_sodium_current__g_Na", std::vector<double>(n, 0.0));
    int thrown = 0;
    try { monodomain_problem.SetCellParameter("no_such_parameter", std::vector<double>(n, 1.0)); }
    catch (const Exception& e) { ++thrown; }
    monodomain_problem.Solve();
    const double v7 = monodomain_problem.GetSolution()[7];
    const bool ok = v7 < -40.0 && thrown == 1;   // measured -60.5 mV; an action potential would be above 0
    if (!ok) ++failures;
    std::printf("{\"case\": \"cell_parameter\", \"V7\": %.4f, \"thrown\": %d, \"pass\": %s}\n", v7, thrown,
                ok ? "true" : "false");
    std::fflush(stdout);
}

// argv[1]: folder holding the mesh files; argv[2]: output root.  Results
// files go to <root>/MonodomainCreatesGeometry/monodomain3d.h5,
// <root>/MonoProblem1d/MonodomainLR91_1d.h5,
// <root>/MonoProblem1dRequestedTimes/mono_testPrintTimes.h5 and
// <root>/MonoProblem1dShortLastInterval/mono_shortLastInterval.h5.
int main(int argc, char** argv)
{
    const std::string out_root = argc > 2 ? argv[2] : "";
    const auto out = [&](const char* sub) { return out_root.empty() ? out_root : out_root + "/" + sub; };
    HeartConfig::Reset();
    HeartConfig::Instance()->SetSpaceDimension(1);
    HeartConfig::Instance()->SetFibreLength(1, 0.01);
    const double g1[4] = {13.9682, 13.9149, 13.8216, 13.7275};
    RunCase<1>("1d", g1, 5.0);

    HeartConfig::Reset();
    HeartConfig::Instance()->SetSpaceDimension(2);
    HeartConfig::Instance()->SetSheetDimensions(0.3, 0.3, 0.01);
    const double g2[4] = {17.5728, 17.4562, 17.2559, 17.0440};
    RunCase<2>("2d", g2, 4.0);

    HeartConfig::Reset();
    HeartConfig::Instance()->SetSpaceDimension(3);
    HeartConfig::Instance()->SetSlabDimensions(0.1, 0.1, 0.1, 0.01);
    const double g3[4] = {31.8958, 32.1109, 32.7315, 33.7011};
    RunCase<3>("3d", g3, 2.0, out("MonodomainCreatesGeometry"), "monodomain3d");

    // Error behaviour: Solve before Initialise, bad time steps, bad slab step.
    HeartConfig::Reset();
    HeartConfig::Instance()->SetSpaceDimension(2);
    HeartConfig::Instance()->SetSheetDimensions(0.05, 0.05, 0.01);
    PlaneStimulusCellFactory<CellLuoRudy1991FromCellML, 2> factory(-600 * 5000);
    MonodomainProblem<2> problem(&factory);
    int thrown = 0;
    try { problem.Solve(); } catch (const Exception& e) { ++thrown; }
    try { HeartConfig::Instance()->SetOdePdeAndPrintingTimeSteps(0.02, 0.01, 0.01); } catch (const Exception& e) { ++thrown; }
    try { TetrahedralMesh<2> m; m.ConstructRegularSlabMesh(0.03, 0.1, 0.1); } catch (const Exception& e) { ++thrown; }
    const bool errors_ok = thrown == 3;
    if (argc > 1) RunMeshFileCase(argv[1], out("MonoProblem1d"));
    if (argc > 1) RunTwoHalvesCase(argv[1]);
    if (argc > 1) RunSolverContractCase(argv[1]);
    if (argc > 2) RunPrintingTimesCase(argv[1], out("MonoProblem1dRequestedTimes"));
    if (argc > 2) RunShortLastIntervalCase(argv[1], out("MonoProblem1dShortLastInterval"));
    RunParameterCase();
    RunMixedCellTypesCase();
    if (!errors_ok) ++failures;
    std::printf("{\"case\": \"errors\", \"thrown\": %d, \"pass\": %s}\n", thrown, errors_ok ? "true" : "false");
    return failures == 0 ? 0 : 1;
}
