// The cell-model table.  Every model the runtime knows sits behind one
// ModelOps entry: its metadata, and launchers of its tissue cell kernel, Iion
// kernel, state-variable interpolation kernel and batched single-cell
// kernels, instantiated in the model's own translation unit
// (cmx_model_*.hip, through cmx_model_impl.h), so the runtime dispatches on a
// model id without compiling every model into one unit.
#pragma once
#include <vector>

#include "cmx_adjoint_kernels.h"
#include "cmx_cell_kernels.h"
#include "cmx_cells_kernels.h"

namespace cmx {

// Metadata of a model id (CMX_MODEL_*).
struct ModelMeta {
    int S, P;
    const double* init;
    const double* pdef;
    const char* const* state_names;
    const char* const* param_names;
    const char* const* param_aliases;   // the reference's names of the parameters ("" where none)
    bool implicit;       // backward-Euler and generalized Rush-Larsen forms available
};

// Row layouts a model's SVI kernel is built for (index into ModelOps::svi*).
template <class Rows> struct RowsSlot;
template <> struct RowsSlot<StencilRows<4>> { static constexpr int k = 0; };
template <> struct RowsSlot<StencilRows<8>> { static constexpr int k = 1; };
template <> struct RowsSlot<StencilRows<16>> { static constexpr int k = 2; };
template <> struct RowsSlot<EllRows> { static constexpr int k = 3; };
constexpr int kRowsSlots = 4;

template <class RS, class RV>
using CellLaunch = void (*)(cmx_ctx*, int integrator, int mode, int kind, int G, const CellArgs<RS, RV>&);
template <class RS, class RV>
using SviLaunch = void (*)(cmx_ctx*, const void* rows, const SviArgs<RS, RV>&, int G);
using IionicLaunch = void (*)(hipStream_t, int G, int n_cells, const int32_t* node, int ps, const double* V,
                              const void* state, const void* params, const void* pdef, int n_param_planes, double cm,
                              double* out);
template <class RS>
using BatchLaunch = void (*)(const CellsArgs<RS>&, bool rush_larsen, int k0, int k1, hipStream_t, int lanes);

struct ModelOps {
    ModelMeta meta;
    CellLaunch<double, double> cells64;
    CellLaunch<float, float> cells32;
    SviLaunch<double, double> svi64[kRowsSlots];
    SviLaunch<float, float> svi32[kRowsSlots];
    IionicLaunch iionic64, iionic32;
    BatchLaunch<double> batch64;
    BatchLaunch<float> batch32;
    // the model's gate lookup table for ODE steps of length dt (error message or null)
    const char* (*gate_table)(double dt, double cm, std::vector<float>& table, int* rows);
    // the adjoint's cell step (FP64, forward Euler or Rush-Larsen), or null
    void (*adjoint)(cmx_ctx*, const CellAdjArgs&, bool rush_larsen);
};

// The entry of a model id; null for an unknown id.
const ModelOps* model_ops(int model);

// Models loaded at run time (cmx_load_model) take ids kPluginModelBase + k.
constexpr int kPluginModelBase = 64;
constexpr int kMaxPluginModels = 64;
constexpr int kMaxModelStates = 64;

// Stamp a plugin reports from cmx_model_plugin_abi(): a plugin built against
// headers whose shared structures differ from the library's is refused.
constexpr uint32_t kModelPluginAbi =
    (1u << 24) | (uint32_t)((sizeof(ModelOps) * 131u + sizeof(cmx_ctx) * 7u + sizeof(CellArgs<double, double>) * 5u +
                             sizeof(CellArgs<float, float>) * 3u + sizeof(SviArgs<double, double>) * 11u +
                             sizeof(CellsArgs<double>) * 13u + sizeof(CellsArgs<float>) * 17u) &
                            0xFFFFFFu);

inline bool model_meta(int model, ModelMeta* m)
{
    const ModelOps* ops = model_ops(model);
    if (!ops) return false;
    *m = ops->meta;
    return true;
}

// The gate table of a model id for ODE steps of length dt; returns an error
// message, or null.
inline const char* gate_table(int model, double dt, double cm, std::vector<float>& table, int* rows)
{
    const ModelOps* ops = model_ops(model);
    if (!ops) return "unknown cell model";
    return ops->gate_table(dt, cm, table, rows);
}

template <class RS, class RV> CellLaunch<RS, RV> cell_launch(const ModelOps* o);
template <> inline CellLaunch<double, double> cell_launch<double, double>(const ModelOps* o) { return o->cells64; }
template <> inline CellLaunch<float, float> cell_launch<float, float>(const ModelOps* o) { return o->cells32; }
template <class RS, class RV> SviLaunch<RS, RV> svi_launch(const ModelOps* o, int slot);
template <> inline SviLaunch<double, double> svi_launch<double, double>(const ModelOps* o, int k) { return o->svi64[k]; }
template <> inline SviLaunch<float, float> svi_launch<float, float>(const ModelOps* o, int k) { return o->svi32[k]; }
template <class RS> BatchLaunch<RS> batch_launch(const ModelOps* o);
template <> inline BatchLaunch<double> batch_launch<double>(const ModelOps* o) { return o->batch64; }
template <> inline BatchLaunch<float> batch_launch<float>(const ModelOps* o) { return o->batch32; }

}  // namespace cmx
