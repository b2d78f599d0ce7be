// Builds a model's ModelOps entry: included by the model's translation unit
// only, which instantiates every kernel of that model.
#pragma once
#include "cmx_models.h"

namespace cmx {

template <class Model, bool IMPLICIT, bool ADJOINT = IMPLICIT>
struct ModelImpl {
    template <class RS, class RV>
    static void cells(cmx_ctx* c, int integrator, int mode, int kind, int G, const CellArgs<RS, RV>& a)
    {
        launch_cell_integrator<Model, RS, RV, IMPLICIT>(c, integrator, mode, kind, G, a);
    }
    template <class RS>
    static void batch(const CellsArgs<RS>& a, bool rl, int k0, int k1, hipStream_t s, int lanes)
    {
        if (rl) launch_cells<Model, RS, CMX_RUSH_LARSEN>(a, k0, k1, s, lanes);
        else launch_cells<Model, RS, CMX_FORWARD_EULER>(a, k0, k1, s, lanes);
    }
    static const char* table(double dt, double cm, std::vector<float>& t, int* rows)
    {
        return fill_gate_table<Model>(dt, cm, t, rows);
    }
    static void adjoint(cmx_ctx* c, const CellAdjArgs& a, bool rl)
    {
        if (rl) launch_cell_adjoint<Model, CMX_RUSH_LARSEN>(c, a);
        else launch_cell_adjoint<Model, CMX_FORWARD_EULER>(c, a);
    }
    static ModelOps make()
    {
        ModelOps o{};
        o.meta = {Model::NS, Model::NP, Model::initial_state(), Model::param_defaults(), Model::state_names(),
                  Model::param_names(), Model::param_aliases(), IMPLICIT};
        o.cells64 = &cells<double, double>;
        o.cells32 = &cells<float, float>;
        o.svi64[0] = &launch_svi_rows<Model, double, double, StencilRows<4>>;
        o.svi64[1] = &launch_svi_rows<Model, double, double, StencilRows<8>>;
        o.svi64[2] = &launch_svi_rows<Model, double, double, StencilRows<16>>;
        o.svi64[3] = &launch_svi_rows<Model, double, double, EllRows>;
        o.svi32[0] = &launch_svi_rows<Model, float, float, StencilRows<4>>;
        o.svi32[1] = &launch_svi_rows<Model, float, float, StencilRows<8>>;
        o.svi32[2] = &launch_svi_rows<Model, float, float, StencilRows<16>>;
        o.svi32[3] = &launch_svi_rows<Model, float, float, EllRows>;
        o.iionic64 = &launch_iionic_model<Model, double>;
        o.iionic32 = &launch_iionic_model<Model, float>;
        o.batch64 = &batch<double>;
        o.batch32 = &batch<float>;
        o.gate_table = &table;
        if constexpr (ADJOINT) o.adjoint = &adjoint;
        return o;
    }
};

}  // namespace cmx

// Defines const cmx::ModelOps* cmx::model_ops_<name>() for model type Model
// (implicit: backward-Euler and GRL forms; adjoint: the adjoint's cell step).
#define CMX_DEFINE_MODEL_OPS(name, Model, implicit, adjoint)                        \
    namespace cmx {                                                                   \
    const ModelOps* model_ops_##name()                                                \
    {                                                                                 \
        static const ModelOps ops = ModelImpl<Model, implicit, adjoint>::make();      \
        return &ops;                                                                  \
    }                                                                                 \
    }
