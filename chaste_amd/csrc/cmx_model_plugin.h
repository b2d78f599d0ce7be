// A cell model built as a plugin of its own (the counterpart of the
// reference's dynamically loaded cell models, heart/src/odes/DynamicCellModelLoader.cpp:55-63):
// a shared library holding one generated model's kernels and its ModelOps
// entry, loaded at run time by cmx_load_model.  Its source is
//
//     #include "cmx_model_plugin.h"
//     #include "<generated header>"          // namespace cmx::<ns>
//     CMX_MODEL_PLUGIN(<ns>)
//
// compiled for gfx950 against these headers and linked to libcmx.so
// (chaste_amd.build.build_model_plugin writes and builds it from a CellML file).
#pragma once
#include "cmx_model_impl.h"

#define CMX_MODEL_PLUGIN(ns)                                                                     \
    namespace cmx {                                                                              \
    CMX_GENERATED_MODEL(ModelPlugin_##ns, ns)                                                    \
    }                                                                                            \
    extern "C" __attribute__((visibility("default"))) uint32_t cmx_model_plugin_abi()            \
    {                                                                                            \
        return cmx::kModelPluginAbi;                                                             \
    }                                                                                            \
    extern "C" __attribute__((visibility("default"))) const cmx::ModelOps* cmx_model_plugin_ops() \
    {                                                                                            \
        static const cmx::ModelOps ops = cmx::ModelImpl<cmx::ModelPlugin_##ns, true>::make();   \
        return &ops;                                                                             \
    }
