// Luo-Rudy 1991 cells: every kernel of the model and its entry in the model table.
#include "cmx_model_impl.h"

CMX_DEFINE_MODEL_OPS(lr91, ModelLr91, true, true)
