// Hodgkin-Huxley 1952 cells: every kernel of the model and its entry in the model table.
#include "cmx_model_impl.h"

CMX_DEFINE_MODEL_OPS(hh52, ModelHh52, true, true)
