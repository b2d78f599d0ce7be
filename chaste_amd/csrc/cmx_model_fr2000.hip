// Faber-Rudy 2000 ventricular cells: every kernel of the model and its entry in the model table.
#include "cmx_model_impl.h"

CMX_DEFINE_MODEL_OPS(fr2000, ModelFr2000, true, true)
