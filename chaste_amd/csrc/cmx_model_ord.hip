// O'Hara-Rudy 2011 endocardial cells: every kernel of the model and its entry in the model table.
#include "cmx_model_impl.h"

CMX_DEFINE_MODEL_OPS(ord, ModelOrd, false, false)
