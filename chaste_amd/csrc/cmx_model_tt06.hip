// ten Tusscher 2006 epicardial cells: every kernel of the model and its entry in the model table.
#include "cmx_model_impl.h"

CMX_DEFINE_MODEL_OPS(tt06, ModelTt06, true, true)
