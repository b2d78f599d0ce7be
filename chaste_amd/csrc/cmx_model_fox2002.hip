// Fox 2002 canine ventricular cells: every kernel of the model and its entry in the model table.
#include "cmx_model_impl.h"

CMX_DEFINE_MODEL_OPS(fox2002, ModelFox2002, true, true)
