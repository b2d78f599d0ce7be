// Maleckar 2008 human atrial cells: every kernel of the model and its entry in the model table.
#include "cmx_model_impl.h"

CMX_DEFINE_MODEL_OPS(maleckar2008, ModelMaleckar2008, true)
