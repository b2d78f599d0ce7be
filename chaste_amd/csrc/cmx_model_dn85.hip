// DiFrancesco-Noble 1985 Purkinje cells: every kernel of the model and its entry in the model table.
#include "cmx_model_impl.h"

CMX_DEFINE_MODEL_OPS(dn85, ModelDn85, true, true)
