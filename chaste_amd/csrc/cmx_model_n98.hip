// Noble-Varghese-Kohl-Noble 1998 guinea-pig ventricular cells: every kernel of the model and its entry in the
// model table.
#include "cmx_model_impl.h"

CMX_DEFINE_MODEL_OPS(n98, ModelN98, true, true)
