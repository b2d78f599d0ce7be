// Mahajan 2008 rabbit ventricular cells: every kernel of the model and its entry in the model table.
#include "cmx_model_impl.h"

CMX_DEFINE_MODEL_OPS(mahajan2008, ModelMahajan2008, true, true)
