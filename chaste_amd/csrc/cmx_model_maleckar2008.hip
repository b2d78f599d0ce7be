// Maleckar 2008 human atrial cells: every kernel of the model and its entry in the model table.
#include "cmx_model_impl.h"

// No adjoint: on a stimulated slab its gradient disagreed with central
// differences (the forward-mode derivatives of a single cell with respect to
// the dyadic calcium Ca_d grow past 1e60 through a strongly driven upstroke).

CMX_DEFINE_MODEL_OPS(maleckar2008, ModelMaleckar2008, true, false)
