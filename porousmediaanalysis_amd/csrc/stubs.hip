// Stages not yet implemented on the GPU fail loudly.
#include "pma_internal.h"
namespace pma {
void run_pores(pma_ctx*) { throw Error(PMA_ERR_INTERNAL, "pma_pores: not implemented yet"); }
void run_remove_empty_pores(pma_ctx*) { throw Error(PMA_ERR_INTERNAL, "pma_remove_empty_pores: not implemented yet"); }

}
