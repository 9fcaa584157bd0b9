// dense.hip -- dense pair-tile layout for the all-pairs unit-square packing.
#include "engine_state.h"
#include "paradmm_gpu.h"

namespace pdm {

int dense_try_finalize(DevGraph*, graph*, const std::vector<int>&) { return -1; }
int dense_alloc_state(DevGraph*) { return PARADMM_ERR_STATE; }
int dense_iteration(DevGraph*, int) { return PARADMM_ERR_STATE; }
int dense_legacy_xpass(DevGraph*) { return PARADMM_ERR_STATE; }
int dense_legacy_zpass(DevGraph*) { return PARADMM_ERR_STATE; }
int dense_residual_reduce(DevGraph*) { return PARADMM_ERR_STATE; }
void dense_free(DevGraph*) {}

}  // namespace pdm
