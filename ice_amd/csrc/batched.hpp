// Batched path: every block and family of a stage in one persistent launch
// fed from a tile work queue, the stage sequence captured in a hipGraph.
#pragma once
struct ice_gpu_ctx;
struct BatchedPlan;
int batched_step(ice_gpu_ctx* c);
void batched_destroy(ice_gpu_ctx* c);
