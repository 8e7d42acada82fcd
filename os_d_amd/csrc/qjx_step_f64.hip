// qjx_step_f64.hip -- the FP64 step GEMM (gemm_f64 and its launch and occupancy
// helpers) in a translation unit of its own, built with the default instruction
// scheduler.  The rest of qjx_step.hip is built with memory clauses kept together,
// which speeds the FP32 and 3xBF16 GEMMs and slows this one.
#define QJX_STEP_F64 1
#include "qjx_step.hip"
