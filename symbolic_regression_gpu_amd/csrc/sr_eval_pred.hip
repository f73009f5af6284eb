// Batched interpreter instantiations for output mode MODE_PRED.
#define SR_MODE MODE_PRED
#define SR_LAUNCH_NAME launch_batched_pred
#define SR_OCC_NAME batched_occupancy_pred
#include "sr_eval_mode.inc"
