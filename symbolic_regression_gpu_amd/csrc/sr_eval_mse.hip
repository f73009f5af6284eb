// Batched interpreter instantiations for output mode MODE_MSE.
#define SR_MODE MODE_MSE
#define SR_LAUNCH_NAME launch_batched_mse
#define SR_OCC_NAME batched_occupancy_mse
#include "sr_eval_mode.inc"
