// Batched interpreter instantiations for output mode MODE_COLUMN.
#define SR_MODE MODE_COLUMN
#define SR_LAUNCH_NAME launch_batched_col
#define SR_OCC_NAME batched_occupancy_col
#include "sr_eval_mode.inc"
