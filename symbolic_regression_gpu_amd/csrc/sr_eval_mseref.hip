// Batched interpreter instantiations for output mode MODE_MSE_REF (MSE of
// skeletons that read cached subtree columns).
#define SR_MODE MODE_MSE_REF
#define SR_LAUNCH_NAME launch_batched_mseref
#define SR_OCC_NAME batched_occupancy_mseref
#include "sr_eval_mode.inc"
