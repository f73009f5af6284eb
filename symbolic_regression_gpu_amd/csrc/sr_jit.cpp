// sr_jit.cpp -- hipRTC backend (filled in below).
#include "sr_ctx.h"

namespace sr {
int eval_jit(Ctx &c, const sr_population &pop, float *mse_out, int64_t *valid_out, sr_stats *st)
{
    (void)pop;
    (void)mse_out;
    (void)valid_out;
    (void)st;
    return c.fail(SR_ERR_UNSUPPORTED, "jit mode not built yet");
}
}  // namespace sr
