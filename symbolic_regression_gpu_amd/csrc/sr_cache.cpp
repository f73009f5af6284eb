// sr_cache.cpp -- cross-population subtree cache (filled in below).
#include "sr_ctx.h"

namespace sr {
int eval_cached(Ctx &c, const sr_population &pop, sr_stats *st)
{
    (void)pop;
    (void)st;
    return c.fail(SR_ERR_UNSUPPORTED, "cached mode not built yet");
}
}  // namespace sr
