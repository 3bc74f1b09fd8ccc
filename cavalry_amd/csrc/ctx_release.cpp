// Context teardown hook: releases state other subsystems hang off a context.
#include "capi_internal.h"

namespace cav {
void render_release(cav_ctx* ctx);
void ctx_release(cav_ctx* ctx) { render_release(ctx); }
}  // namespace cav
