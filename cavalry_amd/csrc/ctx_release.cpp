// Context teardown hook: releases state other subsystems hang off a context.
#include "capi_internal.h"

namespace cav {
void ctx_release(cav_ctx* ctx) { ctx->render = nullptr; }
}  // namespace cav
