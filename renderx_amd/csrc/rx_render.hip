// Scene loading and the wavefront path tracer (placeholder until the kernels land).
#include "rx_context.h"

extern "C" {
void rx_render_state_free(rx_ctx* c) { (void)c; }
int rx_load_scene(rx_ctx* c, uint32_t, const float*, uint32_t, const uint32_t*, uint32_t, const rx_object_desc*,
                  uint32_t) {
  return rx::set_err(c, RX_ERR_STATE, "not built yet");
}
int rx_render(rx_ctx* c, const rx_camera_desc*, const rx_render_params*, float*, double*, double*,
              rx_render_stats*) {
  return rx::set_err(c, RX_ERR_STATE, "not built yet");
}
}
