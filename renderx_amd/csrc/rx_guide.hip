// Path-guiding network (placeholder until the kernels land).
#include "rx_context.h"

extern "C" {
void rx_guide_state_free(rx_ctx* c) { (void)c; }
int rx_guide_init(rx_ctx* c, const rx_guide_params*) { return rx::set_err(c, RX_ERR_STATE, "not built yet"); }
}
