"""renderx_amd: MI355X-native path-tracing hot path of RenderX."""
