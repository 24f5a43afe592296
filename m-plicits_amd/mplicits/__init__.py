"""M-plicits hot path on MI355X (gfx950): drop-in for the reference's i3d model,
extraction, tracing and residual-training entry points, computed by the HIP
library libmplicits_hip.so (C ABI: include/mplicits.h)."""

__version__ = "0.1.0"
