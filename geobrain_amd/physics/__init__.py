"""Physics namespace mirroring ``geobrain.physics`` (wave propagation only)."""

from . import wave  # noqa: F401
