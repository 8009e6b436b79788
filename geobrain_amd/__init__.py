"""geobrain_amd -- MI355X-native 2D acoustic propagator with GeoBrain's operator API.

``geobrain_amd.physics.wave`` mirrors ``geobrain.physics.wave`` for the 2D
acoustic path (``AcousticPropagator``, ``acoustic_forward_kernel`` and the
model/survey/boundary/wavelet types they take); the time stepping and its
adjoint run in hand-written gfx950 HIP kernels (``libgbw_acoustic2d.so``).
"""

__version__ = "0.1.0"
