"""MI355X-native RadioBEAR absorption and ray integration.

The host-side classes mirror the reference's call surface
(``Planet.run``, ``Alpha.get_layers``, ``Brightness.single``,
``raypath.compute_ds`` and the constituent plug-ins); the arithmetic runs in
hand-written HIP kernels for gfx950 behind the C-ABI library ``librbx.so``.
"""
__version__ = '0.1.0'
