"""MI355X-native spectral path tracer with the scene API of
Xyzonox/taichi-spectral-render.

Modules:
  data     -- spectral table loaders (data.py formats), analytic CMFs
  scene    -- scene dict -> flat arrays (setup_scene semantics)
  configs  -- frozen synthetic benchmark scenes
  render   -- Renderer: the reference's render protocol on the HIP library
  _lib     -- ctypes binding of libspectral_render.so
"""
from .data import ColourSystem, EEMMatrix, RefractiveIndex, SpecularVec, srgb_colour_system  # noqa: F401
from .scene import FlatScene, flatten_scene, scene_dimensions  # noqa: F401


def __getattr__(name):
    # Renderer pulls in torch and the HIP library; load it on first use.
    if name == "Renderer":
        from .render import Renderer

        return Renderer
    raise AttributeError(name)
