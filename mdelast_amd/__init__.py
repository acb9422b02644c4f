"""mdelast-mi355x: the per-step EAM path of mdelast on MI355X (gfx950).

Modules:
  md_driver_module   the reference's pybind11 API, backed by libmdelast_hip.so
  get_potential      SETFL loaders returning driver objects
  setfl              SETFL parsing to CGS node arrays
  units              the driver's CGS constants
  build              hipcc build of libmdelast_hip.so
"""

__all__ = ["md_driver_module", "get_potential", "setfl", "units", "build"]
