"""MI355X-native explicit solid-mechanics step with the superelastic SMA law
(Alya solidz stress model 160), behind a C ABI (include/alya_sma.h).

  alya_sma_amd.solid     host mirror of the solidz operator surface
  alya_sma_amd.problems  seeded synthetic problems (bar, cantilever, ...)
  alya_sma_amd.build     in-tree build of the HIP library and Fortran host
"""
from ._lib import HipError, lib  # noqa: F401
