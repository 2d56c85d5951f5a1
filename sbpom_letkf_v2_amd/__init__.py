"""MI355X LETKF local-analysis engine (sbPOM-LETKF v2 da_main drop-in).

Modules:
  _lib     ctypes binding of libletkf_engine.so (include/letkf_engine.h)
  engine   da_main-style host mirror of the C ABI
  synth    synthetic analysis problems in the reference's module-state form
"""
__all__ = ["_lib", "engine", "synth"]
