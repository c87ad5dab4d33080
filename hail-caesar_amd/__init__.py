"""MI355X-native hot path of the HAIL-CAESAR / CAESAR-Lisflood catchment model.

The directory name carries a hyphen, so it is loaded by path and registered
as ``hail_caesar_amd`` (see ``load_package`` in bench.py / __graft_entry__.py).
Submodules:
  inputs   seeded synthetic inputs in the reference's file formats
  hc       ctypes binding of the C-ABI library libhc.so (include/hc.h)
"""
