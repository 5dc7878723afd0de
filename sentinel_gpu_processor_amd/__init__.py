"""MI355X-native interferogram formation for Sentinel-1 TOPS bursts.

Hot path (HIP kernels in ``csrc/``, C ABI in ``include/s1ifg.h``):
geo2rdr offsets from orbit + DEM -> TOPS deramp -> resampling -> reramp and
range flattening -> crossmul with multilook -> Goldstein filter ->
coherence.  Python drop-in modules keep the call signatures of the
reference's ``s1proc.geometry``, ``s1proc.interfere`` and
``s1proc.coherence``.
"""

__version__ = "0.1.0"
