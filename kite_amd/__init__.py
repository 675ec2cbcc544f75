"""kite_amd -- MI355X-native (gfx950) exact k-mer counting engine.

Python mirror of the C ABI in include/kite_gpu.h, used by the tests and the
benchmark.  The command-line drop-in for `kite count|dump|merge|histo|
analyze|info` is the C++ host binary kite_amd/lib/kite.
"""
from ._native import (BATCH_PIPELINED, BATCH_PLAIN, BATCH_SERIAL, CountResult, CountStats, GpuContext, KiteError,
                      PinnedArray, SynthSpec, gpu_lib_path, load_gpu, synth_genome, synth_reads)

__all__ = ["BATCH_PIPELINED", "BATCH_PLAIN", "BATCH_SERIAL", "CountResult", "CountStats", "GpuContext", "KiteError",
           "PinnedArray", "SynthSpec", "gpu_lib_path", "load_gpu", "synth_genome", "synth_reads"]
__version__ = "0.1.0"
