"""MI355X-native simulation core for Perceval's Naive, SLOS and
CliffordClifford2017 backends (HIP kernels behind a C ABI, loaded by ctypes)."""
__version__ = "0.1.0"
