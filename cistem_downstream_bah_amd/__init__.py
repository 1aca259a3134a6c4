"""MI355X (gfx950) template-matching inner loop with cisTEM's TemplateMatchingCore boundary.

The compute path is the HIP library ``libtm_hip.so`` (C ABI in include/tm_hip.h); ``core`` binds
it with ctypes, ``synth`` makes the seeded synthetic workloads and ``mrc`` reads/writes MRC files.
"""
__all__ = ["core", "synth", "mrc", "build"]
