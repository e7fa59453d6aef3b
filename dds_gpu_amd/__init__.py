"""ddx: exact double-dummy solver for the AMD Instinct MI355X (gfx950).

The solver is native code (include/ddx.h, libddx.so); this package holds its build script and
a ctypes mirror of the C ABI (dds_gpu_amd.ddx).
"""
