"""ParFlex (chunked FlexDTW) on AMD MI355X (gfx950).

Drop-in modules mirroring the reference's GPU interface:
    parflexdtw_amd.gpu_flexdtw   <- gpu_flexdtw.py
    parflexdtw_amd.Parflex_gpu   <- Parflex_gpu.py
The kernels live in libparflex_hip.so (C ABI in include/parflex_hip.h).
"""
__version__ = "0.1.0"
