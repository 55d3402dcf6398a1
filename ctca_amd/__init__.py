"""MI355X particle hot path of EMSES (push, deposit, zigzag current, sort).

libemgpu.so (ctca_amd/csrc, C ABI in include/emgpu.h) holds the HIP kernels;
emg.py binds it, units.py holds the workloads and the reference's unit
conversions, fieldsolve.py and driver.py are the host side of a run.
"""
