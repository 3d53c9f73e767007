"""MI355X-native body-of-revolution MoM hot path (fill -> per-mode solve -> far field).

Modules:
  bor_solver     solve_bor / generatrix helpers (GPU drop-in for the reference's bor_solver)
  bor_streaming  the bor_stream_kernel.c sampler entry points, HIP-backed
  _lib           ctypes binding of libghost_bor.so (built in-tree)
"""
