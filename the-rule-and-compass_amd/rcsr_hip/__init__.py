"""rcsr device arm for AMD Instinct MI355X (gfx950).

The product is librcsrhip.so, a C ABI (include/rcsr_hip.h) that is a drop-in
for the reference engine's librcsrcuda.so on the batched bytecode-evaluation
and batched LM/IRLS constant-fitting path.  This package is its host-side
binding: `lib` (ctypes), `host` (the reference host interface: compile_program,
gpu_fit_plan, GpuEvaluator, BatchedGpu, fit_batch_verified), `expr` (trees)
and `workloads` (seeded synthetic configurations).
"""
from . import expr, lib  # noqa: F401
