"""gsr: MI355X (gfx950) differentiable Gaussian-splat rasterizer, Python host side.

``gsr.lib`` binds libgsr.so (include/gsr.h); ``gsr.renderer`` mirrors the
renderer object of trainer/gpu.js; ``gsr.scenes`` builds the seeded workloads.
"""
