"""MI355X-native differentiable Gaussian path tracer (drop-in for PTIR's threedgptir_tracer).

Public surface:
  Tracer, HipTracer, LightBuffers   -- the reference's plugin/operator API (tracer.py)
  envmap                            -- environment importance tables
  scenes                            -- seeded synthetic workloads
"""

from .tracer import HipTracer, LightBuffers, Tracer  # noqa: F401

__version__ = "0.1.0"
