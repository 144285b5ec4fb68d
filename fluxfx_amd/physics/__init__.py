"""Host-side parameter contract of the smoke step (mirrors the reference's
``physics`` package: settings dataclasses, timestep root, source packing)."""
