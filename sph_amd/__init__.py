"""sph_amd: gfx950 SPH solver hot path (WCSPH / PCISPH / DFSPH with Akinci
boundaries) behind the C ABI of libsphg (include/sphg.h)."""
from .sphg import (  # noqa: F401
    DFSPH, PCISPH, WCSPH, Context, Scene, SceneParams, SphgError, load, sample_box, sample_sphere,
    fluid_block, fluid_spheres,
)
