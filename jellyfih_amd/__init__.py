"""jellyfih_amd — MI355X-native batched 2D MLS-MPM engine (HIP, gfx950).

Drop-in for the simulation core of c-m-x-r/jellyfih (``mpm_sim.py``):

* :mod:`jellyfih_amd.mpm_sim` — the reference module's names backed by the HIP engine;
* :class:`jellyfih_amd.engine.Engine` — the engine itself over the C ABI in
  ``include/jellymx.h`` (library ``libjellymx.so``, built by :mod:`jellyfih_amd.build`);
* :mod:`jellyfih_amd.phenotype` — the particle-set generator the engine is fed with.
"""

__all__ = ["engine", "mpm_sim", "phenotype", "build"]
