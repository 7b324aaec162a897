"""MI355X-native SG-MCMC hot path: jax-sgmc's stochastic potential, its gradient, the
langevin_diffusion / friction_leapfrog / obabo steps and the rms_prop preconditioner,
behind the reference's alias / potential call signatures, on hand-written HIP kernels
for gfx950 (libsgmc.so, built from csrc/; see include/sgmc.h for the C-ABI).

The directory name is not a Python identifier, so the package is imported as
``jax_sgmc_amd`` through ``load()`` in ``__graft_entry__.py`` at the repository root.
"""
__version__ = "0.1.0"
