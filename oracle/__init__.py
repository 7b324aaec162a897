"""CPU oracle for the SG-MCMC hot path — TEST INFRASTRUCTURE ONLY.

This package restates, on the CPU and in numpy, the arithmetic that the
reference (jax-sgmc 0.1.5 on JAX 0.4.x) performs on its hot path:

- ``jaxrandom``   — jax.random's legacy Threefry-2x32 derivation (PRNGKey, split,
                    random_bits, uniform, normal, randint) as the reference calls it
                    (integrator.py:119-135, data/numpy_loader.py:128-141,
                    solver.py:254, 283-284, 508, 528).
- ``model``       — the likelihood families (linear-Gaussian, logistic, MLP
                    softmax-CE), the stochastic potential (potential.py:159-214)
                    and the full potential (potential.py:219-293).
- ``steps``       — langevin_diffusion (integrator.py:859-922), rms_prop
                    (adaption.py:225-293), friction_leapfrog (integrator.py:603-757),
                    obabo (integrator.py:177-273), the polynomial step-size schedule
                    (scheduler.py:447-542), the reSGLD swap (solver.py:264-293) and
                    the SGGMC Metropolis-Hastings step (solver.py:500-567).
- ``posterior``   — closed-form Gaussian posterior of Bayesian linear regression and
                    the exact stationary covariance of SGLD at fixed step size.
- ``cport``       — a ctypes handle on ``oracle/_build/liboracle_port.so``, a C port of
                    the multi-chain SGHMC logistic-regression loop, used as the timed
                    CPU baseline.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg may
import anything from here, and only as the checker: the product package never does.

Parity status of each function is stated in its docstring; a function whose result
no golden vector, known answer or mathematical identity checks says
"parity unpinned".
"""
