"""MI355X-native operator-apply and column-solve path of LFRic Core.

Host-side modules (numpy, no GPU needed): ``rng``, ``mesh``,
``function_space``.  Device path (needs liblfr.so and a HIP GPU): ``api``,
``solvers``, ``autograd``.
"""

__all__ = ["rng", "mesh", "function_space"]
