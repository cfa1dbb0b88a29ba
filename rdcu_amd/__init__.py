"""rdcu_amd -- MI355X (gfx950) integrator for rdcu's explicit path.

Dormand-Prince 4/5 with pseudospectral derivatives of a monomial dictionary
RHS, behind the C ABI of include/rdcu_dp45.h (librdcu_hip.so) and the drop-in
``rdcu`` command line (rdcu_amd/bin/rdcu).  The HIP library is required:
nothing here falls back to a CPU path.
"""
from .abi import (HIP_LIB, Batch, CallbackIntegrator, Integrator, Problem, RdcuError, Term,
                  glibc_rand_ic, hip_lib, load, loss_grad_batch)
from .dictionaries import cgle_terms, fhn_terms, to_cli

__all__ = ["HIP_LIB", "Batch", "CallbackIntegrator", "Integrator", "Problem", "RdcuError", "Term", "glibc_rand_ic", "hip_lib",
           "load", "loss_grad_batch", "cgle_terms", "fhn_terms", "to_cli"]
