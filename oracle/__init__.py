"""CPU oracle for the FDTD time-stepping path -- TEST INFRASTRUCTURE ONLY.

A restatement, in PyTorch eager on the CPU, of the reference algorithm for the
hot path (2D TM/TE and 3D Yee leapfrog, CPML recursions, soft sources, DFT
flux/field accumulation, the checkpointed gradient).  Every function names the
reference file:line it follows.  The restatement keeps the reference's
floating-point operation order so that an FP32 run rounds exactly like the
reference's own FP32 PyTorch stepper; that is what the golden fixtures under
``tests/golden`` check.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker or the reported
CPU baseline.  The product (``meep_torch_amd``) never imports it: the engine
runs on the HIP kernels or raises.

Pinning status per function (what fails if a term, sign or index is wrong):
  cpml_profiles       golden vectors (pml.npz, bit-exact) + closed-form checks
  step_tm / step_te   golden trajectories (tm_small_*, te_small_*, c1_*), the
                      Yee energy invariant, TE/TM duality
  flux / DFT monitors golden accumulators + closed-form DFT of a cosine
  step_3d / lean      golden trajectories (d3_small_*)
  tm_gradient         golden autograd gradients (grad_tm_*) + central FD
  filters             golden vectors (filters.npz)
"""
