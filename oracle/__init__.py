"""CPU oracle for the DiReBM step engine -- TEST INFRASTRUCTURE ONLY.

This package restates, on the CPU, the algorithms of the reference (berfandil/DiReBM) that the
HIP engine in ``direbm_amd`` replaces.  It is the checker, never the product: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may import it, and the
product path (``direbm_amd``) never calls into it.

Modules
-------
``constants``  D2Q7 / D3Q13-ico velocity sets and defaults (direbm/constants.py, lattices.py).
``physics``    FP64 recover / equilibrium / BGK collide (direbm/physics.py:16-55).
``v2``         FP64 restatement of the GPU (Warp) step, direbm/warp/{physics,propagation,
               simulator}.py, with every *discrete* decision (cell keys, radius predicates)
               evaluated in the same fp32 arithmetic the reference kernels use, so control-point
               selection is bit-exact; continuous sums are FP64.
``lbm``        D2Q7 hex LBM (direbm/lbm.py:26-66).
``fields``     bin_fields (direbm/fields.py:17-41).
``v1``         the per-object v1 reference solver (direbm/reference/simulator.py, grid.py),
               used as the timed "direbm/reference" CPU baseline and for the macroscopic
               cross-check.

Pinning (how each function is checked, in ``tests/test_oracle_*.py``):

* ``physics``: closed forms and invariants (mass / momentum conservation, f_eq recovers its own
  moments, rest equilibrium is rho*W) -- the reference's tests/test_physics.py known answers.
* ``v2.disperse``: the reference's closed form (tests/test_warp_propagation.py:43-52).
* ``v2.thin``: the reference's own np.unique restatement (tests/test_warp_propagation.py:27-36)
  plus a dict-based brute force and exact count / membership invariants.
* ``v2.refine``: the reference's hand-built ring known answers
  (tests/test_warp_propagation.py:79-112) plus an O(P*nc) pure-Python brute force.
* ``v2.resample``: O(nc*P) pure-Python brute force, mass balance of the scatter (sum of control
  point f equals delivered component mass), isolated-point gap-fill closed form (f = f_rest).
* ``lbm``: exact mass conservation, rest steady state, single-direction translation by one node.
* ``fields``: uniform-density / uniform-velocity known answers (tests/test_fields.py).
* ``v1``: golden vectors written by tests/golden/make_v1_golden.py from the reference package.
"""
