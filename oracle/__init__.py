"""CPU oracle of StructSplat's render path.  TEST INFRASTRUCTURE ONLY.

Nothing in the product package (``structsplat_amd``) imports this package.  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may use it, and only as the
expected results that the HIP path is compared with (or, for ``cpu_baseline``, as the timed CPU
restatement of the reference path).

Modules and what pins each of them (the tests named here live in ``tests/``):

* ``reference_render`` - torch restatement of the reference renderer ``render.py:25-263``
  (``_tile_bounds``, ``_flat_tile_slices``, ``_tile_coords``, ``_support_weight``,
  ``_accumulate``, ``render``, ``render_additive``, ``gaussian_activity``).  Pinned by the
  reference's own closed-form and behavioural tests restated in ``test_oracle_render.py``
  (``test_render.py:17-133, 535-593``; ``test_normalization_epsilon.py:39-88``): the weak-support
  closed form ``c w / (w + eps)``, the dense O(N H W) naive renderer at sigma_cutoff 6, the fade
  zero at exactly 3 sigma, empty fields, eps validation, slice-size invariance.
* ``tile_index`` - NumPy integer restatement of the tile index (``cuda_render.py:177-226``,
  ``render_ext.cu:13-218``): pinned by the reference golden ``offsets=[0,1,1,1,1], gids=[0]``
  (``test_render.py:167-172``), a brute-force loop over every (tile, Gaussian) pair, and, for the
  ellipse cull, the exact-zero property it claims (every culled pair has zero faded weight on
  every pixel of its tile, checked in float64).
* ``exact64`` - NumPy float64 forward and analytic backward (the formulas of
  ``render_ext.cu:485-530`` / ``:784-818``).  The forward is pinned against ``reference_render``
  in float64 (independent implementation: ``bincount`` instead of ``index_add``, no slices); the
  backward against float64 autograd of ``reference_render`` and against float64 central finite
  differences with the support rectangles frozen at the base point.
* ``synth`` - seeded synthetic fields of the benchmark configurations (no reference counterpart;
  the recipes are the benchmark definitions).
"""
