"""FP64 CPU oracle for the RadioBEAR absorption + ray-integration path.

TEST INFRASTRUCTURE ONLY.  Nothing in ``radiobear_amd`` imports, calls or
links anything under this directory: only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` use it, and only as the checker
(or the timed CPU baseline), never as the product path.

Each function restates the reference algorithm (RadioBEAR 2.0.1) in numpy
float64 and cites the reference ``file:line`` it follows.  Modules:

* ``absorbers``  -- the constituent plug-ins (nh3_hs_sjs, nh3_hs, nh3_sjs,
  h2o_bk, h2s_ddb, ph3_jh, h2_jj_ddb) and line-list truncation.
* ``layers``     -- ``Alpha.get_layers`` / ``total_layer_alpha`` / scales.
* ``geometry``   -- ``raypath.compute_ds`` / ``findEdge`` / ``Shape._calcEllipse``.
* ``rt``         -- ``Brightness.single`` (point and disc mode).
* ``atmosphere`` -- the file-atmosphere pipeline used by the reference's
  benchmark (``readGas`` / ``regrid`` / ``JupiterNoTweak`` / ``computeProp``).
* ``jacobian``   -- float64 torch-autograd restatement of the forward model,
  the gradient reference for the adjoint kernels.

How each function is pinned (golden vectors, closed forms, reference-generated
fixtures) is listed in the tests that exercise it; a function with no such
check would be marked "parity unpinned" here.  None currently is.
"""
