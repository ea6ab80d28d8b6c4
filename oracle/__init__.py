"""CPU oracle: a NumPy restatement of the reference FDM + Parareal path.

TEST INFRASTRUCTURE ONLY.  Nothing in ``pararealml_amd`` imports, calls or
links anything in this package.  Only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` use it, and only as the checker
(or, for ``bench.py``, as the timed CPU baseline), never as the thing that is
measured or shipped.

Each function cites the reference file:line it restates.  Pinning status per
function (what independent evidence checks it) is listed in the module
docstring of :mod:`oracle.fdm_numpy`.
"""
