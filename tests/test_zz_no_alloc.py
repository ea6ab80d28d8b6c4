"""The allocation-free stepping entry point (pml_fdm_solve_noalloc, work-unit
tables in a static device array) against pml_fdm_solve, through the C ABI:
two calls in a row on one stream -- a ragged Cahn-Hilliard batch, then a
smaller batch -- each bitwise equal to the existing entry point's result
(same kernels), for the streaming two-step, single-step and LDS tile
variants; a batch whose tables exceed the array is refused before any
device work."""

import ctypes

import numpy as np
import pytest

from tests.helpers import cahn_hilliard, shallow_water

pytestmark = pytest.mark.gpu


def _batch(sizes, prec, integ, seed0=0, make=cahn_hilliard):
    from pararealml_amd import _device as dev

    problems, y0s = [], []
    for i, n in enumerate(sizes):
        cp, ic = make(n, seed=seed0 + i)
        problems.append(dev.device_problem(cp))
        y0s.append(ic.discrete_y_0(True))
    b = dev.FDMBatch(problems, list(range(len(sizes))), prec, integ, dev.solver_stream())
    b.load_host(y0s)
    return b


def _noalloc(b, d_t, n_steps, tile_config, flags=0):
    from pararealml_amd import _native as nat

    ptr = lambda t: None if t is None else t.data_ptr()  # noqa: E731
    return b.lib.pml_fdm_solve_noalloc(
        b.eq_id, b._params, len(b.params), b.integrator, b.prec, b.ndim,
        ctypes.byref(b.grids[0]), len(b.grid_problem), b.bc.data_ptr(), ptr(b.hi), ptr(b.lo),
        ptr(b.work_hi), ptr(b.work_lo), b.total, float(d_t), int(n_steps), None, None, 0,
        int(tile_config), flags, ctypes.c_void_p(b.stream.cuda_stream))


def _variant(b, name):
    names = [b.lib.pml_fdm_config_name(b.eq_id, b.integrator, b.prec, i).decode()
             for i in range(b.lib.pml_fdm_configs(b.eq_id, b.integrator, b.prec))]
    return names.index(name) if name in names else 0


@pytest.mark.parametrize("variant", ["stream2-rw8-pd2-s2", "stream-rw8-pd1-full",
                                     "tile-ox64-oy16-s1"])
def test_noalloc_entry_matches_pml_fdm_solve(variant):
    import torch

    from pararealml_amd import _native as nat

    print(f"case: no-alloc entry {variant}", flush=True)
    for call, (sizes, prec) in enumerate([((96, 64, 130, 257), "float32"),
                                          ((40, 33), "float32")]):
        ref = _batch(sizes, prec, nat.PML_INT_RK4, seed0=10 * call)
        got = _batch(sizes, prec, nat.PML_INT_RK4, seed0=10 * call)
        cfg = _variant(ref, variant)
        ref.advance(1e-4, 37, tile_config=cfg, use_graph=True)
        nat.check(_noalloc(got, 1e-4, 37, cfg))
        torch.cuda.synchronize()
        a = ref.state_f64().cpu().numpy()
        b = got.state_f64().cpu().numpy()
        assert np.all(np.isfinite(a))
        np.testing.assert_array_equal(b, a)
    # three fields, compensated FP32, the second call smaller again
    for call, sizes in enumerate([(65, 48), (33,)]):
        ref = _batch(sizes, "float32c", nat.PML_INT_RK4, seed0=call, make=shallow_water)
        got = _batch(sizes, "float32c", nat.PML_INT_RK4, seed0=call, make=shallow_water)
        ref.advance(1e-3, 11, use_graph=False)
        nat.check(_noalloc(got, 1e-3, 11, 0))
        torch.cuda.synchronize()
        np.testing.assert_array_equal(got.state_f64().cpu().numpy(),
                                      ref.state_f64().cpu().numpy())


def test_noalloc_entry_refuses_oversized_tables_and_prepare():
    from pararealml_amd import _native as nat

    print("case: no-alloc limits", flush=True)
    b = _batch((24,), "float32", nat.PML_INT_FE)
    lib = b.lib
    # a grid whose work-unit table alone exceeds the array: refused while
    # checking the tables, before any copy or launch
    big = (nat.PmlGrid * 1)()
    big[0].n0, big[0].n1, big[0].dx0, big[0].dx1 = 30000, 30000, 0.1, 0.1
    big[0].state_offset, big[0].bc_offset = 0, 0
    rc = lib.pml_fdm_solve_noalloc(
        b.eq_id, b._params, len(b.params), b.integrator, b.prec, b.ndim, big, 1,
        b.bc.data_ptr(), b.hi.data_ptr(), None, b.work_hi.data_ptr(), None,
        2 * 30000 * 30000, 1e-4, 4, None, None, 0, 0, 0, ctypes.c_void_p(b.stream.cuda_stream))
    assert rc == nat.PML_E_INVALID
    assert "bytes of work-unit tables" in lib.pml_last_error().decode()
    assert _noalloc(b, 1e-4, 4, 0, flags=nat.PML_FLAG_PREPARE) == nat.PML_E_INVALID
