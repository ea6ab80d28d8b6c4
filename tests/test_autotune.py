"""The variant tuner's cache: size-class keys, persistence, defaults (CPU),
and tuning on the device (GPU)."""

import json

import numpy as np
import pytest

import pararealml_amd as pml
from pararealml_amd import autotune


class _P:
    eq_name = "diffusion"
    ndim = 2
    plane = 257 * 257


@pytest.fixture
def fresh_cache(tmp_path, monkeypatch):
    path = tmp_path / "tune.json"
    monkeypatch.setenv("PML_AUTOTUNE_CACHE", str(path))
    monkeypatch.setattr(autotune, "_cache", None)
    yield path
    autotune._cache = None


def test_keys_follow_the_batch_size_class(fresh_cache):
    p = _P()
    single = autotune.key_for(p, 2, "float32c")
    batch = autotune.key_for(p, 2, "float32c", cells=p.plane * 128)
    assert single.endswith("/s") and batch.endswith("/xl")
    assert single != batch
    assert autotune.key_for(p, 0, "float64", cells=1 << 14).endswith("/xs")
    assert autotune.key_for(p, 0, "float64", cells=(1 << 14) + 1).endswith("/s")


NAMES = ["stream-a", "stream-b", "tile-c", "tile-d"]


def test_lookup_defaults_to_variant_zero_and_reads_the_cache_file(fresh_cache):
    p = _P()
    assert autotune.lookup(p, 2, "float32c", names=NAMES) == 0
    fresh_cache.write_text(json.dumps({autotune.key_for(p, 2, "float32c"): "tile-d"}))
    autotune._cache = None
    assert autotune.lookup(p, 2, "float32c", names=NAMES) == 3
    assert autotune.lookup(p, 2, "float32c", cells=p.plane * 128, names=NAMES) == 0


def test_stale_cache_entries_fall_back_to_the_default_variant(fresh_cache):
    p = _P()
    key = autotune.key_for(p, 2, "float32c")
    # a variant another build offered, and a bare index from an old cache
    fresh_cache.write_text(json.dumps({key: "pruned-variant"}))
    autotune._cache = None
    assert autotune.lookup(p, 2, "float32c", names=NAMES) == 0
    fresh_cache.write_text(json.dumps({key: 99}))
    autotune._cache = None
    assert autotune.lookup(p, 2, "float32c", names=NAMES) == 0
    # reordered variants: the name, not the position, selects the kernel
    fresh_cache.write_text(json.dumps({key: "tile-c"}))
    autotune._cache = None
    assert autotune.lookup(p, 2, "float32c", names=list(reversed(NAMES))) == 1


def test_lookup_resolves_names_against_the_loaded_library(fresh_cache):
    from pararealml_amd import _native as nat

    p = _P()
    names = autotune.variant_names(nat.PML_EQ["diffusion"], nat.PML_INT_RK4, nat.PML_F32C)
    assert len(names) >= 2 and len(set(names)) == len(names)
    key = autotune.key_for(p, nat.PML_INT_RK4, "float32c")
    fresh_cache.write_text(json.dumps({key: names[-1]}))
    autotune._cache = None
    assert autotune.lookup(p, nat.PML_INT_RK4, "float32c") == len(names) - 1


@pytest.mark.gpu
def test_operator_tuning_picks_a_variant_used_by_solve(fresh_cache):
    from pararealml_amd import _native as nat
    from pararealml_amd.operators.fdm import RK4, FDMOperator, ThreePointCentralDifferenceMethod

    print("case: autotune", flush=True)
    eq = pml.DiffusionEquation(2, 0.5)
    mesh = pml.Mesh([(0.0, 1.0), (0.0, 1.0)], [1.0 / 32, 1.0 / 32])
    bc = pml.DirichletBoundaryCondition(lambda x, t: np.zeros((len(x), 1)), is_static=True)
    cp = pml.ConstrainedProblem(eq, mesh, [(bc, bc), (bc, bc)])
    op = FDMOperator(RK4(), ThreePointCentralDifferenceMethod(), 1e-4, precision="float32")
    best = op.tune(cp, n_grids=3, n_steps=16)
    lib = nat.device_lib()
    n_cfg = lib.pml_fdm_configs(nat.PML_EQ["diffusion"], nat.PML_INT_RK4, nat.PML_F32)
    assert 0 <= best < n_cfg
    problem = op.device_problem(cp)
    assert op.tile_config_for(problem, 3) == best
    assert json.loads(fresh_cache.read_text())  # persisted
    ic = pml.GaussianInitialCondition(cp, [(np.array([0.5, 0.5]), 0.01 * np.eye(2))])
    ivp = pml.InitialValueProblem(cp, (0.0, 0.01), ic)
    tuned = op.solve(ivp).discrete_y()
    default = FDMOperator(RK4(), ThreePointCentralDifferenceMethod(), 1e-4, precision="float32",
                          tile_config=0).solve(ivp).discrete_y()
    np.testing.assert_array_equal(tuned, default)
