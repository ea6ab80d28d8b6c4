"""The SymPy -> device program compiler (CPU): every reference equation's
right-hand sides, compiled and run by the host restatement of the
interpreter's instruction semantics, equal sympy.lambdify's NumPy values
(what the reference evaluates, operators/symbol_mapper.py:222-253) to
rounding; operand indices stay within the declared counts; registers are
reused; unsupported constructs are refused.  The library's validation of
programs is checked without a GPU where it needs no device memory."""

import ctypes

import numpy as np
import pytest
import sympy as sp

import pararealml_amd as pml
from pararealml_amd.operators import device_program as dp


def _equations():
    return {
        "population": pml.PopulationGrowthEquation(0.02),
        "lotka_volterra": pml.LotkaVolterraEquation(),
        "lorenz": pml.LorenzEquation(),
        "sir": pml.SIREquation(),
        "van_der_pol": pml.VanDerPolEquation(2.0),
        "n_body_2d": pml.NBodyGravitationalEquation(2, [5.0, 3.0, 1.0], g=1.0),
        "n_body_3d": pml.NBodyGravitationalEquation(3, [5e4, 5e8, 2e6, 7e5], g=6.6743e-11),
        "diffusion": pml.DiffusionEquation(2, 0.7),
        "convection_diffusion": pml.ConvectionDiffusionEquation(2, [0.3, -0.2], 0.05),
        "wave": pml.WaveEquation(2, 1.2),
        "cahn_hilliard": pml.CahnHilliardEquation(2, 0.1, 0.01),
        "burgers": pml.BurgersEquation(2, 50.0),
        "shallow_water": pml.ShallowWaterEquation(0.5, b=0.02, v=0.05, f=0.3, g=9.81),
        "navier_stokes": pml.NavierStokesEquation(100.0),
    }


def _compile_all(eq):
    exprs = [sp.sympify(e) for e in eq.symbolic_equation_system.rhs]
    syms = sorted(set().union(*[e.free_symbols for e in exprs]), key=lambda s: s.name)
    inputs = [s for s in syms if s.name != "t"]
    index = {s: k for k, s in enumerate(inputs)}
    return exprs, syms, inputs, dp.compile_expressions(exprs, index.__getitem__, inputs)


@pytest.mark.parametrize("name", sorted(_equations()))
def test_programs_match_lambdify(name):
    eq = _equations()[name]
    exprs, syms, inputs, prog = _compile_all(eq)
    rng = np.random.RandomState(len(name))
    values = {s: rng.uniform(0.5, 2.0, 37) for s in inputs}
    t = 0.37
    fn = sp.lambdify([syms], exprs, "numpy")
    want = fn([t if s.name == "t" else values[s] for s in syms])
    got = dp.evaluate_on_host(prog, [values[s] for s in inputs], t)
    assert len(got) == len(exprs)
    for g, w in zip(got, want):
        w = np.broadcast_to(w, (37,))
        # operation order may differ from lambdify's printer: rounding-level
        # differences, relative to the size of the result
        np.testing.assert_allclose(np.broadcast_to(g, (37,)), w, rtol=1e-13,
                                   atol=1e-13 * float(np.max(np.abs(w))))
    code = prog.code
    assert code.shape[1] == 4 and prog.n_regs <= dp.MAX_REGS
    ops = code[:, 0]
    regs = np.concatenate([code[ops != dp.OP["STORE"], 1], code[ops >= dp.OP["ADD"], 2]])
    assert regs.min() >= 0 and regs.max() < prog.n_regs
    assert np.all(code[ops == dp.OP["LOAD"], 2] < len(inputs))
    assert np.all(code[ops == dp.OP["CONST"], 2] < prog.consts.size)
    assert sorted(code[ops == dp.OP["STORE"], 1].tolist()) == list(range(len(exprs)))


def test_ode_program_inputs_are_the_state():
    eq = pml.NBodyGravitationalEquation(3, [1.0, 2.0, 3.0, 4.0, 5.0], g=1.0)
    prog = dp.ode_program(eq)
    assert prog.inputs == list(eq.symbols.y) and prog.n_outputs == eq.y_dimension
    # many more values than registers: registers are reused
    assert prog.code.shape[0] > 4 * prog.n_regs


def test_user_expressions_with_t_x_functions_and_powers():
    t, x0, x1, y0 = sp.symbols("t x_0 x_1 y_0")
    exprs = [
        t / 100.0 * (x0 + x1) ** 2,
        sp.sin(x0) * sp.exp(-t) - sp.sqrt(sp.Abs(y0)) / (1 + x1**2),
        sp.Max(x0, x1) - sp.Min(y0, 1.0) + sp.atan2(x1, x0) + sp.tanh(y0) ** -3,
        y0 ** 1.5 + sp.log(x0) * sp.cosh(x1) - sp.sign(x0 - 1.0),
        sp.Integer(3),
    ]
    inputs = [x0, x1, y0]
    prog = dp.compile_expressions(exprs, {s: k for k, s in enumerate(inputs)}.__getitem__,
                                  inputs)
    rng = np.random.RandomState(0)
    vals = [rng.uniform(0.2, 3.0, 25) for _ in inputs]
    got = dp.evaluate_on_host(prog, vals, 1.7)
    want = sp.lambdify([[t] + inputs], exprs, "numpy")([1.7] + vals)
    for g, w in zip(got, want):
        np.testing.assert_allclose(np.broadcast_to(g, (25,)), np.broadcast_to(w, (25,)),
                                   rtol=1e-13)


def test_unsupported_constructs_are_refused():
    x = sp.Symbol("x_0")
    with pytest.raises(NotImplementedError):
        dp.compile_expressions([sp.Piecewise((x, x > 0), (0, True))], lambda s: 0, [x])
    with pytest.raises(NotImplementedError):
        dp.compile_expressions([sp.gamma(x)], lambda s: 0, [x])
    with pytest.raises(NotImplementedError):
        dp.compile_expressions([sp.zoo * x], lambda s: 0, [x])


def test_library_rejects_out_of_range_operands():
    from pararealml_amd import _native as nat

    lib = nat.load_library()
    out = ctypes.c_void_p()

    def create(code, n_consts=1, n_regs=2, n_in=1, n_out=1):
        code = np.ascontiguousarray(np.array(code, dtype=np.int32).reshape(-1, 4))
        consts = np.zeros(max(n_consts, 1))
        return lib.pml_program_create(code.ctypes.data_as(ctypes.c_void_p), code.shape[0],
                                      consts.ctypes.data_as(ctypes.c_void_p), n_consts,
                                      n_regs, n_in, n_out, ctypes.byref(out))

    bad = [
        [[dp.OP["LOAD"], 0, 1, 0], [dp.OP["STORE"], 0, 0, 0]],  # input 1 of 1
        [[dp.OP["CONST"], 0, 1, 0], [dp.OP["STORE"], 0, 0, 0]],  # constant 1 of 1
        [[dp.OP["LOAD"], 2, 0, 0], [dp.OP["STORE"], 0, 2, 0]],  # register 2 of 2
        [[dp.OP["LOAD"], 0, 0, 0], [dp.OP["ADD"], 1, 0, 5]],  # operand register 5
        [[dp.OP["LOAD"], 0, 0, 0], [dp.OP["STORE"], 1, 0, 0]],  # output 1 of 1
        [[99, 0, 0, 0]],  # unknown op
    ]
    for code in bad:
        assert create(code) == nat.PML_E_INVALID, code
        assert out.value is None
    assert create([[dp.OP["LOAD"], 0, 0, 0]], n_regs=200) == nat.PML_E_INVALID
