"""Symbolic right-hand sides compiled for the device.

The reference evaluates a ``SymbolicEquationSystem`` by lambdifying its
right-hand sides over the free symbols and mapping every symbol to a NumPy
array (pararealml/operators/symbol_mapper.py:160-271,
operators/fdm/fdm_symbol_mapper.py:45-158).  Here the same expressions are
compiled once into a register-machine program that ``libpml.so`` validates,
keeps in device memory and interprets per element (``pml_expr_eval``) or
inside the batched ODE time loop (``pml_ode_solve``):

* every free symbol other than ``t`` is a per-element *input* (a state
  component, a coordinate grid or a derivative array the differentiator
  computed), ``t`` is the time operand of the call;
* numbers are FP64 constants; ``+ - * /``, integer and real powers,
  ``sqrt``, ``exp``, ``log``, the trigonometric and hyperbolic functions and
  their inverses, ``Abs``, ``Min``, ``Max``, ``sign`` and ``atan2`` are
  instructions; anything else raises ``NotImplementedError``;
* identical subexpressions are computed once; registers are reused as soon
  as their value is dead.

Instruction encoding (four int32 per instruction, ``{op, dst, a, b}``) and
op codes are those of ``pml_expr.hip``.
"""

import ctypes
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional, Sequence

import numpy as np
import sympy as sp

OP = {
    "CONST": 0, "LOAD": 1, "TIME": 2, "ADD": 3, "SUB": 4, "MUL": 5, "DIV": 6, "POW": 7,
    "POWI": 8, "NEG": 9, "STORE": 10, "SQRT": 11, "EXP": 12, "LOG": 13, "SIN": 14, "COS": 15,
    "TAN": 16, "TANH": 17, "ABS": 18, "ASIN": 19, "ACOS": 20, "ATAN": 21, "SINH": 22,
    "COSH": 23, "ATAN2": 24, "MIN": 25, "MAX": 26, "SIGN": 27,
}
MAX_REGS = 96  # PML_EXPR_MAX_REGS
MAX_IO = 48  # PML_EXPR_MAX_IO

_UNARY = {
    sp.exp: "EXP", sp.log: "LOG", sp.sin: "SIN", sp.cos: "COS", sp.tan: "TAN",
    sp.tanh: "TANH", sp.Abs: "ABS", sp.asin: "ASIN", sp.acos: "ACOS", sp.atan: "ATAN",
    sp.sinh: "SINH", sp.cosh: "COSH", sp.sign: "SIGN",
}


@dataclass
class Program:
    """A compiled program: ``code`` (n, 4) int32, FP64 ``consts``, the
    register count, the input symbols (input k is ``inputs[k]``) and the
    number of outputs (output k is expression k)."""

    code: np.ndarray
    consts: np.ndarray
    n_regs: int
    inputs: List[sp.Symbol]
    n_outputs: int
    _handles: Dict[int, int] = field(default_factory=dict, repr=False)

    def handle(self) -> int:
        """The program uploaded to the current device (created once per
        device, released with the program)."""
        import torch

        from pararealml_amd import _native as nat

        dev = torch.cuda.current_device()
        if dev not in self._handles:
            lib = nat.device_lib()
            out = ctypes.c_void_p()
            code = np.ascontiguousarray(self.code, dtype=np.int32)
            consts = np.ascontiguousarray(self.consts, dtype=np.float64)
            nat.check(lib.pml_program_create(
                code.ctypes.data_as(ctypes.c_void_p), int(code.shape[0]),
                consts.ctypes.data_as(ctypes.c_void_p) if consts.size else None,
                int(consts.size), int(self.n_regs), len(self.inputs), int(self.n_outputs),
                ctypes.byref(out)))
            self._handles[dev] = out.value
        return self._handles[dev]

    def __del__(self):
        if not self._handles:
            return
        try:
            from pararealml_amd import _native as nat

            lib = nat.load_library()
            for h in self._handles.values():
                lib.pml_program_destroy(h)
        except Exception:  # interpreter shutdown
            pass
        self._handles.clear()


class _Builder:
    def __init__(self, input_of: Callable[[sp.Symbol], int]):
        self.input_of = input_of
        self.ins: List[List[int]] = []  # [op, dst_vreg, a, b]; a/b vregs or immediates
        self.consts: List[float] = []
        self.const_idx: Dict[float, int] = {}
        self.memo: Dict[sp.Basic, int] = {}
        self.n_v = 0
        self.reads: List[tuple] = []  # per instruction: vregs it reads

    def _new(self, op: str, a: int = 0, b: int = 0, reads=()) -> int:
        v = self.n_v
        self.n_v += 1
        self.ins.append([OP[op], v, a, b])
        self.reads.append(tuple(reads))
        return v

    def const(self, value: float) -> int:
        value = float(value)
        key = (value, np.signbit(value))
        if key not in self.const_idx:
            self.const_idx[key] = len(self.consts)
            self.consts.append(value)
        return self._new("CONST", self.const_idx[key])

    def bin(self, op: str, a: int, b: int) -> int:
        return self._new(op, a, b, reads=(a, b))

    def un(self, op: str, a: int, imm: int = 0) -> int:
        return self._new(op, a, imm, reads=(a,))

    def expr(self, e) -> int:
        e = sp.sympify(e)
        if e in self.memo:
            return self.memo[e]
        v = self._expr(e)
        self.memo[e] = v
        return v

    def _expr(self, e) -> int:
        if e.is_Number or e.is_NumberSymbol:
            if e.has(sp.zoo, sp.nan, sp.oo, -sp.oo):
                raise NotImplementedError(f"non-finite constant {e} in a right-hand side")
            return self.const(float(e))
        if e.is_Symbol:
            if e.name == "t":
                return self._new("TIME")
            return self._new("LOAD", self.input_of(e))
        if isinstance(e, sp.Add):
            pos, neg = [], []
            for term in e.args:
                if term.could_extract_minus_sign():
                    neg.append(-term)
                else:
                    pos.append(term)
            if pos:
                acc = self.expr(pos[0])
                rest = pos[1:]
            else:
                acc = self.un("NEG", self.expr(neg[0]))
                neg = neg[1:]
                rest = []
            for term in rest:
                acc = self.bin("ADD", acc, self.expr(term))
            for term in neg:
                acc = self.bin("SUB", acc, self.expr(term))
            return acc
        if isinstance(e, sp.Mul):
            coeff, factors = e.as_coeff_mul()
            num, den = [], []
            for f in factors:
                if isinstance(f, sp.Pow) and f.exp.is_Number and f.exp < 0:
                    den.append(sp.Pow(f.base, -f.exp, evaluate=False) if f.exp != -1 else f.base)
                else:
                    num.append(f)
            negate = coeff < 0
            mag = -coeff if negate else coeff
            acc = None
            if mag != 1 or not num:
                acc = self.const(float(mag))
            for f in num:
                fv = self.expr(f)
                acc = fv if acc is None else self.bin("MUL", acc, fv)
            if den:
                d = None
                for f in den:
                    fv = self.expr(f)
                    d = fv if d is None else self.bin("MUL", d, fv)
                acc = self.bin("DIV", acc, d)
            if negate:
                acc = self.un("NEG", acc)
            return acc
        if isinstance(e, sp.Pow):
            base, ex = e.base, e.exp
            if base == sp.E:
                return self.un("EXP", self.expr(ex))
            if ex.is_Integer:
                return self.un("POWI", self.expr(base), int(ex))
            if ex.is_Number and float(ex) == 0.5:
                return self.un("SQRT", self.expr(base))
            if ex.is_Number and float(ex) == -0.5:
                one = self.const(1.0)
                return self.bin("DIV", one, self.un("SQRT", self.expr(base)))
            if ex.is_Number and float(ex).is_integer() and abs(float(ex)) < 2**31:
                return self.un("POWI", self.expr(base), int(float(ex)))
            return self.bin("POW", self.expr(base), self.expr(ex))
        for cls, op in _UNARY.items():
            if isinstance(e, cls):
                if len(e.args) != 1:
                    raise NotImplementedError(f"{cls.__name__} with {len(e.args)} arguments")
                return self.un(op, self.expr(e.args[0]))
        if isinstance(e, sp.atan2):
            return self.bin("ATAN2", self.expr(e.args[0]), self.expr(e.args[1]))
        if isinstance(e, (sp.Min, sp.Max)):
            op = "MIN" if isinstance(e, sp.Min) else "MAX"
            acc = self.expr(e.args[0])
            for a in e.args[1:]:
                acc = self.bin(op, acc, self.expr(a))
            return acc
        raise NotImplementedError(
            f"{type(e).__name__} ({e}) has no device instruction in right-hand sides"
        )


def compile_expressions(exprs: Sequence, input_of: Callable[[sp.Symbol], int],
                        inputs: Sequence[sp.Symbol]) -> Program:
    """Compiles ``exprs`` (output k = ``exprs[k]``); ``input_of`` maps a
    non-time symbol to its input index among ``inputs``."""
    b = _Builder(input_of)
    outs = [b.expr(e) for e in exprs]
    if len(outs) > MAX_IO or len(inputs) > MAX_IO:
        raise NotImplementedError(
            f"at most {MAX_IO} outputs and inputs per program (got {len(outs)}, {len(inputs)})"
        )
    ins = b.ins
    reads = list(b.reads)
    for k, v in enumerate(outs):
        ins.append([OP["STORE"], k, v, 0])
        reads.append((v,))
    # liveness: last instruction reading each virtual register
    last = {}
    for i, rd in enumerate(reads):
        for v in rd:
            last[v] = i
    phys: Dict[int, int] = {}
    free: List[int] = []
    n_regs = 0
    code = np.zeros((len(ins), 4), dtype=np.int32)
    for i, (op, dst, a, imm) in enumerate(ins):
        rd = reads[i]
        # operands in physical registers (read before the destination is
        # written, so a dying operand's register may be the destination)
        mapped = {v: phys[v] for v in rd}
        for v in set(rd):
            if last[v] == i:
                free.append(phys.pop(v))
        if op == OP["STORE"]:
            code[i] = (op, dst, mapped[a], 0)
            continue
        if dst not in last:  # dead value (never read): still needs a register
            pass
        if free:
            free.sort(reverse=True)
            r = free.pop()
        else:
            r = n_regs
            n_regs += 1
        if dst in last:
            phys[dst] = r
        else:
            free.append(r)
        if op in (OP["CONST"], OP["LOAD"], OP["TIME"]):
            code[i] = (op, r, a, 0)
        elif op == OP["POWI"]:
            code[i] = (op, r, mapped[a], imm)
        elif len(rd) == 2:
            code[i] = (op, r, mapped[a], mapped[imm])
        else:
            code[i] = (op, r, mapped[a], 0)
    if n_regs > MAX_REGS:
        raise NotImplementedError(
            f"right-hand sides need {n_regs} device registers (at most {MAX_REGS})"
        )
    return Program(code, np.array(b.consts, dtype=np.float64), max(n_regs, 1), list(inputs),
                   len(outs))


def ode_program(diff_eq) -> Program:
    """The program of an ODE system: input k is y_k, output k is dy_k/dt."""
    system = diff_eq.symbolic_equation_system
    ys = list(diff_eq.symbols.y)
    index = {s: k for k, s in enumerate(ys)}

    def input_of(sym):
        if sym not in index:
            raise ValueError(f"unknown symbol {sym} in an ODE right-hand side")
        return index[sym]

    return compile_expressions(list(system.rhs), input_of, ys)


def evaluate_on_host(program: Program, inputs: Sequence[np.ndarray], t: float) -> List[np.ndarray]:
    """Runs ``program`` with NumPy (one array per input, broadcast together):
    the instruction semantics of the device interpreter, for inspection."""
    regs: Dict[int, np.ndarray] = {}
    outs: List[Optional[np.ndarray]] = [None] * program.n_outputs
    inv = {v: k for k, v in OP.items()}
    fn = {"SQRT": np.sqrt, "EXP": np.exp, "LOG": np.log, "SIN": np.sin, "COS": np.cos,
          "TAN": np.tan, "TANH": np.tanh, "ABS": np.abs, "ASIN": np.arcsin, "ACOS": np.arccos,
          "ATAN": np.arctan, "SINH": np.sinh, "COSH": np.cosh, "NEG": np.negative,
          "SIGN": np.sign}
    with np.errstate(all="ignore"):
        for op, dst, a, b in program.code.tolist():
            name = inv[op]
            if name == "CONST":
                regs[dst] = np.float64(program.consts[a])
            elif name == "LOAD":
                regs[dst] = np.asarray(inputs[a], dtype=np.float64)
            elif name == "TIME":
                regs[dst] = np.float64(t)
            elif name == "STORE":
                outs[dst] = regs[a]
            elif name == "POWI":
                regs[dst] = np.power(regs[a], float(b)) if b < 0 else regs[a] ** b
            elif name in fn:
                regs[dst] = fn[name](regs[a])
            else:
                x, y = regs[a], regs[b]
                regs[dst] = {
                    "ADD": lambda: x + y, "SUB": lambda: x - y, "MUL": lambda: x * y,
                    "DIV": lambda: x / y, "POW": lambda: np.power(x, y),
                    "ATAN2": lambda: np.arctan2(x, y), "MIN": lambda: np.minimum(x, y),
                    "MAX": lambda: np.maximum(x, y),
                }[name]()
    return outs
