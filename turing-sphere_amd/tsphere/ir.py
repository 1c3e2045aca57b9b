"""The IR surface the backend consumes (src/mgpu/numbl.d.ts:19-153).

numbl lowers a MATLAB function to statements whose expressions are trees of
``NumLit`` / ``Var`` / ``Binary`` / ``Unary`` / ``Call`` nodes, each carrying
a type; its inline pass folds single-use temps back, so one source line of an
element-wise computation becomes one ``Assign``.  These dataclasses have the
same fields and meaning (snake_case), so IR built by hand in tests, by the
MATLAB-subset front end in :mod:`tsphere.mfront`, or translated from numbl's
JSON, all drive the same emitter and planner.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional, Union


@dataclass(frozen=True)
class Span:
    file: str = "model.m"
    start: int = 0
    end: int = 0


NO_SPAN = Span()


@dataclass(frozen=True)
class NumericType:
    """numbl NumericType (numbl.d.ts:32-41); ``shape`` None = not all exact."""
    elem: str = "double"
    is_complex: bool = False
    shape: Optional[tuple] = None
    sign: str = "unknown"
    exact: Optional[float] = None
    kind: str = "Numeric"


@dataclass(frozen=True)
class NonNumericType:
    kind: str = "Unknown"


Type = Union[NumericType, NonNumericType]


def tensor_double(shape) -> NumericType:
    return NumericType(shape=tuple(int(d) for d in shape))


def scalar_double(sign: str = "unknown", exact: Optional[float] = None) -> NumericType:
    return NumericType(shape=(1, 1), sign=sign, exact=exact)


def is_numeric(t) -> bool:
    return getattr(t, "kind", None) == "Numeric"


def is_multi_element(t: NumericType) -> bool:
    """numbl isMultiElement: not provably a single element."""
    if t.shape is None:
        return True
    n = 1
    for d in t.shape:
        n *= d
    return n != 1


def is_tensor(t) -> bool:
    return is_numeric(t) and is_multi_element(t)


def numel(t: NumericType) -> int:
    n = 1
    for d in t.shape or ():
        n *= d
    return n


# ------------------------------------------------------------ expressions
@dataclass
class NumLit:
    value: float
    ty: Type
    span: Span = NO_SPAN
    kind: str = "NumLit"


@dataclass
class Var:
    name: str
    c_name: str
    ty: Type
    span: Span = NO_SPAN
    kind: str = "Var"


@dataclass
class Binary:
    builtin: str
    left: "IRExpr"
    right: "IRExpr"
    ty: Type
    span: Span = NO_SPAN
    kind: str = "Binary"


@dataclass
class Unary:
    builtin: str
    operand: "IRExpr"
    ty: Type
    span: Span = NO_SPAN
    kind: str = "Unary"


@dataclass
class Call:
    c_name: str
    name: str
    args: list
    ty: Type
    span: Span = NO_SPAN
    kind: str = "Call"


@dataclass
class OtherExpr:
    """Any other IR expression kind (IndexLoad, TensorConcat, ...): rejected."""
    kind: str
    ty: Type
    span: Span = NO_SPAN


IRExpr = Union[NumLit, Var, Binary, Unary, Call, OtherExpr]


# ------------------------------------------------------------- statements
@dataclass
class Assign:
    name: str
    c_name: str
    ty: Type
    expr: IRExpr
    span: Span = NO_SPAN
    kind: str = "Assign"


@dataclass
class OtherStmt:
    """If / For / IndexStore / ...: rejected by the planner."""
    kind: str
    span: Span = NO_SPAN


IRStmt = Union[Assign, OtherStmt]


# --------------------------------------------------- the planner's input
@dataclass
class Binding:
    """What the host supplies for an argument (src/mgpu/compile.ts:20-30):
    ``tensor`` (a device buffer of `shape`), ``param`` (a tunable scalar that
    must never force a recompile) or ``const`` (an exact scalar)."""
    kind: str
    shape: Optional[tuple] = None
    value: Optional[float] = None


@dataclass
class Param:
    name: str
    c_name: str
    binding: Binding


@dataclass
class Output:
    name: str
    c_name: str
    ty: Type


@dataclass
class CompiledFunction:
    """One specialised function, as the planner consumes it (compile.ts:49-59)."""
    name: str
    params: list = field(default_factory=list)   # list[Param]
    outputs: list = field(default_factory=list)  # list[Output]
    body: list = field(default_factory=list)     # list[IRStmt]


def walk(e: IRExpr):
    """Pre-order traversal of an expression tree."""
    yield e
    if isinstance(e, Binary):
        yield from walk(e.left)
        yield from walk(e.right)
    elif isinstance(e, Unary):
        yield from walk(e.operand)
    elif isinstance(e, Call):
        for a in e.args:
            yield from walk(a)
