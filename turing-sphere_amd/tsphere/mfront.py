"""MATLAB-subset front end: ``.m`` model source -> the backend's IR.

The reference lowers ``models/*.m`` with numbl (src/mgpu/compile.ts:83-190):
``specializeUserFunction`` per function for the host's argument types, then
``inlinePass``, which folds single-use temps back so each element-wise source
line becomes ONE statement whose right-hand side is an expression tree.
numbl (a separate project) is not part of this library, so this module parses the subset the
models use and produces that folded IR directly:

* files of ``function [outs] = name(params) ... end`` blocks, ``%`` comments,
  ``...`` continuations; statements ``x = expr;``;
* expressions with ``+ - * / .* ./ .^ ^``, unary ``+ -``, parentheses,
  numeric literals, calls of element-wise builtins, ``ones``/``zeros``,
  ``numel``, and the external transforms ``synth``/``analys``;
* MATLAB precedence (power > unary > multiplicative > additive) with the
  ``a^-b`` exception;
* an external call inside an expression is hoisted into its own statement
  (its argument too, when it is not a plain variable) -- the transform is not
  element-wise, so it bounds the fused expression on either side, exactly
  where numbl's ANF + inline pass leave a statement boundary.

Types follow numbl's lattice as far as the backend needs: tensors carry a
static shape, runtime parameters are scalars without an exact value (so a
parameter change never recompiles, compile.ts:34-37), ``const`` bindings and
literal arithmetic fold to exact scalars.  Anything else is refused with a
source span (``ModelCompileError``, src/mgpu/errors.ts:14-90).
"""
from __future__ import annotations

import re
from dataclasses import dataclass

from .ir import (Assign, Binary, Binding, Call, CompiledFunction, NumericType, NumLit, Output,
                 OtherStmt, Param, Span, Unary, Var, is_tensor, scalar_double, tensor_double)

ELEMENTWISE = {"abs", "acos", "asin", "atan", "atan2", "ceil", "cos", "cosh", "exp", "floor",
               "log", "log2", "max", "min", "round", "sign", "sin", "sinh", "sqrt", "tan",
               "tanh"}
EXTERNALS = ("synth", "analys")


class ModelCompileError(Exception):
    """errors.ts:14-90: a compile failure with source offsets and the function."""

    def __init__(self, message: str, start: int | None = None, end: int | None = None,
                 fn: str | None = None):
        super().__init__(message)
        self.start, self.end, self.fn = start, end, fn


# ------------------------------------------------------------------ lexer
_TOKEN = re.compile(r"""
    (?P<ws>[ \t]+)
  | (?P<cont>\.\.\.[^\n]*\n)
  | (?P<comment>%[^\n]*)
  | (?P<nl>\n)
  | (?P<num>(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?)
  | (?P<id>[A-Za-z_][A-Za-z0-9_]*)
  | (?P<op>\.\*|\./|\.\^|\.\\|\.'|==|~=|<=|>=|&&|\|\||[-+*/\\^()\[\],;=<>&|~':])
""", re.X)


@dataclass
class Tok:
    kind: str
    text: str
    start: int
    end: int


def tokenize(src: str) -> list[Tok]:
    out, pos = [], 0
    while pos < len(src):
        m = _TOKEN.match(src, pos)
        if not m:
            raise ModelCompileError(f"unexpected character {src[pos]!r}", pos, pos + 1)
        kind = m.lastgroup
        if kind in ("ws", "cont", "comment"):
            pass
        else:
            out.append(Tok(kind, m.group(), m.start(), m.end()))
        pos = m.end()
    out.append(Tok("eof", "", len(src), len(src)))
    return out


# ----------------------------------------------------------------- parser
@dataclass
class Node:
    kind: str          # num | id | bin | un | call
    start: int
    end: int
    value: float = 0.0
    name: str = ""
    op: str = ""
    args: tuple = ()


@dataclass
class Stmt:
    target: str
    expr: Node
    start: int
    end: int


@dataclass
class FuncDecl:
    name: str
    params: list
    outputs: list
    body: list
    start: int
    end: int


_BIN = {"+": "plus", "-": "minus", ".*": "times", "./": "rdivide", "*": "mtimes",
        "/": "mrdivide", ".^": "power", "^": "mpower"}


class Parser:
    def __init__(self, src: str):
        self.src = src
        self.toks = tokenize(src)
        self.i = 0

    def peek(self, k=0) -> Tok:
        return self.toks[min(self.i + k, len(self.toks) - 1)]

    def next(self) -> Tok:
        t = self.toks[self.i]
        self.i += 1
        return t

    def expect(self, text: str) -> Tok:
        t = self.next()
        if t.text != text:
            raise ModelCompileError(f"expected '{text}' but found '{t.text or 'end of file'}'",
                                    t.start, t.end)
        return t

    def skip_nl(self):
        while self.peek().kind == "nl" or self.peek().text in (";", ","):
            self.next()

    def file(self) -> list[FuncDecl]:
        fns = []
        self.skip_nl()
        while self.peek().kind != "eof":
            t = self.peek()
            if t.text != "function":
                raise ModelCompileError("only function definitions may appear at the top level "
                                        "of a model file", t.start, t.end)
            fns.append(self.function())
            self.skip_nl()
        return fns

    def function(self) -> FuncDecl:
        start = self.expect("function").start
        outs = []
        if self.peek().text == "[":
            self.next()
            while self.peek().text != "]":
                t = self.next()
                if t.kind == "id":
                    outs.append(t.text)
                elif t.text != ",":
                    raise ModelCompileError(f"bad output list near '{t.text}'", t.start, t.end)
            self.next()
            self.expect("=")
            name = self.next()
        else:
            first = self.next()
            if self.peek().text == "=":
                self.next()
                outs = [first.text]
                name = self.next()
            else:
                name = first
        if name.kind != "id":
            raise ModelCompileError("expected a function name", name.start, name.end)
        params = []
        if self.peek().text == "(":
            self.next()
            while self.peek().text != ")":
                t = self.next()
                if t.kind == "id":
                    params.append(t.text)
                elif t.text != ",":
                    raise ModelCompileError(f"bad parameter list near '{t.text}'", t.start, t.end)
            self.next()
        body = []
        self.skip_nl()
        while self.peek().text not in ("end", "function") and self.peek().kind != "eof":
            body.append(self.statement())
            self.skip_nl()
        end = self.peek().end
        if self.peek().text == "end":
            self.next()
        return FuncDecl(name.text, params, outs, body, start, end)

    def statement(self) -> Stmt:
        t = self.peek()
        if t.kind == "id" and t.text in ("if", "for", "while", "switch", "try", "return",
                                         "break", "continue", "global", "persistent"):
            kind = {"if": "If", "for": "For", "while": "While"}.get(t.text, t.text)
            # plan.ts:289-295: the backend runs straight-line assignments only
            raise ModelCompileError(
                f"a model function body may only contain assignments (found '{kind}')",
                t.start, t.end)
        if t.kind != "id" or self.peek(1).text != "=":
            e = self.expr()
            raise ModelCompileError(
                "a model function body may only contain assignments (found 'ExprStmt')",
                e.start, e.end)
        self.next()
        self.next()
        e = self.expr()
        return Stmt(t.text, e, t.start, e.end)

    # precedence climbing
    def expr(self) -> Node:
        return self.additive()

    def additive(self) -> Node:
        left = self.multiplicative()
        while self.peek().text in ("+", "-"):
            op = self.next().text
            right = self.multiplicative()
            left = Node("bin", left.start, right.end, op=op, args=(left, right))
        return left

    def multiplicative(self) -> Node:
        left = self.unary()
        while self.peek().text in ("*", "/", ".*", "./", "\\", ".\\"):
            t = self.next()
            if t.text in ("\\", ".\\"):
                raise ModelCompileError(f"operator '{t.text}' is not supported", t.start, t.end)
            right = self.unary()
            left = Node("bin", left.start, right.end, op=t.text, args=(left, right))
        return left

    def unary(self) -> Node:
        t = self.peek()
        if t.text in ("+", "-"):
            self.next()
            operand = self.unary()
            return Node("un", t.start, operand.end, op=t.text, args=(operand,))
        return self.power()

    def power(self) -> Node:
        base = self.postfix()
        while self.peek().text in ("^", ".^"):
            op = self.next().text
            if self.peek().text in ("+", "-"):  # a^-b
                u = self.next()
                operand = self.postfix()
                exp = Node("un", u.start, operand.end, op=u.text, args=(operand,))
            else:
                exp = self.postfix()
            base = Node("bin", base.start, exp.end, op=op, args=(base, exp))
        return base

    def postfix(self) -> Node:
        t = self.next()
        if t.kind == "num":
            node = Node("num", t.start, t.end, value=float(t.text))
        elif t.kind == "id":
            if self.peek().text == "(":
                self.next()
                args = []
                while self.peek().text != ")":
                    args.append(self.expr())
                    if self.peek().text == ",":
                        self.next()
                    elif self.peek().text != ")":
                        u = self.peek()
                        raise ModelCompileError(f"expected ',' or ')' but found '{u.text}'",
                                                u.start, u.end)
                end = self.next().end
                node = Node("call", t.start, end, name=t.text, args=tuple(args))
            else:
                node = Node("id", t.start, t.end, name=t.text)
        elif t.text == "(":
            inner = self.expr()
            end = self.expect(")").end
            node = Node(inner.kind, t.start, end, inner.value, inner.name, inner.op, inner.args)
        else:
            raise ModelCompileError(f"unexpected '{t.text or 'end of file'}'", t.start, t.end)
        if self.peek().text in ("'", ".'"):
            u = self.next()
            raise ModelCompileError("transpose is not supported", u.start, u.end)
        return node


# ---------------------------------------------------------------- lowering
def _sign_of(v: float) -> str:
    return "positive" if v > 0 else "negative" if v < 0 else "zero"


class _Lowerer:
    def __init__(self, decl: FuncDecl, bindings: dict, grid: dict, file: str):
        self.decl, self.grid, self.file = decl, grid, file
        self.env: dict[str, Var] = {}
        self.body: list = []
        self.ntemp = 0
        self.params: list[Param] = []
        for p in decl.params:
            b = bindings.get(p)
            if b is None:
                offered = ", ".join(bindings)
                raise ModelCompileError(
                    f"'{decl.name}' takes an argument named '{p}', which this app does not "
                    f"provide. Available: {offered}.", fn=decl.name)
            if b.kind == "tensor":
                ty = tensor_double(b.shape)
            elif b.kind == "param":
                ty = scalar_double("unknown")
            else:
                ty = scalar_double(_sign_of(b.value), float(b.value))
            self.env[p] = Var(p, p, ty, self.span(decl.start, decl.start))
            self.params.append(Param(p, p, b))

    def span(self, s, e) -> Span:
        return Span(self.file, s, e)

    def err(self, msg, node) -> ModelCompileError:
        return ModelCompileError(msg, node.start, node.end, self.decl.name)

    def temp(self, expr, node) -> Var:
        self.ntemp += 1
        name = f"t{self.ntemp}_"
        var = Var(name, name, expr.ty, self.span(node.start, node.end))
        self.body.append(Assign(name, name, expr.ty, expr, var.span))
        return var

    def lower(self, node: Node):
        k = node.kind
        sp = self.span(node.start, node.end)
        if k == "num":
            return NumLit(node.value, scalar_double(_sign_of(node.value), node.value), sp)
        if k == "id":
            v = self.env.get(node.name)
            if v is None:
                raise self.err(f"'{node.name}' is used before it is assigned", node)
            return Var(v.name, v.c_name, v.ty, sp)
        if k == "un":
            a = self.lower(node.args[0])
            if a.ty.exact is not None:
                val = -a.ty.exact if node.op == "-" else a.ty.exact
                return NumLit(val, scalar_double(_sign_of(val), val), sp)
            return Unary("uminus" if node.op == "-" else "uplus", a, a.ty, sp)
        if k == "bin":
            a = self.lower(node.args[0])
            b = self.lower(node.args[1])
            builtin = _BIN[node.op]
            if a.ty.exact is not None and b.ty.exact is not None and builtin not in (
                    "power", "mpower"):
                x, y = a.ty.exact, b.ty.exact
                val = {"plus": x + y, "minus": x - y, "times": x * y, "mtimes": x * y,
                       "rdivide": x / y if y else float("inf"),
                       "mrdivide": x / y if y else float("inf")}[builtin]
                return NumLit(val, scalar_double(_sign_of(val), val), sp)
            return Binary(builtin, a, b, self.result_type(a.ty, b.ty, builtin), sp)
        if k == "call":
            return self.lower_call(node, sp)
        raise self.err(f"'{k}' is not supported", node)

    def result_type(self, ta: NumericType, tb: NumericType, builtin: str) -> NumericType:
        if not is_tensor(ta) and not is_tensor(tb):
            return scalar_double()
        if builtin == "mtimes" and is_tensor(ta) and is_tensor(tb):
            return tensor_double((ta.shape[0], tb.shape[1]))
        if is_tensor(ta) and is_tensor(tb):
            # MATLAB implicit expansion: the planner's shape check refuses it
            shape = tuple(max(x, y) for x, y in zip(ta.shape, tb.shape))
            return tensor_double(shape)
        return tensor_double((ta if is_tensor(ta) else tb).shape)

    def lower_call(self, node: Node, sp: Span):
        name = node.name
        if name in self.env and not node.args:
            return self.lower(Node("id", node.start, node.end, name=name))
        if name in EXTERNALS:
            if len(node.args) != 1:
                raise self.err(f"{name} takes exactly one argument, got {len(node.args)}", node)
            arg = self.lower(node.args[0])
            need = (2, self.grid["nlm"]) if name == "synth" else (self.grid["npts"], 1)
            out = (self.grid["npts"], 1) if name == "synth" else (2, self.grid["nlm"])
            if not is_tensor(arg.ty) or tuple(arg.ty.shape) != need:
                got = "x".join(map(str, arg.ty.shape)) if arg.ty.shape else "unknown shape"
                raise self.err(f"{name} requires a {need[0]}x{need[1]} array, got {got}", node)
            if arg.kind != "Var":
                arg = self.temp(arg, node.args[0])
            call = Call(name, name, [arg], tensor_double(out), sp)
            return self.temp(call, node)
        if name == "numel":
            if len(node.args) != 1:
                raise self.err("numel takes one argument", node)
            a = self.lower(node.args[0])
            n = 1
            for d in a.ty.shape or ():
                n *= d
            return NumLit(float(n), scalar_double(_sign_of(n), float(n)), sp)
        if name in ("ones", "zeros"):
            args = [self.lower(a) for a in node.args]
            dims = []
            for a, n in zip(args, node.args):
                if a.ty.exact is None:
                    raise self.err(f"'{name}' needs sizes known at compile time", n)
                dims.append(int(a.ty.exact))
            if len(dims) == 1:
                dims = [dims[0], dims[0]]
            return Call(name, name, args, tensor_double(dims), sp)
        if name in ELEMENTWISE:
            args = [self.lower(a) for a in node.args]
            ty = scalar_double()
            for a in args:
                if is_tensor(a.ty):
                    ty = tensor_double(a.ty.shape)
            return Call(name, name, args, ty, sp)
        if any(f == name for f in self.known_functions):
            return Call(name + "_spec", name, [self.lower(a) for a in node.args],
                        scalar_double(), sp)
        raise self.err(f"'{name}' is not a known function", node)

    known_functions: tuple = ()

    def run(self, nargout: int) -> CompiledFunction:
        for st in self.decl.body:
            if st.expr.kind.startswith("other:"):
                kind = st.expr.kind.split(":", 1)[1]
                self.body.append(OtherStmt(kind, self.span(st.start, st.end)))
                continue
            e = self.lower(st.expr)
            sp = self.span(st.start, st.end)
            c_name = st.target
            # a transform result already in a temp: rename the temp's statement
            if e.kind == "Var" and self.body and isinstance(self.body[-1], Assign) \
                    and self.body[-1].c_name == e.c_name and e.c_name.endswith("_"):
                e = self.body.pop().expr
            self.body.append(Assign(st.target, c_name, e.ty, e, sp))
            self.env[st.target] = Var(st.target, c_name, e.ty, sp)
        outs = []
        for o in self.decl.outputs[:nargout]:
            v = self.env.get(o)
            if v is None:
                raise ModelCompileError(f"'{self.decl.name}' declares the output '{o}' but "
                                        f"never assigns it", fn=self.decl.name)
            outs.append(Output(o, v.c_name, v.ty))
        return CompiledFunction(self.decl.name, self.params, outs, self.body)


class CompiledModel:
    """compile.ts:83-190: parse once, specialise the functions the host
    needs, then ``finish()``."""

    def __init__(self, source: str, bindings: dict, grid: dict, file_name: str = "model.m"):
        try:
            self._decls = {f.name: f for f in Parser(source).file()}
        except ModelCompileError:
            raise
        self._bindings = bindings
        self._grid = grid
        self._file = file_name

    def function_names(self) -> list[str]:
        return list(self._decls)

    def specialize(self, name: str, nargout: int) -> CompiledFunction:
        decl = self._decls.get(name)
        if decl is None:
            defined = self.function_names()
            listing = (f" (it defines {', '.join(repr(n) for n in defined)})" if defined
                       else " (it defines no functions)")
            raise ModelCompileError(f"the model must define a function named '{name}'{listing}")
        if len(decl.outputs) < nargout:
            raise ModelCompileError(
                f"'{name}' must return {nargout} value{'' if nargout == 1 else 's'}, but "
                f"declares {len(decl.outputs)}", fn=name)
        low = _Lowerer(decl, self._bindings, self._grid, self._file)
        low.known_functions = tuple(n for n in self._decls if n not in ("init", "step"))
        return low.run(nargout)

    def finish(self) -> None:
        """numbl's inline pass: the lowering above already emits folded trees."""
        return None


def model_bindings(state, param_names, npts: int, nlm: int) -> dict:
    """What a model may ask for by parameter name (model.ts:108-117)."""
    b = {
        "lam": Binding("tensor", (2, nlm)),
        "noise": Binding("tensor", (npts, 1)),
        "npts": Binding("const", value=float(npts)),
        "nlm": Binding("const", value=float(nlm)),
    }
    for s in state:
        b[s] = Binding("tensor", (2, nlm))
    for p in param_names:
        b[p] = Binding("param")
    return b
