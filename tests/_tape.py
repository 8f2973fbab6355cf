"""A minimal scalar reverse-mode automatic differentiation tape.

Used only by the tests, on tiny inputs, to derive GLM gradients in a way
that is independent of the hand-written adjoints in the oracle and the
kernels (the reference obtains gradients from Stan Math's reverse mode).
"""
import math


class Var:
    __slots__ = ("val", "parents", "adj")

    def __init__(self, val, parents=()):
        self.val = float(val)
        self.parents = parents  # tuples (var, local partial)
        self.adj = 0.0

    def __add__(self, o):
        o = o if isinstance(o, Var) else Var(o)
        return Var(self.val + o.val, ((self, 1.0), (o, 1.0)))

    __radd__ = __add__

    def __sub__(self, o):
        o = o if isinstance(o, Var) else Var(o)
        return Var(self.val - o.val, ((self, 1.0), (o, -1.0)))

    def __rsub__(self, o):
        return Var(o) - self

    def __mul__(self, o):
        o = o if isinstance(o, Var) else Var(o)
        return Var(self.val * o.val, ((self, o.val), (o, self.val)))

    __rmul__ = __mul__

    def __truediv__(self, o):
        o = o if isinstance(o, Var) else Var(o)
        return Var(self.val / o.val, ((self, 1.0 / o.val), (o, -self.val / o.val ** 2)))

    def __neg__(self):
        return Var(-self.val, ((self, -1.0),))


def exp(x):
    v = math.exp(x.val)
    return Var(v, ((x, v),))


def log(x):
    return Var(math.log(x.val), ((x, 1.0 / x.val),))


def log1p(x):
    return Var(math.log1p(x.val), ((x, 1.0 / (1.0 + x.val)),))


def grad(out, inputs):
    order, seen = [], set()

    def visit(v):
        stack = [(v, False)]
        while stack:
            node, done = stack.pop()
            if done:
                order.append(node)
                continue
            if id(node) in seen:
                continue
            seen.add(id(node))
            stack.append((node, True))
            for p, _ in node.parents:
                stack.append((p, False))

    visit(out)
    out.adj = 1.0
    for node in reversed(order):
        for p, d in node.parents:
            p.adj += node.adj * d
    return [x.adj for x in inputs]
