"""The reference's optimisers, learning-rate decay and weight decay, on the host parameter vector (FP64).

Every step follows src/opt/algo.cpp (file:line at each class) on an Objective with the fields the
reference's algorithms read: p (parameters), g (gradient), g_old, val, val_old, gamma.  The training loop's
bookkeeping (src/nnp/nnptefr.cpp:1031-1066) is `Trainer.update`: gamma <- decay(gamma), algorithm step,
weight decay on the weights (not the biases) from the previous parameters, then val_old, p_old, g_old.

Where the reference cannot run an algorithm as written, the restatement says so:
  * BFGS::resize calls itself (algo.cpp:598-604), so BFGS never starts there; here B starts as the identity.
  * CG reads an eps that its constructor leaves unset (algo.hpp:558-565); here eps = 1e-16 like the others.
"""
from __future__ import annotations

import numpy as np


class Objective:
    def __init__(self, p, gamma=1e-3):
        self.p = np.array(p, dtype=np.float64)
        self.p_old = self.p.copy()
        self.g = np.zeros_like(self.p)
        self.g_old = np.zeros_like(self.p)
        self.val = 0.0
        self.val_old = 0.0
        self.gamma = gamma


def _sgn(x):
    return np.where(x > 0, 1.0, np.where(x < 0, -1.0, 0.0))


class SGD:                       # algo.cpp:102-105
    def __init__(self, dim):
        pass

    def step(self, o):
        o.p -= o.gamma * o.g


class SDM:                       # algo.cpp:125-131
    def __init__(self, dim, eta=0.9):
        if not 0.0 <= eta <= 1.0:
            raise ValueError("SDM: invalid eta")
        self.eta, self.dx = eta, np.zeros(dim)

    def step(self, o):
        self.dx = self.dx * self.eta + o.gamma * o.g
        o.p -= self.dx


class NAG:                       # algo.cpp:151-160
    def __init__(self, dim, eta=0.9):
        if not 0.0 <= eta <= 1.0:
            raise ValueError("NAG: invalid eta")
        self.eta, self.dx = eta, np.zeros(dim)

    def step(self, o):
        self.dx = self.dx * self.eta + o.gamma * (self.eta + 1.0) * o.g
        o.p -= self.dx


class ADAGRAD:                   # algo.cpp:180-188
    def __init__(self, dim, eps=1e-16):
        self.eps, self.g2 = eps, np.zeros(dim)

    def step(self, o):
        self.g2 += o.g * o.g
        o.p -= o.gamma * o.g / (np.sqrt(self.g2) + self.eps)


class ADADELTA:                  # algo.cpp:212-225
    def __init__(self, dim, eta=0.9, eps=1e-16):
        self.eta, self.eps = eta, eps
        self.g2, self.dx2 = np.zeros(dim), np.zeros(dim)

    def step(self, o):
        self.g2 = self.g2 * self.eta + (1.0 - self.eta) * o.g * o.g
        dx = o.g * np.sqrt(self.dx2) / (np.sqrt(self.g2) + self.eps)
        o.p -= dx
        self.dx2 = self.dx2 * self.eta + (1.0 - self.eta) * dx * dx


class RMSPROP:                   # algo.cpp:245-254
    def __init__(self, dim, eps=1e-16):
        self.eps, self.g2 = eps, np.zeros(dim)

    def step(self, o):
        self.g2 = self.g2 * 0.9 + 0.1 * o.g * o.g
        o.p -= o.gamma * o.g / (np.sqrt(self.g2) + self.eps)


class _Moments:
    """First and second moment averages with the bias-correction powers b1^t, b2^t (algo.cpp:269-318)."""
    b1, b2 = 0.9, 0.999

    def __init__(self, dim, eps=1e-16):
        self.eps = eps
        self.m, self.v = np.zeros(dim), np.zeros(dim)
        self.b1t, self.b2t = self.b1, self.b2

    def corr(self):
        return 1.0 / (1.0 - self.b1t), 1.0 / (1.0 - self.b2t)

    def advance(self):
        self.b1t *= self.b1
        self.b2t *= self.b2


class ADAM(_Moments):            # algo.cpp:280-296
    def step(self, o):
        c1, c2 = self.corr()
        self.m = self.m * self.b1 + (1.0 - self.b1) * o.g
        self.v = self.v * self.b2 + (1.0 - self.b2) * o.g * o.g
        o.p -= o.gamma * c1 * self.m / (np.sqrt(c2 * self.v) + self.eps)
        self.advance()


class ADAMW(_Moments):           # algo.cpp:324-340
    def __init__(self, dim, w=1e-3, eps=1e-16):
        super().__init__(dim, eps)
        if not 0.0 <= w <= 1.0:
            raise ValueError("ADAMW: invalid w")
        self.w = w

    def step(self, o):
        c1, c2 = self.corr()
        self.m = self.m * self.b1 + (1.0 - self.b1) * o.g
        self.v = self.v * self.b2 + (1.0 - self.b2) * o.g * o.g
        o.p -= o.gamma * (c1 * self.m / (np.sqrt(c2 * self.v) + self.eps) + self.w * o.p)
        self.advance()


class ADAB(_Moments):            # algo.cpp:366-382
    def step(self, o):
        c1, c2 = self.corr()
        self.m = self.m * self.b1 + (1.0 - self.b1) * o.g
        self.v = self.v * self.b2 + (1.0 - self.b2) * (o.g - self.m) * (o.g - self.m) + self.eps
        o.p -= o.gamma * c1 * self.m / (np.sqrt(c2 * self.v) + self.eps)
        self.advance()


class YOGI(_Moments):            # algo.cpp:408-425
    def step(self, o):
        c1, c2 = self.corr()
        self.m = self.m * self.b1 + (1.0 - self.b1) * o.g
        g2 = o.g * o.g
        self.v = self.v - (1.0 - self.b2) * _sgn(self.v - g2) * g2
        o.p -= o.gamma * c1 * self.m / (np.sqrt(c2 * self.v) + self.eps)
        self.advance()


class NOGI(_Moments):            # algo.cpp:451-468
    def step(self, o):
        c1, c2 = self.corr()
        self.m = self.m * self.b1 + (1.0 - self.b1) * o.g
        g2 = o.g * o.g
        self.v = self.v - (1.0 - self.b2) * _sgn(self.v - g2) * g2
        o.p -= o.gamma * c1 * (self.b1 * self.m + (1.0 - self.b1) * o.g) / (np.sqrt(c2 * self.v) + self.eps)
        self.advance()


class NADAM(_Moments):           # algo.cpp:495-511
    def step(self, o):
        c1, c2 = self.corr()
        self.m = self.m * self.b1 + (1.0 - self.b1) * o.g
        self.v = self.v * self.b2 + (1.0 - self.b2) * o.g * o.g
        o.p -= o.gamma * c1 * (self.b1 * self.m + (1.0 - self.b1) * o.g) / (np.sqrt(c2 * self.v) + self.eps)
        self.advance()


class AMSGRAD(_Moments):         # algo.cpp:538-556
    def __init__(self, dim, eps=1e-16):
        super().__init__(dim, eps)
        self.vmax = np.zeros(dim)

    def step(self, o):
        c1, c2 = self.corr()
        self.m = self.m * self.b1 + (1.0 - self.b1) * o.g
        self.v = self.v * self.b2 + (1.0 - self.b2) * o.g * o.g
        self.vmax = np.maximum(self.vmax, self.v)
        o.p -= o.gamma * c1 * self.m / (np.sqrt(c2 * self.vmax) + self.eps)
        self.advance()


class BFGS:                      # algo.cpp:566-596
    def __init__(self, dim):
        self.B = np.eye(dim)
        self.B_old = np.eye(dim)

    def step(self, o):
        s = np.linalg.solve(self.B, o.g)          # B is kept symmetric positive definite (LLT in the reference)
        s *= -o.gamma / np.linalg.norm(s)
        o.p += s
        y = o.gamma * (o.g - o.g_old)
        self.B = self.B + np.outer(y, y) / (y @ s)
        y = self.B_old @ s
        self.B = self.B - np.outer(y, y) / (s @ y)
        self.B_old = self.B.copy()


class RPROP:                     # algo.cpp:610-633 (iRprop+)
    eta_p, eta_m, d_max, d_min = 1.2, 0.5, 50.0, 1e-14

    def __init__(self, dim):
        self.delta = np.full(dim, 0.1)
        self.dx = np.zeros(dim)

    def step(self, o):
        inc = o.val > o.val_old
        s = o.g * o.g_old
        pos, neg, zero = s > 0, s < 0, s == 0
        self.delta[pos] = np.minimum(self.delta[pos] * self.eta_p, self.d_max)
        self.delta[neg] = np.maximum(self.delta[neg] * self.eta_m, self.d_min)
        upd = pos | zero
        self.dx[upd] = -_sgn(o.g[upd]) * self.delta[upd]
        o.p[upd] += self.dx[upd]
        if inc:
            o.p[neg] -= self.dx[neg]
        o.g[neg] = 0.0                             # the reference zeroes these gradient entries in place


class CG:                        # algo.cpp:643-652
    def __init__(self, dim, eps=1e-16):
        self.eps, self.d = eps, np.zeros(dim)

    def step(self, o):
        beta = o.g @ (o.g - o.g_old) / (o.g_old @ o.g_old + self.eps)
        self.d = self.d * beta - o.g
        o.p += o.gamma * self.d


ALGORITHMS = {c.__name__: c for c in (SGD, SDM, NAG, ADAGRAD, ADADELTA, RMSPROP, ADAM, ADAMW, ADAB, YOGI, NOGI,
                                      NADAM, AMSGRAD, BFGS, RPROP, CG)}


def make(name, dim, *args):
    """opt::algo::make by name (algo.cpp:690-712)."""
    key = name.upper()
    if key not in ALGORITHMS:
        raise ValueError("opt::algo::make: invalid algorithm name '%s'" % name)
    return ALGORITHMS[key](dim, *args)


class Decay:
    """opt::Decay::step (src/opt/decay.cpp:58-80): the new gamma from the old one at iteration `step`."""
    NAMES = ("CONST", "EXP", "GAUSS", "SQRT", "INV", "STEP")

    def __init__(self, name="CONST", alpha=0.0, period=1):
        self.name = name.upper()
        if self.name not in self.NAMES:
            raise ValueError("opt::Decay::read: invalid name")
        if alpha < 0:
            raise ValueError("opt::Decay::read: invalid alpha")
        self.alpha, self.period = alpha, period

    def step(self, gamma, step):
        a = self.alpha
        if self.name == "CONST":
            return gamma
        if self.name == "EXP":
            return gamma * (1.0 - a)
        if self.name == "GAUSS":
            return gamma * (1.0 - 2.0 * a * step)
        if self.name == "SQRT":
            return gamma * (1.0 - 0.5 * a / (1.0 + a * step))
        if self.name == "INV":
            return gamma * (1.0 - a / (1.0 + a * step))
        return gamma * a if (step > 0 and step % self.period == 0) else gamma


REGULARIZATION = ("NONE", "LASSO", "RIDGE", "HUBER", "ASINH")


def weight_decay(o: Objective, mask, eta, kind):
    """The training loop's weight decay on the weights (mask 1) from the previous parameters
    (nnptefr.cpp:1036-1058)."""
    k = kind.upper()
    q = o.p_old
    if k == "NONE":
        return
    if k == "LASSO":
        d = _sgn(q)
    elif k == "RIDGE":
        d = q
    elif k == "HUBER":
        d = q / np.sqrt(1.0 + q * q)
    elif k == "ASINH":
        d = np.arcsinh(q)
    else:
        raise ValueError("invalid regularization '%s'" % kind)
    o.p -= o.gamma * eta * mask * d
