"""DeformationState: the optimizer snapshot Deformetrica writes with ``--output-state-file``.

The reference class (src/support/utilities/SerializeDeformationState.h:43-290) holds one
FIFO per value type; ``<<`` appends, ``>>`` pops the front, and the whole object goes
through a Boost text archive.  This module reads and writes that archive byte for byte:

* header ``22 serialization::archive <library version>``, then the object's class record
  ``tracking version`` (DeformationState is class version 1, SerializeDeformationState.h:291);
* the 21 queues in the order of ``serialize`` (SerializeDeformationState.h:211-233), each as
  ``count item_version`` followed by the items;
* a class record ``tracking version`` (``0 0``) the first time a class-level type appears
  (the Armadillo and linear-variable wrappers, std::map / std::pair, and every queue or
  std::vector except those of arithmetic items, which Boost serializes without one --
  a queue of strings has one);
* matrices as ``rows cols`` + the column-major doubles as a binary object
  (ArmadilloMatrixWrapper.h:349-366), vectors as ``rows`` + doubles
  (ArmadilloVectorWrapper.h:316-330): base64 in lines of 76 characters, ``=`` padded,
  opened by two newlines, the next token following one newline;
* linear variables as ``which`` + the alternative (LinearVariableWrapper.h:112, 378-421):
  0 scalar, 1 vector, 2 matrix, 3 matrix list;
* doubles with 17 significant digits, floats with 8, strings as ``length text``.

Python values: bool/int/float/str, 1-D float64 arrays (VectorType), 2-D float64 arrays
(MatrixType), lists of 2-D arrays (MatrixListType), a linear variable is one of
float / 1-D array / 2-D array / list of 2-D arrays, a linear-variable map is a dict
(std::map order: sorted by key bytes), a linear-variables map is a dict of lists.
"""
from __future__ import annotations

import base64
import re
from collections import deque

import numpy as np

# the queues, in archive order (SerializeDeformationState.h:236-256)
KINDS = (
    "bool", "int", "uint", "float", "double", "string", "vector", "matrix", "matrix_list",
    "linear_variable", "linear_variable_map", "linear_variables", "linear_variables_map",
    "vec_string", "vec_int", "vec_uint", "vec_float", "vec_double", "vec_vector",
    "vec_matrix", "vec_matrix_list",
)
_PRIMITIVE = {"bool", "int", "uint", "float", "double", "string"}
# collections of these carry no class record (Boost's collection traits cover arithmetic types)
_ARITHMETIC = {"bool", "int", "uint", "float", "double"}
_SIGNATURE = "serialization::archive"
_CLASS_VERSION = 1
_LIBRARY_VERSION = 10


class StateFormatError(ValueError):
    pass


def _fmt_double(x):
    s = "%.17g" % float(x)
    return s


def _fmt_float(x):
    return "%.8g" % float(np.float32(x))


class _Reader:
    def __init__(self, text):
        self.toks = text.split()
        self.i = 0
        self.seen = set()

    def tok(self):
        if self.i >= len(self.toks):
            raise StateFormatError("unexpected end of archive")
        t = self.toks[self.i]
        self.i += 1
        return t

    def int(self):
        t = self.tok()
        if not re.fullmatch(r"-?\d+", t):
            raise StateFormatError(f"expected an integer, found {t!r} (token {self.i - 1})")
        return int(t)

    def double(self):
        return float(self.tok())

    def string(self):
        n = self.int()
        parts, got = [], 0
        # strings are whitespace-free in the state files written by the optimizers; a string
        # with spaces is rebuilt from the tokens that make up its length
        while got < n:
            t = self.tok()
            parts.append(t)
            got += len(t) + (1 if len(parts) > 1 else 0)
        s = " ".join(parts)
        if len(s) != n:
            raise StateFormatError(f"string length {n} does not match {s!r}")
        return s

    def binary(self, nbytes):
        if nbytes == 0:
            return b""
        need = 4 * ((nbytes + 2) // 3)
        chars = ""
        while len(chars) < need:
            chars += self.tok()
        if len(chars) != need:
            raise StateFormatError("binary object length mismatch")
        data = base64.b64decode(chars)
        if len(data) != nbytes:
            raise StateFormatError("binary object length mismatch")
        return data

    def class_info(self, cls):
        if cls in self.seen:
            return
        self.seen.add(cls)
        tracking, _version = self.int(), self.int()
        if tracking != 0:
            raise StateFormatError(f"object tracking is not supported ({cls})")

    def collection(self, item):
        n = self.int()
        self.int()  # item version
        return [item() for _ in range(n)]

    # value types
    def vector(self):
        self.class_info("vector")
        rows = self.int()
        return np.frombuffer(self.binary(8 * rows), dtype="<f8").copy()

    def matrix(self):
        self.class_info("matrix")
        rows, cols = self.int(), self.int()
        data = np.frombuffer(self.binary(8 * rows * cols), dtype="<f8")
        return data.reshape(cols, rows).T.copy()

    def matrix_list(self):
        self.class_info("matrix_list")
        self.class_info("std::vector<matrix>")
        return self.collection(self.matrix)

    def linear_variable(self):
        self.class_info("linear_variable")
        which = self.int()
        if which == 0:
            return self.double()
        if which == 1:
            return self.vector()
        if which == 2:
            return self.matrix()
        if which == 3:
            return self.matrix_list()
        raise StateFormatError(f"linear variable alternative {which}")

    def linear_variable_map(self):
        self.class_info("linear_variable_map")
        self.class_info("std::map<string,linear_variable>")

        def pair():
            self.class_info("std::pair<string,linear_variable>")
            return self.string(), self.linear_variable()

        return dict(self.collection(pair))

    def linear_variables(self):
        self.class_info("linear_variables")
        self.class_info("std::vector<linear_variable>")
        return self.collection(self.linear_variable)

    def linear_variables_map(self):
        self.class_info("linear_variables_map")
        self.class_info("std::map<string,linear_variables>")

        def pair():
            self.class_info("std::pair<string,linear_variables>")
            return self.string(), self.linear_variables()

        return dict(self.collection(pair))

    def primitive(self, kind):
        if kind == "bool":
            v = self.int()
            if v not in (0, 1):
                raise StateFormatError(f"bool value {v}")
            return bool(v)
        if kind in ("int", "uint"):
            return self.int()
        if kind in ("float", "double"):
            return self.double()
        return self.string()

    def item(self, kind):
        return (lambda: self.primitive(kind)) if kind in _PRIMITIVE else getattr(self, kind)

    def vec_of(self, elem):
        # std::vector<elem>: a class record unless elem is arithmetic
        if elem not in _ARITHMETIC:
            self.class_info(f"std::vector<{elem}>")
        return self.collection(self.item(elem))

    def queue(self, kind):
        if kind not in _ARITHMETIC:
            self.class_info(f"std::deque<{kind}>")
        if kind.startswith("vec_"):
            elem = kind[4:]
            return self.collection(lambda: self.vec_of(elem))
        return self.collection(self.item(kind))


class _Writer:
    def __init__(self):
        self.parts = []
        self.sep = ""  # delimiter before the next token
        self.seen = set()

    def tok(self, t):
        self.parts.append(self.sep + t)
        self.sep = " "

    def int(self, v):
        self.tok(str(int(v)))

    def string(self, s):
        self.int(len(s))
        self.tok(s)

    def binary(self, data):
        # text_oarchive: newline, then (non-empty data) newline + base64 in 76-column lines
        # with the '=' padding appended; the next token follows a newline
        self.parts.append("\n")
        if data:
            enc = base64.b64encode(data).decode("ascii")
            body = enc.rstrip("=")
            pad = len(enc) - len(body)
            lines = [body[i:i + 76] for i in range(0, len(body), 76)]
            self.parts.append("\n" + "\n".join(lines) + "=" * pad)
        self.sep = "\n"

    def class_info(self, cls):
        if cls in self.seen:
            return
        self.seen.add(cls)
        self.int(0)
        self.int(0)

    def collection(self, items, put):
        self.int(len(items))
        self.int(0)
        for it in items:
            put(it)

    def vector(self, v):
        v = np.ascontiguousarray(v, dtype="<f8")
        if v.ndim != 1:
            raise StateFormatError("a vector must be 1-D")
        self.class_info("vector")
        self.int(v.shape[0])
        self.binary(v.tobytes())

    def matrix(self, m):
        m = np.asarray(m, dtype="<f8")
        if m.ndim != 2:
            raise StateFormatError("a matrix must be 2-D")
        self.class_info("matrix")
        self.int(m.shape[0])
        self.int(m.shape[1])
        self.binary(np.asfortranarray(m).tobytes(order="F"))

    def matrix_list(self, ml):
        self.class_info("matrix_list")
        self.class_info("std::vector<matrix>")
        self.collection(list(ml), self.matrix)

    def linear_variable(self, v):
        self.class_info("linear_variable")
        if isinstance(v, (list, tuple)):
            self.int(3)
            self.matrix_list(v)
        elif np.ndim(v) == 0:
            self.int(0)
            self.tok(_fmt_double(v))
        elif np.ndim(v) == 1:
            self.int(1)
            self.vector(v)
        else:
            self.int(2)
            self.matrix(v)

    def linear_variable_map(self, m):
        self.class_info("linear_variable_map")
        self.class_info("std::map<string,linear_variable>")

        def pair(kv):
            self.class_info("std::pair<string,linear_variable>")
            self.string(kv[0])
            self.linear_variable(kv[1])

        self.collection(sorted(m.items(), key=lambda kv: kv[0].encode()), pair)

    def linear_variables(self, lst):
        self.class_info("linear_variables")
        self.class_info("std::vector<linear_variable>")
        self.collection(list(lst), self.linear_variable)

    def linear_variables_map(self, m):
        self.class_info("linear_variables_map")
        self.class_info("std::map<string,linear_variables>")

        def pair(kv):
            self.class_info("std::pair<string,linear_variables>")
            self.string(kv[0])
            self.linear_variables(kv[1])

        self.collection(sorted(m.items(), key=lambda kv: kv[0].encode()), pair)

    def primitive(self, kind, v):
        if kind == "bool":
            self.int(1 if v else 0)
        elif kind in ("int", "uint"):
            self.int(v)
        elif kind == "double":
            self.tok(_fmt_double(v))
        elif kind == "float":
            self.tok(_fmt_float(v))
        else:
            self.string(v)

    def item(self, kind):
        return (lambda v: self.primitive(kind, v)) if kind in _PRIMITIVE else getattr(self, kind)

    def vec_of(self, elem, items):
        if elem not in _ARITHMETIC:
            self.class_info(f"std::vector<{elem}>")
        self.collection(list(items), self.item(elem))

    def queue(self, kind, items):
        if kind not in _ARITHMETIC:
            self.class_info(f"std::deque<{kind}>")
        if kind.startswith("vec_"):
            elem = kind[4:]
            self.collection(list(items), lambda v: self.vec_of(elem, v))
        else:
            self.collection(list(items), self.item(kind))

    def text(self):
        return "".join(self.parts) + "\n"


class DeformationState:
    """One queue per value kind (``KINDS``); ``push`` appends, ``pop`` takes the front, as
    the reference's ``<<`` / ``>>`` (SerializeDeformationState.h:20-34)."""

    def __init__(self):
        self.queues = {k: deque() for k in KINDS}
        self.library_version = _LIBRARY_VERSION

    def push(self, kind, value):
        if kind not in self.queues:
            raise KeyError(f"unknown value kind {kind!r}")
        self.queues[kind].append(value)
        return self

    def pop(self, kind):
        q = self.queues[kind]
        if not q:
            raise IndexError(f"no {kind} value left in the state")
        return q.popleft()

    def __eq__(self, other):
        """operator== (SerializeDeformationState.h:120-145): every queue equal, exactly."""
        if not isinstance(other, DeformationState):
            return NotImplemented
        return states_close(self, other, 0.0, exact=True)

    def reset(self):
        for q in self.queues.values():
            q.clear()

    # archive text
    @classmethod
    def loads(cls, text):
        r = _Reader(text)
        if r.int() != len(_SIGNATURE) or r.tok() != _SIGNATURE:
            raise StateFormatError("not a Boost text archive")
        st = cls()
        st.library_version = r.int()
        tracking, version = r.int(), r.int()
        if tracking != 0 or version != _CLASS_VERSION:
            raise StateFormatError(f"DeformationState class record {tracking} {version}")
        for kind in KINDS:
            st.queues[kind] = deque(r.queue(kind))
        if r.i != len(r.toks):
            raise StateFormatError(f"{len(r.toks) - r.i} trailing tokens")
        return st

    def dumps(self):
        w = _Writer()
        w.int(len(_SIGNATURE))
        w.tok(_SIGNATURE)
        w.int(self.library_version)
        w.int(0)
        w.int(_CLASS_VERSION)
        for kind in KINDS:
            w.queue(kind, self.queues[kind])
        return w.text()

    def load(self, path):
        """Replaces the contents by the archive at ``path`` (DeformationState::load)."""
        with open(path) as f:
            other = DeformationState.loads(f.read())
        self.queues = other.queues
        self.library_version = other.library_version

    def save(self, path):
        with open(path, "w") as f:
            f.write(self.dumps())

    def save_and_reset(self, path):
        self.save(path)
        self.reset()


# FastGradientAscent / GradientAscent snapshot (FastGradientAscent.cxx:86-91, 275-281)
_OPTIMIZER_FIELDS = (
    ("computation_end_state", "bool"), ("iter", "uint"), ("freeze_direction_counter", "uint"),
    ("tau", "double"), ("step", "vector"), ("grad_sq_norms", "vector"),
    ("log_likelihood_history", "vec_vector"), ("fixed_effects", "linear_variable_map"),
    ("aux_fixed_effects", "linear_variable_map"), ("population_rer", "linear_variable_map"),
    ("aux_population_rer", "linear_variable_map"),
    ("individual_rer", "linear_variables_map"), ("aux_individual_rer", "linear_variables_map"),
    ("pop_grad", "linear_variable_map"), ("ind_grad", "linear_variables_map"),
)


def pop_fast_gradient_ascent(state):
    """Pops the FastGradientAscent snapshot in the order the estimator reads it
    (FastGradientAscent.cxx:78-91) and returns it as a dict."""
    return {name: state.pop(kind) for name, kind in _OPTIMIZER_FIELDS}


def push_fast_gradient_ascent(state, fields):
    """Pushes a FastGradientAscent snapshot in the estimator's save order
    (FastGradientAscent.cxx:295-300)."""
    for name, kind in _OPTIMIZER_FIELDS:
        state.push(kind, fields[name])
    return state


def scalar_close(x, y, tolerance, rtol=0.0):
    """ScalarCompare (Tolerance.hpp:20-31): |x - y|, rounded to float, at most the tolerance
    (an absolute criterion); ``rtol`` adds rtol * |y| to the bound (0: the reference's)."""
    bound = np.float32(tolerance) + np.float32(rtol * abs(float(y)))
    return bool(abs(np.float32(float(x) - float(y))) <= bound)


# GradientAscent snapshot (GradientAscent.cxx:78-80, 217-219)
_GRADIENT_ASCENT_FIELDS = (
    ("computation_end_state", "bool"), ("iter", "uint"), ("step", "vector"),
    ("log_likelihood_history", "vec_vector"), ("fixed_effects", "linear_variable_map"),
    ("population_rer", "linear_variable_map"), ("individual_rer", "linear_variables_map"),
    ("pop_grad", "linear_variable_map"), ("ind_grad", "linear_variables_map"),
)


def pop_gradient_ascent(state):
    """Pops the GradientAscent snapshot in the order the estimator reads it
    (GradientAscent.cxx:71-80)."""
    return {name: state.pop(kind) for name, kind in _GRADIENT_ASCENT_FIELDS}


def push_gradient_ascent(state, fields):
    for name, kind in _GRADIENT_ASCENT_FIELDS:
        state.push(kind, fields[name])
    return state


def states_close(a, b, tolerance, exact=False, rtol=0.0):
    """DeformationState::compare (SerializeDeformationState.h:151-200): integers, bools and
    strings equal; floating values within ``tolerance`` as ScalarCompare judges them (or
    equal, with ``exact``).  ``rtol`` widens each bound by rtol * |value of b|, for values
    whose magnitude puts one rounding step above an absolute tolerance."""

    def scalar(x, y):
        return float(x) == float(y) if exact else scalar_close(x, y, tolerance, rtol)

    def close(x, y):
        if isinstance(x, dict):
            return (isinstance(y, dict) and sorted(x) == sorted(y)  # std::map: keyed
                    and all(close(x[k], y[k]) for k in x))
        if isinstance(x, (list, tuple)):
            return (isinstance(y, (list, tuple)) and len(x) == len(y)
                    and all(close(u, v) for u, v in zip(x, y)))
        if isinstance(x, np.ndarray):
            return (isinstance(y, np.ndarray) and x.shape == y.shape
                    and all(scalar(u, v) for u, v in zip(x.ravel(), y.ravel())))
        if isinstance(x, (bool, int, str)) and not isinstance(x, float):
            return x == y
        return scalar(x, y)

    return all(close(list(a.queues[k]), list(b.queues[k])) for k in KINDS)
