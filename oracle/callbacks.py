"""Host-compiled callback systems for the oracle (test infrastructure only).

The oracle's own systems carry hand-written closed-form Jacobians, but a
forward-sensitivity solve also needs second derivatives of the callback
(J_y s + J_p pd, its df/dt and its y-derivative), which no closed form here
provides for an arbitrary system. So for those solves the oracle takes a
callback system: the traced callback's graph and its forward-mode derivative
graphs, emitted as plain C and compiled with gcc (contraction off, so each
operation rounds once, as on the GPU). What the oracle then restates
independently is the integration: the joint Rodas5P step, its
block-triangular stage solve, the error norm over the augmented vector. The
derivative graphs themselves are pinned apart from it -- against the
closed-form Jacobians bit for bit, against central differences, and, for
Robertson and a Van der Pol oscillator, against hand-derived closed-form
sensitivity callbacks (``oracle.ClosedFormCallbacks``) pointwise and through
a whole joint integration (tests/test_sensitivity.py).
"""

from __future__ import annotations

import ctypes
import hashlib
import os
import subprocess
import tempfile

from menodax_amd import _codegen as cg

_BUILD = os.path.join(tempfile.gettempdir(), "menodax_oracle_callbacks")


def _function(name, args, nodes_and_outs, g):
    roots = [n for nodes, _ in nodes_and_outs for n in nodes if n is not None]
    lines, _ = cg.Emitter(g).body(roots)
    body = list(lines)
    for nodes, target in nodes_and_outs:
        body += [f"  {target(j)} = {cg.value_of(x)};" for j, x in enumerate(nodes)]
    unused = [f"  (void){a.split()[-1].lstrip('*')};" for a in args.split(", ")]
    return "\n".join([f"void {name}({args}) {{"] + unused + body + ["}"])


def host_source(traced) -> str:
    g = traced.graph
    n = traced.n_vars
    parts = ["#include <math.h>", "#include <stdbool.h>"]
    parts.append(_function("mdx_f", "const double* y, double t, const double* p, double* out",
                           [(traced.outputs, lambda j: f"out[{j}]")], g))
    cols = [cg.tangents(g, traced.outputs, {g.y(c).id: 1.0}) for c in range(n)]
    dt = cg.tangents(g, traced.outputs, {g.t().id: 1.0})
    jac = [(col, (lambda c: (lambda r: f"J[{r * n + c}]"))(c)) for c, col in enumerate(cols)]
    parts.append(_function("mdx_jac", "const double* y, double t, const double* p, double* J, "
                           "double* dT", jac + [(dt, lambda j: f"dT[{j}]")], g))
    sn = cg.sensitivity_nodes(traced)
    base = "const double* y, double t, const double* p, const double* s, const double* pd"
    parts.append(_function("mdx_sens_f", base + ", double* out", [(sn.rhs, lambda j: f"out[{j}]")], g))
    parts.append(_function("mdx_sens_dt", base + ", double* out", [(sn.dt, lambda j: f"out[{j}]")], g))
    parts.append(_function("mdx_sens_cpl", base + ", const double* v, double* out",
                           [(sn.coupling, lambda j: f"out[{j}]")], g))
    return "\n\n".join(parts) + "\n"


class HostCallbacks:
    """A traced callback compiled for the host; ``addresses`` feeds the
    oracle's callback solver."""

    def __init__(self, ode_fn, n_vars, n_params):
        self.traced = cg.trace(ode_fn, n_vars, n_params)
        self.n, self.n_p = n_vars, n_params
        src = host_source(self.traced)
        key = hashlib.sha256(src.encode()).hexdigest()[:24]
        os.makedirs(_BUILD, exist_ok=True)
        so = os.path.join(_BUILD, f"cb_{key}.so")
        if not os.path.exists(so):
            c = os.path.join(_BUILD, f"cb_{key}.c")
            with open(c, "w") as f:
                f.write(src)
            tmp = so + f".{os.getpid()}"
            subprocess.run(["gcc", "-O1", "-std=c11", "-fPIC", "-shared", "-ffp-contract=off",
                            "-o", tmp, c, "-lm"], check=True)
            os.replace(tmp, so)
        self.lib = ctypes.CDLL(so)
        self.addresses = tuple(ctypes.cast(getattr(self.lib, name), ctypes.c_void_p).value
                               for name in ("mdx_f", "mdx_jac", "mdx_sens_f", "mdx_sens_dt",
                                            "mdx_sens_cpl"))
