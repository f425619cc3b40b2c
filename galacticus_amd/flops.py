"""Static algorithmic FLOP count of the device RHS and of one Cash-Karp attempt.

The count is taken from the source text of ``rhs_part()`` in
``galacticus_amd/csrc/physics.hpp`` (both halves of the RHS, plus the sum
that joins them, and the inlined helpers), so it moves with the code.  The
halo scales, which a two-wave launch computes once per half, count once.
Weights:

* ``+ - * /`` on doubles, ``sqrt``, ``rsqrt``, ``fmin``, ``fmax``, ``fabs``: 1
* ``exp``, ``log``, ``log1p``, ``cbrt``, ``atan``, ``pow``, ``sinh``: 20 each
  (a fixed weight for a transcendental; the device implementations cost
  20-80 FP64 instructions)

Integer index arithmetic, comparisons and selects are not counted.  The RHS is
straight-line code, so every counted operation executes on every evaluation.

One Cash-Karp attempt on N=30 variables costs
``FLOP_step = 6*F_rhs + F_comb`` with ``F_comb = 52*N``: stage inputs
(1+2+3+4+5 products and sums per variable, times 2) plus the 5th-order
solution (4 terms) and the error (5 terms), times 2, plus 4 operations per
variable for the error norm.
"""
from __future__ import annotations

import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
PHYSICS = os.path.join(HERE, "csrc", "physics.hpp")
NVAR = 30
TRANSCENDENTAL_WEIGHT = 20
TRANSCENDENTALS = ("exp", "log", "log1p", "cbrt", "atan", "pow", "sinh")
UNIT_CALLS = ("sqrt", "rsqrt", "fmin", "fmax", "fabs")


def _function_body(src: str, signature_start: str) -> str:
    i = src.index(signature_start)
    j = src.index("{", i)
    depth = 0
    for k in range(j, len(src)):
        if src[k] == "{":
            depth += 1
        elif src[k] == "}":
            depth -= 1
            if depth == 0:
                return src[j:k + 1]
    raise ValueError("unbalanced braces")


def _strip(body: str) -> str:
    body = re.sub(r"//[^\n]*", "", body)
    body = re.sub(r"#pragma[^\n]*", "", body)
    return body


def _count_ops(body: str) -> int:
    b = _strip(body)
    # drop the test-mode branch and the zero-initialisation loop
    b = re.sub(r"if \(mode == GAL_PHYSICS_TEST\) \{.*?return;\s*\}", "", b, flags=re.S)
    b = re.sub(r"for \(int i = 0; i < kNV; \+\+i\) f\[i\] = 0\.0;", "", b)
    b = re.sub(r"\[[^\]]*\]", "", b)                 # array indices are not flops
    b = re.sub(r"\+\+|--", "", b)
    b = re.sub(r"<=|>=|==|!=|&&|\|\||->", " ", b)
    arith = len(re.findall(r"(?<![eE])[+\-*/](?!=)", b)) + len(re.findall(r"[+\-*/]=", b))
    trans = sum(len(re.findall(r"\b%s\(" % f, b)) for f in TRANSCENDENTALS)
    unit = sum(len(re.findall(r"\b%s\(" % f, b)) for f in UNIT_CALLS)
    return arith + unit + TRANSCENDENTAL_WEIGHT * trans


def rhs_flops() -> int:
    src = open(PHYSICS, encoding="utf-8").read()
    total = _count_ops(_function_body(src, "__device__ __forceinline__ void rhs_part("))
    total += NVAR    # f = f_galaxy + f_halo
    # helpers inlined once per evaluation
    total += _count_ops(_function_body(src, "__device__ __forceinline__ Halo halo("))
    total += _count_ops(_function_body(src, "__device__ __forceinline__ void isco("))
    total += _count_ops(_function_body(src, "__device__ __forceinline__ double cooling_function_ln("))
    total += 2 * _count_ops(_function_body(src, "__device__ __forceinline__ double hernquist_density("))
    total += 4 * 3   # pow_half: three multiplies (and one divide) per power law
    return total


def combination_flops(nvar: int = NVAR) -> int:
    return 52 * nvar


def attempt_flops() -> int:
    return 6 * rhs_flops() + combination_flops()


def evaluated_flops(rhs_evaluations: int, attempts: int) -> int:
    """Algorithmic FLOPs of a launch from its measured counters: every RHS
    evaluation (six for an accepted attempt, five for a rejected one, whose
    end-point derivative is never taken, plus the first of each node) and the
    stage combinations and error norm of every attempt."""
    return int(rhs_evaluations) * rhs_flops() + int(attempts) * combination_flops()


if __name__ == "__main__":
    print("F_rhs", rhs_flops(), "F_comb", combination_flops(), "FLOP_step", attempt_flops())
