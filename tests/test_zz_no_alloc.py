"""kg_batch_step_noalloc: the batched step with its launch descriptors in the
library's own fixed device array.  Its results must equal kg_batch_step's bit
for bit, call after call, when a later call is smaller than an earlier one
(the array keeps the earlier call's descriptors beyond the new ones)."""
import numpy as np
import pytest

from kg_amd.batch import MemberSpec, batch_step, make_member
from kg_amd.initial_data import random_lm_coeffs

pytestmark = pytest.mark.gpu


def _specs():
    return [
        MemberSpec(24, 8, 16, 0.1, 1.0, 1.8, 1.8 + 24 * 0.35, 5.0, 1.5, random_lm_coeffs(1, 2, 0.3), 1),
        MemberSpec(40, 16, 32, 0.4, 0.6, 1.08, 1.08 + 40 * 0.3, 6.0, 1.5, random_lm_coeffs(2, 3, 0.3), 2),
        MemberSpec(17, 6, 12, 0.0, 1.9, 3.42, 3.42 + 17 * 0.4, 7.0, 1.2, random_lm_coeffs(3, 2, 0.3), 3),
        MemberSpec(70, 8, 20, 0.25, 1.3, 2.34, 2.34 + 70 * 0.2, 9.0, 2.0, random_lm_coeffs(4, 4, 0.3), 4),
    ]


@pytest.mark.parametrize("scheme,operators", [("rk4", "ref2"), ("rk4", "fd4"), ("pirk3", "ref2")])
def test_noalloc_batch_equals_batch_on_two_calls(scheme, operators):
    specs = _specs()
    print("no-alloc batch:", scheme, operators, flush=True)
    a = [make_member(s, scheme=scheme, operators=operators) for s in specs]
    b = [make_member(s, scheme=scheme, operators=operators) for s in specs]
    # first call: all four members, 3 steps
    batch_step(a, 3, noalloc=True)
    batch_step(b, 3)
    for q in range(len(specs)):
        assert np.array_equal(a[q].u_soa(), b[q].u_soa()), ("call 1", q)
        assert a[q].t == b[q].t
    # second call: a smaller batch (the two smallest grids), 2 steps
    sub = [2, 0]
    batch_step([a[q] for q in sub], 2, noalloc=True)
    batch_step([b[q] for q in sub], 2)
    for q in range(len(specs)):
        assert np.array_equal(a[q].u_soa(), b[q].u_soa()), ("call 2", q)
        assert a[q].t == b[q].t
    for m in a + b:
        m.close()


def test_noalloc_batch_after_a_larger_batch_matches_solo_runs():
    """A member stepped through the no-allocation entry point equals the same
    grid stepped alone, after a larger earlier call left more descriptors in
    the array."""
    specs = _specs()
    big = [make_member(s) for s in specs]
    batch_step(big, 1, noalloc=True)
    small = [make_member(specs[2])]
    alone = make_member(specs[2])
    batch_step(small, 4, noalloc=True)
    alone.step(4)
    assert np.array_equal(small[0].u_soa(), alone.u_soa())
    for m in big + small + [alone]:
        m.close()
