"""thy_mass_noalloc (the allocation-free entry point) against thy_mass through the C-ABI:
bit-identical on two calls in a row, the second on a smaller grid, and again after
steps have changed the state."""
import pytest

from conftest import make_problem

pytestmark = pytest.mark.gpu


def test_mass_noalloc_equals_mass():
    from thysea import native
    big = make_problem(nx=300, ny=220, borders=(-1, -1, -1, -1), seed=31)
    small = make_problem(nx=70, ny=52, seed=32)
    solvers = []
    for pr in (big, small):  # the second call on a smaller input
        s = native.Solver(pr)
        solvers.append(s)
        m_alloc = s.mass()
        m_free = s.mass_noalloc()
        assert m_free == m_alloc and m_free > 0
    for s in solvers:  # the shared array keeps the last call's contents: results still match
        s.okada(s.prob.faults[0])
        dt = s.initial_dt()
        for _ in range(5):
            dt = s.step(dt)
        assert s.mass_noalloc() == s.mass()
        s.close()
