"""The continuous-batching serving loop: every LP's result is the same alone,
in the mix and in a single solve (bitwise: each LP keeps its own SpMV plan and
its own reductions), infeasible and failing LPs retire without disturbing the
others, and admission respects the HBM budget."""
import numpy as np
import pytest

from conftest import netlib
from oracle import oracle as O
from saddlepoint_amd import generate
from saddlepoint_amd.backend import default_params, solve
from saddlepoint_amd.problem import LpProblem
from saddlepoint_amd.serving import DONE, FAILED, Server

pytestmark = pytest.mark.gpu

MIX = ["afiro", "adlittle", "israel", "sctest", "etamacro", "standata", "forest6", "woodinfe"]


def problems():
    ps = [(name, netlib(name)) for name in MIX]
    ps.append(("config0", generate.config0()))
    ps.append(("serving-7", generate.serving_lp(7, m_lo=1000, m_hi=5000)))
    return ps


@pytest.mark.parametrize("precision", [0, 1])
def test_alone_and_in_the_mix_are_bitwise_equal(gpu, precision):
    items = problems()
    mix = Server(gpu)
    ids = [mix.submit(p, default_params(precision=precision)) for _, p in items]
    mix.drain(ids)
    together = {name: mix.result(i) for (name, _), i in zip(items, ids)}
    mix.close()
    for name, p in items:
        alone_srv = Server(gpu)
        j = alone_srv.submit(p, default_params(precision=precision))
        alone_srv.drain([j])
        a = alone_srv.result(j).result
        alone_srv.close()
        t = together[name].result
        print(f"{name:>10}: {a.status} {a.iterations} it alone; {t.status} {t.iterations} it in the mix", flush=True)
        assert a.status == t.status and a.iterations == t.iterations and a.restarts == t.restarts
        assert np.array_equal(a.x, t.x) and np.array_equal(a.y, t.y)


@pytest.mark.parametrize("precision", [0, 1])
def test_served_results_are_bitwise_single_solves(gpu, precision):
    """The batched plain iterations and the batched check phases run each LP's
    own partition and summation order, so a served LP takes exactly the
    iterations of its single solve: same status, counts and bits."""
    items = problems()
    s = Server(gpu)
    ids = [s.submit(p, default_params(precision=precision)) for _, p in items]
    s.drain(ids)
    for (name, p), i in zip(items, ids):
        r = s.result(i).result
        g = solve(gpu, p, default_params(precision=precision))
        print(f"{name:>10}: served {r.status} {r.iterations} it; solve {g.status} {g.iterations} it; "
              f"bitwise {np.array_equal(r.x, g.x)}", flush=True)
        assert r.status == g.status and r.iterations == g.iterations and r.restarts == g.restarts, name
        assert np.array_equal(r.x, g.x) and np.array_equal(r.y, g.y), name
        assert r.objective == g.objective
        if r.status == "Optimal":
            O.check(p, r.x, r.y).verify(2e-6)
    s.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_chunk_schedule_changes_timing_not_results(gpu, precision):
    """With chunk_work = 1 a chunk runs the cheapest LP and those left out of
    the last 8 chunks, so the batch membership changes from chunk to chunk
    (tables rebuilt, graphs captured and dropped); every LP still ends with the
    bits it gets when every LP runs in every chunk."""
    items = problems()
    res = {}
    for work in (1, -1):
        s = Server(gpu, chunk_work=work)
        ids = [s.submit(p, default_params(precision=precision)) for _, p in items]
        s.drain(ids)
        res[work] = [s.result(i).result for i in ids]
        s.close()
    for (name, _), a, b in zip(items, res[1], res[-1]):
        print(f"{name:>10}: {a.status} {a.iterations} it (chunk_work 1) / {b.iterations} it (every chunk)", flush=True)
        assert a.status == b.status and a.iterations == b.iterations and a.restarts == b.restarts, name
        assert np.array_equal(a.x, b.x) and np.array_equal(a.y, b.y), name


def test_budget_limits_admission(gpu):
    items = problems()[:4]
    s = Server(gpu, budget_bytes=1, max_active=8)  # one LP at a time (an empty server always admits one)
    ids = [s.submit(p) for _, p in items]
    seen = 0
    for _ in range(100000):
        active, waiting, finished = s.run_chunk()
        assert active <= 1
        seen = max(seen, active)
        if finished == len(ids):
            break
    assert all(s.state(i) == DONE for i in ids)
    s.close()


def test_failures_are_isolated(gpu):
    good = netlib("afiro")
    bad = netlib("afiro")
    bad.c = bad.c.copy()
    bad.c[0] = np.nan  # ProblemError::NaN at validation
    s = Server(gpu)
    i_good, i_bad = s.submit(good), s.submit(bad)
    s.drain([i_good, i_bad])
    assert s.state(i_bad) == FAILED and s.state(i_good) == DONE
    from saddlepoint_amd._lib import SP_ERR_PROBLEM, BackendError

    with pytest.raises(BackendError) as e:
        s.result(i_bad)
    assert e.value.code == SP_ERR_PROBLEM
    assert s.result(i_good).result.status == "Optimal"
    s.close()


def test_graphs_and_eager_launches_agree(gpu):
    items = problems()[:5]
    res = []
    for graphs in (True, False):
        s = Server(gpu, use_graphs=graphs)
        ids = [s.submit(p) for _, p in items]
        s.drain(ids)
        res.append([s.result(i).result for i in ids])
        s.close()
    for a, b in zip(*res):
        assert a.iterations == b.iterations and np.array_equal(a.x, b.x)
