"""A Jacobi sweep's rounds captured once as a HIP graph and replayed each sweep
(`qh_ctx_set_graphs`, the default) compute exactly what the same rounds launched
one by one compute: the replays run between the sweep's eager convergence update
and its host read-back, several per SVD, and every result is compared bit for bit
-- workload 1's 60 per-call updates and two batched layers at chi = 256."""
import numpy as np
import pytest

from oracle import workloads as wl

pytestmark = pytest.mark.gpu


def _run(qh, graphs, arrays, layers, maxdim, batched):
    ctx = qh.Context()
    ctx.set_graphs(graphs)
    psi = qh.from_internal(arrays, context=ctx)
    for layer in layers:
        if batched:
            qh.evolve_layer_([(psi, qh.gate(g, b, b + 1)) for b, g in layer], maxdim=maxdim, context=ctx)
        else:
            for b, g in layer:
                qh.evolve_(psi, qh.gate(g, b, b + 1), maxdim=maxdim)
    return psi.internal_arrays(), psi.lambdas()


def test_graph_replay_equals_direct_launches_config1(qh):
    arrays, schedule, tr = wl.config1()
    layers = [[(b, g)] for b, g in schedule]
    a = _run(qh, True, arrays, layers, tr["maxdim"], False)
    b = _run(qh, False, arrays, layers, tr["maxdim"], False)
    for x, y in zip(a[0], b[0]):
        assert np.array_equal(x, y)
    for x, y in zip(a[1], b[1]):
        assert (x is None and y is None) or np.array_equal(x, y)


def test_graph_replay_equals_direct_launches_chi256(qh):
    rng = np.random.default_rng(81)
    n = 18
    arrays = wl.left_canonize_normalize(wl.random_mps_arrays(rng, n, 2, wl.max_profile(n, 2, 256)))
    layers = [[(b, wl.haar_gate4(rng)) for b in wl.brickwork_bonds(n, k)] for k in range(2)]
    a = _run(qh, True, arrays, layers, 256, True)
    b = _run(qh, False, arrays, layers, 256, True)
    for x, y in zip(a[0], b[0]):
        assert np.array_equal(x, y)
