"""The benchmark's input generators (qrochet_hip.workloads) produce exactly the
arrays of the test-side generators (oracle.workloads) for every workload."""
import numpy as np

from oracle import workloads as ow
from qrochet_hip import workloads as pw


def _same_arrays(a, b):
    assert len(a) == len(b)
    for x, y in zip(a, b):
        assert x.shape == y.shape and np.array_equal(x, y)


def _same_layers(a, b):
    assert len(a) == len(b)
    for la, lb in zip(a, b):
        assert [x for x, _ in la] == [y for y, _ in lb]
        for (_, ga), (_, gb) in zip(la, lb):
            assert np.array_equal(ga, gb)


def test_config1_inputs_identical():
    a1, s1, t1 = ow.config1()
    a2, s2, t2 = pw.config1()
    _same_arrays(a1, a2)
    _same_layers([s1], [s2])
    assert t1 == t2


def test_config2_inputs_identical():
    a1, l1, t1, (j1, h1) = ow.config2(steps=3)
    a2, l2, t2, (j2, h2) = pw.config2(steps=3)
    _same_arrays(a1, a2)
    _same_layers(l1, l2)
    assert t1 == t2 and np.array_equal(j1, j2) and np.array_equal(h1, h2)


def test_config3_inputs_identical():
    i1 = ow.config3(count=6, depth=2, nmax=12)
    i2 = pw.config3(count=6, depth=2, nmax=12)
    for x, y in zip(i1, i2):
        assert (x["n"], x["chi"]) == (y["n"], y["chi"])
        _same_arrays(x["arrays"], y["arrays"])
        _same_layers(x["layers"], y["layers"])


def test_mpo_generator_identical():
    from oracle import mpo as om

    a = om.random_hermitian_mpo(np.random.default_rng(9), 5, 2, 3)
    b = pw.random_hermitian_mpo(np.random.default_rng(9), 5, 2, 3)
    _same_arrays(a, b)
