"""The product's optimisers (asc_pso, asc_simplex; C ABI, CPU) against the
oracle's restatement, bit-exact, on the oracle's exported test objectives.
The reference order (asynchronous PSO) is also pinned against the
reference's own compiled code in test_oracle_vs_reference.py; the
synchronous (batched-epoch) PSO is this library's variant and is pinned here
against the oracle's synchronous mode and by its single-particle identity.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O

CENTRE = np.array([0.7, -1.1, 0.4, 2.0, -1.5, 0.9])


def bits(a):
    a = np.asarray(a)
    return a.view(np.uint64 if a.dtype == np.float64 else np.uint32)


def point_cost(name):
    f = getattr(O.lib(), "orc_cost_" + name)

    def cost(pts):
        return [f(np.ascontiguousarray(p).ctypes.data_as(C.POINTER(C.c_double)), CENTRE.ctypes.data) for p in pts]

    return cost


@pytest.mark.parametrize("synchronous", [False, True])
@pytest.mark.parametrize("name,seed,epochs,stall,particles", [("ripple", 3, 60, 10, 100), ("bowl", 11, 40, 40, 37),
                                                              ("ripple", 5, 0, 10, 8)])
def test_pso_matches_oracle(asc, name, seed, epochs, stall, particles, synchronous):
    best, cost, ep = asc.optimizers.pso(point_cost(name), seed, -3.0, 3.0, epochs, stall, particles,
                                        synchronous=synchronous)
    ob, oc, oep = O.pso_fn(name, CENTRE, seed, -3.0, 3.0, epochs, stall, particles, synchronous=synchronous)
    assert ep == oep
    assert (bits(best) == bits(ob)).all()
    assert bits(np.float32(cost)) == bits(np.float32(oc))


def test_synchronous_pso_batches_whole_epochs(asc):
    calls = []

    def cost(pts):
        calls.append(len(pts))
        return point_cost("bowl")(pts)

    _, _, ep = asc.optimizers.pso(cost, 2, -3.0, 3.0, 5, 100, 64, synchronous=True)
    assert calls == [64] * ep


@pytest.mark.parametrize("name,ndim,ftol", [("ripple", 6, 1e-4), ("sinc", 2, 1e-9)])
def test_simplex_matches_oracle(asc, name, ndim, ftol):
    P0 = np.zeros((ndim + 1, ndim))
    for i in range(ndim):
        P0[i + 1, i] = 0.3
    if name == "sinc":
        P0 += 2.5
    cost = point_cost(name)
    Y0 = np.array(cost(P0))
    P, Y, it = asc.optimizers.simplex(cost, P0, Y0, ftol)
    Po, Yo, ito = O.amoeba_fn(name, CENTRE, P0, ftol)
    assert it == ito
    assert (bits(P) == bits(Po)).all() and (bits(Y) == bits(Yo)).all()


@pytest.mark.parametrize("method,max_iter,particles,synchronous", [(1, 100, 100, False), (0, 3, 24, False),
                                                                   (0, 3, 24, True)])
def test_frame_driver_matches_oracle(asc, orc, tiny, method, max_iter, particles, synchronous):
    """asc_optimize_frame_fn (the product's Tracker::optimize frame driver)
    fed the oracle's own minimizationFunc reproduces the oracle's
    optimize_frame bit for bit: PSO, store update, simplex, final pose."""
    from autoscoper_amd import synth

    wl = tiny["wl"]
    recs = [(asc.filters.to_records(wl.drr_filters), asc.filters.to_records(wl.rad_filters)) for _ in tiny["cams"]]
    S = orc.Scene(tiny["prep"], (synth.DRR_STEP, synth.DRR_INTENSITY, 0.0), tiny["cams"],
                  [r[1] for r in tiny["rads"]], recs, tiny["render"])
    pivot = orc.frame_from_xyzypr(list(tiny["prep"]["trans_center"]) + [0, 0, 0])
    store0 = orc.pose_store_set(wl.gt_poses[0] + np.array([1.5, -1.2, 0.8, 2.0, -1.5, 1.0]))

    def cost(manips, store):
        return [S.min_func(store, pivot, m) for m in manips]

    opts = dict(method=method, max_iter=max_iter, max_stall=3, seed=5)
    s_prod, st = asc.optimize_frame_fn(cost, store0, pivot, particles=particles, synchronous=synchronous, **opts)
    s_orc, ost, n = S.optimize_frame(store0, pivot, nparticles=particles, synchronous=synchronous, **opts)
    assert st.evaluations == n
    assert (bits(s_prod) == bits(s_orc)).all()


@pytest.mark.parametrize("method", [0, 1])
def test_speculative_simplex_keeps_the_trajectory(asc, orc, tiny, method):
    """Evaluating each simplex iteration's candidates (and a shrink's
    vertices) together gives the same final pose as the reference's
    one-at-a-time order, with more evaluations."""
    from autoscoper_amd import synth

    wl = tiny["wl"]
    recs = [(asc.filters.to_records(wl.drr_filters), asc.filters.to_records(wl.rad_filters)) for _ in tiny["cams"]]
    S = orc.Scene(tiny["prep"], (synth.DRR_STEP, synth.DRR_INTENSITY, 0.0), tiny["cams"],
                  [r[1] for r in tiny["rads"]], recs, tiny["render"])
    pivot = orc.frame_from_xyzypr(list(tiny["prep"]["trans_center"]) + [0, 0, 0])
    store0 = orc.pose_store_set(wl.gt_poses[0] + np.array([1.5, -1.2, 0.8, 2.0, -1.5, 1.0]))
    calls = []

    def cost(manips, store):
        calls.append(len(manips))
        return [S.min_func(store, pivot, m) for m in manips]

    opts = dict(method=method, max_iter=3, max_stall=3, seed=5, particles=24, synchronous=True)
    s_seq, st_seq = asc.optimize_frame_fn(cost, store0, pivot, **opts)
    n_seq = len(calls)
    s_spec, st_spec = asc.optimize_frame_fn(cost, store0, pivot, speculative=True, **opts)
    n_spec = len(calls) - n_seq
    assert (bits(s_spec) == bits(s_seq)).all()
    assert st_spec.simplex_iterations == st_seq.simplex_iterations
    assert st_spec.evaluations > st_seq.evaluations and n_spec < n_seq  # more points, fewer calls
