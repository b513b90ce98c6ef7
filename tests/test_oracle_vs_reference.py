"""The oracle's pose math and optimisers against the reference's own host
C++ (CoordFrame.cpp, PSO.cpp, DownhillSimplex.cpp, math/Quaternion.hpp),
compiled from its sources into oracle/_ref by `make -C oracle ref`.

Everything here is bit-exact: the oracle restates that code operation for
operation, and the product's host geometry (C ABI, CPU functions) is held
to the same bar. The reference sources exist only in the build container,
so these tests skip where oracle/_ref was not built.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O



class _LazyRef:
    """oracle/_ref/libref_host.so, opened by the first test that calls into
    it: importing or collecting this module maps nothing (a `-m gpu` run
    collects it too, and that library is built from the reference's
    sources, which stay in the build container)."""

    def __getattr__(self, name):
        lib = O.ref_lib()
        if lib is None:
            pytest.skip("oracle/_ref/libref_host.so not built (no reference sources)")
        return getattr(lib, name)


REF = _LazyRef()

_d = C.c_double
_P = C.POINTER


def _ptr(a):
    return a.ctypes.data_as(_P(_d))


def ref_call(name, *arrays, n_out=12):
    args = [np.ascontiguousarray(a, np.float64) for a in arrays]
    out = np.zeros(n_out)
    getattr(REF, name)(*[_ptr(a) for a in args], _ptr(out))
    return out


def bits(a):
    a = np.asarray(a)
    return a.view(np.uint64 if a.dtype == np.float64 else np.uint32)


def poses(n, seed, trans=60.0, rot=179.0):
    rng = np.random.default_rng(seed)
    p = np.concatenate([rng.uniform(-trans, trans, (n, 3)), rng.uniform(-rot, rot, (n, 3))], 1)
    # gimbal lock, zero and tiny rotations, half turns
    p[0, 3:] = (0, 90, 0)
    p[1, 3:] = (30, -90, 10)
    p[2, 3:] = (0, 0, 0)
    p[3, 3:] = (1e-9, 0, 0)
    p[4, 3:] = (180, 0, 180)
    return p


def test_frame_math_matches_reference(asc):
    g = asc.geometry
    for k, p in enumerate(poses(300, 11)):
        f_ref = ref_call("ref_frame_from_xyzypr", p)
        assert (bits(O.frame_from_xyzypr(p)) == bits(f_ref)).all()
        assert (bits(g.frame_from_xyzypr(p)) == bits(f_ref)).all()
        a_ref = ref_call("ref_frame_from_xyz_axis_angle", p)
        assert (bits(O.frame_from_xyz_axis_angle(p)) == bits(a_ref)).all()
        assert (bits(g.frame_from_xyz_axis_angle(p)) == bits(a_ref)).all()
        e_ref = ref_call("ref_frame_to_xyzypr", f_ref, n_out=6)
        assert (bits(O.frame_to_xyzypr(f_ref)) == bits(e_ref)).all()
        assert (bits(g.frame_to_xyzypr(f_ref)) == bits(e_ref)).all()
        m_ref = ref_call("ref_frame_mul", f_ref, a_ref)
        assert (bits(O.frame_mul(f_ref, a_ref)) == bits(m_ref)).all()
        assert (bits(g.frame_mul(f_ref, a_ref)) == bits(m_ref)).all()
        i_ref = ref_call("ref_frame_inverse", m_ref)
        assert (bits(O.frame_inverse(m_ref)) == bits(i_ref)).all()
        assert (bits(g.frame_inverse(m_ref)) == bits(i_ref)).all()
        if k == 4:
            # relative rotation of exactly pi: CoordFrame::linear_extrap then
            # rotates about an uninitialised axis (CoordFrame.cpp:456-463,
            # undefined behaviour); the oracle uses the x axis there
            continue
        x_ref = ref_call("ref_frame_linear_extrap", f_ref, a_ref)
        assert (bits(O.frame_linear_extrap(f_ref, a_ref)) == bits(x_ref)).all()
        assert (bits(g.frame_linear_extrap(f_ref, a_ref)) == bits(x_ref)).all()


def test_axis_angle_zero_rotation_matches_reference():
    for p in ([1.0, 2.0, 3.0, 0.0, 0.0, 0.0], [0.0, 0.0, 0.0, 1e-12, 0.0, 0.0], [5.0, 0.0, 0.0, 0.0, -720.0, 0.0]):
        p = np.array(p)
        assert (bits(O.frame_from_xyz_axis_angle(p)) == bits(ref_call("ref_frame_from_xyz_axis_angle", p))).all()


def test_float_pose_store_matches_reference(asc):
    """Single-key float curves + Quatf(yaw, pitch, roll).toEuler()."""
    for p in poses(300, 12):
        r = ref_call("ref_pose_store_roundtrip", p, n_out=6)
        assert (bits(O.pose_store_get(O.pose_store_set(p))) == bits(r)).all()
        assert (bits(asc.geometry.pose_store_get(asc.geometry.pose_store_set(p))) == bits(r)).all()


def test_min_func_composition_matches_reference(asc):
    rng = np.random.default_rng(13)
    for p in poses(200, 14):
        store = O.pose_store_set(p)
        base = ref_call("ref_pose_store_roundtrip", p, n_out=6)
        pivot = O.frame_from_xyzypr(np.concatenate([rng.uniform(-50, 50, 3), rng.uniform(-20, 20, 3)]))
        manip = np.concatenate([rng.uniform(-3, 3, 3), rng.uniform(-5, 5, 3)])
        r = ref_call("ref_compose_pose", base, pivot, manip, n_out=6)
        assert (bits(O.compose_pose(store, pivot, manip)) == bits(r)).all()
        assert (bits(asc.geometry.compose_pose(store, pivot, manip)) == bits(r)).all()


@pytest.mark.parametrize("name,seed,epochs,stall", [("ripple", 1, 100, 25), ("ripple", 77, 30, 5),
                                                    ("bowl", 5, 100, 25), ("ripple", 123456, 0, 25)])
def test_pso_matches_reference(name, seed, epochs, stall):
    centre = np.array([0.7, -1.1, 0.4, 2.0, -1.5, 0.9])
    best_r = np.zeros(6, np.float32)
    ncc_r = np.zeros(1, np.float32)
    calls = REF.ref_pso(O.cost_fn(name), centre.ctypes.data, seed, -3.0, 3.0, epochs, stall,
                        best_r.ctypes.data_as(_P(C.c_float)), ncc_r.ctypes.data_as(_P(C.c_float)))
    best_o, ncc_o, ep = O.pso_fn(name, centre, seed, -3.0, 3.0, epochs, stall, 100)
    assert calls == 100 * ep
    assert (bits(best_o) == bits(best_r)).all()
    assert bits(np.float32(ncc_o)) == bits(ncc_r[0])


@pytest.mark.parametrize("name,ndim,offset,ftol", [("ripple", 6, 0.1, 1e-3), ("ripple", 6, 0.5, 1e-6),
                                                   ("bowl", 6, 1.0, 1e-8), ("sinc", 2, 0.7, 1e-10)])
def test_simplex_matches_reference(name, ndim, offset, ftol):
    centre = np.array([0.7, -1.1, 0.4, 2.0, -1.5, 0.9])
    # Tracker::optimize's start simplex: vertex 0 at the origin, vertex i+1 at offset * e_i
    P0 = np.zeros((ndim + 1, ndim))
    for i in range(ndim):
        P0[i + 1, i] = offset
    if name == "sinc":
        P0 += 3.0
    Po, Yo, ito = O.amoeba_fn(name, centre, P0, ftol)
    Pr = P0.copy()
    Yr = np.array([getattr(O.lib(), "orc_cost_" + name)(_ptr(np.ascontiguousarray(P0[i])), centre.ctypes.data)
                   for i in range(ndim + 1)])
    itr = REF.ref_amoeba(O.cost_fn(name), centre.ctypes.data, _ptr(Pr), _ptr(Yr), ndim, ftol)
    assert ito == itr
    assert (bits(Po) == bits(Pr)).all()
    assert (bits(Yo) == bits(Yr)).all()


def _point_cost(name, centre):
    """Batch cost for the product optimisers: the oracle's exported objective
    evaluated point by point (the same function the reference calls)."""
    f = getattr(O.lib(), "orc_cost_" + name)

    def cost(pts):
        return [f(_ptr(np.ascontiguousarray(p)), centre.ctypes.data) for p in pts]

    return cost


@pytest.mark.parametrize("name,seed,epochs,stall", [("ripple", 1, 100, 25), ("ripple", 77, 30, 5),
                                                    ("bowl", 5, 100, 25)])
def test_product_pso_matches_reference(asc, name, seed, epochs, stall):
    centre = np.array([0.7, -1.1, 0.4, 2.0, -1.5, 0.9])
    best_r = np.zeros(6, np.float32)
    ncc_r = np.zeros(1, np.float32)
    calls = REF.ref_pso(O.cost_fn(name), centre.ctypes.data, seed, -3.0, 3.0, epochs, stall,
                        best_r.ctypes.data_as(_P(C.c_float)), ncc_r.ctypes.data_as(_P(C.c_float)))
    best, cost, ep = asc.optimizers.pso(_point_cost(name, centre), seed, -3.0, 3.0, epochs, stall, 100)
    assert calls == 100 * ep
    assert (bits(best) == bits(best_r)).all()
    assert bits(np.float32(cost)) == bits(ncc_r[0])


def test_product_synchronous_pso_single_particle_is_reference_order(asc):
    """With one particle the synchronous epoch is the asynchronous one."""
    centre = np.array([0.7, -1.1, 0.4, 2.0, -1.5, 0.9])
    a = asc.optimizers.pso(_point_cost("ripple", centre), 9, -3.0, 3.0, 40, 10, 1, synchronous=False)
    s = asc.optimizers.pso(_point_cost("ripple", centre), 9, -3.0, 3.0, 40, 10, 1, synchronous=True)
    assert (bits(a[0]) == bits(s[0])).all() and a[1:] == s[1:]


@pytest.mark.parametrize("name,ndim,offset,ftol", [("ripple", 6, 0.1, 1e-3), ("ripple", 6, 0.5, 1e-6),
                                                   ("bowl", 6, 1.0, 1e-8), ("sinc", 2, 0.7, 1e-10)])
def test_product_simplex_matches_reference(asc, name, ndim, offset, ftol):
    centre = np.array([0.7, -1.1, 0.4, 2.0, -1.5, 0.9])
    P0 = np.zeros((ndim + 1, ndim))
    for i in range(ndim):
        P0[i + 1, i] = offset
    if name == "sinc":
        P0 += 3.0
    cost = _point_cost(name, centre)
    Y0 = np.array(cost(P0))
    Pr, Yr = P0.copy(), Y0.copy()
    itr = REF.ref_amoeba(O.cost_fn(name), centre.ctypes.data, _ptr(Pr), _ptr(Yr), ndim, ftol)
    P, Y, it = asc.optimizers.simplex(cost, P0, Y0, ftol)
    assert it == itr
    assert (bits(P) == bits(Pr)).all()
    assert (bits(Y) == bits(Yr)).all()


def test_product_optimiser_cost_error_propagates(asc):
    def bad(pts):
        raise RuntimeError("cost failed")

    with pytest.raises(Exception):
        asc.optimizers.pso(bad, 1, -1.0, 1.0, 3, 3, 4)
