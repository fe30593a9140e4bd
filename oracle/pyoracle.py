"""ctypes front end of oracle/_build/liboracle.so -- TEST INFRASTRUCTURE ONLY.

Used by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline legs
(cpu_baseline, cpu_table) as the parity checker and the timed CPU path; never
by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path
from typing import Optional

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "_build" / "liboracle.so"

_f32p = C.POINTER(C.c_float)
_f64p = C.POINTER(C.c_double)
_i32p = C.POINTER(C.c_int)
_u8p = C.POINTER(C.c_uint8)


class OrcRobot(C.Structure):
    _fields_ = [("n_dof", C.c_int), ("n_joints", C.c_int),
                ("fixed_transforms", _f32p), ("joint_types", _i32p), ("joint_parents", _i32p),
                ("joint_id_to_dof", _i32p), ("dfs_order", _i32p),
                ("n_spheres", C.c_int), ("spheres", _f32p), ("sphere_to_joint", _i32p),
                ("n_approx", C.c_int), ("approx_spheres", _f32p), ("approx_sphere_to_joint", _i32p),
                ("approx_fixed_transforms", _f32p), ("approx_joint_types", _i32p),
                ("n_self_ranges", C.c_int), ("self_ranges", _i32p),
                ("n_approx_self_ranges", C.c_int), ("approx_self_ranges", _i32p)]


class OrcScene(C.Structure):
    _fields_ = [("n_obbs", C.c_int), ("obb_dims", _f32p), ("obb_pose", _f32p), ("obb_enable", _u8p),
                ("sphere_obb_acm", _u8p), ("sphere_obb_acm_approx", _u8p),
                ("n_voxel_layers", C.c_int), ("voxel_params", _f32p), ("voxel_pose", _f32p),
                ("voxel_enable", _u8p), ("voxel_features", _f32p), ("max_voxels_per_layer", C.c_int),
                ("sphere_acm", _u8p), ("esdf_interp", C.c_int)]


class OrcCostParams(C.Structure):
    _fields_ = [("world_margin", C.c_float), ("self_margin", C.c_float), ("weight", C.c_float),
                ("use_obb", C.c_int), ("use_esdf", C.c_int), ("use_self", C.c_int)]


class OrcEdgeParams(C.Structure):
    _fields_ = [("granularity", C.c_int), ("collision_margin", C.c_float), ("check_self", C.c_int),
                ("compute_clearance", C.c_int), ("resolution", C.c_float)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run `make -C oracle`)")
        _lib = C.CDLL(str(LIB_PATH))
    return _lib


def _ptr(a, t):
    return None if a is None else a.ctypes.data_as(t)


def _c(a, dtype):
    return np.ascontiguousarray(np.asarray(a, dtype=dtype))


def cost_params(world_margin=0.02, self_margin=0.01, weight=1.0, use_obb=True, use_esdf=True,
                use_self=True) -> OrcCostParams:
    return OrcCostParams(world_margin, self_margin, weight, int(use_obb), int(use_esdf), int(use_self))


class Oracle:
    """Binds one robot and one scene; every call takes FP32 inputs."""

    def __init__(self, robot, scene=None, nthreads: int = 0):
        self.robot = robot
        self.nthreads = int(nthreads)
        self._keep = []
        k = self._keep.append
        self.r = OrcRobot()
        r = self.r
        r.n_dof, r.n_joints = robot.n_dof, robot.n_joints
        for name, arr, t in (("fixed_transforms", robot.fixed_transforms, _f32p),
                             ("joint_types", robot.joint_types, _i32p),
                             ("joint_parents", robot.joint_parents, _i32p),
                             ("joint_id_to_dof", robot.joint_id_to_dof, _i32p),
                             ("dfs_order", robot.dfs_order, _i32p),
                             ("spheres", robot.spheres, _f32p),
                             ("sphere_to_joint", robot.sphere_to_joint, _i32p),
                             ("approx_spheres", robot.approx_spheres, _f32p),
                             ("approx_sphere_to_joint", robot.approx_sphere_to_joint, _i32p),
                             ("self_ranges", robot.self_ranges, _i32p),
                             ("approx_self_ranges", robot.approx_self_ranges, _i32p)):
            a = _c(arr, np.float32 if t is _f32p else np.int32)
            k(a)
            setattr(r, name, _ptr(a, t))
        r.n_spheres, r.n_approx = robot.n_spheres, robot.n_approx
        r.n_self_ranges = int(robot.self_ranges.shape[0])
        r.n_approx_self_ranges = int(robot.approx_self_ranges.shape[0])
        aft = getattr(robot, "approx_fixed_transforms", None)
        aty = getattr(robot, "approx_joint_types", None)
        r.approx_fixed_transforms = None
        r.approx_joint_types = None
        if aft is not None:
            a = _c(aft, np.float32)
            k(a)
            r.approx_fixed_transforms = _ptr(a, _f32p)
        if aty is not None:
            a = _c(aty, np.int32)
            k(a)
            r.approx_joint_types = _ptr(a, _i32p)
        self.set_scene(scene)

    def set_scene(self, scene):
        from prismplan_amd.planning_scene import HostScene
        scene = scene if scene is not None else HostScene()
        self.scene = scene
        s = OrcScene()
        keep = []
        def put(name, arr, dtype, t):
            if arr is None:
                setattr(s, name, None)
                return
            a = _c(arr, dtype)
            keep.append(a)
            setattr(s, name, _ptr(a, t))
        s.n_obbs = scene.n_obbs
        put("obb_dims", scene.obb_dims, np.float32, _f32p)
        put("obb_pose", scene.obb_pose, np.float32, _f32p)
        put("obb_enable", scene.obb_enable, np.uint8, _u8p)
        put("sphere_obb_acm", scene.sphere_obb_acm, np.uint8, _u8p)
        put("sphere_obb_acm_approx", scene.sphere_obb_acm_approx, np.uint8, _u8p)
        s.n_voxel_layers = scene.n_voxel_layers
        put("voxel_params", scene.voxel_params, np.float32, _f32p)
        put("voxel_pose", scene.voxel_pose, np.float32, _f32p)
        put("voxel_enable", scene.voxel_enable, np.uint8, _u8p)
        put("voxel_features", scene.voxel_features, np.float32, _f32p)
        s.max_voxels_per_layer = scene.max_voxels_per_layer
        put("sphere_acm", scene.sphere_acm, np.uint8, _u8p)
        s.esdf_interp = int(scene.esdf_interp)
        self.s = s
        self._scene_keep = keep

    # ---- kinematics ----
    def fk_spheres(self, q, precision="f64") -> np.ndarray:
        q = _c(q, np.float32).reshape(-1, self.robot.n_dof)
        if precision == "f32":
            out = np.zeros((q.shape[0], self.robot.n_spheres, 3), np.float32)
            lib().orc_fk_spheres_f32(C.byref(self.r), _ptr(q, _f32p), q.shape[0], _ptr(out, _f32p))
            return out
        out = np.zeros((q.shape[0], self.robot.n_spheres, 3), np.float64)
        lib().orc_fk_spheres_f64(C.byref(self.r), _ptr(q, _f32p), q.shape[0], _ptr(out, _f64p))
        return out

    def fk_frames(self, q) -> np.ndarray:
        q = _c(q, np.float32).reshape(-1, self.robot.n_dof)
        out = np.zeros((q.shape[0], self.robot.n_joints, 3, 4), np.float64)
        lib().orc_fk_frames_f64(C.byref(self.r), _ptr(q, _f32p), q.shape[0], _ptr(out, _f64p))
        return out

    # ---- predicates ----
    def check_collision(self, q, margin=0.0, check_self=False, mode=0, precision="f64"):
        q = _c(q, np.float32).reshape(-1, self.robot.n_dof)
        n = q.shape[0]
        out = np.zeros(n, np.uint8)
        if precision == "f64":
            clr = np.zeros(n, np.float64)
            lib().orc_check_collision_f64(C.byref(self.r), C.byref(self.s), _ptr(q, _f32p), n,
                                          C.c_float(margin), int(check_self), int(mode),
                                          _ptr(out, _u8p), _ptr(clr, _f64p), self.nthreads)
        else:
            clr = np.zeros(n, np.float32)
            lib().orc_check_collision_f32(C.byref(self.r), C.byref(self.s), _ptr(q, _f32p), n,
                                          C.c_float(margin), int(check_self), int(mode),
                                          _ptr(out, _u8p), _ptr(clr, _f32p), self.nthreads)
        return out.astype(bool), clr

    def check_collision_f32_timed(self, q, margin=0.0, check_self=False):
        """FP32 two-phase check without the clearance pass (the timed CPU path)."""
        q = _c(q, np.float32).reshape(-1, self.robot.n_dof)
        out = np.zeros(q.shape[0], np.uint8)
        lib().orc_check_collision_f32(C.byref(self.r), C.byref(self.s), _ptr(q, _f32p), q.shape[0],
                                      C.c_float(margin), int(check_self), 0, _ptr(out, _u8p), None,
                                      self.nthreads)
        return out.astype(bool)

    def gated_flops(self, q, margin=0.0):
        q = _c(q, np.float32).reshape(-1, self.robot.n_dof)
        out = np.zeros(q.shape[0], np.float64)
        lib().orc_gated_flops_f64(C.byref(self.r), C.byref(self.s), _ptr(q, _f32p), q.shape[0], C.c_float(margin),
                                  _ptr(out, _f64p), self.nthreads)
        return out

    def min_clearance(self, q, margin=0.0, include_obb=False, with_amb=False, with_low=False):
        """Per-configuration min clearance; with_amb also returns the value-
        ambiguity flags (an ESDF lookup at a face where the value jumps and
        the jump can move the minimum), with_low the lowest minimum the other
        sides of those jumps can give."""
        q = _c(q, np.float32).reshape(-1, self.robot.n_dof)
        out = np.zeros(q.shape[0], np.float64)
        amb = np.zeros(q.shape[0], np.uint8)
        low = np.zeros(q.shape[0], np.float64)
        lib().orc_min_clearance_amb_f64(C.byref(self.r), C.byref(self.s), _ptr(q, _f32p), q.shape[0],
                                        C.c_float(margin), int(include_obb), _ptr(out, _f64p), _ptr(amb, _u8p),
                                        _ptr(low, _f64p), self.nthreads)
        if with_low:
            return out, amb.astype(bool), low
        return (out, amb.astype(bool)) if with_amb else out

    def hinge_cost_amb(self, q, params: OrcCostParams, with_terms=False):
        """FP64 state cost and its value-ambiguity flags (and, with_terms,
        the number of non-zero hinge terms in each cost)."""
        q = _c(q, np.float32).reshape(-1, self.robot.n_dof)
        out = np.zeros(q.shape[0], np.float64)
        amb = np.zeros(q.shape[0], np.uint8)
        nterm = np.zeros(q.shape[0], np.int32)
        lib().orc_hinge_cost_amb_f64(C.byref(self.r), C.byref(self.s), _ptr(q, _f32p), q.shape[0], C.byref(params),
                                     _ptr(out, _f64p), _ptr(amb, _u8p), _ptr(nterm, _i32p), self.nthreads)
        return (out, amb.astype(bool), nterm) if with_terms else (out, amb.astype(bool))

    def hinge_terms(self, q, params: OrcCostParams):
        """Number of non-zero hinge terms (OBB, ESDF, self) in each FP64 state cost."""
        return self.hinge_cost_amb(q, params, with_terms=True)[2]

    def hinge_cost(self, q, params: OrcCostParams, precision="f64"):
        q = _c(q, np.float32).reshape(-1, self.robot.n_dof)
        n = q.shape[0]
        if precision == "f64":
            out = np.zeros(n, np.float64)
            lib().orc_hinge_cost_f64(C.byref(self.r), C.byref(self.s), _ptr(q, _f32p), n, C.byref(params),
                                     _ptr(out, _f64p), self.nthreads)
        else:
            out = np.zeros(n, np.float32)
            lib().orc_hinge_cost_f32(C.byref(self.r), C.byref(self.s), _ptr(q, _f32p), n, C.byref(params),
                                     _ptr(out, _f32p), self.nthreads)
        return out

    def hinge_grad(self, q, params: OrcCostParams, tol=1e-5):
        q = _c(q, np.float32).reshape(-1, self.robot.n_dof)
        n = q.shape[0]
        cost = np.zeros(n, np.float64)
        grad = np.zeros((n, self.robot.n_dof), np.float64)
        kink = np.zeros(n, np.int32)
        lib().orc_hinge_grad_f64(C.byref(self.r), C.byref(self.s), _ptr(q, _f32p), n, C.byref(params),
                                 _ptr(cost, _f64p), _ptr(grad, _f64p), _ptr(kink, _i32p), C.c_double(tol),
                                 self.nthreads)
        return cost, grad, kink.astype(bool)

    def hinge_grad_f32_timed(self, q, params: OrcCostParams):
        """FP32 hinge cost and gradient per configuration (the timed CPU path)."""
        q = _c(q, np.float32).reshape(-1, self.robot.n_dof)
        n = q.shape[0]
        cost = np.zeros(n, np.float32)
        grad = np.zeros((n, self.robot.n_dof), np.float32)
        lib().orc_hinge_grad_f32(C.byref(self.r), C.byref(self.s), _ptr(q, _f32p), n, C.byref(params),
                                 _ptr(cost, _f32p), _ptr(grad, _f32p), self.nthreads)
        return cost, grad

    # ---- trajectories (ragged [B] x [D][T_b]) ----
    def traj_cost_grad(self, theta_flat, off, params: OrcCostParams, tol=1e-5):
        th = _c(theta_flat, np.float32)
        off = _c(off, np.int32)
        B = off.shape[0] - 1
        total = np.zeros(B, np.float64)
        grad = np.zeros(th.shape, np.float64)
        kink = np.zeros(int(off[-1]), np.int32)
        lib().orc_traj_cost_grad_f64(C.byref(self.r), C.byref(self.s), _ptr(th, _f32p), _ptr(off, _i32p), B,
                                     C.byref(params), _ptr(total, _f64p), _ptr(grad, _f64p), _ptr(kink, _i32p),
                                     C.c_double(tol), self.nthreads)
        return total, grad, kink.astype(bool)

    def chomp_fd_grad(self, theta_flat, off, params: OrcCostParams, eps=1e-3, precision="f64"):
        th = _c(theta_flat, np.float32)
        off = _c(off, np.int32)
        B = off.shape[0] - 1
        if precision == "f64":
            grad = np.zeros(th.shape, np.float64)
            lib().orc_chomp_fd_grad_f64(C.byref(self.r), C.byref(self.s), _ptr(th, _f32p), _ptr(off, _i32p), B,
                                        C.byref(params), C.c_float(eps), _ptr(grad, _f64p), self.nthreads)
        else:
            grad = np.zeros(th.shape, np.float32)
            lib().orc_chomp_fd_grad_f32(C.byref(self.r), C.byref(self.s), _ptr(th, _f32p), _ptr(off, _i32p), B,
                                        C.byref(params), C.c_float(eps), _ptr(grad, _f32p), self.nthreads)
        return grad

    # ---- edges ----
    def edge_eval(self, frm, to, granularity=32, margin=0.0, check_self=False, compute_clearance=False,
                  resolution=0.0):
        frm = _c(frm, np.float32).reshape(-1, self.robot.n_dof)
        to = _c(to, np.float32).reshape(-1, self.robot.n_dof)
        E = frm.shape[0]
        p = OrcEdgeParams(int(granularity), float(margin), int(check_self), int(compute_clearance),
                          float(resolution))
        valid = np.zeros(E, np.uint8)
        cost = np.zeros(E, np.float32)
        clr = np.zeros(E, np.float64)
        lib().orc_edge_eval_f64(C.byref(self.r), C.byref(self.s), _ptr(frm, _f32p), _ptr(to, _f32p), E,
                                C.byref(p), _ptr(valid, _u8p), _ptr(cost, _f32p), _ptr(clr, _f64p), self.nthreads)
        return valid.astype(bool), cost, clr

    def edge_batch(self, frm, to, n_states, skip_first=False, margin=0.0, check_self=False, precision="f64"):
        frm = _c(frm, np.float32).reshape(-1, self.robot.n_dof)
        to = _c(to, np.float32).reshape(-1, self.robot.n_dof)
        ns = _c(n_states, np.int32)
        E = frm.shape[0]
        fh = np.zeros(E, np.int32)
        fn = lib().orc_edge_batch_f64 if precision == "f64" else lib().orc_edge_batch_f32
        fn(C.byref(self.r), C.byref(self.s), _ptr(frm, _f32p), _ptr(to, _f32p), _ptr(ns, _i32p), E,
           int(skip_first), C.c_float(margin), int(check_self), _ptr(fh, _i32p), self.nthreads)
        return fh

    def stomp_costs(self, all_scene, obb_off, rollouts, T_p, roff, R, params, w_u, dt, precision="f64",
                    with_terms=False):
        """all_scene: HostScene with every problem's OBBs concatenated.
        with_terms (FP64 only): also the non-zero hinge terms per state."""
        self.set_scene(all_scene)
        ro = _c(rollouts, np.float32)
        obb_off = _c(obb_off, np.int32)
        T_p = _c(T_p, np.int32)
        roff = _c(roff, np.int32)
        P = T_p.shape[0]
        n_wp = int(roff[-1]) if roff.shape[0] > P else int(np.sum(T_p) * R)
        dt_ = np.float64 if precision == "f64" else np.float32
        state = np.zeros(n_wp, dt_)
        total = np.zeros(P * R, dt_)
        tp = _f64p if precision == "f64" else _f32p
        args = (C.byref(self.r), C.byref(self.s), _ptr(obb_off, _i32p), _ptr(ro, _f32p), _ptr(T_p, _i32p),
                _ptr(roff, _i32p), P, int(R), C.byref(params), C.c_double(w_u), C.c_double(dt),
                _ptr(state, tp), _ptr(total, tp), self.nthreads)
        if precision == "f64":
            nterm = np.zeros(n_wp, np.int32)
            lib().orc_stomp_costs_f64(*args, _ptr(nterm, _i32p))
            return (state, total, nterm) if with_terms else (state, total)
        lib().orc_stomp_costs_f32(*args)
        return state, total


def edge_states(frm, to, n_states, skip_first=False) -> np.ndarray:
    frm = _c(frm, np.float32)
    to = _c(to, np.float32)
    D = frm.shape[1]
    ns = _c(n_states, np.int32)
    k0 = 1 if skip_first else 0
    total = int(np.sum(np.maximum(ns - k0, 0)))
    out = np.zeros((total, D), np.float32)
    lib().orc_edge_states_f32(_ptr(frm, _f32p), _ptr(to, _f32p), _ptr(ns, _i32p), frm.shape[0], D,
                              int(skip_first), _ptr(out, _f32p))
    return out


def edge_granularity(frm, to, granularity=32, resolution=0.0):
    frm = _c(frm, np.float32)
    to = _c(to, np.float32)
    ln = C.c_float(0)
    g = lib().orc_edge_granularity(_ptr(frm, _f32p), _ptr(to, _f32p), frm.shape[0], int(granularity),
                                   C.c_float(resolution), C.byref(ln))
    return int(g), float(ln.value)


def control_R(T: int, dt: float) -> np.ndarray:
    out = np.zeros((T, T), np.float64)
    lib().orc_control_R_f64(int(T), C.c_double(dt), _ptr(out, _f64p))
    return out


def control_cost(theta_flat, off, D, dt, w) -> np.ndarray:
    th = _c(theta_flat, np.float32)
    off = _c(off, np.int32)
    B = off.shape[0] - 1
    out = np.zeros(B, np.float64)
    lib().orc_control_cost_f64(_ptr(th, _f32p), _ptr(off, _i32p), B, int(D), C.c_double(dt), C.c_double(w),
                               _ptr(out, _f64p))
    return out


def esdf_from_obbs(obb_dims, obb_pose, voxel_params, voxel_pose, nthreads=0) -> np.ndarray:
    from prismplan_amd.planning_scene import grid_dims
    d = _c(obb_dims, np.float32)
    p = _c(obb_pose, np.float32)
    vp = _c(voxel_params, np.float32).reshape(4)
    vpo = _c(voxel_pose, np.float32).reshape(8)
    nx, ny, nz = grid_dims(vp)
    out = np.zeros(nx * ny * nz, np.float64)
    lib().orc_esdf_from_obbs_f64(_ptr(d, _f32p), _ptr(p, _f32p), d.shape[0], _ptr(vp, _f32p), _ptr(vpo, _f32p),
                                 _ptr(out, _f64p), int(nthreads))
    return out


def _setup_restypes():
    L = lib()
    for name in ("orc_sphere_obb_clearance_f64", "orc_sphere_esdf_clearance_f64"):
        getattr(L, name).restype = C.c_double
    for name in ("orc_sphere_obb_clearance_f64",):
        getattr(L, name).argtypes = [_f32p, _f32p, C.c_double, C.c_double, C.c_double, C.c_double]
    L.orc_sphere_obb_collide_f64.argtypes = [_f32p, _f32p, C.c_double, C.c_double, C.c_double, C.c_double]
    L.orc_sphere_obb_collide_f64.restype = C.c_int
    L.orc_sphere_esdf_clearance_f64.argtypes = [C.POINTER(OrcScene), C.c_double, C.c_double, C.c_double,
                                                C.c_double, C.c_int]
    L.orc_edge_granularity.restype = C.c_int
    L.orc_sphere_esdf_grad_f64.restype = C.c_double
    L.orc_sphere_esdf_grad_f64.argtypes = [C.POINTER(OrcScene), C.c_double, C.c_double, C.c_double, C.c_double,
                                           C.c_int, _f64p, C.POINTER(C.c_int)]


def sphere_obb_clearance(dims, pose, p, r) -> float:
    _setup_restypes()
    d = _c(dims, np.float32)
    q = _c(pose, np.float32)
    return lib().orc_sphere_obb_clearance_f64(_ptr(d, _f32p), _ptr(q, _f32p), *map(float, p), float(r))


def sphere_obb_collide(dims, pose, p, r) -> bool:
    _setup_restypes()
    d = _c(dims, np.float32)
    q = _c(pose, np.float32)
    return bool(lib().orc_sphere_obb_collide_f64(_ptr(d, _f32p), _ptr(q, _f32p), *map(float, p), float(r)))


def sphere_esdf_clearance(orc: Oracle, p, r, sphere_idx=-1) -> float:
    _setup_restypes()
    return lib().orc_sphere_esdf_clearance_f64(C.byref(orc.s), *map(float, p), float(r), int(sphere_idx))


def sphere_esdf_grad(orc: Oracle, p, r, sphere_idx=-1):
    """(clearance, d clearance / d centre [3], ambiguous) of one sphere."""
    _setup_restypes()
    g = np.zeros(3, np.float64)
    amb = C.c_int(0)
    c = lib().orc_sphere_esdf_grad_f64(C.byref(orc.s), *map(float, p), float(r), int(sphere_idx), _ptr(g, _f64p),
                                       C.byref(amb))
    return c, g, bool(amb.value)
