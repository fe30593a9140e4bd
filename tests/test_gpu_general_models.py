"""Robot models whose two-phase gate decides verdicts, against the FP64 oracle
(needs an MI355X).

The reference's any-collision check is two-phase (two_phase_cc.cuh:57-138):
approximate spheres, walked on the approximate chain
(runtime_kinematics.cuh:176-243), flag joints against the scene; only the
flagged joints' exact spheres are then checked; the exact self pairs count
only after an approximate self pair collides. For the loader's models the
gate is implied by the exact check. Here it is not (tests/parity_common.
model_variant):
  * "shrunk": approximate radii x 0.3, not enclosing the exact spheres;
  * "own_chain": the approximate spheres ride their own chain (every fixed
    translation moved by 12 cm);
  * "sparse_self": one approximate self range, most exact pairs uncovered.
The oracle's gate is pinned by test_oracle_reductions.
test_two_phase_gate_recomposed. Verdicts must match exactly except for
states with an exact or approximate sphere (or self pair) within 1e-5 m of
contact (parity_common.contact_band, counted): with a gate, a verdict can
flip at any such sphere, not only at the state's minimum.
"""
import numpy as np
import pytest
import torch

import pyoracle as PO
from parity_common import (approx_model, assert_values, banner, clearance_den, contact_band, cost_den,
                           model_variant)
from prismplan_amd import _ppl
from prismplan_amd import substrate as S
from prismplan_amd import workloads as W
from prismplan_amd.planning_scene import HostScene

pytestmark = pytest.mark.gpu
DEV = "cuda"
KINDS = ["shrunk", "own_chain", "sparse_self"]


def world(seed, esdf=False):
    rng = np.random.default_rng(seed)
    sc = W.random_obbs(rng, 12, lo=(0.1, -0.8, 0.0), hi=(1.0, 0.8, 1.2), half_hi=0.12)
    if not esdf:
        return sc
    src = W.random_obbs(rng, 6, lo=(0.3, -0.5, 0.2), hi=(0.8, 0.5, 0.9), half_hi=0.1)
    params, pose = W.esdf_grid(center=(0.55, 0.0, 0.55), n=40, voxel=0.025)
    vals = PO.esdf_from_obbs(src.obb_dims, src.obb_pose, params, pose).astype(np.float32)
    return HostScene(obb_dims=sc.obb_dims[:6], obb_pose=sc.obb_pose[:6], obb_enable=sc.obb_enable[:6],
                     voxel_params=params, voxel_pose=pose, voxel_enable=np.ones(1, np.uint8),
                     voxel_features=vals.reshape(1, -1), esdf_interp=0)


class Case:
    def __init__(self, kind, seed, n=3000, esdf=False):
        self.rob = S.Robot(model_variant("panda", kind))
        self.r = self.rob.arrays
        self.ra = approx_model(self.r)
        self.scene = world(seed, esdf)
        self.q = W.uniform_configs(np.random.default_rng(seed + 1), self.r, n)
        self.o = PO.Oracle(self.r, self.scene, nthreads=8)
        self.oa = PO.Oracle(self.ra, self.scene, nthreads=8)
        self.ds = S.DeviceScene.from_host(self.scene)

    def band(self, q, margin, check_self):
        return (contact_band(self.o, self.r, self.scene, q, margin, check_self) |
                contact_band(self.oa, self.ra, self.scene, q, margin, check_self, esdf_acm=False))


def assert_gated(got, ref, amb, what):
    got = np.asarray(got).astype(bool)
    bad = (got != ref) & ~amb
    banner(f"{what}: n={got.size} colliding={ref.mean():.3f} in band={int(amb.sum())} "
           f"mismatch in band={int(((got != ref) & amb).sum())} outside={int(bad.sum())}")
    assert not bad.any()
    assert amb.mean() < 0.02


@pytest.mark.parametrize("kind", KINDS)
def test_gated_verdicts(kind):
    c = Case(kind, 140)
    banner(f"model '{kind}': {c.r.n_spheres} spheres, {c.r.n_approx} approximate, "
           f"own chain {c.r.approx_fixed_transforms is not None}")
    q = c.q
    for margin in (0.0, 0.01):
        ref, _ = c.o.check_collision(q, margin=margin)
        ungated, _ = c.o.check_collision(q, margin=margin, mode=1)
        amb = c.band(q, margin, False)
        assert_gated(S.check_collision_scene(c.rob, c.ds, q, collision_margin=margin), ref, amb,
                     f"  plain, margin {margin} (ungated would differ on {int((ref != ungated).sum())})")
    plan = S.Plan.configs(c.rob, len(q))
    x = torch.from_numpy(q).to(DEV)
    ref, _ = c.o.check_collision(q, margin=0.01, check_self=True)
    amb = c.band(q, 0.01, True)
    got = S.check_collision_batched(plan, c.ds, x, collision_margin=0.01, check_self=True).cpu().numpy()
    assert_gated(got, ref, amb, "  batched + self, margin 0.01")
    cc = S.mitstar_batch_cc(c.rob, c.ds, x, 0.01).cpu().numpy()
    assert_gated(cc, ref, amb, "  MIT* batch_cc (+ self)")


@pytest.mark.parametrize("kind", KINDS)
def test_gated_edges_first_hit(kind):
    c = Case(kind, 150, n=500)
    rng = np.random.default_rng(151)
    frm = c.q
    to = np.clip(frm + rng.normal(scale=0.25, size=frm.shape), c.r.joint_lower, c.r.joint_upper).astype(np.float32)
    n_states = rng.integers(8, 48, len(frm)).astype(np.int32)
    plan = S.Plan.edges(c.rob, n_states)
    fh, hits = S.edge_check_batched(plan, c.ds, torch.from_numpy(frm).to(DEV), torch.from_numpy(to).to(DEV),
                                    check_self=True, skip_resolved=False, state_hits=True)
    fh, hits = fh.cpu().numpy(), hits.cpu().numpy()
    states = PO.edge_states(frm, to, n_states)
    ref, _ = c.o.check_collision(states, check_self=True)
    amb = c.band(states, 0.0, True)
    assert_gated(hits, ref, amb, f"edges '{kind}': per-state verdicts")
    off = np.concatenate([[0], np.cumsum(n_states)])
    for k in range(len(frm)):
        hk = ref[off[k]:off[k + 1]]
        want = int(np.argmax(hk)) if hk.any() else -1
        if fh[k] != want:
            lo = min(x for x in (fh[k], want) if x >= 0)
            assert amb[off[k] + lo], f"edge {k}: first hit {fh[k]} vs {want}"


@pytest.mark.parametrize("kind", ["shrunk", "own_chain"])
def test_gated_esdf_scene(kind):
    c = Case(kind, 160, n=1500, esdf=True)
    ref, _ = c.o.check_collision(c.q, check_self=True)
    ungated, _ = c.o.check_collision(c.q, check_self=True, mode=1)
    amb = c.band(c.q, 0.0, True)
    plan = S.Plan.configs(c.rob, len(c.q))
    got = S.check_collision_batched(plan, c.ds, torch.from_numpy(c.q).to(DEV), check_self=True).cpu().numpy()
    assert_gated(got, ref, amb, f"ESDF + boxes '{kind}' (ungated would differ on {int((ref != ungated).sum())})")


@pytest.mark.parametrize("kind", KINDS)
def test_gated_device_inline_api(kind):
    """two_phase_cc of the device-inline header (both overloads) on the same
    models: its gate and approximate chain follow the model too."""
    import ctypes as C
    from test_gpu_parity import DEVTEST
    assert DEVTEST.exists(), f"{DEVTEST} not built (tests/devapi/build_devtest.py)"
    lib = C.CDLL(str(DEVTEST))
    c = Case(kind, 170, n=1200)
    q, r = c.q, c.r
    ref, _ = c.o.check_collision(q, check_self=True)
    amb = c.band(q, 0.0, True)
    dv = _ppl.RobotDeviceView()
    _ppl.call("ppl_robot_get_device_view", c.rob.handle, C.byref(dv))
    assert (dv.aft is not None) == (r.approx_fixed_transforms is not None)
    view = c.ds.view()
    vp = C.c_void_p
    lib.ppl_devtest_run.restype = C.c_int
    lib.ppl_devtest_run.argtypes = [C.POINTER(_ppl.RobotDeviceView), C.POINTER(_ppl.SceneView), vp, C.c_int, C.c_int,
                                    C.c_float, C.c_int, C.c_float, C.c_float, vp, vp, vp, vp, vp, vp, vp]
    n = len(q)
    dq = torch.from_numpy(q).to(DEV)
    for G, frames in ((4, 0), (16, 1)):
        slots = (n + 64 // G - 1) // (64 // G) * (64 // G)
        hit = torch.zeros(n, dtype=torch.uint8, device=DEV)
        outs = [torch.zeros(n, dtype=torch.float32, device=DEV) for _ in range(4)]
        scratch = torch.zeros(slots * (3 * (r.n_spheres + r.n_approx) + 12 * r.n_joints), dtype=torch.float32,
                              device=DEV)
        st = lib.ppl_devtest_run(C.byref(dv), C.byref(view), dq.data_ptr(), n, G, 0.0, 1 | (2 * frames), 0.03, 0.01,
                                 hit.data_ptr(), *[o.data_ptr() for o in outs], scratch.data_ptr(), None)
        assert st == 0
        torch.cuda.synchronize()
        assert_gated(hit.cpu().numpy(), ref, amb, f"device two_phase_cc '{kind}' G={G} frames-overload={frames}")


@pytest.mark.parametrize("kind", ["shrunk", "own_chain"])
def test_costs_ignore_the_approximate_model(kind):
    """The hinge cost and its adjoint use exact spheres only; their broad
    phase culls with spheres built around each joint's exact spheres, so a
    non-enclosing or displaced approximate model changes nothing."""
    c = Case(kind, 180, n=2000)
    base = S.Robot("panda")
    prm = S.cost_params(world_margin=0.05, self_margin=0.01, weight=2.0)
    prm_o = PO.cost_params(world_margin=0.05, self_margin=0.01, weight=2.0)
    x = torch.from_numpy(c.q).to(DEV)
    st, _, g = S.hinge_cost(S.Plan.configs(c.rob, len(c.q)), c.ds, x, prm, grad=True)
    st0, _, g0 = S.hinge_cost(S.Plan.configs(base, len(c.q)), c.ds, x, prm, grad=True)
    c_ref, g_ref, kink = c.o.hinge_grad(c.q, prm_o)
    rel = np.abs(st.cpu().numpy() - c_ref) / cost_den(c_ref, 2.0, c.o.hinge_terms(c.q, prm_o))
    banner(f"hinge '{kind}': max rel err {rel.max():.2e}, active {np.mean(c_ref > 0):.2f}")
    assert rel.max() < 1e-5 and np.mean(c_ref > 0) > 0.1
    torch.testing.assert_close(st, st0, rtol=0, atol=0)
    torch.testing.assert_close(g, g0, rtol=0, atol=0)
    mref, vamb = c.o.min_clearance(c.q, margin=0.01, with_amb=True)
    assert_values(S.min_clearance_scene(c.rob, c.ds, c.q, collision_margin=0.01), mref, vamb, clearance_den(mref),
                  "  min clearance")
