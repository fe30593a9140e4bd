"""The allocation-free entry point ppl_mitstar_batch_cc_static against
ppl_mitstar_batch_cc, through the C ABI (needs an MI355X).

The static form keeps its scratch (self verdicts, clearances) in one fixed
device array per GPU instead of the robot handle's growable buffer. Both run
the same kernels on the same inputs, so verdicts and clearances must be
identical, bit for bit. Two calls in a row, the second smaller, show that the
array's leftover contents from the first call do not reach the second's
results. A sample count above PPL_BATCH_CC_STATIC_MAX is refused on the host.
"""
import ctypes as C

import numpy as np
import pytest
import torch

import pyoracle as PO
from prismplan_amd import _ppl
from prismplan_amd import substrate as S
from prismplan_amd import workloads as W
from prismplan_amd.planning_scene import HostScene

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")
STATIC_MAX = 262144   # PPL_BATCH_CC_STATIC_MAX (include/ppl/ppl.h)


def world(seed=140):
    rng = np.random.default_rng(seed)
    obbs = W.random_obbs(rng, 12, lo=(0.2, -0.4, 0.1), hi=(0.7, 0.4, 0.8), half_hi=0.12)
    params, pose = W.esdf_grid(center=(0.45, 0.0, 0.45), n=48, voxel=0.02)
    vals = PO.esdf_from_obbs(obbs.obb_dims, obbs.obb_pose, params, pose).astype(np.float32)
    extra = W.random_obbs(rng, 6, lo=(0.3, -0.8, 0.5), hi=(0.9, 0.8, 1.2), half_hi=0.1)
    return HostScene(obb_dims=extra.obb_dims, obb_pose=extra.obb_pose, obb_enable=extra.obb_enable,
                     voxel_params=params, voxel_pose=pose, voxel_enable=np.ones(1, np.uint8),
                     voxel_features=vals.reshape(1, -1), esdf_interp=1)


def batch_cc(entry, robot, view, x, margin, weight):
    n = x.shape[0]
    cc = torch.full((n,), 7, dtype=torch.uint8, device=DEV)
    clr = torch.full((n,), -7.0, dtype=torch.float32, device=DEV)
    st = getattr(_ppl.lib(), entry)(robot.handle, C.byref(view), C.c_void_p(x.data_ptr()), n, C.c_float(margin),
                                    C.c_void_p(cc.data_ptr()), C.c_void_p(clr.data_ptr()), C.c_float(weight), None)
    assert st == _ppl.PPL_OK, _ppl.lib().ppl_last_error()
    torch.cuda.synchronize()
    return cc.cpu().numpy(), clr.cpu().numpy()


def test_static_batch_cc_matches_the_allocating_form():
    panda = S.Robot("panda")
    ds = S.DeviceScene.from_host(world())
    view = ds.view()
    rng = np.random.default_rng(141)
    for n, margin in ((6000, 0.01), (1500, 0.0)):     # the second call is smaller
        x = torch.from_numpy(W.uniform_configs(rng, panda.arrays, n)).to(DEV)
        print(f"[gpu-test] batch_cc_static vs batch_cc: n={n} margin={margin}", flush=True)
        cc_ref, clr_ref = batch_cc("ppl_mitstar_batch_cc", panda, view, x, margin, 1.0)
        cc, clr = batch_cc("ppl_mitstar_batch_cc_static", panda, view, x, margin, 1.0)
        assert 0.1 < cc_ref.mean() < 0.9, "the world should make some samples collide and leave others free"
        np.testing.assert_array_equal(cc, cc_ref)
        np.testing.assert_array_equal(clr, clr_ref)
        assert np.all(clr[cc == 1] == -7.0) and np.all(clr[cc == 0] != -7.0)


def test_static_batch_cc_refuses_more_than_its_scratch():
    panda = S.Robot("panda")
    ds = S.DeviceScene.from_host(world())
    view = ds.view()
    x = torch.zeros((1, panda.n_dof), dtype=torch.float32, device=DEV)   # never read: the count is refused first
    cc = torch.zeros(1, dtype=torch.uint8, device=DEV)
    st = _ppl.lib().ppl_mitstar_batch_cc_static(panda.handle, C.byref(view), C.c_void_p(x.data_ptr()),
                                                STATIC_MAX + 1, C.c_float(0.0), C.c_void_p(cc.data_ptr()),
                                                None, C.c_float(0.0), None)
    assert st == _ppl.PPL_ERR_CAPACITY
