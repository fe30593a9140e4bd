// prrtc_module.cpp -- pybind11 module with the reference's `prrtc` names over the C ABI.
//
// Mirrors the builder / check surface of the reference's pybind module
// (python/bindings.cpp:118-222 RobotModelBuilder, 363-402 and 992-1006
// SceneCollisionDataBuilder, 493-530 and 1033-1037 check_collision_scene):
// the same class and field names, int64 device-pointer fields filled from
// torch.Tensor.data_ptr(), and shape errors raised as RuntimeError. Host C++
// only: every device operation goes through libppl.so (include/ppl/ppl.h).
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "ppl/ppl.h"

namespace py = pybind11;

namespace {

void check(ppl_status s, const char* where) {
    if (s != PPL_OK) throw std::runtime_error(std::string(where) + ": " + ppl_last_error());
}

template <typename T>
const T* or_null(const std::vector<T>& v) {
    return v.empty() ? nullptr : v.data();
}

struct RobotModelBuilder {
    int n_dof = 0, n_joints = 0;
    std::vector<float> fixed_transforms, spheres_flat, approx_spheres_flat, approx_fixed_transforms;
    std::vector<float> joint_lower, joint_upper;
    std::vector<int> joint_types, sphere_to_joint, self_cc_ranges_flat, approx_sphere_to_joint;
    std::vector<int> approx_self_cc_ranges_flat, approx_joint_types;
    std::vector<int> joint_parents, joint_id_to_dof, t_memory_idx, dfs_order;
    ppl_robot robot = nullptr;

    RobotModelBuilder() = default;
    RobotModelBuilder(const RobotModelBuilder&) = delete;
    RobotModelBuilder& operator=(const RobotModelBuilder&) = delete;
    ~RobotModelBuilder() { destroy(); }

    // Uploads the model on first use (the reference builds lazily as well).
    void build() {
        if (robot) return;
        if ((int)fixed_transforms.size() != 16 * n_joints || (int)joint_types.size() != n_joints)
            throw std::runtime_error("fixed_transforms / joint_types size does not match n_joints");
        if (spheres_flat.size() != 4 * sphere_to_joint.size())
            throw std::runtime_error("spheres_flat size != 4 * len(sphere_to_joint)");
        if (approx_spheres_flat.size() != 4 * approx_sphere_to_joint.size())
            throw std::runtime_error("approx_spheres_flat size != 4 * len(approx_sphere_to_joint)");
        if ((int)joint_lower.size() != n_dof || (int)joint_upper.size() != n_dof)
            throw std::runtime_error("joint_lower/joint_upper size != n_dof");
        if (self_cc_ranges_flat.size() % 3 || approx_self_cc_ranges_flat.size() % 3)
            throw std::runtime_error("self_cc_ranges are triples");
        ppl_robot_desc d{};
        d.n_dof = n_dof;
        d.n_joints = n_joints;
        d.fixed_transforms = fixed_transforms.data();
        d.joint_types = joint_types.data();
        d.n_spheres = (int)sphere_to_joint.size();
        d.spheres = or_null(spheres_flat);
        d.sphere_to_joint = or_null(sphere_to_joint);
        d.n_self_cc_ranges = (int)self_cc_ranges_flat.size() / 3;
        d.self_cc_ranges = or_null(self_cc_ranges_flat);
        d.n_approx_spheres = (int)approx_sphere_to_joint.size();
        d.approx_spheres = or_null(approx_spheres_flat);
        d.approx_sphere_to_joint = or_null(approx_sphere_to_joint);
        d.n_approx_self_cc_ranges = (int)approx_self_cc_ranges_flat.size() / 3;
        d.approx_self_cc_ranges = or_null(approx_self_cc_ranges_flat);
        d.approx_fixed_transforms = or_null(approx_fixed_transforms);
        d.approx_joint_types = or_null(approx_joint_types);
        d.joint_lower = joint_lower.data();
        d.joint_upper = joint_upper.data();
        d.joint_parents = or_null(joint_parents);
        d.joint_id_to_dof = or_null(joint_id_to_dof);
        d.t_memory_idx = or_null(t_memory_idx);
        d.dfs_order = or_null(dfs_order);
        check(ppl_robot_create(&d, &robot), "RobotModelBuilder.build");
    }

    void update_joint_limits() {
        if (!robot) {
            build();
            return;
        }
        if ((int)joint_lower.size() != n_dof || (int)joint_upper.size() != n_dof)
            throw std::runtime_error("joint_lower/joint_upper size != n_dof");
        check(ppl_robot_update_joint_limits(robot, joint_lower.data(), joint_upper.data()),
              "RobotModelBuilder.update_joint_limits");
    }

    void destroy() {
        if (robot) ppl_robot_destroy(robot);
        robot = nullptr;
    }
};

struct SceneCollisionDataBuilder {
    int64_t obb_dims_ptr = 0, obb_pose_ptr = 0, obb_enable_ptr = 0;
    int n_obbs = 0;
    int64_t voxel_params_ptr = 0, voxel_pose_ptr = 0, voxel_enable_ptr = 0, voxel_features_ptr = 0;
    int n_voxel_layers = 0, max_voxels_per_layer = 0;
    int64_t sphere_acm_mask_ptr = 0, sphere_obb_acm_mask_ptr = 0, sphere_obb_acm_mask_approx_ptr = 0;
    int esdf_interp = 0;        // 0: the reference's nearest-voxel lookup; 1: trilinear
    py::object owner;           // keeps the tensors the pointers come from alive

    ppl_scene_view view() const {
        ppl_scene_view v{};
        v.obb_dims = reinterpret_cast<const float*>(obb_dims_ptr);
        v.obb_pose = reinterpret_cast<const float*>(obb_pose_ptr);
        v.obb_enable = reinterpret_cast<const uint8_t*>(obb_enable_ptr);
        v.n_obbs = n_obbs;
        v.voxel_params = reinterpret_cast<const float*>(voxel_params_ptr);
        v.voxel_pose = reinterpret_cast<const float*>(voxel_pose_ptr);
        v.voxel_enable = reinterpret_cast<const uint8_t*>(voxel_enable_ptr);
        v.voxel_features = reinterpret_cast<const float*>(voxel_features_ptr);
        v.n_voxel_layers = n_voxel_layers;
        v.max_voxels_per_layer = max_voxels_per_layer;
        v.sphere_acm_mask = reinterpret_cast<const uint8_t*>(sphere_acm_mask_ptr);
        v.sphere_obb_acm_mask = reinterpret_cast<const uint8_t*>(sphere_obb_acm_mask_ptr);
        v.sphere_obb_acm_mask_approx = reinterpret_cast<const uint8_t*>(sphere_obb_acm_mask_approx_ptr);
        v.esdf_interp = esdf_interp;
        return v;
    }
};

// check_collision_scene(configs [N, n_dof] float32, scene, robot, collision_margin) -> bool[N]
py::array_t<bool> check_collision_scene(py::array_t<float, py::array::c_style | py::array::forcecast> configs,
                                        const SceneCollisionDataBuilder& scene, RobotModelBuilder& robot,
                                        float collision_margin) {
    robot.build();
    if (configs.ndim() != 2) throw std::runtime_error("configs must be [N, n_dof]");
    if ((int)configs.shape(1) != robot.n_dof) throw std::runtime_error("configs columns mismatch with robot n_dof");
    const int n = (int)configs.shape(0);
    py::array_t<bool> out(n);
    if (n == 0) return out;
    std::vector<uint8_t> hits(n);
    const ppl_scene_view v = scene.view();
    {
        py::gil_scoped_release nogil;
        check(ppl_check_collision_scene(robot.robot, &v, configs.data(), n, collision_margin, hits.data()),
              "check_collision_scene");
    }
    auto o = out.mutable_unchecked<1>();
    for (int i = 0; i < n; ++i) o(i) = hits[i] != 0;
    return out;
}

}  // namespace

PYBIND11_MODULE(_prrtc, m) {
    m.doc() = "MI355X collision-kinematics substrate: the reference prrtc module's builder and check names";
    py::class_<RobotModelBuilder>(m, "RobotModelBuilder")
        .def(py::init<>())
        .def_readwrite("n_dof", &RobotModelBuilder::n_dof)
        .def_readwrite("n_joints", &RobotModelBuilder::n_joints)
        .def_readwrite("fixed_transforms", &RobotModelBuilder::fixed_transforms)
        .def_readwrite("joint_types", &RobotModelBuilder::joint_types)
        .def_readwrite("spheres_flat", &RobotModelBuilder::spheres_flat)
        .def_readwrite("sphere_to_joint", &RobotModelBuilder::sphere_to_joint)
        .def_readwrite("self_cc_ranges_flat", &RobotModelBuilder::self_cc_ranges_flat)
        .def_readwrite("approx_spheres_flat", &RobotModelBuilder::approx_spheres_flat)
        .def_readwrite("approx_sphere_to_joint", &RobotModelBuilder::approx_sphere_to_joint)
        .def_readwrite("approx_self_cc_ranges_flat", &RobotModelBuilder::approx_self_cc_ranges_flat)
        .def_readwrite("approx_fixed_transforms", &RobotModelBuilder::approx_fixed_transforms)
        .def_readwrite("approx_joint_types", &RobotModelBuilder::approx_joint_types)
        .def_readwrite("joint_lower", &RobotModelBuilder::joint_lower)
        .def_readwrite("joint_upper", &RobotModelBuilder::joint_upper)
        .def_readwrite("joint_parents", &RobotModelBuilder::joint_parents)
        .def_readwrite("joint_id_to_dof", &RobotModelBuilder::joint_id_to_dof)
        .def_readwrite("t_memory_idx", &RobotModelBuilder::t_memory_idx)
        .def_readwrite("dfs_order", &RobotModelBuilder::dfs_order)
        .def("build", &RobotModelBuilder::build)
        .def("update_joint_limits", &RobotModelBuilder::update_joint_limits)
        .def("destroy", &RobotModelBuilder::destroy)
        .def_property_readonly("handle", [](const RobotModelBuilder& b) { return reinterpret_cast<intptr_t>(b.robot); });
    py::class_<SceneCollisionDataBuilder>(m, "SceneCollisionDataBuilder")
        .def(py::init<>())
        .def_readwrite("obb_dims_ptr", &SceneCollisionDataBuilder::obb_dims_ptr)
        .def_readwrite("obb_pose_ptr", &SceneCollisionDataBuilder::obb_pose_ptr)
        .def_readwrite("obb_enable_ptr", &SceneCollisionDataBuilder::obb_enable_ptr)
        .def_readwrite("n_obbs", &SceneCollisionDataBuilder::n_obbs)
        .def_readwrite("voxel_params_ptr", &SceneCollisionDataBuilder::voxel_params_ptr)
        .def_readwrite("voxel_pose_ptr", &SceneCollisionDataBuilder::voxel_pose_ptr)
        .def_readwrite("voxel_enable_ptr", &SceneCollisionDataBuilder::voxel_enable_ptr)
        .def_readwrite("voxel_features_ptr", &SceneCollisionDataBuilder::voxel_features_ptr)
        .def_readwrite("n_voxel_layers", &SceneCollisionDataBuilder::n_voxel_layers)
        .def_readwrite("max_voxels_per_layer", &SceneCollisionDataBuilder::max_voxels_per_layer)
        .def_readwrite("sphere_acm_mask_ptr", &SceneCollisionDataBuilder::sphere_acm_mask_ptr)
        .def_readwrite("sphere_obb_acm_mask_ptr", &SceneCollisionDataBuilder::sphere_obb_acm_mask_ptr)
        .def_readwrite("sphere_obb_acm_mask_approx_ptr", &SceneCollisionDataBuilder::sphere_obb_acm_mask_approx_ptr)
        .def_readwrite("esdf_interp", &SceneCollisionDataBuilder::esdf_interp)
        .def_readwrite("owner", &SceneCollisionDataBuilder::owner);
    m.def("check_collision_scene", &check_collision_scene, py::arg("configs"), py::arg("scene"), py::arg("robot"),
          py::arg("collision_margin") = 0.0f,
          "Batch environment any-collision (two-phase gate) of configs [N, n_dof] -> bool[N]");
}
