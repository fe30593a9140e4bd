// prrtc_module.cpp -- pybind11 module with the reference's `prrtc` names over the C ABI.
//
// Mirrors the builder / check surface of the reference's pybind module
// (python/bindings.cpp:118-222 RobotModelBuilder, 363-402 and 992-1006
// SceneCollisionDataBuilder, 493-530 and 1033-1037 check_collision_scene,
// 764-902 and 1191-1282 STOMP / CHOMP settings, results, buffers and
// stomp_solve_scene / chomp_solve_scene, 225-233, 405-437, 910-938 and
// 1023-1027, 1163-1176 pRRTC's Settings, PlannerResult, SolverBuffers and
// solve_scene):
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

// CHOMP_settings (planning/CHOMP_settings.hh:21-57) with the two fields the
// solver takes from pRRTC_settings (time_limit_ms, collision_margin), same
// names and defaults.
struct CHOMPSettings {
    float time_limit_ms = 4000.0f;
    float collision_margin = 0.0f;
    int num_timesteps = 51;
    int num_iterations = 100;
    int num_iterations_after_valid = 5;
    float step_size = 0.3f;
    float smoothness_weight = 1.0f;
    float collision_cost_weight = 20.0f;
    float fd_epsilon = 1.0e-3f;
    float world_collision_margin = 0.02f;
    float self_collision_margin = 0.01f;
    float delta_t = 0.1f;
    float convergence_eps = 1.0e-4f;
    float collision_free_tol = 1.0e-6f;
    float collision_check_step = 0.05f;
    bool fd_gradient = false;   // true: the reference's central-FD obstacle gradient; false: the fused adjoint
};

// STOMP_settings (planning/STOMP_settings.hh:17-62) with time_limit_ms and
// collision_margin from pRRTC_settings, same names and defaults; `seed` keys
// the sampler.
struct STOMPSettings {
    float time_limit_ms = 4000.0f;
    float collision_margin = 0.0f;
    int num_batch = 1;
    int num_timesteps = 51;
    int num_iterations = 20;
    int num_iterations_after_valid = 3;
    int num_rollouts_new = 30;
    int num_rollouts_old = 20;
    float delta_t = 0.1f;
    float control_cost_weight = 0.1f;
    float collision_cost_weight = 20.0f;
    float exponentiated_cost_sensitivity = 2.0f;
    float world_collision_margin = 0.02f;
    float self_collision_margin = 0.01f;
    float convergence_eps = 1.0e-4f;
    float collision_free_tol = 1.0e-6f;
    float noise_scale = 1.0f;
    float collision_check_step = 0.05f;
    unsigned long long seed = 0x5be5ull;
};

// CHOMPResult / STOMPResult (python/bindings.cpp:764-797, 837-868).
struct OptResult {
    bool solved = false;
    py::array_t<float> path;   // (T, n_dof), start -> goal
    int iterations_run = 0;
    float final_state_cost = 0.0f, final_control_cost = 0.0f, final_total_cost = 0.0f;
    double wall_time_ms = 0.0;
};
struct CHOMPResult : OptResult {};
struct STOMPResult : OptResult {};

// CHOMPBuffers / STOMPBuffers (planning/chomp_buffers.hh:20-80,
// planning/stomp_buffers.hh:9-140): device buffers reused across solves of
// one shape.
struct OptBuffers {
    ppl_chomp_buffers h = nullptr;
    int n_dof = 0, num_timesteps = 0;
    bool owns_memory() const { return h != nullptr; }
    void destroy() {
        if (h) ppl_chomp_buffers_destroy(h);
        h = nullptr;
    }
    ~OptBuffers() { destroy(); }
};
struct CHOMPBuffers : OptBuffers {
    static CHOMPBuffers* create(int n_dof, int num_timesteps) {
        auto* b = new CHOMPBuffers;
        const ppl_status st = ppl_chomp_buffers_create(n_dof, num_timesteps, &b->h);
        if (st != PPL_OK) {
            delete b;
            check(st, "CHOMPBuffers.create");
        }
        b->n_dof = n_dof;
        b->num_timesteps = num_timesteps;
        return b;
    }
};
struct STOMPBuffers : OptBuffers {
    int num_rollouts_new = 0, num_rollouts_old = 0, num_batch = 1;
    int num_rollouts_all() const { return num_rollouts_new + num_rollouts_old + 1; }
    static STOMPBuffers* create(int n_dof, int num_timesteps, int num_rollouts_new, int num_rollouts_old,
                                int num_batch) {
        if (num_batch != 1) throw std::runtime_error("STOMPBuffers.create: num_batch must be 1 (single goal)");
        auto* b = new STOMPBuffers;
        const ppl_status st =
            ppl_stomp_buffers_create(n_dof, num_timesteps, num_rollouts_new, num_rollouts_old, &b->h);
        if (st != PPL_OK) {
            delete b;
            check(st, "STOMPBuffers.create");
        }
        b->n_dof = n_dof;
        b->num_timesteps = num_timesteps;
        b->num_rollouts_new = num_rollouts_new;
        b->num_rollouts_old = num_rollouts_old;
        return b;
    }
};

using F32Array = py::array_t<float, py::array::c_style | py::array::forcecast>;

template <typename Result, typename Solve>
Result solve_scene(F32Array start, F32Array goal, RobotModelBuilder& robot, int T, Solve&& solve, const char* what) {
    robot.build();
    const int D = robot.n_dof;
    if (start.ndim() != 1 || (int)start.shape(0) != D) throw std::runtime_error("start size mismatch with robot n_dof");
    if (goal.ndim() != 1 || (int)goal.shape(0) != D) throw std::runtime_error("goal size mismatch with robot n_dof");
    if (T < 2) throw std::runtime_error("num_timesteps must be >= 2");
    Result res;
    res.path = py::array_t<float>({T, D});
    ppl_opt_result r{};
    float* path = res.path.mutable_data();
    {
        py::gil_scoped_release nogil;
        check(solve(start.data(), goal.data(), path, &r), what);
    }
    res.solved = r.solved != 0;
    res.iterations_run = r.iterations_run;
    res.final_state_cost = r.final_state_cost;
    res.final_control_cost = r.final_control_cost;
    res.final_total_cost = r.final_total_cost;
    res.wall_time_ms = r.wall_time_ms;
    return res;
}

// destroyed or absent buffers: the call allocates its own
template <typename Buffers>
ppl_chomp_buffers buffers_of(const py::object& o) {
    return o.is_none() ? nullptr : o.cast<Buffers*>()->h;
}

// chomp_solve_scene (python/bindings.cpp:872-902): single-goal CHOMP from start to goal.
CHOMPResult chomp_solve_scene(F32Array start, F32Array goal, const SceneCollisionDataBuilder& scene,
                              const CHOMPSettings& st, RobotModelBuilder& robot, py::object bufs_obj) {
    ppl_chomp_settings cs{};
    cs.step_size = st.step_size;
    cs.smoothness_weight = st.smoothness_weight;
    cs.collision_cost_weight = st.collision_cost_weight;
    cs.world_margin = st.world_collision_margin;
    cs.self_margin = st.self_collision_margin;
    cs.dt = st.delta_t;
    cs.collision_check_step = st.collision_check_step;
    cs.collision_margin = st.collision_margin;
    cs.convergence_eps = st.convergence_eps;
    cs.iterations_after_valid = st.num_iterations_after_valid;
    cs.use_self = 1;
    cs.fd_gradient = st.fd_gradient ? 1 : 0;
    cs.fd_eps = st.fd_epsilon;
    const ppl_scene_view v = scene.view();
    ppl_chomp_buffers h = buffers_of<CHOMPBuffers>(bufs_obj);
    auto solve = [&](const float* a, const float* b, float* path, ppl_opt_result* r) {
        return ppl_chomp_solve(robot.robot, h, &v, a, b, st.num_timesteps, st.num_iterations, st.time_limit_ms, &cs,
                               path, r, nullptr);
    };
    return solve_scene<CHOMPResult>(start, goal, robot, st.num_timesteps, solve, "chomp_solve_scene");
}

// stomp_solve_scene (python/bindings.cpp:801-831): single-goal STOMP from start to goal.
STOMPResult stomp_solve_scene(F32Array start, F32Array goal, const SceneCollisionDataBuilder& scene,
                              const STOMPSettings& st, RobotModelBuilder& robot, py::object bufs_obj) {
    if (st.num_batch != 1) throw std::runtime_error("stomp_solve_scene: num_batch must be 1 (single goal)");
    ppl_stomp_settings ss{};
    ss.num_rollouts_new = st.num_rollouts_new;
    ss.num_rollouts_old = st.num_rollouts_old;
    ss.control_cost_weight = st.control_cost_weight;
    ss.collision_cost_weight = st.collision_cost_weight;
    ss.exponentiated_cost_sensitivity = st.exponentiated_cost_sensitivity;
    ss.world_margin = st.world_collision_margin;
    ss.self_margin = st.self_collision_margin;
    ss.dt = st.delta_t;
    ss.noise_scale = st.noise_scale;
    ss.collision_check_step = st.collision_check_step;
    ss.collision_margin = st.collision_margin;
    ss.convergence_eps = st.convergence_eps;
    ss.iterations_after_valid = st.num_iterations_after_valid;
    ss.use_self = 1;
    ss.seed = st.seed;
    const ppl_scene_view v = scene.view();
    ppl_chomp_buffers h = buffers_of<STOMPBuffers>(bufs_obj);
    auto solve = [&](const float* a, const float* b, float* path, ppl_opt_result* r) {
        return ppl_stomp_solve(robot.robot, h, &v, a, b, st.num_timesteps, st.num_iterations, st.time_limit_ms, &ss,
                               path, r, nullptr);
    };
    return solve_scene<STOMPResult>(start, goal, robot, st.num_timesteps, solve, "stomp_solve_scene");
}

// pRRTC_settings (planning/pRRTC_settings.hh:3-32, python/bindings.cpp:919-938):
// the same field names and defaults; `seed` picks the Halton base shuffle.
struct Settings {
    int max_samples = 1000000;
    int max_iters = 1000000;
    float time_limit_ms = 4000.0f;
    int gpu_clock_rate_khz = 0;   // kept for scripts that set it; the device wall clock has a fixed rate
    int num_new_configs = 600;
    int granularity = 16;
    float range = 0.5f;
    int balance = 1;
    float tree_ratio = 0.5f;
    bool dynamic_domain = true;
    float dd_alpha = 0.0001f;
    float dd_radius = 4.0f;
    float dd_min_radius = 1.0f;
    bool enable_mesh_collision = false;
    bool shortcut_path = false;
    float valid_segment_length = 0.025f;
    float collision_margin = 0.0f;
    float shortcut_collision_margin = 0.0005f;
    unsigned long long seed = 1;
};

// PlannerResult (python/bindings.cpp:225-233, 910-916): path (N, n_dof) start -> goal.
struct PlannerResult {
    bool solved = false;
    py::array_t<float> path;
    int iters = 0;
    float cost = 0.0f;
    double wall_time_ms = 0.0;
    int start_tree_size = 0, goal_tree_size = 0;
};

// SolverBuffers (planning/solver_buffers.hh:15-54, python/bindings.cpp:1163-1176).
struct SolverBuffers {
    ppl_rrtc_buffers h = nullptr;
    int max_samples = 0, n_dof = 0, num_new_configs = 0, granularity = 0;
    bool owns_memory() const { return h != nullptr; }
    void destroy() {
        if (h) ppl_rrtc_buffers_destroy(h);
        h = nullptr;
    }
    ~SolverBuffers() { destroy(); }
    static SolverBuffers* create(int max_samples, int n_dof, int num_new_configs, int granularity, int n_joints,
                                 bool enable_mesh) {
        (void)n_joints;
        if (enable_mesh) throw std::runtime_error("SolverBuffers.create: mesh collision is not supported");
        auto* b = new SolverBuffers;
        const ppl_status st = ppl_rrtc_buffers_create(n_dof, max_samples, num_new_configs, granularity, &b->h);
        if (st != PPL_OK) {
            delete b;
            check(st, "SolverBuffers.create");
        }
        b->max_samples = max_samples;
        b->n_dof = n_dof;
        b->num_new_configs = num_new_configs;
        b->granularity = granularity;
        return b;
    }
};

// solve_scene (python/bindings.cpp:405-437): pRRTC from start to any of goals [G, n_dof].
PlannerResult solve_rrtc_scene(F32Array start, F32Array goals, const SceneCollisionDataBuilder& scene,
                               const Settings& st, RobotModelBuilder& robot, py::object bufs_obj) {
    if (st.enable_mesh_collision) throw std::runtime_error("solve_scene: mesh collision is not supported");
    robot.build();
    const int D = robot.n_dof;
    if (start.ndim() != 1 || (int)start.shape(0) != D) throw std::runtime_error("start size mismatch with robot n_dof");
    if (goals.ndim() != 2 || (int)goals.shape(1) != D) throw std::runtime_error("goals columns mismatch with robot n_dof");
    const int G = (int)goals.shape(0);
    if (G < 1) throw std::runtime_error("goals must hold at least one configuration");
    ppl_rrtc_settings cs{};
    cs.max_samples = st.max_samples;
    cs.max_iters = st.max_iters;
    cs.time_limit_ms = st.time_limit_ms;
    cs.num_new_configs = st.num_new_configs;
    cs.granularity = st.granularity;
    cs.range = st.range;
    cs.balance = st.balance;
    cs.tree_ratio = st.tree_ratio;
    cs.dynamic_domain = st.dynamic_domain ? 1 : 0;
    cs.dd_alpha = st.dd_alpha;
    cs.dd_radius = st.dd_radius;
    cs.dd_min_radius = st.dd_min_radius;
    cs.shortcut_path = st.shortcut_path ? 1 : 0;
    cs.valid_segment_length = st.valid_segment_length;
    cs.collision_margin = st.collision_margin;
    cs.shortcut_collision_margin = st.shortcut_collision_margin;
    cs.seed = st.seed;
    const ppl_scene_view v = scene.view();
    ppl_rrtc_buffers h = bufs_obj.is_none() ? nullptr : bufs_obj.cast<SolverBuffers*>()->h;
    std::vector<float> path;
    ppl_rrtc_result r{};
    {
        py::gil_scoped_release nogil;
        path.assign((size_t)PPL_RRTC_MAX_WAYPOINTS * D, 0.0f);   // room for any solution
        check(ppl_rrtc_solve(robot.robot, &v, start.data(), goals.data(), G, &cs, h, path.data(),
                             PPL_RRTC_MAX_WAYPOINTS, &r, nullptr),
              "solve_scene");
    }
    PlannerResult res;
    res.solved = r.solved != 0;
    const int n = res.solved ? r.n_waypoints : 0;
    res.path = py::array_t<float>({n, D});
    if (n) std::copy(path.begin(), path.begin() + (size_t)n * D, res.path.mutable_data());
    res.iters = r.iters;
    res.cost = r.cost;
    res.wall_time_ms = r.wall_ms;
    res.start_tree_size = r.start_tree_size;
    res.goal_tree_size = r.goal_tree_size;
    return res;
}

template <typename R>
void bind_result(py::module_& m, const char* name) {
    py::class_<R>(m, name)
        .def_readonly("solved", &R::solved)
        .def_readonly("path", &R::path)
        .def_readonly("iterations_run", &R::iterations_run)
        .def_readonly("final_state_cost", &R::final_state_cost)
        .def_readonly("final_control_cost", &R::final_control_cost)
        .def_readonly("final_total_cost", &R::final_total_cost)
        .def_readonly("wall_time_ms", &R::wall_time_ms);
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
    py::class_<CHOMPSettings>(m, "CHOMPSettings")
        .def(py::init<>())
        .def_readwrite("time_limit_ms", &CHOMPSettings::time_limit_ms)
        .def_readwrite("collision_margin", &CHOMPSettings::collision_margin)
        .def_readwrite("num_timesteps", &CHOMPSettings::num_timesteps)
        .def_readwrite("num_iterations", &CHOMPSettings::num_iterations)
        .def_readwrite("num_iterations_after_valid", &CHOMPSettings::num_iterations_after_valid)
        .def_readwrite("step_size", &CHOMPSettings::step_size)
        .def_readwrite("smoothness_weight", &CHOMPSettings::smoothness_weight)
        .def_readwrite("collision_cost_weight", &CHOMPSettings::collision_cost_weight)
        .def_readwrite("fd_epsilon", &CHOMPSettings::fd_epsilon)
        .def_readwrite("world_collision_margin", &CHOMPSettings::world_collision_margin)
        .def_readwrite("self_collision_margin", &CHOMPSettings::self_collision_margin)
        .def_readwrite("delta_t", &CHOMPSettings::delta_t)
        .def_readwrite("convergence_eps", &CHOMPSettings::convergence_eps)
        .def_readwrite("collision_free_tol", &CHOMPSettings::collision_free_tol)
        .def_readwrite("collision_check_step", &CHOMPSettings::collision_check_step)
        .def_readwrite("fd_gradient", &CHOMPSettings::fd_gradient);
    bind_result<CHOMPResult>(m, "CHOMPResult");
    bind_result<STOMPResult>(m, "STOMPResult");
    py::class_<CHOMPBuffers>(m, "CHOMPBuffers")
        .def_property_readonly("owns_memory", &CHOMPBuffers::owns_memory)
        .def_readonly("n_dof", &CHOMPBuffers::n_dof)
        .def_readonly("num_timesteps", &CHOMPBuffers::num_timesteps)
        .def_static("create", &CHOMPBuffers::create, py::arg("n_dof"), py::arg("num_timesteps"),
                    py::return_value_policy::take_ownership, "Allocate device buffers for chomp_solve_scene calls.")
        .def("destroy", &CHOMPBuffers::destroy, "Free the device buffers. Safe to call more than once.");
    py::class_<STOMPBuffers>(m, "STOMPBuffers")
        .def_property_readonly("owns_memory", &STOMPBuffers::owns_memory)
        .def_readonly("n_dof", &STOMPBuffers::n_dof)
        .def_readonly("num_timesteps", &STOMPBuffers::num_timesteps)
        .def_readonly("num_rollouts_new", &STOMPBuffers::num_rollouts_new)
        .def_readonly("num_rollouts_old", &STOMPBuffers::num_rollouts_old)
        .def_property_readonly("num_rollouts_all", &STOMPBuffers::num_rollouts_all)
        .def_readonly("num_batch", &STOMPBuffers::num_batch)
        .def_static("create", &STOMPBuffers::create, py::arg("n_dof"), py::arg("num_timesteps"),
                    py::arg("num_rollouts_new"), py::arg("num_rollouts_old"), py::arg("num_batch") = 1,
                    py::return_value_policy::take_ownership, "Allocate device buffers for stomp_solve_scene calls.")
        .def("destroy", &STOMPBuffers::destroy, "Free the device buffers. Safe to call more than once.");
    py::class_<STOMPSettings>(m, "STOMPSettings")
        .def(py::init<>())
        .def_readwrite("time_limit_ms", &STOMPSettings::time_limit_ms)
        .def_readwrite("collision_margin", &STOMPSettings::collision_margin)
        .def_readwrite("num_batch", &STOMPSettings::num_batch)
        .def_readwrite("num_timesteps", &STOMPSettings::num_timesteps)
        .def_readwrite("num_iterations", &STOMPSettings::num_iterations)
        .def_readwrite("num_iterations_after_valid", &STOMPSettings::num_iterations_after_valid)
        .def_readwrite("num_rollouts_new", &STOMPSettings::num_rollouts_new)
        .def_readwrite("num_rollouts_old", &STOMPSettings::num_rollouts_old)
        .def_readwrite("delta_t", &STOMPSettings::delta_t)
        .def_readwrite("control_cost_weight", &STOMPSettings::control_cost_weight)
        .def_readwrite("collision_cost_weight", &STOMPSettings::collision_cost_weight)
        .def_readwrite("exponentiated_cost_sensitivity", &STOMPSettings::exponentiated_cost_sensitivity)
        .def_readwrite("world_collision_margin", &STOMPSettings::world_collision_margin)
        .def_readwrite("self_collision_margin", &STOMPSettings::self_collision_margin)
        .def_readwrite("convergence_eps", &STOMPSettings::convergence_eps)
        .def_readwrite("collision_free_tol", &STOMPSettings::collision_free_tol)
        .def_readwrite("collision_check_step", &STOMPSettings::collision_check_step)
        .def_readwrite("noise_scale", &STOMPSettings::noise_scale)
        .def_readwrite("seed", &STOMPSettings::seed);
    py::class_<Settings>(m, "Settings")
        .def(py::init<>())
        .def_readwrite("max_samples", &Settings::max_samples)
        .def_readwrite("max_iters", &Settings::max_iters)
        .def_readwrite("num_new_configs", &Settings::num_new_configs)
        .def_readwrite("granularity", &Settings::granularity)
        .def_readwrite("range", &Settings::range)
        .def_readwrite("balance", &Settings::balance)
        .def_readwrite("tree_ratio", &Settings::tree_ratio)
        .def_readwrite("dynamic_domain", &Settings::dynamic_domain)
        .def_readwrite("dd_alpha", &Settings::dd_alpha)
        .def_readwrite("dd_radius", &Settings::dd_radius)
        .def_readwrite("dd_min_radius", &Settings::dd_min_radius)
        .def_readwrite("enable_mesh_collision", &Settings::enable_mesh_collision)
        .def_readwrite("time_limit_ms", &Settings::time_limit_ms)
        .def_readwrite("gpu_clock_rate_khz", &Settings::gpu_clock_rate_khz)
        .def_readwrite("shortcut_path", &Settings::shortcut_path)
        .def_readwrite("valid_segment_length", &Settings::valid_segment_length)
        .def_readwrite("collision_margin", &Settings::collision_margin)
        .def_readwrite("shortcut_collision_margin", &Settings::shortcut_collision_margin)
        .def_readwrite("seed", &Settings::seed);
    py::class_<PlannerResult>(m, "PlannerResult")
        .def_readonly("solved", &PlannerResult::solved)
        .def_readonly("path", &PlannerResult::path)
        .def_readonly("iters", &PlannerResult::iters)
        .def_readonly("cost", &PlannerResult::cost)
        .def_readonly("wall_time_ms", &PlannerResult::wall_time_ms)
        .def_readonly("start_tree_size", &PlannerResult::start_tree_size)
        .def_readonly("goal_tree_size", &PlannerResult::goal_tree_size);
    py::class_<SolverBuffers>(m, "SolverBuffers")
        .def_property_readonly("owns_memory", &SolverBuffers::owns_memory)
        .def_readonly("max_samples", &SolverBuffers::max_samples)
        .def_readonly("n_dof", &SolverBuffers::n_dof)
        .def_static("create", &SolverBuffers::create, py::arg("max_samples"), py::arg("n_dof"),
                    py::arg("num_new_configs"), py::arg("granularity"), py::arg("n_joints"), py::arg("enable_mesh"),
                    py::return_value_policy::take_ownership, "Allocate device trees for solve_scene calls.")
        .def("destroy", &SolverBuffers::destroy, "Free the device buffers. Safe to call more than once.");
    m.def("solve_scene", &solve_rrtc_scene, py::arg("start"), py::arg("goals"), py::arg("scene"),
          py::arg("settings"), py::arg("robot"), py::arg("bufs") = py::none(),
          "pRRTC (GPU-parallel RRT-Connect) against the scene; returns PlannerResult with path [N, n_dof].");
    m.def("stomp_solve_scene", &stomp_solve_scene, py::arg("start"), py::arg("goal"), py::arg("scene"),
          py::arg("settings"), py::arg("robot"), py::arg("bufs") = py::none(),
          "Single-goal STOMP against the scene; returns STOMPResult with path [T, n_dof].");
    m.def("chomp_solve_scene", &chomp_solve_scene, py::arg("start"), py::arg("goal"), py::arg("scene"),
          py::arg("settings"), py::arg("robot"), py::arg("bufs") = py::none(),
          "Single-goal CHOMP against the scene; returns CHOMPResult with path [T, n_dof].");
}
