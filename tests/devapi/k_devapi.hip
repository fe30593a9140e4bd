// k_devapi.hip -- TEST INFRASTRUCTURE: a planner-style kernel built only on
// the device-inline API (include/ppl/ppl_device.hpp). Each G-lane group of a
// wave handles one configuration; the test compares every output with the
// FP64 oracle.
#include "ppl/ppl_device.hpp"

namespace {

template <int G>
__global__ __launch_bounds__(64) void k_devapi(ppl_robot_device_view R, ppl_scene_view S, const float* q, int n,
                                               float margin, int check_self, float world_margin, float self_margin,
                                               uint8_t* hit, float* minclr, float* hinge_obb, float* hinge_esdf,
                                               float* self_hinge, float* scratch) {
    using namespace ppl::device;
    const int tid = threadIdx.x;
    const int slot = blockIdx.x * (64 / G) + tid / G;   // one scratch slot per group, padded grid
    const int cfg = slot < n ? slot : n - 1;            // tail groups recompute the last configuration
    float* pos = scratch + (size_t)slot * (3 * (R.n_spheres + R.n_approx_spheres) + 12 * R.n_joints);
    float* apos = pos + 3 * R.n_spheres;
    float* frames = apos + 3 * R.n_approx_spheres;
    const float* qc = q + (size_t)cfg * R.n_dof;
    // check_self bit 0: self check; bit 1: the one-walk overload with joint frames
    const TwoPhaseResult r = (check_self & 2)
                                 ? two_phase_cc<G>(R, S, qc, pos, apos, frames, tid, margin, (check_self & 1) != 0)
                                 : two_phase_cc<G>(R, S, qc, pos, apos, tid, margin, (check_self & 1) != 0);
    fk_runtime<G>(R, qc, pos, tid);
    const float mc = scene_min_clearance_runtime<G>(R, pos, S, tid, margin);
    const float ho = scene_sum_obb_hinge_runtime<G>(R, pos, S, tid, world_margin);
    const float he = scene_sum_esdf_hinge_runtime<G>(R, pos, S, tid, world_margin);
    const float sh = self_min_hinge_runtime<G>(R, pos, tid, self_margin);
    if (slot < n && (tid & (G - 1)) == 0) {
        hit[slot] = r.collision ? 1 : 0;
        minclr[slot] = mc;
        hinge_obb[slot] = ho;
        hinge_esdf[slot] = he;
        self_hinge[slot] = sh;
    }
}

template <int G>
hipError_t launch(const ppl_robot_device_view& R, const ppl_scene_view& S, const float* q, int n, float margin,
                  int check_self, float wm, float sm, uint8_t* hit, float* mc, float* ho, float* he, float* sh,
                  float* scratch, hipStream_t s) {
    const int per = 64 / G;
    const int blocks = (n + per - 1) / per;
    hipLaunchKernelGGL(k_devapi<G>, dim3(blocks), dim3(64), 0, s, R, S, q, n, margin, check_self, wm, sm, hit, mc, ho,
                       he, sh, scratch);
    return hipGetLastError();
}

// One lane per dimension: n_steps samples of the incremental Halton generator.
__global__ void k_halton(const float* bases, int dim, int n_steps, float* out) {
    const int i = threadIdx.x;
    if (i >= dim) return;
    float n = 0.0f, d = 1.0f;
    for (int k = 0; k < n_steps; ++k) out[(size_t)k * dim + i] = ppl::device::halton_step(bases[i], n, d);
}

}  // namespace

extern "C" {

// d_out [n_steps][dim]; dim <= 64.
int ppl_devtest_halton(const float* d_bases, int dim, int n_steps, float* d_out, void* stream) {
    if (!d_bases || !d_out || dim < 1 || dim > 64 || n_steps < 0) return 1;
    hipLaunchKernelGGL(k_halton, dim3(1), dim3(64), 0, (hipStream_t)stream, d_bases, dim, n_steps, d_out);
    return hipGetLastError() == hipSuccess ? 0 : 2;
}

// Scratch: (ceil(n / (64 / G)) * (64 / G)) * (3 * (n_spheres + n_approx_spheres) + 12 * n_joints) floats.
int ppl_devtest_run(const ppl_robot_device_view* R, const ppl_scene_view* S, const float* d_q, int n, int G,
                    float margin, int check_self, float world_margin, float self_margin, uint8_t* d_hit,
                    float* d_minclr, float* d_hinge_obb, float* d_hinge_esdf, float* d_self_hinge, float* d_scratch,
                    void* stream) {
    if (!R || !S || n <= 0) return 1;
    hipStream_t s = (hipStream_t)stream;
    hipError_t e;
    switch (G) {
        case 1: e = launch<1>(*R, *S, d_q, n, margin, check_self, world_margin, self_margin, d_hit, d_minclr,
                              d_hinge_obb, d_hinge_esdf, d_self_hinge, d_scratch, s); break;
        case 4: e = launch<4>(*R, *S, d_q, n, margin, check_self, world_margin, self_margin, d_hit, d_minclr,
                              d_hinge_obb, d_hinge_esdf, d_self_hinge, d_scratch, s); break;
        case 8: e = launch<8>(*R, *S, d_q, n, margin, check_self, world_margin, self_margin, d_hit, d_minclr,
                              d_hinge_obb, d_hinge_esdf, d_self_hinge, d_scratch, s); break;
        case 32: e = launch<32>(*R, *S, d_q, n, margin, check_self, world_margin, self_margin, d_hit, d_minclr,
                                d_hinge_obb, d_hinge_esdf, d_self_hinge, d_scratch, s); break;
        case 16: e = launch<16>(*R, *S, d_q, n, margin, check_self, world_margin, self_margin, d_hit, d_minclr,
                                d_hinge_obb, d_hinge_esdf, d_self_hinge, d_scratch, s); break;
        case 64: e = launch<64>(*R, *S, d_q, n, margin, check_self, world_margin, self_margin, d_hit, d_minclr,
                                d_hinge_obb, d_hinge_esdf, d_self_hinge, d_scratch, s); break;
        default: return 1;
    }
    return e == hipSuccess ? 0 : 2;
}

}  // extern "C"
