/*
 * roofline.h -- algorithmic FP32 flop counts of the substrate's work units.
 *
 * Counting rules: FMA = 2, any other arithmetic op = 1, sqrt / sin / cos = 1,
 * abs = 0 (a source modifier). Counts are exhaustive (no gating, no early
 * exit); gated counts are measured by running the two-phase algorithm.
 * These are the numerators of every roofline fraction the benchmark prints:
 * achieved FLOP/s = units/s x flops per unit, against the FP32 VALU peak.
 */
#ifndef PPL_ROOFLINE_H
#define PPL_ROOFLINE_H

/* Per-item counts. */
#define PPL_FLOPS_REVOLUTE 83          /* sin, cos, rotate FT's two columns (18), compose 3x4 with parent (63) */
#define PPL_FLOPS_PRISMATIC 69         /* translate along the axis, compose */
#define PPL_FLOPS_SPHERE_XFORM 18      /* 3x4 affine map of a sphere centre */
#define PPL_FLOPS_SPHERE_OBB_BOOL 30   /* inverse-pose transform + strict squared-distance test */
#define PPL_FLOPS_SPHERE_OBB_CLR 36    /* inverse-pose transform + Quilez SDF + running min */
#define PPL_FLOPS_HINGE 3              /* relu(m - c) + accumulate */
#define PPL_FLOPS_SELF_BOOL 9          /* squared distance vs (ri + rj)^2 */
#define PPL_FLOPS_SELF_MIN 11          /* distance - ri - rj + running min */
#define PPL_FLOPS_ESDF_TRILINEAR 72    /* trilinear value + gradient */
#define PPL_FLOPS_ADJ_SPHERE 35        /* argmin-pair clearance gradient + force / moment */
#define PPL_FLOPS_ADJ_REVOLUTE 25      /* moment transfer + axis projection */

/* Forward kinematics of one state plus its sphere centres. */
static inline double ppl_flops_fk(int n_revolute, int n_prismatic, int n_spheres) {
    return (double)n_revolute * PPL_FLOPS_REVOLUTE + (double)n_prismatic * PPL_FLOPS_PRISMATIC +
           (double)n_spheres * PPL_FLOPS_SPHERE_XFORM;
}

/* Exhaustive any-collision of one state against n_obbs OBBs. */
static inline double ppl_flops_any_collision(int n_rev, int n_pri, int n_sph, int n_obbs) {
    return ppl_flops_fk(n_rev, n_pri, n_sph) + (double)n_sph * n_obbs * PPL_FLOPS_SPHERE_OBB_BOOL;
}

/* Summed OBB hinge of one state (forward only), optional self min-hinge. */
static inline double ppl_flops_hinge_state(int n_rev, int n_pri, int n_sph, int n_obbs, int n_pairs) {
    return ppl_flops_fk(n_rev, n_pri, n_sph) + (double)n_sph * n_obbs * PPL_FLOPS_SPHERE_OBB_CLR +
           (double)n_sph * PPL_FLOPS_HINGE + (n_pairs > 0 ? (double)n_pairs * PPL_FLOPS_SELF_MIN + 3.0 : 0.0);
}

/* CHOMP obstacle cost + gradient of one T-waypoint trajectory, fused adjoint:
 * forward on every waypoint, backward (argmin pairs, joint moments and a
 * recomputed FK) on the T - 2 interior ones. */
static inline double ppl_flops_chomp_adjoint_traj(int n_rev, int n_pri, int n_sph, int n_obbs, int T) {
    const double fwd = ppl_flops_hinge_state(n_rev, n_pri, n_sph, n_obbs, 0);
    const double bwd = (double)n_sph * PPL_FLOPS_ADJ_SPHERE + (double)(n_rev + n_pri) * PPL_FLOPS_ADJ_REVOLUTE +
                       ppl_flops_fk(n_rev, n_pri, n_sph);
    return T * fwd + (T - 2) * bwd;
}

/* The same gradient by central finite differences (chomp.cu:77-144 semantics):
 * 2 D (T - 2) perturbed evaluations plus the T nominal ones. */
static inline double ppl_flops_chomp_fd_traj(int n_rev, int n_pri, int n_sph, int n_obbs, int T, int D) {
    return (2.0 * D * (T - 2) + T) * ppl_flops_hinge_state(n_rev, n_pri, n_sph, n_obbs, 0);
}

/* STOMP rollout state: world hinge + self min-hinge (control cost excluded). */
static inline double ppl_flops_stomp_state(int n_rev, int n_pri, int n_sph, int n_obbs, int n_pairs) {
    return ppl_flops_hinge_state(n_rev, n_pri, n_sph, n_obbs, n_pairs);
}

/* Control cost per state: 9 stencil taps, 2 flops each, per dof. */
static inline double ppl_flops_control_state(int D) { return 18.0 * D; }

/* Trilinear-ESDF hinge waypoint: forward (FK, value + gradient, hinge) plus
 * backward (adjoint per sphere and joint, recomputed FK). */
static inline double ppl_flops_esdf_waypoint(int n_rev, int n_pri, int n_sph) {
    const double fwd = ppl_flops_fk(n_rev, n_pri, n_sph) + (double)n_sph * (PPL_FLOPS_ESDF_TRILINEAR + PPL_FLOPS_HINGE);
    const double bwd = (double)n_sph * PPL_FLOPS_ADJ_SPHERE + (double)(n_rev + n_pri) * PPL_FLOPS_ADJ_REVOLUTE +
                       ppl_flops_fk(n_rev, n_pri, n_sph);
    return fwd + bwd;
}

#endif /* PPL_ROOFLINE_H */
