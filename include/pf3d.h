/*
 * pf3d.h - C ABI of the MI355X (gfx950) stiffness-assembly-and-solve library.
 *
 * All array arguments are DEVICE pointers (HIP global memory) unless noted;
 * `stream` is a hipStream_t (NULL = default stream).  Every entry point
 * returns 0 on success and a negative PF3D_E* code otherwise; the message
 * of the last failure on the calling thread is in pf3d_last_error().
 * Nothing allocates device memory except where stated.
 *
 * Element conventions follow pyfe3d v0.7.0 (references are to its sources):
 *   - conn[e*nn + i]  = c_I, the global DOF offset of node I (6*node position);
 *                       node coordinates are read at x[c_I/2 + 0..2]
 *                       (quad4.pyx:526-537);
 *   - init_k[e]       = first slot of element e in the r/c/v arrays
 *                       (int64; NULL means size*e);
 *   - values accumulate (v += ...) like the reference unless
 *     PF3D_OVERWRITE; rows/cols are written unless PF3D_VALUES_ONLY.
 *   - shell property record (32 doubles): A11 A12 A16 A22 A26 A66,
 *     B11..B66, D11..D66, E44 E45 E55, scf_k13, scf_k23, h, offset,
 *     intrho, intrhoz, intrhoz2, 4 unused  (shellprop.pxd:38-60)
 *   - beam property record (16 doubles): A E G Iyy Izz Iyz J Ay Az intrho
 *     intrhoy intrhoz intrhoy2 intrhoz2 intrhoyz, 1 unused (beamprop.pyx:15-93);
 *     BeamC and Truss read it
 *   - spring record (6 doubles): kxe kye kze krxe krye krze (spring.pyx:97-125)
 */
#ifndef PF3D_H
#define PF3D_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PF3D_OK 0
#define PF3D_E_ARG -1     /* invalid argument (null pointer, bad size/type) */
#define PF3D_E_HIP -2     /* HIP runtime error */
#define PF3D_E_NOTIMPL -3 /* combination not provided */
#define PF3D_E_SOLVE -4   /* solver did not reach the tolerance */

#define PF3D_QUAD4 0
#define PF3D_QUAD4R 1
#define PF3D_TRIA3R 2
#define PF3D_BEAMC 3
#define PF3D_TRUSS 4
#define PF3D_SPRING 5
#define PF3D_BEAMLR 6

#define PF3D_KC0 0
#define PF3D_KG 1
#define PF3D_KG_STRESS 2
#define PF3D_M 3
#define PF3D_KA_BETA 4  /* piston theory (Quad4, Quad4R) */
#define PF3D_KA_GAMMA 5
#define PF3D_CA 6

#define PF3D_VALUES_ONLY 1 /* update_KC0v_only != 0 (quad4.pyx:1278) */
#define PF3D_OVERWRITE 2   /* store instead of accumulate */

int pf3d_version(void);
const char *pf3d_last_error(void);

/* Validation pass: adds to *count (device int64, caller-zeroed) the
 * number of elements among n_elem whose `size` slots from init_k[e] (NULL:
 * size*e) hold a value that is not finite.  The reference's elements have
 * no such check (cpdef void, cdivision: a zero area writes inf/nan,
 * quad4.pyx:1-7); run it after an assembly to flag degenerate elements. */
int pf3d_nonfinite_elements(int64_t n_elem, int size, const int64_t *init_k, const double *v,
                            int64_t *count, void *stream);

/* Slots written per element: Quad4Data/Quad4RData/Tria3RData/BeamCData
 * KC0_SPARSE_SIZE, KG_SPARSE_SIZE, M_SPARSE_SIZE (quad4.pyx:174-180,
 * quad4r.pyx:140-146, tria3r.pyx:148-151, beamc.pyx:43-46,
 * truss.pyx:45-47, spring.pyx:40-43, beamlr.pyx:43-46; KA_BETA/KA_GAMMA/CA
 * quad4.pyx:157-180); for lumped
 * masses the used prefix of M_SPARSE_SIZE. */
int pf3d_slot_count(int etype, int matrix, int mtype);

/* Element axes, material direction and local coordinates of a batch.
 * geom[e*14 + 0..13] = r11 r12 r13 r21 r22 r23 r31 r32 r33 m11 m12 m21 m22,
 * area (shells) or length (beams); xe[e*3nn ..] = local x, y, z per node.
 * Replaces update_rotation_matrix + update_probe_xe + update_area/length:
 * quad4.pyx:491-752, quad4r.pyx:286-548, tria3r.pyx:294-548,
 * beamc.pyx:158-350, truss.pyx:203-338, spring.pyx:147-206.  evec = xmat
 * (shells, may be NULL), vxy (beams), unused (trusses: axes from the nodes)
 * or x axis + XY-plane vector, 6 per element (springs); springs have no
 * length (0). */
int pf3d_geometry(int etype, int64_t n_elem, const double *x, const int64_t *conn,
                  const double *evec, double *geom, double *xe, void *stream);

/* Shell KC0 of a batch (Quad4, Quad4R, Tria3R), geometry computed from x.
 * Replaces, per element, update_rotation_matrix(x, xmat) ->
 * update_probe_xe(x) -> update_KC0(KC0r, KC0c, KC0v, prop,
 * update_KC0v_only[, hgfactor_*]): quad4.pyx:1204-1313,
 * quad4r.pyx:1145-4619, tria3r.pyx:950-2974.
 * k6rot (Quad4R/Tria3R), alpha (Tria3R alpha_shear_locking) and hg
 * (Quad4R hgfactor_u,v,w,rx,ry; n_elem*5) may be NULL for the defaults. */
int pf3d_shell_kc0(int etype, int64_t n_elem, const double *x, const int64_t *conn,
                   const int64_t *init_k, const double *xmat, const int32_t *pid,
                   const double *props, const double *k6rot, const double *alpha,
                   const double *hg, int64_t *r, int64_t *c, double *v, int flags,
                   void *stream);

/* Shell geometric stiffness from global displacements u (6 per node).
 * Replaces update_probe_ue(u) -> update_probe_xe(x) -> update_KG(...):
 * quad4.pyx:1365-2257, quad4r.pyx:4681-5573, tria3r.pyx:3020-3575. */
int pf3d_shell_kg(int etype, int64_t n_elem, const double *x, const int64_t *conn,
                  const int64_t *init_k, const double *xmat, const int32_t *pid,
                  const double *props, const double *u, int64_t *r, int64_t *c, double *v,
                  int flags, void *stream);

/* Shell geometric stiffness for a given membrane stress state (Nxx, Nyy,
 * Nxy): one triple for all elements (stress_per_elem = 0) or one per
 * element.  Replaces update_KG_given_stress: quad4.pyx:2259-3081,
 * quad4r.pyx:5574-6397, tria3r.pyx:3576-4062. */
int pf3d_shell_kg_stress(int etype, int64_t n_elem, const double *x, const int64_t *conn,
                         const int64_t *init_k, const double *xmat, const double *stress,
                         int stress_per_elem, int64_t *r, int64_t *c, double *v, int flags,
                         void *stream);

/* Shell mass matrix, mtype 0 consistent, 1 reduced, 2 lumped.
 * Replaces update_M: quad4.pyx:3083-9489, quad4r.pyx:6398-12788,
 * tria3r.pyx:4063-7723. */
int pf3d_shell_m(int etype, int64_t n_elem, const double *x, const int64_t *conn,
                 const int64_t *init_k, const double *xmat, const int32_t *pid,
                 const double *props, int mtype, int64_t *r, int64_t *c, double *v, int flags,
                 void *stream);

/* BeamC KC0 (vxy: n_elem*3 orientation vectors).  Replaces
 * update_rotation_matrix(vxyi, vxyj, vxyk, x) -> update_probe_xe(x) ->
 * update_KC0(...): beamc.pyx:158-350, 489-1353. */
int pf3d_beam_kc0(int64_t n_elem, const double *x, const int64_t *conn, const int64_t *init_k,
                  const double *vxy, const int32_t *pid, const double *props, int64_t *r,
                  int64_t *c, double *v, int flags, void *stream);

/* BeamC geometric stiffness from global displacements u (axial force
 * N = A E (u2 - u1)/L in local axes).  Replaces update_probe_ue(u) ->
 * update_probe_xe(x) -> update_KG(...): beamc.pyx:234-286, 1392-2180. */
int pf3d_beam_kg(int64_t n_elem, const double *x, const int64_t *conn, const int64_t *init_k,
                 const double *vxy, const int32_t *pid, const double *props, const double *u,
                 int64_t *r, int64_t *c, double *v, int flags, void *stream);

/* BeamC mass, mtype 0 consistent (144 slots) or 1 lumped (36 slots).
 * Replaces update_M: beamc.pyx:2181-3152. */
int pf3d_beam_m(int64_t n_elem, const double *x, const int64_t *conn, const int64_t *init_k,
                const double *vxy, const int32_t *pid, const double *props, int mtype,
                int64_t *r, int64_t *c, double *v, int flags, void *stream);

/* One group of elements of one type for pf3d_mixed_assemble. */
typedef struct pf3d_elem_group {
    int etype;              /* PF3D_QUAD4, PF3D_QUAD4R, PF3D_TRIA3R or PF3D_BEAMC */
    int64_t n_elem;
    const int64_t *conn;    /* n_elem x nn DOF offsets */
    int64_t k0;             /* added to every slot of the group */
    const int64_t *init_k;  /* first slot per element (+ k0), or NULL for SIZE*e (+ k0) */
    const double *evec;     /* xmat (shells, may be NULL) or vxy (beams, required) */
    const int32_t *pid;     /* property index per element, or NULL for record 0 */
    const double *props;    /* shell (32) or beam (16) records */
    const double *k6rot;    /* shells, may be NULL (100) */
    const double *alpha;    /* Tria3R, may be NULL (0.7) */
    const double *hg;       /* Quad4R, may be NULL (1) */
} pf3d_elem_group;

/* KC0 (matrix PF3D_KC0), KG from displacements u (PF3D_KG) or mass
 * (PF3D_M, mtype 0 or 1; 2 without beams) of up to four groups of different
 * element types (each type at most once) in ONE kernel launch, into the
 * shared COO arrays r, c, v; the slot layouts and values are those of the
 * per-type entry points.  Replaces the caller's loops over mixed element
 * objects, e.g. a stiffened panel's quads, triangles and beams
 * (quad4.pyx:1204-1313, tria3r.pyx:950-2974, beamc.pyx:489-1353 and the
 * KG/M counterparts). */
int pf3d_mixed_assemble(int matrix, int mtype, int n_groups, const pf3d_elem_group *groups,
                        const double *x, const double *u, int64_t *r, int64_t *c, double *v,
                        int flags, void *stream);

/* Truss KC0 (72 slots: axial E A / L and torsion G J / L), props = beam
 * records.  Replaces update_rotation_matrix(x) -> update_probe_xe(x) ->
 * update_length() -> update_KC0(...): truss.pyx:203-338, 386-801. */
int pf3d_truss_kc0(int64_t n_elem, const double *x, const int64_t *conn, const int64_t *init_k,
                   const int32_t *pid, const double *props, int64_t *r, int64_t *c, double *v,
                   int flags, void *stream);

/* Truss mass, mtype 0 consistent (144 slots) or 1 lumped (36 slots).
 * Replaces update_M(..., mtype): truss.pyx:838-1800. */
int pf3d_truss_m(int64_t n_elem, const double *x, const int64_t *conn, const int64_t *init_k,
                 const int32_t *pid, const double *props, int mtype, int64_t *r, int64_t *c,
                 double *v, int flags, void *stream);

/* Spring KC0 (72 slots), axes[e*6..] = x axis (xi, xj, xk) and XY-plane
 * vector (vxyi, vxyj, vxyk), k = spring records (6 doubles, by pid or
 * record 0).  Replaces update_rotation_matrix(xi, xj, xk, vxyi, vxyj,
 * vxyk) -> update_KC0(...): spring.pyx:147-206, 295-707. */
int pf3d_spring_kc0(int64_t n_elem, const double *x, const int64_t *conn, const int64_t *init_k,
                    const double *axes, const int32_t *pid, const double *k, int64_t *r,
                    int64_t *c, double *v, int flags, void *stream);

/* BeamLR (linear Timoshenko beam, reduced integration) KC0 (144 slots,
 * 12x12 row-major).  Replaces update_rotation_matrix(vxyi, vxyj, vxyk, x)
 * -> update_probe_xe(x) -> update_length() -> update_KC0(...):
 * beamlr.pyx:160-353, 407-1187. */
int pf3d_beamlr_kc0(int64_t n_elem, const double *x, const int64_t *conn, const int64_t *init_k,
                    const double *vxy, const int32_t *pid, const double *props, int64_t *r,
                    int64_t *c, double *v, int flags, void *stream);

/* BeamLR KG from global displacements u (36 slots over the rotations of
 * both nodes).  Replaces update_probe_ue(u) -> update_KG(...):
 * beamlr.pyx:236-288, 1226-1463. */
int pf3d_beamlr_kg(int64_t n_elem, const double *x, const int64_t *conn, const int64_t *init_k,
                   const double *vxy, const int32_t *pid, const double *props, const double *u,
                   int64_t *r, int64_t *c, double *v, int flags, void *stream);

/* BeamLR mass, mtype 0 consistent (144 slots) or 1 lumped (36 slots).
 * Replaces update_M(..., mtype): beamlr.pyx:1464-2424. */
int pf3d_beamlr_m(int64_t n_elem, const double *x, const int64_t *conn, const int64_t *init_k,
                  const double *vxy, const int32_t *pid, const double *props, int mtype,
                  int64_t *r, int64_t *c, double *v, int flags, void *stream);

/* Piston-theory aerodynamic matrices of Quad4 / Quad4R (matrix =
 * PF3D_KA_BETA, PF3D_KA_GAMMA or PF3D_CA; 144 slots over the translations
 * (I, a, J, b), values P_IJ n_a n_b with the element normal n):
 * KA_beta = -int N_I dN_J/dX dA (X: global x, the flow direction),
 * KA_gamma = int N_I N_J dA, CA = -int N_I N_J dA, 2x2 Gauss.  Replaces
 * update_KA_beta / update_KA_gamma / update_CA (quad4.pyx:9491-11917,
 * quad4r.pyx:12789-15200); the caller's init_k is init_k_KA_beta,
 * init_k_KA_gamma or init_k_CA. */
int pf3d_shell_piston(int etype, int matrix, int64_t n_elem, const double *x,
                      const int64_t *conn, const int64_t *init_k, int64_t *r, int64_t *c,
                      double *v, int flags, void *stream);

/* Shell property records (32 doubles each) of laminates, one per element,
 * from per-element ply stacks: theta_deg and plyt (n x n_plies, device;
 * plies of zero thickness are padding), ply_mat (n x n_plies indices into
 * mats, or NULL for material 0), mats (n_mat x 8, device: e1 e2 nu12 g12
 * g13 g23 rho, unused), offset (n, or NULL for 0); calc_scf != 0 computes
 * the shear correction factors, else 5/6.  Replaces one
 * laminated_plate(stack, plyts, laminaprops, rhos, offset, calc_scf) per
 * element on the caller side (shellprop_utils.py:96-180:
 * Lamina.rebuild shellprop.pyx:278-337, calc_constitutive_matrix
 * shellprop.pyx:568-621, calc_scf shellprop.pyx:485-548). */
int pf3d_laminate_records(int64_t n, int n_plies, const double *theta_deg, const double *plyt,
                          const int32_t *ply_mat, int n_mat, const double *mats,
                          const double *offset, int calc_scf, double *out, void *stream);

/* Shell property records (32 doubles each) from per-element lamination
 * parameters: thickness[n], xi[n][14] (xiA1..4, xiB1..4, xiD1..4, xiE1,
 * xiE2), material mat_id[n] (NULL: material 0) among n_mat rows of
 * mats[n_mat][8] = (U1..U7 Tsai-Pagano invariants, rho), all on the device.
 * A = h (U0 + G xiA), B = h^2/4 G xiB, D = h^3/12 (U0 + G xiD),
 * E = h (U6 + U7 xiE1, -U7 xiE2, U6 - U7 xiE1); shear factors 5/6, offset 0,
 * intrho = rho h, intrhoz2 = rho h^3 / 12.  Replaces one
 * shellprop_from_LaminationParameters(thickness, mat, lp) per element on the
 * caller side (shellprop.pyx:767-865). */
int pf3d_lp_records(int64_t n, const double *thickness, const double *xi, const int32_t *mat_id,
                    int n_mat, const double *mats, double *out, void *stream);

/* Shell property records (32 doubles each) of isotropic plates, one per
 * element, from a thickness field h[n] (device) and the record of the same
 * material at unit thickness (host, 32 doubles, h field = 1, no offset):
 * A, E44..E55, intrho scale with h, D and intrhoz2 with h^3.  Replaces
 * one isotropic_plate(h, E, nu, rho=...) per element on the caller side
 * (shellprop_utils.py:182-206; shellprop.pyx:568-621). */
int pf3d_isotropic_records(int64_t n, const double *h, const double *unit_record, double *out,
                           void *stream);

/* ------------------------------------------------------------------ */
/* Internal forces fint = sum_e R6 K_e ue (update_fint)                */
/* ------------------------------------------------------------------ */

/* Element internal forces in global axes, fe[(e*nn + i)*6 + d] for node i
 * of element e (nn = 4 or 3), from the displacements u (6*n_nodes).
 * Replaces update_probe_ue + update_probe_finte + the rotation in
 * update_fint: quad4.pyx:627-680,1174-1202,1316-1363; quad4r.pyx:421,549,
 * 4620-4680; tria3r.pyx:426,549-948,2975-3030 (the Tria3R internal forces
 * keep all drilling couplings, unlike its KC0 scatter).  Optional arrays
 * (xmat, pid, k6rot, alpha, hg) as in pf3d_shell_kc0. */
int pf3d_shell_fe(int etype, int64_t n_elem, const double *x, const int64_t *conn,
                  const double *xmat, const int32_t *pid, const double *props,
                  const double *k6rot, const double *alpha, const double *hg, const double *u,
                  double *fe, void *stream);

/* BeamC element internal forces, fe[(e*2 + i)*6 + d].
 * Replaces beamc.pyx:234-286 (ue), 352-488 (finte), 1354-1391 (rotation). */
int pf3d_beam_fe(int64_t n_elem, const double *x, const int64_t *conn, const double *vxy,
                 const int32_t *pid, const double *props, const double *u, double *fe,
                 void *stream);

/* BeamLR element internal forces, fe[(e*2 + i)*6 + d].  Replaces
 * beamlr.pyx:236-288 (ue), 354-406 (finte), 1188-1225 (rotation). */
int pf3d_beamlr_fe(int64_t n_elem, const double *x, const int64_t *conn, const double *vxy,
                   const int32_t *pid, const double *props, const double *u, double *fe,
                   void *stream);

/* Truss (etype 4, props = beam records, axes unused) or Spring (etype 5,
 * props = spring records, axes as in pf3d_spring_kc0) element internal
 * forces, fe[(e*2 + i)*6 + d].  Replaces update_probe_ue +
 * update_probe_finte + update_fint: truss.pyx:150-202, 339-385, 802-837;
 * spring.pyx:208-294, 708-739. */
int pf3d_rod_fe(int etype, int64_t n_elem, const double *x, const int64_t *conn,
                const double *axes, const int32_t *pid, const double *props, const double *u,
                double *fe, void *stream);

/* Node incidence plan of n_entries (element, node) entries whose DOF
 * offsets are conn[0..n_entries): ids (n_entries) lists the entries node by
 * node, in ascending entry order within a node, and offsets (n_nodes + 1)
 * delimits each node's run.  Entries whose offset is not 6*node with
 * node < n_nodes are left out of every run.  The workspace size comes from
 * pf3d_incidence_bytes.  (Built once per mesh; no reference counterpart:
 * the reference scatters element by element, e.g. quad4.pyx:1338-1361.) */
int pf3d_incidence_bytes(int64_t n_entries, int64_t n_nodes, int64_t *bytes);
int pf3d_incidence_plan(int64_t n_entries, const int64_t *conn, int64_t n_nodes, void *work,
                        int64_t work_bytes, int64_t *offsets, int64_t *ids, void *stream);

/* fint[6p + d] += sum over node p's entries of fe[id*6 + d], in ascending
 * element order: the order in which the reference's per-element loop of
 * update_fint accumulates (quad4.pyx:1338-1361, beamc.pyx:1376-1391). */
int pf3d_gather_fint(int64_t n_nodes, const int64_t *offsets, const int64_t *ids,
                     const double *fe, double *fint, void *stream);

/* ------------------------------------------------------------------ */
/* COO -> 6x6 BSR (replaces coo_matrix((v, (r, c))).tocsc() on the     */
/* reference's caller side, e.g. tests/test_quad4_static_point_load.py:80) */
/* ------------------------------------------------------------------ */

/* Workspace bytes of the plan for n_slots COO slots and n_nodes nodes
 * (matrix order 6*n_nodes).  The only call that sizes device memory. */
int pf3d_bsr_plan_bytes(int64_t n_slots, int64_t n_nodes, int64_t *bytes);

/* Sort the slots by (node row, node col, dof row, dof col); counts (host)
 * receives {unique (row, col) entries, 6x6 blocks}. */
int pf3d_bsr_plan_sort(int64_t n_slots, const int64_t *r, const int64_t *c, int64_t n_nodes,
                       void *work, int64_t work_bytes, int64_t *counts, void *stream);

/* Emit the plan: perm[n_slots], seg[n_unique+1], dest[n_unique] and the
 * BSR structure row_ptr[n_nodes+1], col_idx[nnzb] (int32). */
int pf3d_bsr_plan_emit(int64_t n_slots, int64_t n_nodes, void *work, int64_t work_bytes,
                       int64_t n_unique, int64_t nnzb, int64_t *perm, int64_t *seg,
                       int64_t *dest, int32_t *row_ptr, int32_t *col_idx, void *stream);

/* BSR values (nnzb*36, row-major 6x6 blocks) = per-entry sums of v in slot
 * order (deterministic); entries absent from the COO are zero. */
int pf3d_bsr_apply(int64_t n_unique, const int64_t *perm, const int64_t *seg,
                   const int64_t *dest, const double *v, int64_t nnzb, double *bsr_val,
                   void *stream);

/* Element-block plan: the same BSR from the element topology instead of
 * the r/c arrays, for COO values written by the element entry points
 * (block (I, J) of element e at init_k[e] + 36*nn*I + 6*J, row stride
 * 6*nn: quad4.pyx:1287-1293).  One int64 per element block instead of one
 * per slot; values are summed in ascending slot order, so the result is
 * bit-identical to pf3d_bsr_apply on the same COO.  init_k NULL means
 * 36*nn*nn*e. */
int pf3d_block_plan_bytes(int64_t n_elem, int nn, int64_t n_nodes, int64_t *bytes);
int pf3d_block_plan_sort(int64_t n_elem, int nn, const int64_t *conn, const int64_t *init_k,
                         int64_t n_nodes, void *work, int64_t work_bytes, int64_t *nnzb,
                         void *stream);
/* src[n_elem*nn*nn] (first slot of each sorted element block),
 * src_ptr[nnzb+1], row_ptr[n_nodes+1], col_idx[nnzb] */
int pf3d_block_plan_emit(int64_t n_elem, int nn, const int64_t *init_k, int64_t n_nodes,
                         void *work, int64_t work_bytes, int64_t nnzb, int64_t *src,
                         int64_t *src_ptr, int32_t *row_ptr, int32_t *col_idx, void *stream);
/* bsr_val (=, or += with accumulate) sums of the element blocks; a mesh of
 * several element types applies one plan per type, the later ones with
 * accumulate = 1 (their plans must share the BSR structure). */
int pf3d_block_apply(int64_t nnzb, const int64_t *src_ptr, const int64_t *src, int nn,
                     const double *v, double *bsr_val, int accumulate, void *stream);

/* ------------------------------------------------------------------ */
/* Static solve (replaces KC0[bu, :][:, bu] + scipy spsolve/cg,          */
/* tests/test_quad4_static_point_load.py:84-102)                         */
/* ------------------------------------------------------------------ */

/* r = b - A x (or r = A x when b is NULL) with compensated accumulation
 * (FMA-split products, TwoSum): accurate to ~eps of the result instead of
 * ~eps |A| |x|, which matters where the row sums cancel by many digits
 * (shell stiffness: membrane and shear against bending and drilling).
 * The true residuals of the multigrid CG and the products of the block
 * eigensolver use it (no reference counterpart). */
int pf3d_bsr_residual_dd(int64_t nb, const int32_t *row_ptr, const int32_t *col_idx,
                         const double *val, const double *x, const double *b, double *r,
                         void *stream);

/* Smoothed-aggregation multigrid setup on the GPU (no reference
 * counterpart: pyfe3d's callers factorise with SuperLU).  6x6 BSR, int32
 * row pointers and block columns ascending within each row, FP64 values;
 * fixed summation order (repeatable bits).
 *
 * C = X Y in two steps: the symbolic step writes C's row pointer (nr + 1)
 * and returns the block count through *nnz (synchronises the stream; at
 * most 128 block columns per row of C, else PF3D_E_ARG); the numeric step
 * writes C's columns and values (nnz blocks). */
int pf3d_bsr_spgemm_bytes(int64_t nr, int64_t *bytes);
int pf3d_bsr_spgemm_symbolic(int64_t nr, const int32_t *x_row_ptr, const int32_t *x_col_idx,
                             const int32_t *y_row_ptr, const int32_t *y_col_idx,
                             int32_t *c_row_ptr, void *work, int64_t work_bytes, int64_t *nnz,
                             void *stream);
int pf3d_bsr_spgemm_numeric(int64_t nr, const int32_t *x_row_ptr, const int32_t *x_col_idx,
                            const double *x_val, const int32_t *y_row_ptr,
                            const int32_t *y_col_idx, const double *y_val,
                            const int32_t *c_row_ptr, int32_t *c_col_idx, double *c_val,
                            int64_t nnz, void *stream);
/* T = X^T for X with nr block rows, nc block columns and nnzb blocks. */
int pf3d_bsr_transpose_bytes(int64_t nr, int64_t nc, int64_t nnzb, int64_t *bytes);
int pf3d_bsr_transpose(int64_t nr, int64_t nc, int64_t nnzb, const int32_t *row_ptr,
                       const int32_t *col_idx, const double *val, int32_t *t_row_ptr,
                       int32_t *t_col_idx, double *t_val, void *work, int64_t work_bytes,
                       void *stream);
/* P = P0 - omega D^-1 (A P0) on the structure of T = A P0 (values tv),
 * where block row i of the tentative P0 is q[i] (36) at column agg[i] and
 * dinv holds the inverted 6x6 diagonal blocks of A.  *missing counts rows
 * of T without column agg[i] (P0's block would be lost). */
int pf3d_amg_smooth_prolongator(int64_t nr, const int32_t *t_row_ptr, const int32_t *t_col_idx,
                                const double *tv, const int32_t *agg, const double *q,
                                const double *dinv, double omega, double *pv, int *missing,
                                void *stream);
/* out = (V + V^T) / 2 for a square BSR matrix of symmetric structure,
 * plus 1 on the diagonal of each DOF with dead[dof] != 0 (dead may be
 * NULL); *unpaired counts blocks whose transpose partner is absent. */
int pf3d_bsr_symmetrize(int64_t nr, int64_t nnzb, const int32_t *row_ptr, const int32_t *col_idx,
                        const double *val, const uint8_t *dead, double *out, int *unpaired,
                        void *stream);

/* Conjugate-gradient vector updates with the scalars on the device (no
 * reference counterpart), so that an iteration needs no host round trip
 * and can be captured in a graph.  s holds r.z at s[0], p.q at s[1], the
 * new r.z at s[3], beta at s[4] and r.z_old at s[6] (flexible CG):
 *   pf3d_cg_update_xr: alpha = s[0] / s[1]; x += alpha p; r -= alpha q
 *   pf3d_cg_beta:      s[4] = (s[3] - flexible * s[6]) / s[0]; s[0] = s[3]
 *   pf3d_cg_update_p:  p = z + s[4] p
 * (a zero denominator gives a zero step). */
int pf3d_cg_update_xr(int64_t n, const double *s, const double *p, const double *q, double *x,
                      double *r, void *stream);
int pf3d_cg_beta(double *s, int flexible, void *stream);
int pf3d_cg_update_p(int64_t n, const double *s, const double *z, double *p, void *stream);

/* Fused multigrid vector step over the nb block rows of a 6x6 BSR matrix
 * A (no reference counterpart): q = A xin (xin NULL: q = 0),
 * r = sa * (b ? b : r) + sb * q, then dnew = c1 * Dinv r + c2 * dold (when
 * dinv, nb x 36 inverted diagonal blocks, is given; dold may be NULL) and
 * xacc += dold (when xacc is given).  One pass for a Chebyshev smoothing
 * step, a residual b - A x, or y += P x.  xin must not alias r, dnew or
 * xacc, nor dold dnew. */
int pf3d_bsr_fused(int64_t nb, const int32_t *row_ptr, const int32_t *col_idx, const double *val,
                   const double *xin, const double *b, double *r, const double *dinv,
                   const double *dold, double *dnew, double *xacc, double sa, double sb,
                   double c1, double c2, void *stream);
int pf3d_bsr_fused_f32(int64_t nb, const int32_t *row_ptr, const int32_t *col_idx,
                       const float *val, const float *xin, const float *b, float *r,
                       const float *dinv, const float *dold, float *dnew, float *xacc, double sa,
                       double sb, double c1, double c2, void *stream);

/* pf3d_bsr_fused for nv <= 8 vectors at once, vector v of xin at offset
 * v * ldx and of every other vector operand at v * ld (ld >= 6 nb; the
 * two differ for a rectangular operator such as a prolongator): each block
 * of the matrix is read once for all of them (the block eigensolver's
 * preconditioner).  sa == 0 reads neither b nor r. */
int pf3d_bsr_fused_mv(int64_t nb, const int32_t *row_ptr, const int32_t *col_idx,
                      const double *val, int nv, int64_t ldx, const double *xin, int64_t ld,
                      const double *b, double *r, const double *dinv, const double *dold,
                      double *dnew, double *xacc, double sa, double sb, double c1, double c2,
                      void *stream);
int pf3d_bsr_fused_mv_f32(int64_t nb, const int32_t *row_ptr, const int32_t *col_idx,
                          const float *val, int nv, int64_t ldx, const float *xin, int64_t ld,
                          const float *b, float *r, const float *dinv, const float *dold,
                          float *dnew, float *xacc, double sa, double sb, double c1, double c2,
                          void *stream);

/* y[v] = A x[v] for nv <= 8 vectors (ld apart, ld >= 6 nb) with the
 * compensated accumulation of pf3d_bsr_residual_dd (b = NULL), each
 * matrix block read once for all of them; the same bits per vector. */
int pf3d_bsr_product_dd_mv(int64_t nb, const int32_t *row_ptr, const int32_t *col_idx,
                           const double *val, int nv, int64_t ld, const double *x, double *y,
                           void *stream);

/* y = K x for a 6x6 BSR matrix with nb block rows (FP64 / FP32). */
int pf3d_bsr_spmv(int64_t nb, const int32_t *row_ptr, const int32_t *col_idx, const double *val,
                  const double *x, double *y, void *stream);
int pf3d_bsr_spmv_f32(int64_t nb, const int32_t *row_ptr, const int32_t *col_idx,
                      const float *val, const float *x, float *y, void *stream);

/* In place K <- P K P + (I - P), P = diag(fixed == 0): the free-DOF block
 * is the reference's KC0[bu, :][:, bu]; fixed DOFs solve to zero. */
int pf3d_bsr_dirichlet(int64_t nb, const int32_t *row_ptr, const int32_t *col_idx, double *val,
                       const uint8_t *fixed, void *stream);

/* Same with a chosen diagonal on the fixed DOFs: 0 for a mass or
 * geometric-stiffness matrix (P B P), so the fixed DOFs drop out of
 * A x = lambda B x (their eigenvalues move to infinity). */
int pf3d_bsr_dirichlet_diag(int64_t nb, const int32_t *row_ptr, const int32_t *col_idx,
                            double *val, const uint8_t *fixed, double diag, void *stream);

/* b[i] = (float)a[i] */
int pf3d_to_f32(int64_t n, const double *a, float *b, void *stream);

/* Inverses of the 6x6 diagonal blocks (FP64 and, if dinv32, FP32);
 * *singular (device int) counts blocks without a pivot. */
int pf3d_bjacobi(int64_t nb, const int32_t *row_ptr, const int32_t *col_idx, const double *val,
                 double *dinv, float *dinv32, int *singular, void *stream);

/* FP64 block-Jacobi PCG to ||r|| <= rtol ||b||; x is the initial guess when
 * x0 != 0.  work: pf3d_pcg_bytes(nb, 0) bytes. */
int pf3d_pcg_bytes(int64_t nb, int fp32, int64_t *bytes);
int pf3d_pcg(int64_t nb, const int32_t *row_ptr, const int32_t *col_idx, const double *val,
             const double *dinv, const double *b, double *x, double rtol, int maxiter, int x0,
             void *work, int64_t work_bytes, int64_t *iters, double *relres, void *stream);

/* pf3d_pcg without a workspace argument and without any allocation: no
 * counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc).  The PCG vectors live in a fixed device array of the
 * library (one per GPU, ~24 MiB, held until the process ends), so nb is
 * limited to PF3D_PCG_STATIC_MAX_ROWS block rows (786,432 DOF); larger
 * systems get PF3D_E_ARG.  Calls that overlap on one GPU share that array
 * and race: serialise them (one stream, or events).  Same kernels, same
 * results (bit for bit) as pf3d_pcg. */
#define PF3D_PCG_STATIC_MAX_ROWS 131072
int pf3d_pcg_static(int64_t nb, const int32_t *row_ptr, const int32_t *col_idx,
                    const double *val, const double *dinv, const double *b, double *x,
                    double rtol, int maxiter, int x0, int64_t *iters, double *relres,
                    void *stream);

/* FP64 iterative refinement around a block-Jacobi PCG on the FP32-stored
 * matrix and block inverses (FP64 vectors and arithmetic): the outer
 * residual b - K64 x and the update x += d are FP64; stops at
 * ||b - K64 x|| <= rtol ||b||.  stats = {outer steps, inner iterations}.
 * work: pf3d_ir_bytes(nb) bytes. */
int pf3d_ir_bytes(int64_t nb, int64_t *bytes);
int pf3d_ir_solve(int64_t nb, const int32_t *row_ptr, const int32_t *col_idx,
                  const double *val64, const float *val32, const float *dinv32,
                  const double *b, double *x, double rtol, double inner_rtol, int max_outer,
                  int max_inner, int x0, void *work, int64_t work_bytes, int64_t *stats,
                  double *relres, void *stream);

/* Mixed-precision PCG with reliable updates: the Krylov recursion reads K
 * and the block-Jacobi inverses in FP32 with FP64 vectors and arithmetic;
 * every `replace` iterations the recursive residual is replaced by the true
 * FP64 residual b - K64 x (search direction kept), and after 20
 * replacements without halving the true residual the solve finishes with
 * FP64 PCG (dinv64) from the current x.  Stops at ||b - K64 x|| <= rtol ||b||.
 * stats[3]: iterations, replacements, FP64 finishing iterations.  work:
 * pf3d_ir_bytes(nb).  Replaces the same spsolve / cg call as pf3d_pcg. */
int pf3d_pcg_mixed(int64_t nb, const int32_t *row_ptr, const int32_t *col_idx,
                   const double *val64, const double *dinv64, const float *val32,
                   const float *dinv32, const double *b, double *x, double rtol, int64_t maxiter,
                   int replace, int x0, void *work, int64_t work_bytes, int64_t *stats,
                   double *relres, void *stream);

/* ------------------------------------------------------------------ */
/* Batched solves of a design sweep: one launch for many independent   */
/* systems (the reference's tutorials re-assemble and spsolve per      */
/* design, doc/source/tut-opt-lightweight-smeared-deflection-SLSQP.ipynb) */
/* ------------------------------------------------------------------ */

/* The systems are the diagonal blocks of one block-diagonal BSR matrix of
 * nb block rows: system v owns rows [row0[v], row0[v+1]) (row0 on the
 * device).  Each persistent workgroup (`threads` = 256, 512 or 1024; 0 =
 * 1024) takes the next system (in `order`, or 0..n_var-1 if NULL) and runs
 * a 6x6 block-Jacobi PCG on it alone, FP64 vectors and arithmetic:
 *   val32 == NULL : operator and inverses in FP64 (val64, dinv64);
 *   val32 != NULL : operator and inverses read in FP32 (val32, dinv32) with
 *                   reliable updates: r <- b - K64 x every `replace`
 *                   iterations (0: only at convergence checks), keeping the
 *                   search direction.
 * Each system stops when ||b - K64 x|| <= rtol ||b|| (FP64 true residual)
 * or after max_iter iterations.  x0 != 0 starts from the x given.  Per
 * system: iters (PCG iterations), outer (true-residual evaluations),
 * relres (final true relative residual).  n_groups workgroups (0 = one
 * per CU-slot).  No host synchronisation: the call may be captured in a
 * graph.  work: pf3d_batch_bytes(nb). */
int pf3d_batch_bytes(int64_t nb, int64_t *bytes);
int pf3d_batch_solve(int64_t n_var, int64_t nb, const int64_t *row0, const int32_t *order,
                     const int32_t *row_ptr, const int32_t *col_idx, const double *val64,
                     const double *dinv64, const float *val32, const float *dinv32,
                     const double *b, double *x, int x0, double rtol, int replace, int max_iter,
                     void *work, int64_t work_bytes, int32_t *iters, int32_t *outer,
                     double *relres, int n_groups, int threads, void *stream);

/* ------------------------------------------------------------------ */
/* Batched banded direct solves (the design-sweep counterpart of the     */
/* callers' scipy spsolve, tests/test_quad4_static_point_load.py:84-102) */
/* ------------------------------------------------------------------ */

/* Largest bandwidth (scalars below the diagonal) the kernels accept. */
int pf3d_band_max_width(void);

/* Copy the systems of a block-diagonal BSR matrix (after pf3d_bsr_dirichlet)
 * into lower band storage and factorise each in place (Cholesky), one
 * workgroup per system.  System v owns DOFs [dof0[v], dof0[v+1]) of the
 * permuted numbering; var[node] is the system of each node, pinv[node] its
 * permuted (global) node index, bw[v] the bandwidth in scalars and
 * boff[v] the offset of the system's band (dof count x (bw+1) doubles) in
 * band (band_size doubles in all).  *bad (device int) counts systems with
 * a non-positive pivot or a bandwidth above max_b.  Launch configuration:
 * threads per workgroup (256, 512 or 1024; 0 = 1024) and the bandwidth the
 * LDS panel is sized for, max_b (256, 448 or 576; 0 = 576): a narrower
 * panel fits more workgroups per CU (chosen by pyfe3d_amd.autotune). */
int pf3d_band_factor(int64_t n_var, int64_t n_nodes, const int32_t *order, const int64_t *dof0,
                     const int32_t *bw, const int64_t *boff, int64_t band_size, double *band,
                     const int32_t *row_ptr, const int32_t *col_idx, const double *val,
                     const int32_t *var, const int64_t *pinv, int *bad, int threads, int max_b,
                     void *stream);

/* The two halves of pf3d_band_factor, for callers that factorise groups
 * of systems with different launch configurations (a design sweep whose
 * narrow-band plates fit a narrower LDS panel, and so more workgroups per
 * CU, than its widest): pf3d_band_load zeroes band and copies every
 * system's entries into it (arguments as pf3d_band_factor); pf3d_band_chol
 * factorises the n_sys systems order[0 .. n_sys) in place (order null:
 * systems 0 .. n_sys - 1), counting failures and refusals in *bad as
 * pf3d_band_factor.  With work non-null, pf3d_band_chol also runs the
 * forward substitution L y = f of each system from the panels it holds in
 * LDS (pnode and f as in pf3d_band_solve; y left in work), so that
 * pf3d_band_backsolve below finishes the solve reading the factor once.
 * Same replaced reference step: the factorisation inside spsolve
 * (tests/test_quad4_static_point_load.py:102). */
int pf3d_band_load(int64_t n_nodes, const int64_t *dof0, const int32_t *bw, const int64_t *boff,
                   int64_t band_size, double *band, const int32_t *row_ptr,
                   const int32_t *col_idx, const double *val, const int32_t *var,
                   const int64_t *pinv, void *stream);
int pf3d_band_chol(int64_t n_sys, const int32_t *order, const int64_t *dof0, const int32_t *bw,
                   const int64_t *boff, double *band, int *bad, int threads, int max_b,
                   const int64_t *pnode, const double *f, double *work, void *stream);

/* x = A^-1 f per system with the factors of pf3d_band_factor; pnode[p] is
 * the original node of permuted node p; work: one double per DOF; threads
 * per workgroup 256, 512 or 1024 (0 = 1024). */
int pf3d_band_solve(int64_t n_var, const int32_t *order, const int64_t *dof0, const int32_t *bw,
                    const int64_t *boff, const double *band, const int64_t *pnode,
                    const double *f, double *x, double *work, int threads, void *stream);

/* The backward half of pf3d_band_solve (L^T x = y, then x scattered to the
 * original numbering) for systems whose forward result y is already in
 * work (pf3d_band_chol with work).  Same replaced step as pf3d_band_solve. */
int pf3d_band_backsolve(int64_t n_var, const int32_t *order, const int64_t *dof0,
                        const int32_t *bw, const int64_t *boff, const double *band,
                        const int64_t *pnode, double *work, double *x, int threads, void *stream);

/* ------------------------------------------------------------------ */
/* FP64 dense helpers of the eigen-solvers (Krylov bases, column-major  */
/* n x m with leading dimension ld; fixed-order reductions)             */
/* ------------------------------------------------------------------ */
/* C (p x q, row-major) = A B^T for p vectors A[i*lda ..] and q vectors
 * B[j*ldb ..] of length n (p, q <= 32), fixed reduction order; work:
 * 256*p*q doubles.  The Gram matrices of the block eigensolver (LOBPCG)
 * that replaces eigsh for large models (no reference counterpart: the
 * reference calls scipy, tests/test_quad4_natural_freq.py:120-123). */
int pf3d_block_dot(int64_t n, int p, const double *A, int64_t lda, int q, const double *B,
                   int64_t ldb, double *C, double *work, void *stream);

/* Y[j*ldy + t] = sum_i C[i*k + j] S[i*lds + t], j < k, for m vectors S of
 * length n (m, k <= 32; C on the device): the block eigensolver's basis
 * updates, reading S once. */
int pf3d_block_comb(int64_t n, int m, const double *S, int64_t lds, int k, const double *C,
                    double *Y, int64_t ldy, void *stream);

/* y = a x + b y on FP32 vectors, and FP32 -> FP64 widening: the
 * single-precision multigrid preconditioner inside FP64 CG (no reference
 * counterpart). */
int pf3d_axpby_f32(int64_t n, float a, const float *x, float b, float *y, void *stream);
int pf3d_to_f64(int64_t n, const float *a, double *b, void *stream);

/* out[0] = a . b (FP64, fixed reduction order: reruns give the same bits);
 * work: 1024 doubles of device scratch.  The inner products of the
 * multigrid-preconditioned CG (no reference counterpart: the reference
 * calls scipy's solvers, tests/test_quad4_static_point_load.py:84-102). */
int pf3d_dot(int64_t n, const double *a, const double *b, double *out, double *work,
             void *stream);

/* c = alpha V^T x + beta c */
int pf3d_gemv_t(int64_t n, int m, const double *V, int64_t ld, const double *x, double *c,
                double alpha, double beta, void *stream);
/* y = alpha V c + beta y */
int pf3d_gemv_n(int64_t n, int m, const double *V, int64_t ld, const double *c, double *y,
                double alpha, double beta, void *stream);
/* y = a x + b y */
int pf3d_axpby(int64_t n, double a, const double *x, double b, double *y, void *stream);

/* ------------------------------------------------------------------ */
/* Warm-start network kernels (no reference counterpart: pyfe3d starts   */
/* every solve from zero).  FP32 tensors, NHWC activations, weights      */
/* [c_out][(ky, kx, c_in)]; products on the matrix cores in bfloat16     */
/* with FP32 accumulation.                                               */
/* ------------------------------------------------------------------ */

/* C[m*ldc + n] (+)= sum_k A[m*sam + k*sak] B[n*sbn + k*sbk] (+ bias[n],
 * ReLU).  splits > 1 divides K into slices summed in slice order through
 * partial (splits*M*N floats). */
int pf3d_nn_gemm(int M, int N, int K, const float *A, int64_t sam, int64_t sak, const float *B,
                 int64_t sbn, int64_t sbk, float *C, int64_t ldc, int accumulate,
                 const float *bias, int relu, int splits, float *partial, void *stream);
/* 3x3 patches with zero padding 1 and stride 1 or 2, and the adjoint */
int pf3d_nn_im2col(int B, int H, int W, int C, int stride, const float *x, float *col,
                   void *stream);
int pf3d_nn_col2im(int B, int H, int W, int C, int stride, const float *dcol, float *dx,
                   int accumulate, void *stream);
int pf3d_nn_relu_bwd(int64_t n, const float *y, float *dy, void *stream);
/* db[c] = sum over the rows of dy[r][c] (fixed order: chunks of rows summed
 * into work, then the chunks in a fixed order); work holds work_floats >= C
 * floats (1024 C is enough for the most parallel split). */
int pf3d_nn_bias_grad(int64_t rows, int C, const float *dy, float *db, float *work,
                      int64_t work_floats, void *stream);
/* y = nearest-upsample(lo) + skip on an H x W grid, and d(lo) */
int pf3d_nn_up2_add(int B, int H, int W, int C, const float *lo, const float *skip, float *y,
                    void *stream);
int pf3d_nn_up2_bwd(int B, int H, int W, int C, const float *dy, float *dlo, void *stream);
/* element-centre values (B, n, n, C) -> interior node means on the
 * (n+1)^2 grid, and the gradient of 0.5*scale*sum_c wc[c] (u - t)^2 */
int pf3d_nn_node_avg(int B, int n, int C, const float *e, float *u, void *stream);
int pf3d_nn_node_loss_grad(int B, int n, int C, const float *u, const float *target,
                           const float *wc, float scale, float *de, void *stream);
/* Energy-norm loss of an initial guess (per plate b of B, per = 6*nodes
 * DOFs each): pf3d_nn_to_dofs scales the network's node values per channel
 * into an FP64 DOF vector x (0 on fixed DOFs); with kx = K x (the plates'
 * BSR product), f the loads and s[b] = u*.f the solution's energy,
 * pf3d_nn_energy_dots writes out[3b..] = (a = x.f, c = x.K x,
 * L = 1 - a^2 / (c s)), the relative energy-norm error of the best multiple
 * of x; pf3d_nn_energy_grad writes gscale * dL/d(element outputs) through
 * the adjoint of pf3d_nn_node_avg. */
int pf3d_nn_to_dofs(int64_t n, int C, const float *u, const float *scale, const uint8_t *fixed,
                    double *x, void *stream);
int pf3d_nn_energy_dots(int B, int64_t per, const double *x, const double *kx, const double *f,
                        const double *s, double *out, void *stream);
int pf3d_nn_energy_grad(int B, int n, int C, const double *kx, const double *f,
                        const double *acl, const double *s, const float *scale,
                        const uint8_t *fixed, float gscale, float *de, void *stream);
/* Adam step `step` (1-based) with bias correction */
int pf3d_nn_adam(int64_t n, float *p, const float *g, float *m, float *v, float lr, float b1,
                 float b2, float eps, int step, void *stream);

#ifdef __cplusplus
}
#endif
#endif
