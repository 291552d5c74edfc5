/*
 * ngpb.h — C ABI of the MI355X (gfx950) linearized Poisson-Boltzmann hot path.
 *
 * One opaque context per system: it owns resident device state (mesh, ray table,
 * markers, assembled operator, solver vectors) and one HIP stream. Inputs are
 * caller-owned host arrays that are copied to the device; outputs are copied into
 * caller-owned host arrays. A context is not thread-safe. Every function returns
 * an NGPB_* status code; ngpb_last_error() describes the most recent failure.
 *
 * The entry points replace these reference interfaces (NextGenPB v1.0.0):
 *   ngpb_set_physics      model keys parsed at src/pb_class.cpp:940-945
 *   ngpb_set_mesh         the tmesh_3d quadrant sweep: p(d,i), gt(i), is_hanging(i),
 *                         num_parents(i), gparent(j,i) (src/pb_class.cpp:1838-1852, 2024-2036)
 *   ngpb_build_octree     init_tmesh / refine_surface on p4est (src/pb_class.cpp:1289-1391,
 *                         1467-1745) producing the same views
 *   ngpb_set_rays         ray_cache_t::rays, one map per axis (include/raytracer.h:40-54)
 *   ngpb_cast_rays        the casting that fills it (src/raytracer.cpp:22-63, 337-467)
 *   ngpb_locate_atoms     search_points / cerca_atomo (src/pb_class.cpp:3633-3729)
 *   ngpb_create_markers   poisson_boltzmann::create_markers (include/pb_class.h:716-717,
 *                         src/pb_class.cpp:1792-1941) with is_in_ns_surf (:768-805)
 *   ngpb_set_markers      the members create_markers fills (include/pb_class.h:168-188)
 *   ngpb_set_charges      pos_atoms/charge_atoms + lookup_table from search_points
 *                         (src/pb_class.cpp:3633-3729)
 *   ngpb_set_dirichlet    bcells / dirichlet_bcs3 (src/pb_class.cpp:2108-2149)
 *   ngpb_dirichlet_values bc_type 1/2/3 data (src/pb_class.cpp:2121-2149, 3530-3571)
 *   ngpb_assemble         create_density_map + assemple_system_matrix
 *                         (include/pb_class.h:731-735, src/pb_class.cpp:1943-2159)
 *   ngpb_export_csr       distributed_sparse_matrix::csr as used at src/pb_class.cpp:2255
 *   ngpb_set_csr          lis_matrix_set_csr + lis_vector_set_value (src/pb_class.cpp:2263-2304)
 *   ngpb_parse_lis_options lis_solver_set_option (src/pb_class.cpp:2311-2313)
 *   ngpb_solve            lis_compute_electric_potential / lis_solve
 *                         (include/pb_class.h:740-741, src/pb_class.cpp:2243-2329)
 *   ngpb_energy           poisson_boltzmann::energy / energy_fast
 *                         (include/pb_class.h:758-762, src/pb_class.cpp:2790-3338)
 *   ngpb_phi_on_atoms     write_potential_on_atoms_fast (src/pb_class.cpp:2333-2368)
 *   ngpb_pot_field        pot_field / pot_field_fast (src/pb_class.cpp:3731-5060)
 */
#ifndef NGPB_H
#define NGPB_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ngpb_ctx ngpb_ctx;

enum ngpb_status_code {
  NGPB_OK = 0,
  NGPB_E_ARG = 1,      /* invalid argument or inconsistent sizes */
  NGPB_E_STATE = 2,    /* a required earlier step has not been done */
  NGPB_E_HIP = 3,      /* HIP runtime error */
  NGPB_E_NOCONV = 4,   /* Krylov solver reached maxiter */
  NGPB_E_NUMERIC = 5,  /* NaN/Inf or non-positive diagonal in the operator */
  NGPB_E_NOMEM = 6,    /* device or host allocation failed */
  NGPB_E_NODEV = 7     /* no HIP device */
};

enum ngpb_precision {
  NGPB_PREC_FP64 = 0,  /* plain method: FP64 CSR, FP64 vectors, Jacobi PCG (LIS -i cg -p jacobi) */
  NGPB_PREC_MIXED = 1  /* FP32 compact operator + FP32 inner CG, FP64 residual/reductions/update */
};

/* Model parameters (relative permittivities, temperature in K, ionic strength in M). */
typedef struct {
  double e_in;
  double e_out;
  double T;
  double ionic_strength;
  int calc_energy; /* 0 none, 1 polarization only, 2 polarization + direct ionic */
  int calc_coulombic; /* 1: ngpb_energy also sums the Coulomb pair energy (calc_coulombic,
                         src/pb_class.cpp:946, 2842-2857); 0 (default): not computed */
} ngpb_physics;

/* Krylov method and preconditioner, as LIS's "-i" and "-p" options. */
enum ngpb_solver { NGPB_SOLVER_CG = 0, NGPB_SOLVER_CGS = 1, NGPB_SOLVER_BICGSTAB = 2 };
enum ngpb_precon { NGPB_PRECON_JACOBI = 0, NGPB_PRECON_NONE = 1, NGPB_PRECON_SSOR = 2 };

typedef struct {
  int precision;     /* enum ngpb_precision */
  double tol;        /* stop when ||b - A x||_2 <= tol * ||b||_2 (x0 = 0) */
  int maxiter;       /* limit on iterations (FP64) or inner iterations (mixed) */
  double drop;       /* mixed: FP64 residual replacement once the inner residual has
                        dropped by this factor since the last replacement (default 1e-3) */
  int check_every;   /* iterations launched between host polls of the stop flag (>= 1) */
  int solver;        /* enum ngpb_solver (FP64 only; MIXED runs CG) */
  int precon;        /* enum ngpb_precon (FP64 only; MIXED uses symmetric Jacobi scaling) */
  double omega;      /* SSOR relaxation in (0, 2); <= 0 means LIS's default 1.0 */
} ngpb_solve_opts;

typedef struct {
  int status;          /* NGPB_OK or NGPB_E_NOCONV */
  int iterations;      /* Krylov iterations (matrix-vector products of the iteration) */
  int replacements;    /* mixed: FP64 residual replacements */
  double relres;       /* ||b - A x|| / ||b|| : recursive (FP64) or true FP64 (mixed) */
  double ms_solve;     /* wall time of the solve on the device stream */
  double bytes_per_iter;  /* algorithmic HBM bytes of one iteration */
} ngpb_solve_stats;

typedef struct {
  double energy_pol;    /* polarization energy [kT] */
  double energy_react;  /* direct ionic energy [kT] (calc_energy = 2, kappa > 0) */
  double charge_pol;    /* sum of fluxes; flux charge = charge_pol / (4 pi) [e] */
  int64_t n_border;     /* border leaves (marker 0.5) */
  int64_t n_flux_edges; /* edges with eps(i1) != eps(i2) on border leaves */
  int64_t n_triangles;  /* marching-cubes triangles (ionic term) */
  double coul_energy;   /* calc_coulombic = 1: sum over pairs i < j of charged atoms
                           (|q| > 1e-5) of q_i q_j / r_ij, times 1 / eps_in_hat [kT] */
} ngpb_energy_out;

/* Energies printed by pot_field / pot_field_fast (src/pb_class.cpp:4253-4274). */
typedef struct {
  double energy_pol;    /* 0.5 constant_pol sum_a q_a sum_e flux_e / r_ae (0 in branch 0) */
  double energy_react;  /* branches 3-5: 0.5 (second_int - first_int constant_react) */
  double coul_energy;   /* sum over atom pairs q_i q_j / (r_ij eps_in_hat) */
  double charge_pol;    /* sum of the edge fluxes (flux charge times 4 pi) */
  int32_t branch;       /* which columns were computed, see ngpb_pot_field */
  int32_t reserved;
} ngpb_pot_field_out;

/* ---- library / context ------------------------------------------------------- */
const char* ngpb_version(void);
int ngpb_device_count(int* count);
int ngpb_create(ngpb_ctx** ctx, int device);
int ngpb_destroy(ngpb_ctx* ctx);
const char* ngpb_last_error(const ngpb_ctx* ctx);
/* out[0..5] = eps_in_hat, eps_out_hat, k2 [A^-2], k [A^-1], constant_pol, constant_react
   (include/pb_class.h:47-54, src/pb_class.cpp:1800-1806, 2803-2814). Host only. */
int ngpb_constants(const ngpb_physics* physics, double* out6);
/* Bytes of device memory held by the context. */
int ngpb_device_bytes(const ngpb_ctx* ctx, int64_t* bytes);

/* ---- inputs -------------------------------------------------------------------- */
int ngpb_set_physics(ngpb_ctx* ctx, const ngpb_physics* physics);

/* Leaves with 8 corners in p4est order (bit0 = x, bit1 = y, bit2 = z). leaf_corner[8l+i]
   >= 0 is a node id (gt), < 0 is hanging node -1-c whose value is the mean of its
   hang_np[k] (2 or 4) parents hang_par[4k+j]. Coordinates are the p(d,i) values. */
int ngpb_set_mesh(ngpb_ctx* ctx, int64_t n_nodes, const double* node_xyz, int64_t n_hang,
                  const double* hang_xyz, const int32_t* hang_np, const int32_t* hang_par,
                  int64_t n_leaves, const int32_t* leaf_corner);
/* Builds the mesh on the device instead and installs it as ngpb_set_mesh does, with
   the nodes on the faces of the box as the Dirichlet list (values to follow from
   ngpb_dirichlet_values, or ngpb_set_dirichlet to replace them). Replaces
   init_tmesh / refine_surface and the tmesh_3d views above (src/pb_class.cpp:
   1289-1391, 1467-1745, 1838-1852) with a distance-band rule, as p4est is not part
   of this library: the box [origin, origin + n hf] is tiled by root cells of edge
   2^smax hf (n[d] positive multiples of 2^smax, at most 2^21, with fewer than 2^32 grid
   vertices (n[0]+1)(n[1]+1)(n[2]+1); smax in 0..10); a cell
   of level s >= 1 (edge 2^s hf) is split while some ball a has
   |centre - xyz_a| - r_a < band[s-1] + sqrt(3)/2 * edge (band[0..smax-1] finite and
   non-decreasing); then leaves of level s >= 2 next to (faces, edges, corners) a leaf
   below level s - 1 are split until none is (2:1 balance). Leaves are in Morton order
   of their anchors; nodes and hanging nodes are numbered in lexicographic order (x
   fastest) of their grid position; hanging nodes are edge midpoints (2 parents, the
   edge ends) and face centres (4 parents, the face corners in corner order) of a
   leaf. sizes (may be NULL) receives {n_nodes, n_hang, n_leaves, n_boundary}. */
int ngpb_build_octree(ngpb_ctx* ctx, const double* origin, double hf, const int64_t* n, int smax,
                      const double* band, int64_t n_balls, const double* xyz, const double* r,
                      int64_t* sizes);
/* As ngpb_build_octree with the reference's marker-driven rule (refine_surface,
   src/pb_class.cpp:1467-1745, after init_tmesh, :1289-1296): the box is tiled by
   uniform leaves of level s_uni (maxlevel - unilevel in the reference's terms; smax =
   maxlevel - minlevel, 0 <= s_uni <= smax), then `cycles` times (cycles < 0: the
   reference's max(s_uni, smax - s_uni)):
     refine   every leaf of level s >= 1 whose regular (non-hanging) corners are on
              both sides of the surface is split once, then 2:1 balance;
     coarsen  a family of 8 sibling leaves of level s < smax is merged when each of
              them has all its regular corners on one side, then 2:1 balance.
   A vertex is inside when an odd number of the crossings of the z-line through it lie
   at or below it (is_in_ns_surf, src/pb_class.cpp:769-806); the crossings are those
   of ngpb_cast_rays (probe = 0: union of the balls) or ngpb_cast_rays_ses (probe > 0)
   on every grid z-line. The atom-box test of the reference's markers is left out: it
   never fires there (the atom list is still empty). Merging a family only when all 8
   members are marked is this library's reading of tmesh_3d's coarsening callback,
   which is not part of the reference. Leaves, nodes, numbering and the boundary list
   as in ngpb_build_octree; the context's ray table is left empty (cast the mesh's
   rays next). */
int ngpb_build_octree_surface(ngpb_ctx* ctx, const double* origin, double hf, const int64_t* n, int smax,
                              int s_uni, int cycles, int64_t n_balls, const double* xyz, const double* r,
                              double probe, int64_t* sizes);
/* Reads the context's mesh back in ngpb_set_mesh's layout: sizes receives {n_nodes,
   n_hang, n_leaves, n_boundary}; the arrays (each may be NULL) receive
   node_xyz[3 N], hang_xyz[3 H], hang_np[H], hang_par[4 H], leaf_corner[8 L] and the
   Dirichlet node list bnd_nodes[n_boundary]. No counterpart in the reference. */
int ngpb_get_mesh(ngpb_ctx* ctx, int64_t* sizes, double* node_xyz, double* hang_xyz, int32_t* hang_np,
                  int32_t* hang_par, int32_t* leaf_corner, int32_t* bnd_nodes);

/* Rays of one axis: key[2r], key[2r+1] are the two other coordinates in increasing
   axis order ((y,z) for x rays, (x,z) for y rays, (x,y) for z rays), keys strictly
   increasing lexicographically; crossings inters[ptr[r] .. ptr[r+1]) increasing,
   normals 3 per crossing. Lines without crossings may be omitted. */
int ngpb_set_rays(ngpb_ctx* ctx, int axis, int64_t n_rays, const double* key, const int64_t* ptr,
                  const double* inters, const double* normals);
/* Casts the rays on the device instead: for each axis, every grid line through a
   vertex of the mesh (nodes and hanging nodes) against the union of the balls (centre
   xyz[3a..3a+2], radius r[a]). Crossings are the end points of the merged covering
   intervals along the line, with the outward normal of the ball owning each end point;
   a line that only touches a ball does not cross it; lines without crossings are
   omitted. Replaces the casting behind ray_cache_t (src/raytracer.cpp:22-63, the
   castAxisOrientedRay calls at :337-467) for the analytic union of atom spheres. */
int ngpb_cast_rays(ngpb_ctx* ctx, int64_t n_balls, const double* xyz, const double* r);
/* The same lines cast against the solvent-excluded surface (SES) of the balls with
   probe radius probe_radius (the reference's default surface: surface_type 0,
   probe 1.4 A, src/pb_class.cpp:956-963, cast by NanoShaper at src/raytracer.cpp:337-376).
   A point is inside the SES iff no probe ball centred outside every solvent-accessible
   ball (radius r + probe; interior cavities count as solvent) reaches it. Crossings
   lie on atom spheres, re-entrant tori and probe spheres at SAS vertices; normals
   point out of the molecule. Torus intersections are found by sampling the quartic
   at 65 points and bisecting each sign change, so two torus crossings closer than
   1/64 of the torus extent along the line are not resolved. */
int ngpb_cast_rays_ses(ngpb_ctx* ctx, int64_t n_balls, const double* xyz, const double* r, double probe_radius);
/* Owner leaf of each atom, located on the device as search_points / cerca_atomo /
   controlla_coordinate build lookup_table (src/pb_class.cpp:3633-3729): the first
   leaf, in leaf order, with corner0 - tol < x <= corner7 - tol on every axis; -1 when
   none (tol = 2^-P8EST_QMAXLEVEL = 2^-18 of the domain edge in the reference,
   include/pb_class.h:30, src/pb_class.cpp:3637-3640). The
   result is written to owner_leaf[n_atoms] and can be passed to ngpb_set_charges. */
int ngpb_locate_atoms(ngpb_ctx* ctx, int64_t n_atoms, const double* xyz, double tol, int32_t* owner_leaf);
/* Reads the context's ray table of one axis: sizes[0] rays, sizes[1] crossings; the
   arrays (each may be NULL) receive key[2 sizes[0]], ptr[sizes[0] + 1],
   inters[sizes[1]], normals[3 sizes[1]]. No counterpart in the reference. */
int ngpb_get_rays(ngpb_ctx* ctx, int axis, int64_t* sizes, double* key, int64_t* ptr, double* inters,
                  double* normals);

/* In/out markers on the device from the z rays. */
int ngpb_create_markers(ngpb_ctx* ctx);
/* Caller-computed markers (leaf_marker: 0 inside, 1 border, 2 outside). */
int ngpb_set_markers(ngpb_ctx* ctx, const double* eps_node, const double* reac_node,
                     const uint8_t* leaf_marker);
/* Any pointer may be NULL. border receives n_border leaf ids in leaf order. */
int ngpb_get_markers(ngpb_ctx* ctx, double* eps_node, double* reac_node, uint8_t* leaf_marker,
                     int64_t* n_border, int32_t* border);
/* 12 edge fractions per leaf (0.5 for non-border leaves), computed by ngpb_assemble. */
int ngpb_get_fractions(ngpb_ctx* ctx, double* frac);

/* Atoms with their owner leaf (-1: not in the domain). */
int ngpb_set_charges(ngpb_ctx* ctx, int64_t n_atoms, const double* xyz, const double* q,
                     const int32_t* owner_leaf);
/* Dirichlet node list; values may be NULL (then ngpb_dirichlet_values must follow). */
int ngpb_set_dirichlet(ngpb_ctx* ctx, int64_t n_bnd, const int32_t* nodes, const double* values);
/* Device-computed Dirichlet data: 1 zero, 2 Coulomb sum over the charges, 3 Born sphere.
   The values depend on the charges and the physics: after ngpb_set_charges or
   ngpb_set_physics, ngpb_assemble recomputes them for the same bc_type. Values passed
   to ngpb_set_dirichlet are used as given. */
int ngpb_dirichlet_values(ngpb_ctx* ctx, int bc_type);

/* ---- assembly / solve / energy ---------------------------------------------------- */
/* Edge fractions, Laplacian with harmonic-mean face coefficients, kappa^2 mass,
   hanging-node constraints, trilinear charge projection and Dirichlet elimination. */
int ngpb_assemble(ngpb_ctx* ctx);
/* ngpb_assemble of n contexts on one device (a batch's systems): each stage of the
   symbolic and numeric assembly runs as one launch over a table of the systems, on
   ctx[0]'s stream, instead of one launch per system. Every context's system is
   bit-identical to its own ngpb_assemble. No counterpart in the reference, which
   assembles one system per run (create_density_map + assemple_system_matrix,
   src/pb_class.cpp:1943-2159). On an error in a context's own preparation that
   context holds the message; on an error in the shared stages every context does, and
   none holds an assembled system. The systems are taken in groups of at most 16M
   leaves (the environment variable NGPB_GROUP_LEAVES sets another bound). */
int ngpb_assemble_many(ngpb_ctx* const* ctx, int n);
int ngpb_system_size(ngpb_ctx* ctx, int64_t* n, int64_t* nnz);
int ngpb_export_csr(ngpb_ctx* ctx, int32_t* rowptr, int32_t* col, double* val, double* rhs);
/* Load an external CSR system (0-based, columns ascending, n+1 row pointers).
   NGPB_E_NUMERIC for NaN/Inf entries in val or rhs; ngpb_solve also returns
   NGPB_E_NUMERIC for a zero or negative diagonal entry. The next ngpb_assemble
   rebuilds the mesh's system from scratch. ngpb_energy, ngpb_phi_on_atoms and
   ngpb_pot_field need a potential with one value per mesh node (n == n_nodes). */
int ngpb_set_csr(ngpb_ctx* ctx, int64_t n, int64_t nnz, const int32_t* rowptr, const int32_t* col,
                 const double* val, const double* rhs);
/* LIS-style option string, e.g. "-i cg -p jacobi -tol 1e-10 -maxiter 100000" or the
   reference's default "-p ssor -ssor_omega 0.51 -i cgs -tol 1.e-6" (src/pb_class.cpp:970).
   Understood: -i cg|cgs|bicgstab (or LIS's numbers 1|3|4), -p none|jacobi|ssor (0|1|3),
   -ssor_omega, -tol, -maxiter, -f/-precision; -print -conv_cond -initx_zeros -scale
   -tol_w take a value and have no effect. Anything else (other LIS solvers or
   preconditioners such as -p ilu, unknown options) returns NGPB_E_ARG and leaves *out
   holding the LIS defaults for the options read so far: a caller should then fall
   back to LIS (lis_compute_electric_potential) and say so, rather than solve with
   those defaults. */
int ngpb_parse_lis_options(const char* options, ngpb_solve_opts* out);
/* Solves the resident system. FP64: CG with Jacobi runs the fused CG kernels; every
   other solver/preconditioner pair runs the general Krylov kernels with
   level-scheduled SSOR sweeps (lis_solve with those options, src/pb_class.cpp:2294-2315).
   MIXED needs solver CG and precon Jacobi (NGPB_E_ARG otherwise). */
int ngpb_solve(ngpb_ctx* ctx, const ngpb_solve_opts* opts, double* phi_out, ngpb_solve_stats* stats);
/* z = M^-1 r for the resident operator's Jacobi or SSOR preconditioner (relaxation
   omega); host vectors of n entries. The preconditioner application inside LIS's
   solvers; no direct counterpart in the reference's own code. */
int ngpb_precond_apply(ngpb_ctx* ctx, int precon, double omega, const double* r, double* z);
/* Allocation-free FP64 solve. No counterpart in the reference: it is for callers that
   cannot allocate device memory (code under HIP graph capture elsewhere in the
   process, tight serving loops, harnesses that forbid hipMalloc). It runs the CG with
   Jacobi preconditioning of ngpb_solve (the same kernels and, under SpMV schedule 0,
   the same result bit for bit) on a CSR system the caller holds in DEVICE memory on
   `device`: rowptr (n+1), col (nnz), val (nnz), rhs (n) in, x (n) out, all device
   pointers; stream is a hipStream_t (NULL: the default stream). opts: precision FP64,
   solver CG, precon Jacobi, tol, maxiter, check_every. All work memory (4 vectors of
   n doubles, partial sums, scalars) lives in a file-scope device array of 32.1 MiB per
   GPU, held until the process ends; n is limited to 1,048,576 rows (NGPB_E_ARG
   beyond). The call makes no hipMalloc/hipFree or pinned-host allocation. Every call
   on one GPU shares that array: calls that overlap on one GPU race. Returns NGPB_OK,
   NGPB_E_NOCONV, NGPB_E_NUMERIC (NaN/Inf, non-positive diagonal), NGPB_E_ARG (also
   for a CSR whose row pointers or columns are out of range), NGPB_E_HIP, NGPB_E_NODEV. */
int ngpb_solve_csr_noalloc(int device, void* stream, int64_t n, int64_t nnz, const int32_t* rowptr,
                           const int32_t* col, const double* val, const double* rhs, double* x,
                           const ngpb_solve_opts* opts, ngpb_solve_stats* stats);
int ngpb_get_phi(ngpb_ctx* ctx, double* phi);
int ngpb_set_phi(ngpb_ctx* ctx, const double* phi);
/* y = A x with the resident operator: precision FP64 (CSR) or MIXED (scaled FP32 form,
   y = D^1/2 Ahat D^1/2 x evaluated in the compact format). */
int ngpb_spmv(ngpb_ctx* ctx, int precision, const double* x, double* y);
int ngpb_energy(ngpb_ctx* ctx, ngpb_energy_out* out);
/* Potential at each atom (n_atoms of ngpb_set_charges), interpolated trilinearly in
   its owner leaf; NaN for atoms outside the domain. Replaces the computation of
   write_potential_on_atoms_fast (src/pb_class.cpp:2333-2368); the phi_on_atoms.txt
   writer is cudanextgenpb_amd.pb.PoissonBoltzmann.write_potential_on_atoms. */
int ngpb_phi_on_atoms(ngpb_ctx* ctx, double* phi_atoms);
/* Per-atom potentials and fields, replacing the computation of pot_field (meshes with
   hanging nodes, src/pb_class.cpp:4389-5060) and pot_field_fast (:3731-4387).
   table (may be NULL): n_atoms rows of 12 doubles, the columns of pot_field.dat in
   order phi_c phi_p phi_i Ex_c Ey_c Ez_c Ex_p Ey_p Ez_p Ex_i Ey_i Ez_i, with the
   reference's write-out corrections of phi_i and E_i applied (:4286-4290). The
   columns computed follow the reference's branches on calc_potential_term P and
   calc_field_term F (out->branch):
     1  (F == 1 || (F == 2 && k < 1e-5)) && P < 2    phi_p, E_p
     2  (P == 1 || (P == 2 && k < 1e-5)) && F == 0    phi_p
     3  F == 2 && k > 1e-5                             phi_p, phi_i, E_p, E_i
     4  P == 2 && k > 1e-5 && F == 0                   phi_p, phi_i
     5  P == 2 && k > 1e-5 && F == 1                   phi_p, phi_i, E_p
     0  otherwise                                      Coulomb columns only
   The Coulomb columns phi_c, E_c are always computed; other columns are 0. Every atom
   of ngpb_set_charges takes part (no charge filter, as in the reference). The
   pot_field.dat writer is cudanextgenpb_amd.pb.PoissonBoltzmann.pot_field. */
int ngpb_pot_field(ngpb_ctx* ctx, int calc_potential_term, int calc_field_term, double* table,
                   ngpb_pot_field_out* out);

/* ---- measurement ----------------------------------------------------------------- */
/* Times `reps` back-to-back launches of the SpMV kernel of the given precision on the
   context's stream with HIP events; returns the mean kernel time in ms and the
   algorithmic bytes of one launch. No counterpart in the reference. */
int ngpb_time_spmv(ngpb_ctx* ctx, int precision, int reps, double* ms_per_launch,
                   double* bytes_per_launch);

/* With profiling on, every SpMV launched by ngpb_solve is bracketed by HIP events on
   the context's stream; ngpb_kernel_stats returns the launch count, their summed time
   and the algorithmic bytes of one launch for FP64 (CSR) or MIXED (compact) SpMV.
   Enabling or disabling resets the counters. No counterpart in the reference. */
int ngpb_set_profiling(ngpb_ctx* ctx, int enable);
int ngpb_kernel_stats(ngpb_ctx* ctx, int precision, int64_t* launches, double* total_ms,
                      double* bytes_per_launch);
/* Mean time of a plain device copy kernel moving `bytes` (read + write each), used to
   calibrate hardware traffic counters. No counterpart in the reference. */
int ngpb_time_copy(ngpb_ctx* ctx, int64_t bytes, int reps, double* ms_per_launch);
/* Layout of the compact mixed-precision operator (built on demand): info[0] slices,
   [1] delta slices (constant ones included), [2] stored slot entries in delta slices
   (constant ones excluded), [3] in offset slices, [4] 32-bit column entries,
   [5] overflow entries, [6] algorithmic bytes of one compact SpMV, [7] algorithmic
   bytes of one FP64 CSR SpMV, [8] packed slices, [9] packed entries, [10] constant
   delta slices (64 rows with the same value in each slot, stored once per slot),
   [11] their slots. No counterpart in the reference. */
int ngpb_operator_info(ngpb_ctx* ctx, int64_t* info12);
/* Slice forms of the compact operator: constant_slices != 0 (the default) stores a
   delta slice whose 64 rows hold the same value in every slot as one value per slot.
   SpMV results are the same bit for bit either way (same values, same order); the
   operator is rebuilt on next use when the setting changes. No counterpart in the
   reference. */
int ngpb_set_compact_options(ngpb_ctx* ctx, int constant_slices);
/* Work schedule of the SpMV kernels. Compact FP32: 0 one kernel per slice form
   (delta / packed / offset, each walking a list of its slices) on up to 2048 blocks,
   1 a single kernel for all forms, 2 as 0 on the resident block count, 3 the delta
   and packed slices in one launch on the resident grid, every wave alternating
   between its share of the two lists so that the two forms run side by side (then
   the offset slices, if any), 4 as 3 with the lists cut into 8 contiguous parts, one
   per XCD (block b runs on XCD b % 8), the blocks of an XCD interleaving within their
   part, so that each XCD's L2 holds the input entries of one row range, 5 as 4 with
   non-temporal operator loads (measured slower), 6 (default) as 4 with fewer vector
   instructions per entry: the input vector carries 64 zero entries on either side so
   gathers are not clamped, delta gathers use scalar bases, and packed entries of rows
   that skip a slot read a stored +0 instead of being selected away (bit-identical to
   4 up to the sign of exact zeros).
   FP64 CSR: 0 grid-stride over row chunks on up to 2048 blocks, 1 one contiguous
   chunk range per block with XCD-aware block numbering on the resident grid
   (measured 1.36x slower), 2, 3 and 4 grid-stride on the resident grid (an
   XCD-segmented FP64 variant measured 7% slower and was dropped). Results differ
   only in the grouping of the partial sums of (p, q). No counterpart in the
   reference. */
int ngpb_set_spmv_schedule(ngpb_ctx* ctx, int schedule);
/* Row order of the compact FP32 operator of the mixed solve. Mesh nodes come in
   lexicographic order, so a row's z neighbours are a z-plane of rows away; band > 0
   (for a context that assembled its own mesh) cuts y into bands of `band`
   finest-grid spacings and orders the rows band by band (node order within a band),
   so that z neighbours are one band-plane apart. 0 (default): node order. The FP32
   iteration runs in the chosen order and maps rows to nodes where it meets the FP64
   residual and solution, so results differ from node order only by rounding (row
   sums are taken in another column order). Measured on the 63.3M-row operator: the
   same SpMV time for bands of 16, 32 and 64 as for node order. A batch uses node
   order. No counterpart in the reference. */
int ngpb_set_row_order(ngpb_ctx* ctx, int band);
/* Band of the compact operator's row order (built on demand): 0 for node order. */
int ngpb_row_order(ngpb_ctx* ctx, int* band);

/* ---- batched solve of many independent systems ------------------------------------
 * No single counterpart in the reference, which solves one system per run
 * (lis_compute_electric_potential, src/pb_class.cpp:2243-2329); a batch runs that
 * solve (mixed precision) for many contexts at once. Each system keeps its own
 * context (assembled with ngpb_assemble); the batch references the context's operator
 * and vectors, so a context in a batch accepts only read calls (ngpb_get_phi,
 * ngpb_energy, ngpb_export_csr ...) until ngpb_batch_remove. Systems join at replay
 * boundaries; one replay of the batch's captured hipGraph is one FP64 residual
 * replacement step plus inner_per_replay FP32 CG iterations over every system. A
 * system's results do not depend on the other systems in the batch or on the tiling. */
typedef struct ngpb_batch ngpb_batch;
int ngpb_batch_create(ngpb_batch** batch, int device, int max_systems);
int ngpb_batch_destroy(ngpb_batch* batch);
const char* ngpb_batch_last_error(const ngpb_batch* batch);
/* tol (||b - A x|| <= tol ||b||, FP64), maxiter (inner iterations), drop; applies to
   systems added afterwards. */
int ngpb_batch_set_options(ngpb_batch* batch, const ngpb_solve_opts* opts, int inner_per_replay);
int ngpb_batch_add(ngpb_batch* batch, ngpb_ctx* ctx, int* slot);
int ngpb_batch_remove(ngpb_batch* batch, int slot);
/* Autotuner: times the batched compact SpMV for tile sizes (slices of 64 rows per
   block) and waves per block over the current systems and keeps the fastest. */
int ngpb_batch_autotune(ngpb_batch* batch, int* tile_slices, int* waves_per_block, double* ms_per_spmv);
int ngpb_batch_set_tiling(ngpb_batch* batch, int tile_slices, int waves_per_block);
/* Up to max_replays replays, returning early once every system is done;
   *running = systems not yet done. */
int ngpb_batch_run(ngpb_batch* batch, int max_replays, int* running);
int ngpb_batch_status(ngpb_batch* batch, int slot, int* done, ngpb_solve_stats* stats);
int ngpb_batch_timing(ngpb_batch* batch, int64_t* replays, double* ms_total);

#ifdef __cplusplus
}
#endif

#endif /* NGPB_H */
