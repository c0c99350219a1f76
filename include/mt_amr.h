/* mt_amr.h -- C-ABI of the MI355X AMR lattice-Boltzmann update.
 *
 * One opaque case handle holds the device-resident state of one AMR flow case:
 * per level, the Morton-sorted cell tables (bricks of 4x4x4 cells, neighbour
 * and parent/child index tables) and the SoA populations f / fcol with the
 * macroscopic fields rho / u.  The entry points are the seams through which the
 * reference's LBM_Manager drives its update; each cites the reference
 * interface it replaces.  Plain pointers and sizes only; every call returns an
 * int status (MT_OK = 0) and mt_last_error() gives the text of the last failure.
 * Host arrays are read/written synchronously; device work runs on the case's
 * own HIP stream.
 */
#ifndef MT_AMR_H
#define MT_AMR_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define MT_OK 0
#define MT_ERR_ARG (-1)    /* bad argument or unsupported configuration        */
#define MT_ERR_HIP (-2)    /* a HIP runtime call failed                         */
#define MT_ERR_STATE (-3)  /* call out of order (e.g. before mt_case_finalize)  */
#define MT_ERR_LOOKUP (-4) /* a required key is absent: the reference's
                              std::out_of_range from D_map_define::at()         */

/* ---- cell kinds (uint16 bit flags) --------------------------------------
 * The level-set contract of header/Lat_Manager.hpp:245-316.                  */
#define MT_KIND_FLUID 0x0001u     /* lat_f                                     */
#define MT_KIND_F2C 0x0002u       /* lat_overlap_F2C                           */
#define MT_KIND_C2F 0x0004u       /* lat_overlap_C2F                           */
#define MT_KIND_GHOST_C2F 0x0008u /* ghost_overlap_C2F                         */
#define MT_KIND_GHOST_F2C 0x0010u /* ghost_overlap_F2C (flag on lat_f cells)   */
#define MT_KIND_SOLID 0x0020u     /* lat_sf                                    */
#define MT_KIND_BCRES 0x0040u     /* lat_bc (frozen level-0 reservoir)         */
#define MT_KIND_DIS 0x0080u       /* has an entry in dis                       */
#define MT_FACE_W 0x0100u         /* level-0 cell lies on face x-min ...       */
#define MT_FACE_E 0x0200u
#define MT_FACE_S 0x0400u
#define MT_FACE_N 0x0800u
#define MT_FACE_B 0x1000u
#define MT_FACE_T 0x2000u

/* ---- face boundary kinds (faces ordered W,E,S,N,Bot,Top) ---------------- */
#define MT_BC_NONE 0      /* missing neighbours keep their f (Stream_Manager.cpp:60-62) */
#define MT_BC_RESERVOIR 1 /* lat_bc reservoir + non-equilibrium extrapolation on its f  */
#define MT_BC_ZOUHE 2     /* wet-node velocity (Zou-He / Hecht-Harting closure)         */
#define MT_BC_OUTFLOW 3   /* zero-gradient copy of the unknown populations              */
#define MT_BC_PERIODIC 4  /* wrapped neighbour tables                                   */
#define MT_BC_NEE 5       /* wet-node velocity wall by non-equilibrium extrapolation from
                             the inward neighbour (the NonEquilibrumExtrapolation rule,
                             BC/NonEquilibrumExtrapolation.cpp:50-66, on the wall cells) */

#define MT_COLLIDE_LBGK 0
#define MT_COLLIDE_CUMULANT 1
#define MT_IB_NONE 0
#define MT_IB_FH 1 /* IB_Method::FH                */
#define MT_IB_DW 2 /* IB_Method::DISTANCE_WEIGHTED */
#define MT_IB_UNIFORM 3 /* UniformBoundary::treat (ComplexBoundary_Manager.cpp:15-40):
                           mt_ib_treat only; mt_step refuses it (the reference
                           never schedules it)                                  */

#define MT_FP64 0          /* populations stored and computed in FP64               */
#define MT_FP32_SHIFTED 1  /* stored as FP32 f_q - w_q, FP32 arithmetic             */
#define MT_FP16_STORAGE 2  /* stored as FP16 f_q - w_q, FP32 arithmetic             */

/* field selectors for mt_get_field / mt_set_field */
#define MT_FIELD_F 0    /* [Q][n]  post-stream populations (user[L].df.f)       */
#define MT_FIELD_FCOL 1 /* [Q][n]  post-collision populations (user[L].df.fcol) */
#define MT_FIELD_RHO 2  /* [n]     density map                                   */
#define MT_FIELD_U 3    /* [3][n]  velocity map                                  */

typedef struct mt_case_desc {
  int32_t Q;          /* 19 or 27                                                 */
  int32_t Lmax;       /* finest level, 0..6                                       */
  int32_t collision;  /* MT_COLLIDE_*                                             */
  int32_t ib;         /* MT_IB_*                                                  */
  int32_t precision;  /* MT_FP64 / MT_FP32_SHIFTED / MT_FP16_STORAGE              */
  int32_t dims0[3];   /* level-0 extent in cells                                  */
  int32_t face_kind[6];
  double face_u[6][3];/* wall / reservoir velocity per face (lattice units)       */
  double tau0;        /* level-0 relaxation time; tau_L = 2^L (tau0-1/2)+1/2     */
  double les_C;       /* Smagorinsky constant, 0 = off                            */
  double rho0;        /* initial density and He-Luo reference density             */
  double u0[3];       /* initial velocity                                         */
  int32_t options;    /* MT_OPT_* bit flags, 0 = reference semantics throughout   */
} mt_case_desc;

/* Optional departures from the reference semantics (off by default):
 * MT_OPT_CONVECTIVE_XFER: rescale f_neq by tau_f/(2 tau_c) coarse->fine and
 *   2 tau_c/tau_f fine->coarse (the convective-scaling ratio of the
 *   non-equilibrium parts) instead of the reference's 2 tau_f/tau_c and
 *   tau_c/(2 tau_f) (LBM_Manager.cpp:1065, 1137).                                */
#define MT_OPT_CONVECTIVE_XFER 0x1
/* MT_OPT_EXPLODE_BOTH: also refill the C2F / ghost_C2F layer from the coarse
 *   level in sub-step B (the reference explodes in sub-step A only,
 *   LBM_Simulator.cpp:47-51 vs 152-217).                                        */
#define MT_OPT_EXPLODE_BOTH 0x2

typedef struct mt_case mt_case;

/* One member of a packed case (independent flows sharing every launch).  Its
 * cells live in the box [origin0, origin0 + dims0) of level-0 cells (and the
 * box scaled by 2^L on level L); cells outside the box are never its
 * neighbours, and its periodic faces wrap inside the box.  origin0 must be a
 * multiple of 4; periodic extents too. */
typedef struct mt_member_desc {
  int32_t origin0[3];
  int32_t dims0[3];
  int32_t face_kind[6];
  double face_u[6][3];
  double tau0, les_C, rho0, u0[3];
} mt_member_desc;

const char* mt_last_error(void);

/* ---- lifecycle: replaces LBM_Manager::LBM_Manager + allocMacrosAndDDF
 *      (LBM_Manager.cpp:32-95, 482-592) -------------------------------------- */
int mt_case_create(const mt_case_desc* desc, mt_case** out);
int mt_case_destroy(mt_case* c);
/* one level's cell set (any order); kinds use MT_KIND_* (Lat_Manager.hpp:245-316).
 * The arrays are copied to the device before the call returns (the caller may
 * free them); a second upload of the level replaces the first.                */
int mt_case_upload_level(mt_case* c, int level, const uint64_t* keys, const uint16_t* kinds, int64_t n);
/* Lat_Manager::dis: n finest-level keys, n x (Q-1) link fractions (-1 = no
 * wall), copied to the device before the call returns                         */
int mt_case_upload_dis(mt_case* c, const uint64_t* keys, const double* fractions, int64_t n);
/* sorts the keys on the device, builds bricks, neighbour / parent-child tables,
 * sets f = fcol = feq(rho0,u0) (solid 0) -- allocMacrosAndDDF + initialBC       */
int mt_case_finalize(mt_case* c);
/* Packed batching (no counterpart in the reference, which runs one flow per
 * process): replace the case's single member, described by mt_case_desc, by n
 * members sharing its Q, Lmax, collision, IB method, precision and options.
 * Between mt_case_create and mt_case_finalize; keys of every member are
 * uploaded together, in the case's coordinates.                              */
int mt_case_set_members(mt_case* c, const mt_member_desc* members, int32_t n);
/* Lagrangian markers on the finest level (no counterpart in the reference, whose
 * marker interface is a sketch): n markers, each with 27 finest-level cell keys
 * and kernel weights (weight 0 = unused entry), a surface element dV and the body
 * velocity ub[3].  Every finest sub-step of mt_step applies explicit direct
 * forcing to the post-collision populations: u_k and rho_k are the weighted sums
 * of the cells' moments, F_k = dV_k rho_k (ub - u_k), each cell's populations
 * shift by feq(rho, u + sum_k w F_k / rho) - feq(rho, u).  sum_k F_k adds to
 * the force mt_ib_force reports: the body's reaction -sum_k F_k with the sign of
 * the reported velocities flipped back, as in the link-wise force (a drag along
 * the physical flow is positive).  Before mt_case_finalize; single-member cases. */
int mt_case_upload_markers(mt_case* c, int64_t n, const uint64_t* keys27, const double* weights27,
                           const double* dV, const double* ub3);
/* Moving markers: the FSI marker interface the reference sketches (Infor_IB and
 * map_node_IB, the cells within reach of the moving solid points, kept up to
 * date as the solid moves: header/Grid_Class.h:101-109, 250-253,
 * header/Grid_Manager.h:50-52, Grid_Manager_dynamic.cpp:595-803), on the device.
 * Positions are in finest-level cell units (cell i spans [i, i + 1)); a
 * marker's kernel cells are the nodes within 3/2 per axis, weighted by the
 * three-point kernel of Roma, Peskin & Berger (J. Comput. Phys. 153, 1999), and
 * must all be streamed fluid cells of the finest level.
 *   mt_markers_move: new positions xyz[n][3] and velocities ub[n][3] (reported
 *     frame, as mt_get_field's; NULL keeps them) for all n markers;
 *   mt_markers_rigid: moves every marker with a rigid motion over dt finest
 *     time steps: x' = center + vel dt + Rot(omega dt)(x - center) (Rodrigues;
 *     vel, omega physical, lattice units), marker velocity vel + omega x r'
 *     (stored in the reported frame); needs positions from mt_markers_move;
 *   mt_markers_get: the positions (NULL to skip) and velocities.
 * The kernel cells, weights and band are rebuilt on the device (a sort of the
 * (cell, marker) pairs when some marker's cells change, new weights in place
 * otherwise).  The new positions act from the next collision on.  A move that
 * would put a kernel cell outside the streamed fluid cells of the finest level
 * returns MT_ERR_ARG naming the first such marker and changes nothing.  After
 * mt_case_finalize.                                                            */
int mt_markers_move(mt_case* c, int64_t n, const double* xyz, const double* ub);
int mt_markers_rigid(mt_case* c, double dt, const double* center3, const double* vel3, const double* omega3);
int mt_markers_get(mt_case* c, int64_t n, double* xyz, double* ub);
/* Remeshing as a solid moves: the reference adds and removes nodes per level as
 * its solid points move (Grid_Manager::update_nodes_near_solid / update_nodes,
 * Grid_Manager_dynamic.cpp:37-594, 1083-1300) and carries no flow state across.
 * Here a new case built for the new geometry (its level sets from
 * mt_levelset_build, say) takes the state of the old one, cell by cell on the
 * device: a key present in src with populations is copied; else its nearest
 * coarser src ancestor is exploded (feq_HL + s (f - feq_HL), s the product of
 * the transfer factors 2 tau_l / tau_{l-1}, LBM_Manager.cpp:1059-1129); else
 * its src children one level finer are averaged and coalesced (factor
 * tau_L / (2 tau_{L+1}), LBM_Manager.cpp:1131-1169); else it is new to the
 * flow and starts at feq(rho0, u_fill) (u_fill reported frame, NULL = 0).
 * Both f and fcol follow the rule; density and velocity are the moments of the
 * new f.  dst's solid and reservoir cells keep their initial state.  Both
 * cases single-member with the same Q, precision and number of levels, on one
 * device.  counts4 (may be NULL): cells copied, exploded, coalesced, filled.
 * No counterpart in the reference.                                            */
int mt_case_remap(mt_case* dst, mt_case* src, const double* u_fill3, int64_t* counts4);
int mt_case_level_info(mt_case* c, int level, int64_t* n_cells, int64_t* n_bricks);

/* ---- key lookup: replaces D_map_define<T>::find (mortonTrieMap.cpp:116-126,
 *      header/Grid_Class.h:116-137): slot index or -1 per key                   */
int mt_find(mt_case* c, int level, const uint64_t* keys, int64_t n, int32_t* slot_out);

/* ---- mt_find without any allocation.  No counterpart in the reference: for
 *      callers that cannot allocate device memory (code under HIP graph capture,
 *      tight serving loops, harnesses that forbid hipMalloc).  Same inputs, output
 *      and errors as mt_find; it stages the keys and slots in a fixed per-GPU
 *      device array of MT_FIND_NOALLOC_MAX_KEYS keys (3 MiB, held until the
 *      process ends) and returns MT_ERR_ARG for larger n.  Every call on a GPU
 *      shares that array: calls that overlap on one GPU (other threads, other
 *      cases' streams) race and must be serialised by the caller.              */
#define MT_FIND_NOALLOC_MAX_KEYS (1 << 18)
int mt_find_noalloc(mt_case* c, int level, const uint64_t* keys, int64_t n, int32_t* slot_out);

/* ---- state access in the caller's key order (user[L].df / density / velocity,
 *      header/user.h:18-55,154-167); MT_ERR_LOOKUP if a key is absent          */
int mt_get_field(mt_case* c, int level, int field, const uint64_t* keys, int64_t n, double* out);
int mt_set_field(mt_case* c, int level, int field, const uint64_t* keys, int64_t n, const double* in);

/* ---- the seams of LBM_Manager::LBMkernel (LBM_Simulator.cpp:20-229) ---- */
/* LBM_Manager::calculateMacros(level)                 LBM_Manager.cpp:1003-1050 */
int mt_macros(mt_case* c, int level);
/* Collision_Model::collide(level)   lbm_kernel/Collision_Manager.cpp:13-53,
 *                                    lbm_kernel/Collision_Cumulant.cpp:22-159   */
int mt_collide(mt_case* c, int level);
/* Stream_Model::stream(level)       lbm_kernel/Stream_Manager.cpp:43-76         */
int mt_stream(mt_case* c, int level);
/* LBM_Manager::AMR_transDDF_explosion(level)          LBM_Manager.cpp:1059-1129 */
int mt_explode(mt_case* c, int level);
/* LBM_Manager::AMR_transDDF_coalescence(level)        LBM_Manager.cpp:1131-1169 */
int mt_coalesce(mt_case* c, int level);
/* D3Q19_BoundaryCondtion_Manager::applyBoundaryCondition(face, "AfterStream"):
 * all six faces W,E,S,N,Bot,Top in turn (header/BC/D3Q19_BC_Manager.hpp:220-244,
 * LBM_Simulator.cpp:89-96)                                                      */
int mt_apply_bc(mt_case* c);
/* ComplexBoundary_Model::treat() + accumIBForce      ComplexBoundary_Manager.cpp:43-256,
 *                                                     LBM_Manager.cpp:846-852;
 * force3 (may be NULL) receives this call's momentum-exchange force            */
int mt_ib_treat(mt_case* c, double* force3);

/* ---- whole schedule: n coarse steps of LBM_Manager::fluidSimulate's loop body
 *      (LBMkernel(0); calculateMacros()), LBM_Simulator.cpp:11-17.  Runs the
 *      fused per-level kernels, captured as HIP graphs when graphs are on.    */
int mt_step(mt_case* c, int n_coarse);
int mt_set_graphs(mt_case* c, int on);
/* Per-level relaxation times of one member as given, in place of
 * tau_L = 2^L (tau0 - 1/2) + 1/2 from the descriptor.  The reference computes
 * each level's tau from the viscosity and that level's conversion factors
 * (LBM_Manager::setUnit, LBM_Manager.cpp:814-817), which can differ from the
 * doubling rule in the last bit; the transfer factors 2 tau_f / tau_c and
 * tau_c / (2 tau_f) (LBM_Manager.cpp:1065, 1137) follow from the values given.
 * n_levels must be Lmax + 1, each tau > 1/2.                                   */
int mt_set_tau(mt_case* c, int32_t member, const double* tau, int32_t n_levels);
/* How mt_step issues the levels' fused updates (no counterpart in the
 * reference; every schedule computes the same bits):
 *   MT_SCHED_MERGED   one launch per finest sub-step on one stream, carrying
 *                     each coarser level's update whose span of finest
 *                     sub-steps ends with it (the default; where explosions do
 *                     not run inside the updates, as MT_SCHED_SERIAL);
 *   MT_SCHED_STREAMS  each level on its own stream, a level's update
 *                     concurrent with its finer pair, events between streams;
 *   MT_SCHED_SERIAL   the levels' updates one after another on one stream.   */
#define MT_SCHED_MERGED 0
#define MT_SCHED_STREAMS 1
#define MT_SCHED_SERIAL 2
int mt_set_schedule(mt_case* c, int schedule);
/* The schedule mt_step runs (no counterpart in the reference): the one set,
 * except MT_SCHED_MERGED where the explosions do not run inside the updates
 * (an interface whose coalescence cannot be fused, e.g. an F2C parent with a
 * treated child), which runs as MT_SCHED_SERIAL, or as MT_SCHED_STREAMS when
 * no schedule was set.  Valid after mt_case_finalize.                          */
int mt_get_schedule(mt_case* c, int32_t* schedule);
/* The coarse steps one captured HIP graph holds (no counterpart in the
 * reference): mt_step(n) replays graphs of this many steps and one of
 * n mod this many (or of this many) that ends the call; each graph is captured
 * on its first replay, per level buffer parities, length and end flag.        */
int mt_steps_per_graph(void);
/* Launch-configuration autotuner (no counterpart in the reference): tries the
 * fused kernel's block -> brick maps and the schedules MT_SCHED_STREAMS and
 * MT_SCHED_MERGED, twice each, for `steps` coarse steps of the real run (+1
 * untimed) per try, and keeps the fastest (the run advances 8 (steps + 1)
 * coarse steps).  Results do not depend on the choice.  ms_per_step[4] (may
 * be NULL) gets each candidate's faster try, in the order (map 0, streams),
 * (map 1, streams), (map 0, merged), (map 1, merged).                         */
int mt_autotune(mt_case* c, int steps, double* ms_per_step);
/* per-kernel event timing of mt_step (off by default; forces graphs off):
 * ms[level] accumulates the fused update kernel's time, calls[level] its count;
 * a launch that carries several levels' updates counts under the finest.     */
int mt_set_timing(mt_case* c, int on);
int mt_get_timing(mt_case* c, double* ms, int64_t* calls, int max_levels);
int mt_sync(mt_case* c);

/* ---- reductions (FP64, fixed order) ------------------------------------- */
/* mass and momentum over lat_f and F2C of one level, and a NaN count       */
int mt_reduce(mt_case* c, int level, double* mass, double* momentum3, int64_t* nan_count);
/* accumulated IB force since the last reset (LBM_Manager.cpp:846-852, 964-979) */
/* ---- run diagnostics as device reductions (LBM_Manager::infoPrint,
 *      LBM_Manager.cpp:914-962; diagL2MaxF / diagL2RhoMin, :854-912).
 *      mt_set_diagnostics(c, 1) makes every update of mt_step leave its
 *      density and velocity, so the previous sub-step's velocity (the v_old that
 *      calculateMacros keeps, :1003-1050) is at hand; it costs each update 4
 *      reals per cell of extra writes.  mt_diagnostics then reduces, over the
 *      state at the end of the last mt_step (FP64 sums in a fixed order):      */
typedef struct mt_diag {
  double velocity_error;  /* sqrt(sum |v - v_old|^2) / sqrt(sum |v|^2 + 1e-16), lat_f of every level */
  double max_v[8];        /* per level: max |v| over lat_f and lat_overlap_F2C                   */
  int64_t n_big[8];       /* per level: cells with |v| > 1                                       */
  /* finest level, lat_f: diagL2MaxF                                                               */
  double max_f, max_fcol;
  uint64_t max_f_key, max_fcol_key;
  int32_t max_f_q, max_fcol_q;
  /* diagL2RhoMin: minimum sum f and sum fcol, counts of sum f <= 0, sum fcol <= 0, sum f < 0.5,
   * the key of the minimum sum f and whether it is a solid-boundary (dis) cell                  */
  double min_rho_f, min_rho_fcol;
  int64_t n_neg_f, n_neg_fcol, n_low_f;
  uint64_t worst_key;
  int32_t worst_is_dis;
  int32_t valid_v_old;    /* 1 when every update since the previous mt_step left its macros  */
} mt_diag;
int mt_set_diagnostics(mt_case* c, int on);
int mt_diagnostics(mt_case* c, mt_diag* out);

/* ---- the per-level cell sets built on the GPU from a geometry: replaces
 *      Grid_Manager::initial + Lat_Manager::voxelize (Grid_Manager.cpp:40-134,
 *      Lat_Voxelize.cpp:17-1285) for a body given as an analytic sphere or a
 *      closed triangle mesh (an STL surface, Shape.cpp:341-548).  Level L+1
 *      covers the level-L cells inside the body or within shells[L] (level-0
 *      cell units) of its surface; the finest level's cells inside are solid;
 *      link fractions come from the first crossing of the surface along each
 *      link from a cell centre outside (Moller-Trumbore on the triangles,
 *      Shape.cpp:809-843).  The sets (ascending keys) and dis are read back with
 *      the copy calls and feed mt_case_upload_level / mt_case_upload_dis.      */
#define MT_SHAPE_SPHERE 0
#define MT_SHAPE_MESH 1
typedef struct mt_levelset_desc {
  int32_t Q, Lmax;
  int32_t dims0[3];       /* level-0 cells per axis                                       */
  int32_t face_kind[6];   /* MT_BC_*: periodic pairs wrap the level-0 refinement band;
                             the level-0 face cells get the face bits                      */
  int32_t reservoir;      /* 1: a frozen lat_bc shell around level 0 (Lat_Voxelize.cpp:242-288) */
  int32_t shape;          /* MT_SHAPE_*                                                   */
  int32_t solid;          /* 1: the body is solid on the finest level, with link fractions */
  double shells[8];
  double center[3], radius;   /* MT_SHAPE_SPHERE                                          */
  const double* tri;          /* MT_SHAPE_MESH: [n_tri][3 vertices][xyz], level-0 units   */
  int64_t n_tri;
} mt_levelset_desc;
typedef struct mt_levelset mt_levelset;
int mt_levelset_build(const mt_levelset_desc* d, mt_levelset** out);
int mt_levelset_counts(mt_levelset* ls, int64_t* n_cells_per_level, int64_t* n_dis);
int mt_levelset_copy(mt_levelset* ls, int level, uint64_t* keys, uint16_t* kinds);
int mt_levelset_copy_dis(mt_levelset* ls, uint64_t* keys, double* fractions);
int mt_levelset_destroy(mt_levelset* ls);

/* ---- field snapshots and output files (the reference's IO_Manager::writeFlow ->
 *      IO_TECPLOTH::write, tecplot.cpp:16-177, and the AMR node dump of
 *      Obj_Manager::initial, Obj_Manager.cpp:483-489).  mt_snapshot compacts the
 *      cells of one level whose kind has a bit of kind_mask (0: lat_f |
 *      lat_overlap_F2C) on the device, in ascending key order, with density and
 *      velocity (u as [3][n]); with keys, rho and u all NULL it returns the count
 *      in *n, otherwise *n holds the buffers' capacity on entry.
 *      mt_write_tecplot writes the FEBRICK zone of every level's lat_f and F2C
 *      cells: node coordinates in units of dx0 per level-0 cell, U,V,W = velocity
 *      x convert_u.  mt_write_nodes writes the finest level's keys, one 64-bit
 *      binary string per line.                                                  */
int mt_snapshot(mt_case* c, int level, uint32_t kind_mask, int64_t* n, uint64_t* keys, double* rho, double* u);
int mt_write_tecplot(mt_case* c, const char* path, double dx0, double convert_u);
int mt_write_nodes(mt_case* c, const char* path);

/* ---- per-member failure scan: the NaN checks of LBMkernel (LBM_Simulator.cpp:33-41,
 *      52-60, 78-86, 102-111, 128-137) reduced on the device per member of a packed
 *      case: bad_cells[i] = fluid and F2C cells of member i, all levels, holding a
 *      non-finite population.  n_members must equal the case's member count.    */
int mt_scan_members(mt_case* c, int32_t n_members, int64_t* bad_cells);
int mt_ib_force(mt_case* c, double* sum3, int64_t* count, int reset);
/* the same per member: sum3[m][3] over member m's cells, fixed order          */
int mt_ib_force_members(mt_case* c, double* sum3, int32_t n_members, int64_t* count, int reset);
/* Continuous batching (no counterpart in the reference): copy a member's
 * populations and macroscopic maps from one case to a member of another case
 * with the same cells (its box may sit elsewhere, by whole bricks); both cases
 * are settled first.  A server repacks its running members this way when
 * requests arrive or finish.                                                  */
int mt_case_copy_member(mt_case* dst, int32_t dst_member, mt_case* src, int32_t src_member);

/* Continuous-batching server (no counterpart in the reference, which runs one
 * flow per process; these are the submit / poll entry points of a serving
 * loop).  A request is a single-member case built with mt_case_create,
 * mt_case_upload_level and mt_case_upload_dis and not finalized; the server
 * takes it over at submission (the handle is freed then, also on an error).
 * Requests must share Q, Lmax, collision, IB method and options and fit a
 * slot of 96 x 48 x 48 level-0 cells.  mt_server_step runs one chunk of
 * `chunk_steps` coarse steps over every running request, packed into one case;
 * every `admit_every` chunks (or at once when no request is left running) it
 * admits waiting requests in submission order while their device footprint
 * fits `hbm_budget` bytes.  A request finishes at max_steps or when its drag
 * coefficient over the last chunk moved by less than drag_tol from the
 * previous chunk's; Cd = 2 Fx / (U^2 pi d^2 / 4) with U = cd_U, d = cd_d, and
 * each request computes exactly what it computes alone.  After every chunk each
 * member is scanned for non-finite populations (mt_scan_members): a member with
 * any, or with a non-finite drag, ends with status MT_REQ_DIVERGED and leaves
 * the batch at the next step; the others go on unchanged.  mt_server_poll hands
 * out finished requests (id, steps, last chunk's Cd, status) in completion
 * order.  A repack or a step that fails stops the server: later calls return
 * MT_ERR_STATE.                                                             */
#define MT_REQ_FINISHED 0 /* max_steps reached or the drag settled */
#define MT_REQ_DIVERGED 1 /* non-finite state or drag: evicted */
typedef struct mt_server mt_server;
int mt_server_create(int32_t precision, int32_t chunk_steps, int32_t admit_every, int64_t hbm_budget,
                     mt_server** out);
int mt_server_submit(mt_server* s, mt_case* request, int64_t id, int32_t max_steps, double drag_tol,
                     double cd_U, double cd_d);
int mt_server_step(mt_server* s, int32_t* n_running, int32_t* n_waiting);
int mt_server_poll(mt_server* s, int64_t* ids, int32_t* steps, double* cd_last, int32_t* status,
                   int32_t max_results, int32_t* n_results);
int mt_server_destroy(mt_server* s);

#ifdef __cplusplus
}
#endif
#endif /* MT_AMR_H */
