/*
 * mpas_hip.h — C ABI of the MI355X (gfx950) shallow-water RK4 step with FCT
 * tracer advection and del2 mixing on the spherical Voronoi (TRiSK) mesh.
 *
 * The entry points stand in for the FleCSI tasks of MPAS-O-FleCSI 2.0
 * (reference tree ipdemes/mpas-o-flecsi-2.0):
 *   apps/sw/tasks/solve.hh:12-136      shallow-water tasks
 *   apps/ocean/tasks/solve.hh:12-111   FCT / del2 tracer tasks
 *   mpasoflecsi/specialization/mesh.cc:67-116  topology flattening
 * A reference task `f(mesh::accessor m, field accessors..., scalars)` maps to
 * `mpas_<group>_f(ctx, device pointers..., scalars)`: the context plays the
 * mesh accessor (it owns the flattened connectivity and the static mesh
 * fields on the device), the field accessors become device pointers in the
 * reference's own layouts:
 *   vltensor  (per entity, [nLevels])            -> double[n * nLevels]
 *   vltracer  (per cell,  [nLevels][nTracers])   -> double[nCells * nLevels * nTracers]
 * All calls are asynchronous on the context's stream and return 0 on
 * success or a negative mpas_status; mpas_last_error() describes the last
 * failure of the calling thread.  No C++ exception crosses this boundary
 * (the reference aborts through flog_fatal / flog_assert instead,
 * mpasoflecsi/specialization/definition.hh:197-200).  A context is not
 * thread-safe; distinct contexts may be used from distinct threads.
 */
#ifndef MPAS_HIP_H
#define MPAS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mpas_ctx mpas_ctx;
typedef struct mpas_host_mesh mpas_host_mesh;

enum mpas_status {
  MPAS_OK = 0,
  MPAS_EINVAL = -1, /* bad argument, unknown name, inconsistent mesh */
  MPAS_EHIP = -2,   /* HIP runtime error (no device, launch failure, ...) */
  MPAS_ESHAPE = -3, /* buffer size / shape mismatch */
  MPAS_ESTATE = -4, /* call out of order (e.g. step before state upload) */
  MPAS_ENOMEM = -5, /* device allocation failed */
  MPAS_ENONFINITE = -6 /* a state holds NaN or Inf (uploads, mpas_sw_state_check) */
};

/* Tracer transport inside the shallow-water RK4 stage. */
enum mpas_tracer_scheme {
  MPAS_TRACER_NONE = 0,    /* no tracers (nTracers must be 0)            */
  MPAS_TRACER_CENTRED = 1, /* compute_scalar_tend, apps/sw/tasks/solve.cc:237-270 */
  MPAS_TRACER_FCT = 2      /* monotone FCT, apps/ocean/tasks/solve.cc:40-326 */
};

/*
 * Mesh as read from an MPAS mesh file.  Connectivity keeps the file's
 * conventions (1-based ids, 0 = padding, row-major with the given strides);
 * the library drops the zeros and rebases the ids exactly as
 * read_connectivity_2D does (mpasoflecsi/specialization/definition.hh:670-718)
 * and derives the transposed lists in ascending source id as transpose_cnx
 * does (mpasoflecsi/specialization/mesh.cc:108-116).  Everything is copied
 * during mpas_ctx_create; the caller keeps ownership of its arrays.
 */
typedef struct mpas_mesh_desc {
  int32_t nCells, nEdges, nVertices;
  int32_t maxEdges;     /* row stride of verticesOnCell/edgesOnCell/cellsOnCell */
  int32_t maxEdges2;    /* row stride of edgesOnEdge/weightsOnEdge (16: the
                           reference reads 16 columns, io/helpers.hh:172-184) */
  int32_t vertexDegree; /* row stride of kiteAreasOnVertex (3) */
  const int32_t *verticesOnCell; /* [nCells * maxEdges]  */
  const int32_t *edgesOnCell;    /* [nCells * maxEdges]  */
  const int32_t *cellsOnCell;    /* [nCells * maxEdges]  */
  const int32_t *verticesOnEdge; /* [nEdges * 2]         */
  const int32_t *edgesOnEdge;    /* [nEdges * maxEdges2] */
  const double *areaCell;          /* [nCells]                 */
  const double *dvEdge, *dcEdge;   /* [nEdges]                 */
  const double *areaTriangle;      /* [nVertices]              */
  const double *kiteAreasOnVertex; /* [nVertices*vertexDegree] */
  const double *fVertex;           /* [nVertices]              */
  const double *weightsOnEdge;     /* [nEdges * maxEdges2]     */
  /* Cartesian positions, used only by the FCT stencil setup
     (apps/ocean/tasks/init_functions.cc:83-328); may be NULL otherwise. */
  const double *xCell, *yCell, *zCell, *xVertex, *yVertex, *zVertex;
  const double *meshDensity;    /* [nCells]; NULL = 1 everywhere          */
  const int32_t *maxLevelCell;  /* [nCells]; NULL = nLevels everywhere     */
  const int32_t *boundaryCell;  /* [nCells * nLevels]; NULL = 0 everywhere */
} mpas_mesh_desc;

typedef struct mpas_params {
  int32_t nLevels;       /* reference nVertLevels (constants.hh:16)          */
  int32_t nTracers;      /* reference maxTracers  (constants.hh:18)          */
  int32_t tracer_scheme; /* enum mpas_tracer_scheme                           */
  int32_t freeze_u;      /* SW test case 1 keeps u fixed (solve.cc:307-313)   */
  double gravity;        /* mpas_constants::gravity = 9.80616                 */
  double sphere_radius;  /* a, used by the FCT stencil setup                  */
  double coef3rd;        /* 3rd/4th-order blend, 0.25 (ocean/solve.cc:11)     */
  double visc2;          /* del2 viscosity on u (m^2/s); 0 = off              */
  double kappa2;         /* del2 tracer diffusion coefficient; 0 = off        */
} mpas_params;

const char *mpas_last_error(void);

/* Flattens the topology (mesh.cc:67-116), builds the FCT stencils
   (init_functions.cc:18-447) when tracer_scheme == FCT or kappa2 != 0, and
   uploads every static mesh array to the current HIP device. */
int mpas_ctx_create(const mpas_mesh_desc *mesh, const mpas_params *params,
                    mpas_ctx **out);
int mpas_ctx_destroy(mpas_ctx *ctx);
/* Stream for subsequent launches (hipStream_t); NULL = the context's own. */
int mpas_ctx_set_stream(mpas_ctx *ctx, void *hip_stream);
int mpas_ctx_sync(mpas_ctx *ctx);
/* Sizes: out[0..4] = nCells, nEdges, nVertices, nLevels, nTracers. */
int mpas_ctx_dims(const mpas_ctx *ctx, int64_t out[5]);

/* Mesh orders (no counterpart in the reference, whose order is the file's):
   a context with the entities of src renumbered, cells in the given order,
   reverse Cuthill-McKee or along a 3-D Morton curve of the cell centres,
   edges by their (lower, higher) new cell pair, vertices by their sorted
   new cell triple.  Every list keeps its entries' order, and the sums the
   reference forms in ascending edge id stay in the original ids' order, so
   a state advanced on the reordered context equals, entity for entity, the
   one advanced on src bit for bit.  Fields of a reordered context are in
   its numbering; mpas_ctx_order gives the original id of each entity
   (arrays of nCells, nEdges, nVertices; any may be NULL). */
enum { MPAS_ORDER_GIVEN = 0, MPAS_ORDER_RCM = 1, MPAS_ORDER_MORTON = 2 };
int mpas_ctx_create_reordered(const mpas_ctx *src, int order, mpas_ctx **out);
int mpas_ctx_order(const mpas_ctx *ctx, int32_t *cell_of, int32_t *edge_of,
                   int32_t *vertex_of);

/* The context's current stream (hipStream_t). */
void *mpas_ctx_stream(const mpas_ctx *ctx);

/* Per-kernel timing of the fused step (no reference counterpart: the
   reference has no timers).  While enabled, every kernel launch of
   mpas_rk4_step is bracketed by HIP events on the context's stream and
   tagged with its compulsory HBM bytes (every array element the launch
   reads or writes, counted once).  mpas_prof_report drains the stream and
   writes a JSON array of {"kernel", "launches", "total_ms",
   "bytes_per_launch"} into buf (NUL-terminated, at most cap bytes);
   enabling clears earlier records. */
int mpas_prof_enable(mpas_ctx *ctx, int on);
int mpas_prof_report(mpas_ctx *ctx, char *buf, size_t cap);

/* Device buffers for callers without their own allocator (tests, tools). */
int mpas_buf_alloc(mpas_ctx *ctx, size_t bytes, void **dptr);
int mpas_buf_free(mpas_ctx *ctx, void *dptr);
int mpas_buf_upload(mpas_ctx *ctx, void *dptr, const void *host, size_t bytes);
int mpas_buf_download(mpas_ctx *ctx, void *host, const void *dptr, size_t bytes);
/* Device-to-device copy by a plain streaming kernel (k_copy), asynchronous:
   a known-size reference for bandwidth and HBM-counter calibration. */
int mpas_buf_copy(mpas_ctx *ctx, void *dst, const void *src, size_t bytes);

/* Host-side mesh preparation alone (no device needed): the topology
   flattening of mesh::initialize (mesh.cc:67-116) and, when the parameters
   ask for FCT or del2 mixing, the setup of init_functions.cc:18-447.
   mpas_ctx_create runs the same preparation; mpas_ctx_host_mesh returns the
   context's copy (owned by the context). */
int mpas_host_mesh_create(const mpas_mesh_desc *mesh, const mpas_params *params,
                          mpas_host_mesh **out);
int mpas_host_mesh_destroy(mpas_host_mesh *hm);
/* The host mesh renumbered as mpas_ctx_create_reordered does (a new mesh,
   freed with mpas_host_mesh_destroy), and the original id of each entity. */
int mpas_host_mesh_reorder(const mpas_host_mesh *hm, int order, mpas_host_mesh **out);
int mpas_host_mesh_order(const mpas_host_mesh *hm, int32_t *cell_of, int32_t *edge_of,
                         int32_t *vertex_of);
const mpas_host_mesh *mpas_ctx_host_mesh(const mpas_ctx *ctx);

/* Host copies of the flattened connectivity, for checks.  which = "c2v",
   "c2e", "c2c", "e2v", "e2e", "v2c", "e2c", "v2e".  offsets gets rows+1
   entries, indices *count entries (0-based); pass NULL arrays to query the
   sizes in *rows / *count. */
int mpas_mesh_connectivity(const mpas_host_mesh *hm, const char *which,
                           int32_t *offsets, int32_t *indices, int64_t *rows,
                           int64_t *count);
/* Host copies of the one-time setup products in the reference layouts:
   int32: maxLevelEdgeTop[E], nAdvCellsForEdge[E], advCellsForEdge[E*16],
          highOrderAdvMask[E*nLevels];
   double: edgeSignOnCell[C*8], meshScalingDel2[E], derivTwo[E*16*2],
          advCoef[E*16], advCoef_3rd[E*16]. */
int mpas_mesh_setup_field(const mpas_host_mesh *hm, const char *name,
                          void *host, int64_t n);

/* ---- shallow-water tasks (apps/sw/tasks/solve.hh) ----------------------
   Device pointers; sizes follow the context (n*nLevels, tracers
   n*nLevels*nTracers). */

/* solve.hh:12-27 / solve.cc:11-107 */
int mpas_sw_compute_solve_diagnostics(mpas_ctx *ctx, const double *h,
                                      const double *u, double *h_edge,
                                      double *h_vertex, double *circulation,
                                      double *vorticity, double *ke,
                                      double *pv_edge, double *pv_vertex);
/* solve.hh:29-49 / solve.cc:110-141 */
int mpas_sw_init_provis(mpas_ctx *ctx, const double *u_old, const double *h_old,
                        const double *tracers_old, const double *h_edge_old,
                        const double *ke_old, const double *pv_edge_old,
                        double *u_provis, double *h_provis,
                        double *tracers_provis, double *h_edge_provis,
                        double *ke_provis, double *pv_edge_provis);
/* solve.hh:51-61 / solve.cc:144-166 */
int mpas_sw_setup_timestep(mpas_ctx *ctx, const double *u_old,
                           const double *h_old, const double *tracers_old,
                           double *u_new, double *h_new, double *tracers_new);
/* solve.hh:63-76 / solve.cc:169-234 (h_s = bottomDepth, [nCells]) */
int mpas_sw_compute_tend(mpas_ctx *ctx, const double *h_edge, const double *h,
                         const double *h_s, const double *u, const double *ke,
                         const double *pv_edge, double *tend_h, double *tend_u);
/* solve.hh:78-85 / solve.cc:237-270 */
int mpas_sw_compute_scalar_tend(mpas_ctx *ctx, const double *h_edge,
                                const double *u, const double *tracers,
                                double *tracer_tend);
/* solve.hh:87-97 / solve.cc:273-314 */
int mpas_sw_compute_substep(mpas_ctx *ctx, double *u_provis,
                            const double *u_old, const double *u_tend,
                            double *h_provis, const double *h_old,
                            const double *h_tend, double *tracers_provis,
                            const double *tracers_old,
                            const double *tracers_tend, double rkweight);
/* solve.hh:100-107 / solve.cc:317-341 */
int mpas_sw_accumulate_update(mpas_ctx *ctx, double *u_new, const double *u_tend,
                              double *h_new, const double *h_tend,
                              double *tracers_new, const double *tracers_tend,
                              double rkweight);
/* solve.hh:110-115 / solve.cc:344-367 */
int mpas_sw_finalize_timestep(mpas_ctx *ctx, const double *u_old,
                              double *u_new, const double *h_new,
                              double *tracers_new);
/* solve.hh:117-129 / solve.cc:370-401 */
int mpas_sw_timeshift(mpas_ctx *ctx, double *u_old, double *h_old,
                      double *tracers_old, double *h_edge_old, double *ke_old,
                      double *pv_edge_old, const double *u_new,
                      const double *h_new, const double *tracers_new,
                      const double *h_edge_new, const double *ke_new,
                      const double *pv_edge_new);
/* solve.hh:131-136 / solve.cc:404-436: case-1 relative L2 error of h
   (level 0) against the advected cosine bell; *err is written on the host
   after the stream has drained.  latCell/lonCell: device [nCells]. */
int mpas_sw_compute_error(mpas_ctx *ctx, const double *h,
                          const double *latCell, const double *lonCell,
                          double sphere_radius, double elapsed, double *err);
/* No reference counterpart (the reference's compute_tend has no dissipation,
   solve.cc:228-231): del2 viscosity on u after MPAS-Ocean's del2 momentum
   mixing, tend_u += meshScalingDel2 * visc2 * (grad(div) - curl(vorticity)),
   with divergence [nCells*nLevels] written as a by-product. */
int mpas_sw_del2_viscosity(mpas_ctx *ctx, const double *u,
                           const double *vorticity, double *divergence,
                           double *tend_u);

/* ---- FCT / del2 tracer tasks (apps/ocean/tasks/solve.hh) --------------- */

/* solve.hh:12-15 / solve.cc:13-26 */
int mpas_fct_copy_tracers(mpas_ctx *ctx, const double *tracers,
                          double *tracersNew, double *tracersPrev);
/* solve.hh:17-18 / solve.cc:28-38 */
int mpas_fct_reinitialize_tracer_tendencies(mpas_ctx *ctx, double *tend);
/* solve.hh:20-23 / solve.cc:40-57 */
int mpas_fct_calc_normal_thickness_flux(mpas_ctx *ctx,
                                        const double *layerThicknessEdge,
                                        const double *normalVelocity,
                                        double *normalThicknessFlux);
/* solve.hh:25-36 / solve.cc:60-97 (the vertical velocity w is not read) */
int mpas_fct_calc_provisional_layer_thicknesses(
    mpas_ctx *ctx, const double *layerThickness,
    const double *normalThicknessFlux, double *hProv, double *hProvInv,
    double *hNewInv, double dt);
/* solve.hh:38-42 / solve.cc:99-118 (lowOrderFlux: [nEdges*nLevels*nTracers]) */
int mpas_fct_low_order(mpas_ctx *ctx, const double *normalThicknessFlux,
                       const double *tracers, double *lowOrderFlux);
/* solve.hh:44-54 / solve.cc:120-167 */
int mpas_fct_high_order(mpas_ctx *ctx, const double *normalThicknessFlux,
                        const double *tracers, double *highOrderFlux);
/* solve.hh:56-74 / solve.cc:169-285 (tend accumulates, as in the
   reference, so clear it first with reinitialize_tracer_tendencies) */
int mpas_fct_correct_and_rescale(mpas_ctx *ctx, const double *hProvInv,
                                 const double *layerThickness,
                                 const double *normalThicknessFlux,
                                 const double *lowOrderFlux,
                                 const double *tracers, double *highOrderFlux,
                                 double *flxIn, double *flxOut, double *tend,
                                 double *tracerMax, double *tracerMin,
                                 double dt);
/* solve.hh:76-87 / solve.cc:287-326 */
int mpas_fct_accumulate_tend(mpas_ctx *ctx, const double *hProvInv,
                             const double *layerThickness,
                             const double *highOrderFlux, const double *tracers,
                             double *tracersProv, double *tend, double dt);
/* solve.hh:89-99 / solve.cc:328-366 */
int mpas_hmix_del2_tend(mpas_ctx *ctx, const double *layerThicknessEdge,
                        const double *tracers, double *tend, double coef);
/* solve.hh:101-107 / solve.cc:368-399 */
int mpas_fct_step_tracers_forward(mpas_ctx *ctx, const double *tend,
                                  const double *tracers, double *tracersNew,
                                  double *tracersPrev, double rk_weight,
                                  double rk_subweight);
/* solve.hh:109-111 / solve.cc:401-414 */
int mpas_fct_update_tracers(mpas_ctx *ctx, const double *tracersNew,
                            double *tracers);

/* ---- fused, device-resident shallow-water run (apps/sw/run.cc) ---------
   The context owns the prognostic state (u, h, tracers) and every
   diagnostic and tendency field of apps/sw/state.hh:74-87. */

/* Host arrays: h [nCells*nLevels], u [nEdges*nLevels], tracers
   [nCells*nLevels*nTracers] (NULL when nTracers == 0), h_s [nCells].
   Runs the initial compute_solve_diagnostics (apps/sw/run.cc:88-92). */
int mpas_sw_state_upload(mpas_ctx *ctx, const double *h, const double *u,
                         const double *tracers, const double *h_s);
int mpas_sw_state_download(mpas_ctx *ctx, double *h, double *u,
                           double *tracers);
/* Non-finite values: the uploads (mpas_sw_state_upload, mpas_batch_upload)
   refuse h, u, tracers or h_s holding NaN or Inf with MPAS_ENONFINITE,
   naming the field and the first index, as the reference's HDF5 reads abort
   a conversion of NaN or Inf (mpasoflecsi/specialization/definition.hh:
   165-191).  The task and step kernels do not test values: NaN and Inf
   propagate through them by IEEE arithmetic as through the reference's task
   loops (apps/sw/tasks/solve.cc; apps/sw/finalize.cc:24-25 only compares
   the error).  mpas_sw_state_check counts the non-finite values of the
   resident h, u and tracers on the device (count in *nonfinite, may be
   NULL): 0 when every value is finite, else MPAS_ENONFINITE. */
int mpas_sw_state_check(mpas_ctx *ctx, int64_t *nonfinite);
/* Host copy of a field of the current state: "h", "u", "tracers",
   "h_edge", "ke", "pv_edge", "vorticity", "pv_vertex", "h_s". */
int mpas_sw_field_download(mpas_ctx *ctx, const char *name, double *host,
                           int64_t n);
/* nsteps calls of advance_timestep (apps/sw/run.cc:13-70, loop 97-106),
   with the stage-fused kernels; the result equals the per-task sequence. */
int mpas_rk4_step(mpas_ctx *ctx, double dt, int nsteps);
/* ---- allocation-free step ------------------------------------------------
   No counterpart in the reference.  For callers that cannot allocate device
   memory (code under HIP graph capture, serving loops, harnesses that forbid
   hipMalloc): mpas_rk4_step on a state the caller holds in device memory,
   h [nCells*nLevels] and u [nEdges*nLevels] and tracers
   [nCells*nLevels*nTracers] (NULL without tracers) advanced in place, h_s
   [nCells] read.  The initial diagnostics are formed first, so the result
   equals mpas_sw_state_upload of the same state followed by mpas_rk4_step.
   Every intermediate lives in one static device array of
   MPAS_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends; a context
   whose step needs more returns MPAS_ESHAPE (64 MiB holds a level-6 mesh of
   40,962 cells with up to 8 tracers).  All calls on one GPU share that
   array, so calls that overlap in time on one GPU (several streams or
   threads) race.  Asynchronous on the context's stream (mpas_ctx_sync). */
#define MPAS_NOALLOC_SCRATCH_BYTES (64u << 20)
int mpas_rk4_step_noalloc(mpas_ctx *ctx, double dt, int nsteps, double *h, double *u,
                          double *tracers, const double *h_s);
/* nsteps calls of the tracer-only ocean driver's rk4_timestep
   (apps/ocean/run.cc:10-50, loop 70-72): the uploaded
   u (normalVelocity) and h (layerThickness) stay fixed, layerThicknessEdge
   is the h_edge of the upload (init_functions.cc:30-45), and the tracers
   are advanced with FCT and, when kappa2 != 0, del2 mixing, by the same
   stage-fused kernels; the result equals the per-task sequence. */
int mpas_ocean_rk4_step(mpas_ctx *ctx, double dt, int nsteps);

/* ---- tangent-linear model and discrete adjoint of mpas_rk4_step ----------
   No counterpart in the reference (it has no derivative code); the forward
   they differentiate is apps/sw/run.cc:13-70 with the FCT tracer tendency of
   apps/ocean/tasks/solve.cc:40-366.  The FCT limiter factors are frozen at
   the forward pass's values.  Every pointer below is a device array in the
   state layouts of mpas_sw_state_upload (tracer pointers may be NULL when
   nTracers == 0); mpas_sw_state_upload must have been called once (it sets
   h_s).  Supports tracer_scheme FCT (or no tracers) with every level active.
   Calls are ordered on the context's stream. */

/* Allocates the tape: 4*checkpoint_every + 1 stage records. */
int mpas_adj_init(mpas_ctx *ctx, int checkpoint_every);
/* Runs nsteps fused steps from (u0, h0, tr0), keeping a checkpoint of the
   state every checkpoint_every steps; the final state goes to (u1, h1, tr1)
   (each may be NULL) and stays resident as the context's state. */
int mpas_adj_forward(mpas_ctx *ctx, double dt, int nsteps, const double *u0,
                     const double *h0, const double *tr0, double *u1, double *h1,
                     double *tr1);
/* Gradient of a scalar loss with respect to the initial state of the last
   mpas_adj_forward, given its gradient with respect to the final state.
   Replays each checkpoint segment into the tape, then sweeps it in reverse. */
int mpas_adj_backward(mpas_ctx *ctx, const double *u1_bar, const double *h1_bar,
                      const double *tr1_bar, double *u0_bar, double *h0_bar,
                      double *tr0_bar);
/* Tangent-linear run: nsteps steps of the state from (u0, h0, tr0) and of the
   perturbation (du0, dh0, dtr0) along it; outputs may be NULL.  Leaves the
   context's resident state untouched. */
int mpas_tlm_forward(mpas_ctx *ctx, double dt, int nsteps, const double *u0,
                     const double *h0, const double *tr0, const double *du0,
                     const double *dh0, const double *dtr0, double *u1, double *h1,
                     double *tr1, double *du1, double *dh1, double *dtr1);
/* Check hook: nsteps forward steps from (u0, h0, tr0) that use the limiter
   factors held in the tape instead of forming them (the map the TLM
   differentiates).  The tape holds steps 0 .. K-1 of the last forward run
   after mpas_adj_backward; nsteps must not exceed them. */
int mpas_adj_frozen_forward(mpas_ctx *ctx, int nsteps, const double *u0,
                            const double *h0, const double *tr0, double *u1,
                            double *h1, double *tr1);

/* Tile and schedule of the TLM and adjoint (no reference counterpart):
   workgroup sizes of the adjoint and tangent-linear kernels (multiples of
   64 up to 512), whether each segment's replay runs on a second stream
   beside the sweep of the segment after it (a second tape, when the device
   has the memory), and whether each segment's reverse sweep is replayed
   from a hipGraph.  Results do not depend on them, bit for bit.
   mpas_adj_autotune times gradients (forward with checkpoints + sweep) of
   nsteps steps from the resident state for each candidate (schedule, then
   the adjoint kernels' size) and TLM runs for the tangent-linear kernels'
   size, each the best of two runs after a warm-up run, keeps the fastest,
   restores the resident state and writes the trials as JSON into report;
   the tape is left empty. */
int mpas_adj_config(mpas_ctx *ctx, int ad_block, int tl_block, int overlap, int graph);
int mpas_adj_get_config(const mpas_ctx *ctx, int *ad_block, int *tl_block, int *overlap,
                        int *graph);
int mpas_adj_autotune(mpas_ctx *ctx, double dt, int nsteps, char *report, size_t cap);

/* ---- ensembles -------------------------------------------------------------
   No counterpart in the reference (one mesh per run there): many members,
   on the meshes of several contexts, advanced together by the fused step.
   Member i runs on meshes[mesh_of[i]] with time step dt[i]; every context
   must use one device, one tracer count and scheme, and one level.  Members
   sharing a mesh and a time step form a group stored member-minor (member
   g is level g of one state; they must share h_s); each member's state
   evolves exactly as mpas_rk4_step would evolve it alone. */
typedef struct mpas_batch mpas_batch;
enum {
  MPAS_BATCH_GRAPH = 0,      /* one launch per phase for all groups, the step
                                captured in a hipGraph and replayed */
  MPAS_BATCH_LAUNCH = 1,     /* the same launches, issued directly */
  MPAS_BATCH_PERSISTENT = 2, /* every phase of up to 20 steps in one
                                cooperative launch, grid barriers between
                                phases (meshes of at most 6 edges per cell,
                                0, 1, 2 or 4 FCT tracers) */
  MPAS_BATCH_TUNED = 3       /* the mode mpas_batch_autotune found fastest
                                (MPAS_BATCH_GRAPH before it ran) */
};
int mpas_batch_create(mpas_ctx *const *meshes, int nmeshes, int nmembers,
                      const int32_t *mesh_of, const double *dt, mpas_batch **out);
int mpas_batch_destroy(mpas_batch *batch);
/* Number of member groups (distinct mesh and time step pairs). */
int mpas_batch_groups(const mpas_batch *batch);
/* Host arrays as mpas_sw_state_upload; runs the group's diagnostics. */
int mpas_batch_upload(mpas_batch *batch, int member, const double *h, const double *u,
                      const double *tracers, const double *h_s);
int mpas_batch_download(mpas_batch *batch, int member, double *h, double *u,
                        double *tracers);
int mpas_batch_step(mpas_batch *batch, int nsteps, int mode);
int mpas_batch_sync(mpas_batch *batch);
/* Launch configuration: workgroup size of the 6 phase kernels (momentum,
   high-order flux, cells, tracer update, vertex and edge diagnostics;
   multiples of 64 up to 512), members per lane (consecutive members of one
   entity a lane advances), workgroup order (0 group after group, 1 the
   groups' workgroups interleaved), XCD-contiguous placement (0/1) and one
   entity per wave (each entity's lanes padded to whole waves, so the static
   mesh loads are scalar: 0 for no group, 1 for the groups whose padded
   lanes stay at least 3/4 busy, 2 for every group). */
int mpas_batch_config(mpas_batch *batch, const int32_t *block, int members_per_lane,
                      int order, int xcd, int wave_entity);
int mpas_batch_get_config(const mpas_batch *batch, int32_t *block, int *members_per_lane,
                          int *order, int *xcd, int *wave_entity);
/* Workgroup sizes per kernel and group layout, for kernels that run in more
   than one layout: 6 kernels x 6 layouts (MPAS_BATCH_LAYOUTS), row-major by
   kernel; 0 takes the kernel's size from mpas_batch_config (which clears
   them all).  The autotuner sets them; get + set reproduce its choice. */
enum { MPAS_BATCH_KERNELS = 6, MPAS_BATCH_LAYOUTS = 6 };
int mpas_batch_get_layout_blocks(const mpas_batch *batch, int32_t *blocks);
int mpas_batch_set_layout_blocks(mpas_batch *batch, const int32_t *blocks);
/* Times trial_steps graph replays per candidate (entity per wave and
   members per lane, placement and order, then each kernel's workgroup size,
   then that size per layout), keeps the fastest configuration, then times
   the persistent launch against the graph (mpas_batch_tuned_mode returns
   the faster), and restores the members' state; writes the trials as JSON
   into report. */
int mpas_batch_tuned_mode(const mpas_batch *batch);
/* Mesh order of the groups (MPAS_ORDER_*, as mpas_ctx_create_reordered):
   the batch runs on reordered copies of the mesh contexts and permutes the
   members' state; uploads and downloads keep the callers' numbering, and
   every member's result is unchanged bit for bit.  The autotuner tries
   every order first. */
int mpas_batch_set_order(mpas_batch *batch, int order);
int mpas_batch_get_order(const mpas_batch *batch);
int mpas_batch_autotune(mpas_batch *batch, int trial_steps, char *report, size_t cap);

#ifdef __cplusplus
}
#endif
#endif /* MPAS_HIP_H */
