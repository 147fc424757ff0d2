/*
 * athena_gpu.h -- C ABI of the MI355X per-MeshBlock MHD update.
 *
 * The library keeps every MeshBlock register (u, u1, w, w1, bcc, b, b1, fluxes, face
 * and corner EMFs, CT weights) resident in HBM in the reference's AthenaArray layout
 * (athena_arrays.hpp: (n,k,j,i), i fastest, one contiguous array per block), so a host
 * AthenaArray can be copied in or out with one memcpy per register.  Blocks of one
 * mesh are numbered row-major, x1 fastest.  Each pair of opposite mesh faces is either
 * periodic or outflow (ath_mesh_desc.bc).
 *
 * Every function returns 0 on success and a non-zero status otherwise; the message is
 * available from ath_gpu_last_error().  A non-zero status must be mapped to the
 * reference's fatal ATHENA_ERROR (defs.hpp.in:155-159), never to TaskStatus::fail,
 * which means "not ready, retry" (task_list.cpp:43-52).
 *
 * Calls on one mesh are ordered on that mesh's HIP stream; distinct meshes may be
 * driven from distinct host threads.
 */
#ifndef ATHENA_GPU_H_
#define ATHENA_GPU_H_

#ifdef __cplusplus
extern "C" {
#endif

/* Input-file level description of one uniform mesh.  Mirrors the athinput
 * keys read by Mesh (mesh.cpp:72-…), Reconstruction (reconstruction.cpp:39-…:
 * time/xorder), EquationOfState (adiabatic_mhd.cpp:26-31, adiabatic_mhd_sr.cpp:36-48)
 * and the configure switches --flux, -s, --nghost (configure.py:80-353). */
typedef struct ath_mesh_desc {
  int nx[3];          /* <mesh> nx1..nx3 */
  int bnx[3];         /* <meshblock> nx1..nx3 */
  int ng;             /* NGHOST (>= 3 for xorder 3) */
  int xorder;         /* 1 donor cell, 2 PLM, 3 PPM (second VL2 stage; first is DC) */
  int rsolver;        /* 0 HLLD (hlld.cpp; relativistic: hlld_rel.cpp), 1 HLLE */
  int relativistic;   /* 0 Newtonian, 1 special relativistic */
  double xmin[3], xmax[3];
  double gamma;       /* <hydro> gamma */
  double dfloor, pfloor;   /* < 0 selects the reference default sqrt(1024*FLT_MIN) */
  double cfl;         /* <time> cfl_number */
  double sigma_max, beta_min, gamma_max;   /* SR floors; < 0 selects 0, 0, 1000 */
  int bc[6];          /* <mesh> ix1_bc, ox1_bc, ix2_bc, ox2_bc, ix3_bc, ox3_bc: 0 periodic,
                         1 outflow (bvals.cpp:453-…, cc/outflow_cc.cpp, fc/outflow_fc.cpp;
                         a zero-filled descriptor is fully periodic) */
} ath_mesh_desc;

typedef struct ath_mesh *ath_mesh_t;

/* Register codes for ath_gpu_set_block / ath_gpu_get_block. */
enum {
  ATH_U = 0,     /* Hydro::u   (NHYDRO, nc3, nc2, nc1)   */
  ATH_W = 1,     /* Hydro::w   (NHYDRO, nc3, nc2, nc1)   */
  ATH_BCC = 2,   /* Field::bcc (3, nc3, nc2, nc1)        */
  ATH_B1 = 3,    /* Field::b.x1f (nc3, nc2, nc1+1)       */
  ATH_B2 = 4,    /* Field::b.x2f (nc3, nc2+1, nc1)       */
  ATH_B3 = 5,    /* Field::b.x3f (nc3+1, nc2, nc1)       */
  ATH_W1 = 6,    /* Hydro::w1  (NHYDRO, nc3, nc2, nc1)   */
  ATH_FLUX1 = 20, ATH_FLUX2 = 21, ATH_FLUX3 = 22,   /* Hydro::flux[X1DIR..X3DIR] */
  ATH_E1 = 23, ATH_E2 = 24, ATH_E3 = 25,            /* Field::e.x1e .. x3e       */
  ATH_WGHT1 = 26, ATH_WGHT2 = 27, ATH_WGHT3 = 28    /* Field::wght.x1f .. x3f    */
};

/* Creates the device mirror of a mesh (replaces the per-block allocations of the
 * Hydro/Field constructors, hydro.cpp:31-60, field.cpp:27-105). */
int ath_gpu_mesh_create(const ath_mesh_desc *desc, ath_mesh_t *out);

/* Creates the device mirror of a statically refined mesh (<mesh> refinement = static;
 * Mesh::Mesh with <refinementN> regions, mesh.cpp:306-482).  desc describes the root grid
 * (nx: root-grid cells, bnx: MeshBlock cells); locs[4 g .. 4 g + 3] = (level, lx1, lx2,
 * lx3) of MeshBlock g in the tree's order, as ath_gpu_tree_build lists them, root_level
 * the logical level of the root grid.  Neighbours follow BoundaryBase::
 * SearchAndSetNeighbors (bvals_base.cpp:298-700).  Every stage of ath_gpu_cycle then adds,
 * as the multilevel task list does (time_integrator.cpp:922-1185): the restricted fine
 * fluxes on coarse/fine faces (cc/flux_correction_cc.cpp:77-286), the EMF correction
 * (fc/flux_correction_fc.cpp:51-1548), and the boundary exchange between levels with
 * restriction and prolongation (cc/bvals_cc.cpp:287-540, fc/bvals_fc.cpp:344-975,
 * bvals_refine.cpp:98-673) in place of the same-level exchange.  Periodic faces and
 * Newtonian MHD only; nghost even (as MeshRefinement requires, mesh_refinement.cpp:42-48),
 * MeshBlock sizes even and at least 2 nghost; the list must cover the
 * root grid with adjacent MeshBlocks at most one level apart.  The fused flux kernel and
 * the autotuner are refused on such meshes. */
int ath_gpu_smr_mesh_create(const ath_mesh_desc *desc, const long long *locs, int nblocks,
                            int root_level, ath_mesh_t *out);
/* Logical level of MeshBlock gid (0 on a uniform mesh). */
int ath_gpu_smr_block_level(ath_mesh_t mesh, int gid, int *level);
/* Host only, no GPU: checks a refined mesh's MeshBlock list as ath_gpu_smr_mesh_create
 * does and builds its step programs (every box checked against its arrays).  Writes
 * *nnbr and, up to cap entries, block gid's neighbour list in SearchAndSetNeighbors order
 * as 9 ints each (gid, level, ox1, ox2, ox3, type 1 face / 2 edge / 3 corner, fi1, fi2,
 * eid for edges or fid for faces), nsteps[0..2] = the step counts of the flux, EMF and
 * boundary programs of the whole mesh and nsteps[3..5] their numbers of phases (launches
 * per stage). */
int ath_gpu_smr_plan(const ath_mesh_desc *desc, const long long *locs, int nblocks,
                     int root_level, int gid, int *nbr_out, int cap, int *nnbr, long *nsteps);
int ath_gpu_mesh_destroy(ath_mesh_t mesh);
int ath_gpu_nblocks(ath_mesh_t mesh, int *nblocks);
/* Number of doubles of one block's register `what`. */
int ath_gpu_block_size(ath_mesh_t mesh, int what, long *ndoubles);

/* Copy one block's register between a host AthenaArray buffer and HBM. */
int ath_gpu_set_block(ath_mesh_t mesh, int gid, int what, const double *host);
int ath_gpu_get_block(ath_mesh_t mesh, int gid, int what, double *host);

/* Mesh::Initialize after the problem generator (mesh.cpp:1686-1966): same-level ghost
 * exchange of u and b, ConservedToPrimitive(u, w1, b, w, bcc) over every cell,
 * NewBlockTimeStep on every block and Mesh::NewTimeStep. */
int ath_gpu_initialize(ath_mesh_t mesh, double tlim);

/* One VL2 stage of the TimeIntegratorTaskList for all blocks (stage = 1 or 2):
 * CalculateFluxes (DC on stage 1, xorder on stage 2), ComputeCornerE, the same-level
 * EMF flux correction, WeightedAve + AddFluxDivergence, WeightedAve + CT, ghost
 * exchange, ConservedToPrimitive and, on stage 2, NewBlockTimeStep
 * (time_integrator.cpp:1606-2318). */
int ath_gpu_vl2_stage(ath_mesh_t mesh, int stage);

/* Main-loop cycles (main.cpp:457-560): both stages, ncycle++, time += dt,
 * Mesh::NewTimeStep.  Runs at most ncycles cycles while time < tlim; *done receives
 * the number of cycles run.  With tlim = +inf no host synchronisation happens. */
int ath_gpu_cycle(ath_mesh_t mesh, int ncycles, double tlim, int *done);

/* Mesh clock (Mesh::time, Mesh::dt, Mesh::ncycle). */
int ath_gpu_get_clock(ath_mesh_t mesh, double *time, double *dt, int *ncycle);
int ath_gpu_set_clock(ath_mesh_t mesh, double time, double dt, int ncycle);

/* Per-function entry points over all blocks, for function-level parity:
 * Hydro::CalculateFluxes(w, b, bcc, order)  (calculate_fluxes.cpp:38-415)
 * Field::ComputeCornerE(w, bcc)              (calculate_corner_e.cpp:28-236)
 * EMF same-level flux correction is applied inside ath_gpu_ct. */
int ath_gpu_calculate_fluxes(ath_mesh_t mesh, int order);
int ath_gpu_compute_corner_e(ath_mesh_t mesh);

/* Per-kernel timing with HIP events recorded around every launch on the mesh stream
 * (no counterpart in the reference, which only prints zone-cycles/s,
 * main.cpp:634-651).  enable != 0 clears and starts collection; stats return the
 * kernel name, accumulated milliseconds and launch count for kernel id 0..17. */
int ath_gpu_kernel_timing(ath_mesh_t mesh, int enable);
int ath_gpu_kernel_stats(ath_mesh_t mesh, int id, char *name, int namelen,
                         double *total_ms, long *count);

/* History sums (outputs/history.cpp:84-152 for MHD): mass, 1-mom, 2-mom, 3-mom, 1-KE,
 * 2-KE, 3-KE, tot-E, 1-ME, 2-ME, 3-ME, volume-weighted over the active cells of all
 * blocks (11 doubles).  Summed in a fixed order; synchronous. */
int ath_gpu_history(ath_mesh_t mesh, double *sums);

/* Hydro stream the library launches on (a hipStream_t), for event timing. */
int ath_gpu_stream(ath_mesh_t mesh, void **stream);
int ath_gpu_synchronize(ath_mesh_t mesh);

/* Batched single-interface Riemann solves (rotated frame, 7 primitives per side):
 * solver 0 HLLD (hlld.cpp), 1 HLLE (hlle_mhd.cpp), 2 SR HLLE (hlle_mhd_rel.cpp),
 * 3 SR HLLD (hlld_rel.cpp).
 * Host arrays of n*7, n and n*7 doubles. */
int ath_gpu_riemann_batch(int solver, double gamma, int n, const double *wl,
                          const double *wr, const double *bx, double *flux);

/* Batched SR primitive recovery, one cell per entry (adiabatic_mhd_sr.cpp:72-272):
 * cons (n,5) in/out, prim (n,5) seed in/result out, bcc (n,3); iters (n) receives the
 * first-solve Newman-Hamlin iteration count or -1.  Host arrays. */
int ath_gpu_sr_con2prim_batch(int n, double gamma, double dfloor, double pfloor,
                              double sigma_max, double beta_min, double gamma_max,
                              double *cons, double *prim, const double *bcc, int *iters);

const char *ath_gpu_last_error(void);

/* ---- One MeshBlock over caller-owned device arrays -----------------------------------
 * The reference's per-MeshBlock functions with their own signatures, for a caller that
 * keeps a MeshBlock's AthenaArrays in device memory itself (one array per register,
 * AthenaArray layout, shapes as for ath_gpu_set_block with nc = nx + 2 ng).  A block
 * object holds what the MeshBlock's Coordinates and EquationOfState hold; the register
 * arrays stay the caller's.  Every call is asynchronous on `stream` (a hipStream_t; null
 * for the default stream) unless stated otherwise, and acts on that one block only:
 * neighbour exchange and EMF correction between blocks are the caller's, exactly as in
 * the reference, where they are separate tasks. */
typedef struct ath_block_desc {
  int nx[3];                        /* block_size.nx1 .. nx3 */
  int ng;                           /* NGHOST */
  const double *x1f, *x2f, *x3f;    /* host Coordinates::x1f .. x3f (nc + 1 each) */
  const double *x1v, *x2v, *x3v;    /* host Coordinates::x1v .. x3v (nc each) */
  const double *dx1f, *dx2f, *dx3f; /* host Coordinates::dx1f .. dx3f (nc each) */
  int rsolver;                      /* 0 HLLD, 1 HLLE (Newtonian and SR, as ath_mesh_desc) */
  int relativistic;
  double gamma, dfloor, pfloor;     /* < 0 floors select the reference defaults */
  double sigma_max, beta_min, gamma_max;
  double cfl;                       /* Mesh::cfl_number, used by ath_gpu_block_new_dt */
} ath_block_desc;

typedef struct ath_block *ath_block_t;

/* Outputs of the flux and EMF functions: Hydro::flux[X1DIR..X3DIR] (NHYDRO, face array),
 * the face EMFs of Field (e3_x1f = -F(By) and e2_x1f = F(Bz) on x1 faces, e1_x2f and
 * e3_x2f on x2 faces, e2_x3f and e1_x3f on x3 faces: the ey, ez arguments of the
 * RiemannSolver calls at calculate_fluxes.cpp:102, 223, 343), Field::wght.x1f .. x3f and
 * Field::e.x1e .. x3e. */
typedef struct ath_block_regs {
  double *flux[3];
  double *e3_x1f, *e2_x1f, *e1_x2f, *e3_x2f, *e2_x3f, *e1_x3f;
  double *wght[3];
  double *e[3];
} ath_block_regs;

/* Coordinates and EquationOfState of one MeshBlock (coordinates.cpp:129-146,
 * adiabatic_mhd.cpp:26-31, adiabatic_mhd_sr.cpp:36-48); synchronous. */
int ath_gpu_block_create(const ath_block_desc *desc, ath_block_t *out);
int ath_gpu_block_destroy(ath_block_t blk);

/* Reconstruction::DonorCellX{dir+1} / PiecewiseLinearX{dir+1} / PiecewiseParabolicX{dir+1}
 * for order 1 / 2 / 3 (reconstruction.hpp:59-105; dc.cpp, plm.cpp, ppm.cpp) on pencil
 * (k, j), cells il .. iu: w (NHYDRO, nc3, nc2, nc1), bcc (3, nc3, nc2, nc1); wl, wr pencils
 * (NWAVE, nc1).  x1 writes wl(n, i+1), wr(n, i); x2 and x3 write wl(n, i), wr(n, i) for the
 * cell j (x2) or k (x3), as the reference does. */
int ath_gpu_block_reconstruct(ath_block_t blk, int dir, int order, int k, int j, int il,
                              int iu, const double *w, const double *bcc, double *wl,
                              double *wr, void *stream);

/* Hydro::RiemannSolver(k, j, il, iu, ivx, bx, wl, wr, flx, ey, ez, wct, dxw) (hydro.hpp:
 * 82-87; hlld.cpp, hlle_mhd.cpp, hlle_mhd_rel.cpp, hlld_rel.cpp by the block's rsolver
 * and dynamics): ivx = 1, 2, 3 (IVX, IVY, IVZ); bx, ey, ez, wct face arrays and flx the
 * (NHYDRO, face array) of that direction; wl, wr, dxw pencils; dt is Mesh::dt for
 * GetWeightForCT (hydro.cpp:154-158). */
int ath_gpu_block_riemann(ath_block_t blk, int k, int j, int il, int iu, int ivx,
                          const double *bx, const double *wl, const double *wr, double *flx,
                          double *ey, double *ez, double *wct, const double *dxw, double dt,
                          void *stream);

/* Hydro::CalculateFluxes(w, b, bcc, order) (hydro.hpp:72-73, calculate_fluxes.cpp:38-415;
 * MHD ranges): writes out->flux, the face EMFs and out->wght, with Mesh::dt = dt. */
int ath_gpu_block_calculate_fluxes(ath_block_t blk, const double *w, const double *b1,
                                   const double *b2, const double *b3, const double *bcc,
                                   int order, double dt, const ath_block_regs *out,
                                   void *stream);

/* Field::ComputeCornerE(w, bcc) (field.hpp:62, calculate_corner_e.cpp:129-230): reads the
 * face EMFs and weights of r, writes r->e. */
int ath_gpu_block_compute_corner_e(ath_block_t blk, const double *w, const double *bcc,
                                   const ath_block_regs *r, void *stream);

/* Field::CT(wght, b_out) (field.hpp:65, ct.cpp:31-116): b_out -= wght * curl(r->e) on the
 * active faces, in place. */
int ath_gpu_block_ct(ath_block_t blk, double wght, const ath_block_regs *r, double *b1_out,
                     double *b2_out, double *b3_out, void *stream);

/* Hydro::AddFluxDivergence(wght, u_out) (hydro.hpp:67, add_flux_divergence.cpp:39-96):
 * u_out -= wght * div(r->flux) on the active cells, in place. */
int ath_gpu_block_add_flux_divergence(ath_block_t blk, double wght, const ath_block_regs *r,
                                      double *u_out, void *stream);

/* EquationOfState::ConservedToPrimitive(cons, prim_old, b, prim, bcc, pco, il, iu, jl, ju,
 * kl, ku) (eos.hpp:37-40; adiabatic_mhd.cpp:41-90 or adiabatic_mhd_sr.cpp:72-272) with
 * Field::CalculateCellCenteredField over the same range: writes prim and bcc, rewrites
 * cons where a floor applies.  Like the reference, the SR recovery seeds from prim
 * itself; prim_old is not read. */
int ath_gpu_block_cons2prim(ath_block_t blk, double *cons, const double *prim_old,
                            const double *b1, const double *b2, const double *b3,
                            double *prim, double *bcc, int il, int iu, int jl, int ju,
                            int kl, int ku, void *stream);

/* Hydro::NewBlockTimeStep() (hydro.hpp:66, new_blockdt.cpp:45-230, ideal MHD): the block's
 * new_block_dt (cfl applied) from w, bcc and b, into the device double *dt_out.
 * Allocates and frees its per-work-group scratch, so it synchronises `stream`. */
int ath_gpu_block_new_dt(ath_block_t blk, const double *w, const double *bcc,
                         const double *b1, const double *b2, const double *b3,
                         double *dt_out, void *stream);

/* ath_gpu_block_new_dt without any allocation, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  The per-work-group minima go to a
 * fixed device array of 65536 doubles (512 KiB, held on each GPU the library runs on
 * until the process ends), so a MeshBlock may have at most 16,777,216 active cells; a
 * larger one is refused with an error.  Asynchronous on `stream`.  Every call on one GPU
 * shares that array: calls that overlap on one GPU (different streams or host threads)
 * race and must be serialised by the caller. */
int ath_gpu_block_new_dt_noalloc(ath_block_t blk, const double *w, const double *bcc,
                                 const double *b1, const double *b2, const double *b3,
                                 double *dt_out, void *stream);

/* ath_gpu_cycle with on != 0 replays each cycle (both stages and the new dt) as one
 * captured HIP graph, recaptured whenever the mesh's registers, tiles, fused band height
 * or recovery mode differ from the captured ones; the state evolves bit for bit as in the
 * eager cycle.  Kernel timing (ath_gpu_kernel_timing) runs eagerly.  Off by default.  No
 * counterpart in the reference. */
int ath_gpu_mesh_set_graph(ath_mesh_t mesh, int on);

/* Optional fused stage: the fluxes and Hydro::AddFluxDivergence in one kernel that
 * marches up the planes of bands of `rows` active rows of each MeshBlock, one thread per
 * (j, i) column, and keeps the fluxes on chip (Hydro::flux is then not written by
 * ath_gpu_vl2_stage / ath_gpu_cycle; ath_gpu_calculate_fluxes still fills it).  rows = 0
 * (the default) keeps the separate flux and update kernels, which measured faster on the
 * 256^3 HLLD and SR HLLE workloads; -1 picks the tallest band a work-group holds.  Both
 * give the same bits.  No counterpart in the reference. */
int ath_gpu_mesh_set_fused(ath_mesh_t mesh, int rows);
int ath_gpu_mesh_get_fused(ath_mesh_t mesh, int *rows);

/* Host check of that kernel's plan for one MeshBlock shape (nx[3], ng) and band height:
 * enumerates every band and plane with the kernel's own index helpers, fails if an LDS
 * access leaves the work-group's array or an owned cell reads a position that does not
 * hold the face it needs, and lists the faces whose EMFs and weight are written
 * (dir, k, j, i) and the cells updated (k, j, i), with their counts and the LDS doubles of
 * one work-group.  No GPU needed. */
int ath_gpu_flux_div_plan(const int *nx, int ng, int rows, int *faces, long max_faces,
                          long *nfaces, int *cells, long max_cells, long *ncells,
                          long *lds_doubles);

/* ---- Ensembles ------------------------------------------------------------------------
 * Independent meshes of uneven sizes advanced together: each VL2 stage kernel is
 * launched once for all members (persistent work-groups over a (member, work-group)
 * map), the cycle can be replayed as one HIP graph, and the thread-to-cell tile shape and
 * persistent grid of every kernel are chosen by a built-in autotuner.  No counterpart
 * in the reference, which advances one Mesh per process (main.cpp:457-560); each member
 * evolves exactly as ath_gpu_cycle would advance it on its own, with its own dt.
 * Members must share xorder, Riemann solver, dynamics and primitive-recovery mode
 * (ath_gpu_mesh_set_c2p; with the learned guess a seed kernel runs over all members
 * before the recovery) and be initialised (ath_gpu_initialize); while in an ensemble
 * they must not run single stages (ath_gpu_vl2_stage), which the ensemble detects. */
typedef struct ath_ensemble *ath_ensemble_t;

int ath_gpu_ensemble_create(const ath_mesh_t *members, int nmembers, ath_ensemble_t *out);
int ath_gpu_ensemble_destroy(ath_ensemble_t ens);
int ath_gpu_ensemble_size(ath_ensemble_t ens, int *nmembers, long *active_cells);
/* ncycles VL2 cycles of every member (main.cpp:457-560 per member, no tlim check);
 * use_graph != 0 replays one captured HIP graph per cycle.  Synchronous. */
int ath_gpu_ensemble_cycle(ath_ensemble_t ens, int ncycles, int use_graph);
/* Work decomposition per kernel kind (0 fluxes, 1 corner EMFs, 2 update, 3 primitive
 * recovery): tile_idx picks a thread-to-cell tile shape (0 linear chunks of 256 cells,
 * 1..5 the tiles 64x4x1, 32x8x1, 16x16x1, 16x4x4, 8x8x4; for the corner kernel only,
 * 6 and 7 a z-march over segments of 8 and 16 planes with the cell-centred E staged in
 * LDS), persist the launch grid of the persistent kernel (0: one work-group per work
 * item). */
int ath_gpu_ensemble_get_tiles(ath_ensemble_t ens, int *tile_idx, int *persist);
int ath_gpu_ensemble_set_tiles(ath_ensemble_t ens, const int *tile_idx, const int *persist);
/* Times every tile shape and persistent grid per kernel kind on the ensemble's data
 * (reps launches each), keeps the fastest, and leaves the members' state as it was.
 * best_ms (4 doubles, may be null) receives the chosen per-launch times. */
int ath_gpu_ensemble_autotune(ath_ensemble_t ens, int reps, double *best_ms);
int ath_gpu_ensemble_stream(ath_ensemble_t ens, void **stream);

/* Tile shape (index as for ensembles) of each kernel kind of a single mesh, and the
 * same autotuner for a single mesh (its state is left as it was; periodic meshes only,
 * like ensembles).  A single mesh also accepts indices 6 and 7 for the flux kind: the
 * flux z-march for its PLM and PPM launches (each cell reconstructed once along x1 and x3,
 * same bits; donor-cell launches keep linear chunks); the single-mesh autotuner times it
 * against its chosen flux tile at the mesh's reconstruction order. */
int ath_gpu_mesh_set_tiles(ath_mesh_t mesh, const int *tile_idx);
int ath_gpu_mesh_get_tiles(ath_mesh_t mesh, int *tile_idx);
int ath_gpu_mesh_autotune(ath_mesh_t mesh, int reps, double *best_ms);
/* Position box (first index lo[3], extent n[3]) of kernel kind 0..3 in a MeshBlock of
 * nx^3 active cells and ng ghosts, for host-side bounds checks; no device call. */
int ath_gpu_kernel_box(int kind, int nx, int ng, int *lo, int *n);
/* Lane `lane` of wavefront `wave` of the flux kernel's z-march (flux tile options 6, 7)
 * on nblocks MeshBlocks of shape = {nx1, nx2, nx3, ng, nblocks} with segments of about
 * seg planes: pos = {owned or repeated (1) / past the end (0), block, j, i, first plane,
 * last plane, planes of the tallest segment}; *nwave = wavefronts per launch.  Lane 0
 * repeats the previous wavefront's last position.  For host-side checks; no device call;
 * no counterpart in the reference (a launch-geometry query). */
int ath_gpu_flux_march_position(const int *shape, int seg, long wave, int lane, int *pos,
                                long *nwave);
/* Ghost position number t of the same-level exchange kernel in a MeshBlock of nx^3
 * cells and ng ghosts, for array 0 (cell-centred) or 1..3 (x1f, x2f, x3f face field):
 * pos = {i, j, k, source i, source j, source k, neighbour offsets p1, p2, p3}, the
 * source being the neighbour's active position that the reference's same-level buffers
 * (cc/bvals_cc.cpp:287-427, fc/bvals_fc.cpp:344-688) copy there; no device call. */
int ath_gpu_ghost_position(int nx, int ng, int array, int t, int *pos);

/* ---- Static mesh refinement: MeshRefinement operators -------------------------------
 * The 3-D uniform-Cartesian branches of MeshRefinement (mesh_refinement.cpp), on device
 * arrays of nb MeshBlocks of one shape stored back to back, asynchronous on stream.
 * shape = {nx1, nx2, nx3, ng} (even nx): a block has nx + 2 ng cells per direction and
 * its coarse buffer nx/2 + 2 ng (meshblock.cpp:66,93-94); fine index = (coarse index -
 * ng) * 2 + ng.  Cell arrays are (nvar, n3, n2, n1) per block, face component d has one
 * more entry along d.  fcoord / ccoord: per block the fine / coarse coordinate tables
 * x1f (n1+1), x1v (n1), dx1f (n1), then the same for x2 and x3.  range = {si, ei, sj, ej,
 * sk, ek}: the coarse index range, as the reference's arguments; a range whose stencil
 * leaves the arrays is rejected before any launch. */
/* RestrictCellCenteredValues (mesh_refinement.cpp:116-150): volume averages. */
int ath_gpu_restrict_cc(const int *shape, int nb, int nvar, const double *fcoord,
                        const double *fine, double *coarse, const int *range, void *stream);
/* RestrictFieldX1/X2/X3 (mesh_refinement.cpp:272-430), comp 0..2: area averages. */
int ath_gpu_restrict_fc(const int *shape, int nb, int comp, const double *fcoord,
                        const double *fine, double *coarse, const int *range, void *stream);
/* ProlongateCellCenteredValues (mesh_refinement.cpp:464-525): minmod-limited linear. */
int ath_gpu_prolongate_cc(const int *shape, int nb, int nvar, const double *fcoord,
                          const double *ccoord, const double *coarse, double *fine,
                          const int *range, void *stream);
/* ProlongateSharedFieldX1/X2/X3, Cartesian branch (mesh_refinement.cpp:835-872 and the
 * matching blocks of :914-1221): the four fine faces on each coarse face. */
int ath_gpu_prolongate_fc_shared(const int *shape, int nb, int comp, const double *fcoord,
                                 const double *ccoord, const double *coarse, double *fine,
                                 const int *range, void *stream);
/* ProlongateInternalField (mesh_refinement.cpp:1223-1338): the twelve fine faces inside
 * each coarse cell of range, divergence-preserving; b1, b2, b3 the fine face arrays. */
int ath_gpu_prolongate_fc_internal(const int *shape, int nb, const double *fcoord,
                                   double *b1, double *b2, double *b3, const int *range,
                                   void *stream);
/* The EMF restriction of FaceCenteredBoundaryVariable::LoadFluxBoundaryBufferToCoarser
 * (flux_correction_fc.cpp:315-420), comp 0..2: each coarse edge along comp is the
 * length-weighted average of its two fine edges.  Edge arrays per block: along x1
 * (n3+1, n2+1, n1), along x2 (n3+1, n2, n1+1), along x3 (n3, n2+1, n1+1).  (The face-flux
 * restriction of the cell-centred flux correction, flux_correction_cc.cpp:77-120, is
 * ath_gpu_restrict_fc applied to each flux component: the same area average.) */
int ath_gpu_restrict_emf(const int *shape, int nb, int comp, const double *fcoord,
                         const double *fine, double *coarse, const int *range, void *stream);

/* ---- Static mesh refinement: the MeshBlock tree ------------------------------------
 * A refinement region (<refinementN> x1min .. x3max, level; mesh.cpp:351-398): the box
 * inside the mesh and its level relative to the root grid (>= 1). */
typedef struct {
  double xmin[3];
  double xmax[3];
  int level;
} ath_refine_region;

/* MeshBlockTree of a 3-D mesh of nrb[0] x nrb[1] x nrb[2] root MeshBlocks on
 * [xmin, xmax] with static refinement: CreateRootGrid, then AddMeshBlock over every
 * region's logical range (mesh.cpp:306,400-482; meshblock_tree.cpp:87-258), which refines
 * neighbours so that adjacent MeshBlocks (faces, edges, corners; periodic faces wrap)
 * differ by at most one level.  bc as ath_mesh_desc.bc.  Host only, no GPU needed.
 * Writes *nblocks, *root_level (the logical level of the root grid) and, when cap is at
 * least *nblocks: locs[4 g .. 4 g + 3] = (level, lx1, lx2, lx3) of MeshBlock g in the
 * reference's Z order (GetMeshBlockList, meshblock_tree.cpp:336-356), and, if nbr is not
 * null, nbr[27 g + (o3+1) 9 + (o2+1) 3 + (o1+1)] for each offset o in {-1,0,1}^3: the
 * MeshBlock at that side on the same or the next coarser level (FindNeighbor,
 * meshblock_tree.cpp:362-468), -2 when that side is refined further (finer
 * neighbours), -1 across a non-periodic face.  locs or nbr may be null.  Fails when a
 * region is inside-out or outside the mesh, deeper than 40 levels or would need more than
 * 2^22 tree nodes, or cap < *nblocks (a region of zero width is accepted, as the
 * reference accepts it). */
int ath_gpu_tree_build(const int *nrb, const int *bc, const double *xmin, const double *xmax,
                   const ath_refine_region *reg, int nreg, long long *locs, int *nbr,
                   int cap, int *nblocks, int *root_level);

/* ---- Learned initial guess for SR primitive recovery --------------------------------
 * No counterpart in the reference: its SR recovery (adiabatic_mhd_sr.cpp:404-522) seeds
 * the Newman-Hamlin iteration with the previous gas pressure.  Here a 6 -> 64 -> 64 -> 1
 * tanh MLP predicts log10(p_gas / (E - D)) from six rotation- and scale-invariant inputs
 * (E/D, |M|/E, B^2/E, sqrt(E^2-M^2)/D, the M-B angle); the prediction seeds an FP64 Newton
 * iteration on the reference's fixed-point residual, converged to the reference's 1e-12,
 * and the reference iteration is the fallback.  Parameter vector (floats):
 *   W1[64][6] b1[64] W2[64][64] b2[64] W3[64] b3 | x_mean[6] x_scale[6] y_mean y_scale
 * Training is FP32 on the GPU (forward/backward/Adam kernels), inference BF16 MFMA.
 * Errors of the ath_ml_* calls are reported by ath_ml_last_error(). */
typedef struct ath_mlp *ath_mlp_t;

int ath_ml_param_count(void);                         /* floats in the parameter vector */
int ath_ml_create(ath_mlp_t *out);                    /* all parameters zero */
int ath_ml_destroy(ath_mlp_t mlp);
int ath_ml_set_params(ath_mlp_t mlp, const float *params);   /* resets Adam moments */
int ath_ml_get_params(ath_mlp_t mlp, float *params);
/* Gradient of mean((y(x_b) - t_b)^2) over n normalised samples x (n,6), t (n) with
 * respect to the first 4673 parameters (W1..b3), and the loss.  Host arrays. */
int ath_ml_grad(ath_mlp_t mlp, int n, const float *x, const float *t, float *grad,
                float *loss);
/* `steps` Adam steps (beta 0.9/0.999, eps 1e-8, learning rate lr) on minibatches of
 * `batch` samples; order (steps*batch) lists the sample of each minibatch slot;
 * loss_hist (steps, may be null) receives the minibatch losses.  Host arrays. */
int ath_ml_train(ath_mlp_t mlp, int n, const float *x, const float *t, int steps,
                 int batch, const int *order, float lr, float *loss_hist);
/* De-normalised outputs y (n) for normalised features x (n,6): bf16 != 0 runs the
 * MFMA path used by the solver, 0 the FP32 path.  Host arrays. */
int ath_ml_infer(ath_mlp_t mlp, int n, const float *x, float *y, int bf16);
/* The solver's raw (un-normalised) features x (n,6) of cons (n,5), bcc (n,3). */
int ath_ml_features(int n, const double *cons, const double *bcc, float *x);
const char *ath_ml_last_error(void);

/* Batched SR primitive recovery with a choice of first solver: mode 0 the reference
 * iteration (as ath_gpu_sr_con2prim_batch), 1 Newton from the prim seed, 2 Newton from
 * the learned seed with the reference iteration as fallback (mlp required).  iters
 * (n): as ath_gpu_sr_con2prim_batch; evals (n): evaluations of the fixed-point map;
 * fallback (n): 1 where the learned-seed solve fell back.  Output pointers may be null.
 * _modes takes host arrays and is synchronous; _device takes device arrays and is
 * asynchronous on `stream` (a hipStream_t, null for the default stream). */
int ath_gpu_sr_con2prim_modes(ath_mlp_t mlp, int mode, int n, double gamma, double dfloor,
                              double pfloor, double sigma_max, double beta_min,
                              double gamma_max, double *cons, double *prim,
                              const double *bcc, int *iters, int *evals, int *fallback);
int ath_gpu_sr_con2prim_device(ath_mlp_t mlp, int mode, int n, double gamma,
                               double dfloor, double pfloor, double sigma_max,
                               double beta_min, double gamma_max, double *cons,
                               double *prim, const double *bcc, int *iters, int *evals,
                               int *fallback, void *stream);

/* Selects the first solver of a relativistic mesh's ConservedToPrimitive (modes as
 * above; the mesh keeps a reference to mlp, which must outlive it or be replaced). */
int ath_gpu_mesh_set_c2p(ath_mesh_t mesh, ath_mlp_t mlp, int mode);
/* Recovery counters summed over every cell solve of the mesh: totals[0] evaluations
 * of the fixed-point map, [1] learned-seed fallbacks, [2] cells.  enable 1 clears and
 * starts counting, 0 stops, -1 only reads; totals (may be null) receives the counts
 * before the call's change. */
int ath_gpu_mesh_c2p_stats(ath_mesh_t mesh, int enable, unsigned long long *totals);

#ifdef __cplusplus
}
#endif
#endif /* ATHENA_GPU_H_ */
